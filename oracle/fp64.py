"""FP64 oracle for the muFFT hot path (test infrastructure only).

Conventions follow the reference's public API (fft.h:44-71): MUFFT_FORWARD = -1
means exp(-2*pi*i*n*k/N), transforms are unnormalised, r2c returns N/2+1 bins
(N with FULL_R2C), c2r reads N/2+1 bins *including* the imaginary parts of DC
and Nyquist (the reference's resolve, kernel.c:36-45, keeps them: an extra
-Im(X0)*(-1)^n - Im(X_{N/2}) term).  ``np.fft`` in complex128 supplies the
transforms; ``dft_direct`` is the brute-force O(N^2) sum that pins it.
"""
from __future__ import annotations

import numpy as np


def dft_direct(x, direction=-1):
    """Brute-force DFT in FP64, X[k] = sum_n x[n] exp(direction*2*pi*i*n*k/N)."""
    x = np.asarray(x, np.complex128)
    N = x.shape[-1]
    n = np.arange(N)
    W = np.exp(direction * 2j * np.pi * np.outer(n, n) / N)
    return x @ W.T


def c2c(x, direction=-1):
    x = np.asarray(x, np.complex128)
    N = x.shape[-1]
    if direction == -1:
        return np.fft.fft(x, axis=-1)
    return np.fft.ifft(x, axis=-1) * N


def r2c(x, full=False):
    """Real N -> N/2+1 bins (N bins with ``full``)."""
    X = np.fft.fft(np.asarray(x, np.float64), axis=-1)
    N = X.shape[-1]
    return X if full else X[..., : N // 2 + 1]


def c2r(X, N):
    """muFFT c2r semantics on N/2+1 bins: Hermitian part plus the Q2 term."""
    X = np.asarray(X, np.complex128)
    h = N // 2
    Xr = X.copy()
    im0 = Xr[..., 0].imag.copy()
    imh = Xr[..., h].imag.copy()
    Xr[..., 0] = Xr[..., 0].real
    Xr[..., h] = Xr[..., h].real
    y = np.fft.irfft(Xr[..., : h + 1], n=N, axis=-1) * N
    n = np.arange(N)
    sign = np.where(n % 2 == 0, 1.0, -1.0)
    return y - im0[..., None] * sign - imh[..., None]


def c2c_2d(x, direction=-1):
    x = np.asarray(x, np.complex128)
    if direction == -1:
        return np.fft.fft2(x, axes=(-2, -1))
    return np.fft.ifft2(x, axes=(-2, -1)) * (x.shape[-1] * x.shape[-2])


def r2c_2d(x, full=False):
    X = np.fft.fft2(np.asarray(x, np.float64), axes=(-2, -1))
    Nx = X.shape[-1]
    return X if full else X[..., : Nx // 2 + 1]


def c2r_2d(X, Nx):
    """2D c2r of the first Nx/2+1 columns, Hermitian-consistent input assumed
    except for the DC/Nyquist-column imaginary parts, which follow the same Q2
    rule per row after the inverse column transform (fft.c:863-916)."""
    X = np.asarray(X, np.complex128)[..., : Nx // 2 + 1]
    Ny = X.shape[-2]
    cols = np.fft.ifft(X, axis=-2) * Ny
    return c2r(cols, Nx)


def conv_direct(a, b):
    """Linear convolution, test.c:375-403 restated in FP64: length len(a)+len(b)-1,
    returned zero-extended to 2*len(a) like the reference's output buffer."""
    a = np.asarray(a)
    b = np.asarray(b)
    n = a.shape[-1]
    out_len = 2 * n
    dt = np.complex128 if np.iscomplexobj(a) or np.iscomplexobj(b) else np.float64
    out = np.zeros(a.shape[:-1] + (out_len,), dt)
    for x in range(b.shape[-1]):
        out[..., x: x + n] += a[..., :] * b[..., x: x + 1]
    return out


def conv_fft(a, b, N):
    """Circular convolution of zero-padded a, b at length N (FP64 FFT)."""
    A = np.fft.fft(np.asarray(a, np.complex128), n=N, axis=-1)
    B = np.fft.fft(np.asarray(b, np.complex128), n=N, axis=-1)
    y = np.fft.ifft(A * B, axis=-1)
    if not (np.iscomplexobj(a) or np.iscomplexobj(b)):
        return y.real
    return y


def rel_rms(y, ref):
    y = np.asarray(y, np.complex128)
    ref = np.asarray(ref, np.complex128)
    den = np.sqrt(np.mean(np.abs(ref) ** 2))
    return float(np.sqrt(np.mean(np.abs(y - ref) ** 2)) / (den if den > 0 else 1.0))
