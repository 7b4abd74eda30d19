"""Adjoint kernels: the dot-product identity <T x, y> = <x, T* y> for every
plan kind, with <a, b> = sum a*b over reals and Re sum a*conj(b) over complex
(gradient convention dL/dRe + i dL/dIm).  A wrong weight, sign or index in an
adjoint (e.g. a dropped Q2 cross term at the DC/Nyquist bins of c2r) breaks
the identity far beyond FP32 rounding."""
import numpy as np
import pytest

from tests.helpers import cplx, empty_dev, from_dev, real, to_dev

pytestmark = pytest.mark.gpu


def inner(a, b):
    a = np.asarray(a, np.complex128 if np.iscomplexobj(a) else np.float64)
    b = np.asarray(b, np.complex128 if np.iscomplexobj(b) else np.float64)
    return float(np.real(np.sum(a * np.conj(b))))


CASES = [("c2c", -1, 0), ("c2c", 1, 0), ("c2c", -1, "zp"), ("r2c", -1, 0), ("r2c", -1, "full"),
         ("r2c", -1, "zp"), ("c2r", 1, 0)]


@pytest.mark.parametrize("N", [8, 256, 4096, 65536, 262144])
@pytest.mark.parametrize("kind,d,mode", CASES)
def test_dot_product_identity(gpu, N, kind, d, mode):
    import mufft_amd as M
    import torch
    fl = {0: 0, "zp": M.FLAG_ZERO_PAD_UPPER_HALF, "full": M.FLAG_FULL_R2C}[mode]
    p = M.Plan1D(kind, N, d, fl)
    B = 2
    zp = mode == "zp"
    if kind == "c2c":
        nin, nout = (N // 2 if zp else N), N
        x = cplx(N + 1, B * nin).reshape(B, nin)
        y = cplx(N + 2, B * nout).reshape(B, nout)
        Tx = empty_dev((B, nout), np.complex64, gpu)
        Ty = empty_dev((B, N), np.complex64, gpu)
    elif kind == "r2c":
        nin, nout = (N // 2 if zp else N), (N if mode == "full" else N // 2 + 1)
        x = real(N + 3, B * nin).reshape(B, nin)
        y = cplx(N + 4, B * nout).reshape(B, nout)
        Tx = empty_dev((B, nout), np.complex64, gpu)
        Ty = empty_dev((B, N), np.float32, gpu)
    else:
        nin, nout = N // 2 + 1, N
        x = cplx(N + 5, B * nin).reshape(B, nin)
        y = real(N + 6, B * nout).reshape(B, nout)
        Tx = empty_dev((B, nout), np.float32, gpu)
        Ty = empty_dev((B, nin), np.complex64, gpu)
    p.execute_batched(Tx, to_dev(x, gpu), B)
    p.execute_adjoint_batched(Ty, to_dev(y, gpu), B)
    torch.cuda.synchronize()
    tx, ty = from_dev(Tx), from_dev(Ty)
    if zp:
        ty = ty[:, :nin]  # gradient w.r.t. the inputs the transform reads
    lhs, rhs = inner(tx, y), inner(x, ty)
    scale = np.sqrt(inner(tx, tx) * inner(y, y))
    assert abs(lhs - rhs) <= 2e-6 * scale, (N, kind, mode, lhs, rhs)


def test_adjoint_fused_multiply_broadcast(gpu):
    """Adjoint with a fused multiply by a broadcast conjugated spectrum."""
    import mufft_amd as M
    import torch
    N, B = 1024, 3
    p = M.Plan1D("c2c", N, -1)
    y = cplx(1, B * N).reshape(B, N)
    w = cplx(2, N)
    out = empty_dev((B, N), np.complex64, gpu)
    p.execute_mul_batched(out, to_dev(y, gpu), to_dev(w, gpu), B, idist2=0, norm=0.5, conj2=True,
                          adjoint=True)
    torch.cuda.synchronize()
    ref = np.fft.ifft(0.5 * y.astype(np.complex128) * np.conj(w), axis=-1) * N
    assert np.sqrt(np.mean(np.abs(from_dev(out) - ref) ** 2) / np.mean(np.abs(ref) ** 2)) < 2e-6


def test_mul_reduce(gpu):
    import mufft_amd as M
    import torch
    from mufft_amd import api as A
    n, B = 5000, 7
    a = cplx(3, B * n).reshape(B, n)
    b = cplx(4, B * n).reshape(B, n)
    out = empty_dev((n,), np.complex64, gpu)
    A.mul_reduce(out, to_dev(a, gpu), to_dev(b, gpu), n, B, norm=0.25, conj_b=True)
    torch.cuda.synchronize()
    ref = 0.25 * np.sum(a.astype(np.complex128) * np.conj(b), axis=0)
    np.testing.assert_allclose(from_dev(out), ref, rtol=1e-5, atol=1e-5)
    assert M.device_count() >= 1
