"""FP64-on-GPU twiddle generation, pinned bit-for-bit.

The Stockham table entry (m-1)*p + k of pass (r, p) is
exp(dir * 2 pi i * k*m / (r*p)), evaluated like the reference's twiddle()
(fft.c:102-106): phase = (M_PI * dir * k*m) / (r*p/2) in double, then
(float)cos, (float)sin.  The host recomputes every entry with the C library
and counts FP32 bit mismatches: expected 0.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def host_table(N, d, E):
    # Radix sequence of the fused kernels: the remainder first, then E.
    E = min(E, N)
    radix = []
    left = N
    while left > 1:
        r = min(E, left)
        radix.append(r)
        left //= r
    radix = radix[-1:] + radix[:-1]
    out = []
    p = 1
    for s, r in enumerate(radix):
        if s > 0:
            for m in range(1, r):
                for k in range(p):
                    ph = (math.pi * d * (k * m)) / (r * p // 2)
                    out.append(complex(np.float32(math.cos(ph)), np.float32(math.sin(ph))))
        p *= r
    return np.array(out, np.complex64)


@pytest.mark.parametrize("N", [2, 16, 32, 256, 2048, 4096, 8192, 32768])
@pytest.mark.parametrize("d", [-1, 1])
def test_twiddle_table_bit_exact(gpu, N, d):
    import mufft_amd as M
    from mufft_amd import _lib
    E = _lib.load().mufft_hip_default_variant(N, 0) & 0xFF  # points per thread of the default kernel
    t = M.twiddle_table(N, d)
    h = host_table(N, d, E)
    assert t.shape == h.shape
    mism = int(np.count_nonzero(t.view(np.uint32) != h.view(np.uint32)))
    assert mism == 0, f"{mism} FP32 mismatches of {2 * h.size}"


def _table_mode(mode, n, param, d):
    import ctypes

    from mufft_amd import _lib
    L = _lib.load()
    size = L.mufft_hip_twiddle_table_mode(mode, n, param, d, None, 0)
    assert size > 0, _lib.last_error()
    out = np.zeros(size, np.complex64)
    assert L.mufft_hip_twiddle_table_mode(mode, n, param, d, out.ctypes.data_as(ctypes.c_void_p), size) == size
    return out


def _host(phases):
    return np.array([complex(np.float32(math.cos(ph)), np.float32(math.sin(ph))) for ph in phases],
                    np.complex64)


@pytest.mark.parametrize("n", [2, 512, 4096, 32768])
@pytest.mark.parametrize("d", [-1, 1])
def test_resolve_table_bit_exact_vs_restatement(gpu, n, d):
    """W[k] = (0, dir) * twiddle(dir, k, n) (fft.c:441-455), complex multiply in float."""
    from oracle import restate as R
    t = _table_mode(1, n, 0, d)
    h = R.build_r2c_twiddles(d, n)
    assert t.shape == h.shape
    mism = int(np.count_nonzero(t.view(np.uint32) != h.view(np.uint32)))
    assert mism == 0, f"{mism} FP32 mismatches of {2 * h.size}"


@pytest.mark.parametrize("n,count", [(1 << 17, 512), (1 << 20, 1024), (1 << 26, 4096)])
@pytest.mark.parametrize("d", [-1, 1])
def test_four_step_table_bit_exact(gpu, n, count, d):
    """exp(dir*2*pi*i*j/n), j < count: the phase (pi*dir*j)/(n/2) in double, rounded once."""
    t = _table_mode(2, n, count, d)
    h = _host([(math.pi * d * j) / (n // 2) for j in range(count)])
    assert int(np.count_nonzero(t.view(np.uint32) != h.view(np.uint32))) == 0


@pytest.mark.parametrize("n,n2", [(32768, 16384), (65536, 32768), (16384, 4096)])
@pytest.mark.parametrize("d", [-1, 1])
def test_split_table_bit_exact(gpu, n, n2, d):
    """Entry (k1-1)*n2 + j = exp(dir*2*pi*i*j*k1/n) for 0 < k1 < n/n2."""
    t = _table_mode(3, n, n2, d)
    ph = [(math.pi * d * (j * k1)) / (n // 2) for k1 in range(1, n // n2) for j in range(n2)]
    h = _host(ph)
    assert t.shape == h.shape
    assert int(np.count_nonzero(t.view(np.uint32) != h.view(np.uint32))) == 0


def test_table_mode_arguments_refused(gpu):
    from mufft_amd import _lib
    L = _lib.load()
    assert L.mufft_hip_twiddle_table_mode(4, 64, 0, -1, None, 0) == -1
    assert L.mufft_hip_twiddle_table_mode(3, 64, 64, -1, None, 0) == -1
    assert L.mufft_hip_twiddle_table_mode(1, 63, 0, -1, None, 0) == -1
    assert L.mufft_hip_twiddle_table_mode(1, 64, 0, 0, None, 0) == -1
