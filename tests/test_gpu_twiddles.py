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
