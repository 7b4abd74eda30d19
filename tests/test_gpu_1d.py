"""GPU parity of the 1D transforms (c2c, r2c, c2r) through the C ABI.

Mirrors the reference sweep (test.c:490-524: N = 2..65536, both directions,
r2c / zero-padded r2c / c2r) with three checks per case:
  * per-element |delta| < 1e-6*sqrt(N) against FP64 (test.c:58, 195) for
    inputs in [-0.5, 0.5];
  * relative RMS <= 5e-7*log2(N) against the float32 restatement of muFFT's
    C path (oracle/restate.py, itself bit-identical to the compiled reference);
  * the drop-in (host buffers, one transform) and the batched device entry
    point give bit-identical results.
"""
import math

import numpy as np
import pytest

from oracle import fp64
from oracle import restate as R
from tests.helpers import abs_tol, cplx, empty_dev, from_dev, real, rel_rms, rms_tol_vs_mufft, to_dev

pytestmark = pytest.mark.gpu

SIZES = [2 ** k for k in range(1, 17)]


def _c2c_case(gpu, N, direction, flags, batch=3, seed=0x1001):
    import mufft_amd as M
    zp = bool(flags & M.FLAG_ZERO_PAD_UPPER_HALF)
    nin = N // 2 if zp else N
    x = cplx(seed + N, batch * nin, 0.5).reshape(batch, nin)
    plan = M.Plan1D("c2c", N, direction, flags)
    dx = to_dev(x, gpu)
    dy = empty_dev((batch, N), np.complex64, gpu)
    plan.execute_batched(dy, dx, batch)
    import torch
    torch.cuda.synchronize()
    y = from_dev(dy)
    xf = np.zeros((batch, N), np.complex64)
    xf[:, :nin] = x
    ref64 = fp64.c2c(xf, direction)
    assert np.max(np.abs(y - ref64)) < abs_tol(N), (N, direction, flags)
    assert rel_rms(y, ref64) < 2e-6
    orc = R.execute_plan_1d(R.create_plan_1d_c2c(N, direction, flags), x)
    assert rel_rms(y, orc) <= rms_tol_vs_mufft(N)
    # drop-in on host buffers, one transform
    out = np.zeros(N, np.complex64)
    plan.execute(out, np.ascontiguousarray(x[1]))
    np.testing.assert_array_equal(out.view(np.uint32), y[1].view(np.uint32))
    return y


@pytest.mark.parametrize("N", SIZES)
@pytest.mark.parametrize("direction", [-1, 1])
def test_c2c(gpu, N, direction):
    _c2c_case(gpu, N, direction, 0)


@pytest.mark.parametrize("N", [2, 8, 64, 1024, 4096, 32768, 65536])
@pytest.mark.parametrize("direction", [-1, 1])
def test_c2c_zero_pad(gpu, N, direction):
    import mufft_amd as M
    _c2c_case(gpu, N, direction, M.FLAG_ZERO_PAD_UPPER_HALF)


@pytest.mark.parametrize("N", [2 ** 17, 2 ** 20])
def test_c2c_large(gpu, N):
    """Beyond one workgroup: generic passes over global memory."""
    import mufft_amd as M
    import torch
    x = cplx(7 + N, N, 0.5)
    for d in (-1, 1):
        plan = M.Plan1D("c2c", N, d)
        dy = empty_dev((N,), np.complex64, gpu)
        plan.execute_batched(dy, to_dev(x, gpu), 1)
        torch.cuda.synchronize()
        y = from_dev(dy)
        ref = fp64.c2c(x, d)
        assert rel_rms(y, ref) < 2e-6
        assert np.max(np.abs(y - ref)) < abs_tol(N)


@pytest.mark.parametrize("N", [4 * 2 ** k for k in range(0, 16)])
@pytest.mark.parametrize("mode", ["plain", "full", "zp", "zp_full"])
def test_r2c(gpu, N, mode):
    import mufft_amd as M
    import torch
    flags = {"plain": 0, "full": M.FLAG_FULL_R2C, "zp": M.FLAG_ZERO_PAD_UPPER_HALF,
             "zp_full": M.FLAG_FULL_R2C | M.FLAG_ZERO_PAD_UPPER_HALF}[mode]
    zp = mode.startswith("zp")
    full = mode.endswith("full")
    batch = 2
    nin = N // 2 if zp else N
    x = real(0x2002 + N, batch * nin, 0.5).reshape(batch, nin)
    plan = M.Plan1D("r2c", N, flags=flags)
    nout = N if full else N // 2 + 1
    dy = empty_dev((batch, nout), np.complex64, gpu)
    plan.execute_batched(dy, to_dev(x, gpu), batch)
    torch.cuda.synchronize()
    y = from_dev(dy)
    xf = np.zeros((batch, N), np.float32)
    xf[:, :nin] = x
    ref = fp64.r2c(xf, full=full)
    assert np.max(np.abs(y - ref)) < abs_tol(N), (N, mode)
    if N >= 8:
        orc = R.execute_plan_1d(R.create_plan_1d_r2c(N, flags), x)
        assert rel_rms(y, orc[:, :nout]) <= rms_tol_vs_mufft(N)
    if full:  # conjugate symmetry, test.c:295-302
        k = np.arange(1, N // 2)
        assert np.max(np.abs(y[:, k] - np.conj(y[:, N - k]))) < abs_tol(N)
    out = np.zeros(N + 2, np.complex64)
    plan.execute(out, np.ascontiguousarray(x[0]))
    np.testing.assert_array_equal(out[:nout].view(np.uint32), y[0].view(np.uint32))


@pytest.mark.parametrize("N", [4 * 2 ** k for k in range(0, 16)])
@pytest.mark.parametrize("hermitian", [True, False])
def test_c2r(gpu, N, hermitian):
    import mufft_amd as M
    import torch
    batch = 2
    X = cplx(0x2003 + N, batch * (N // 2 + 1), 0.5).reshape(batch, N // 2 + 1)
    if hermitian:  # real DC and Nyquist, test.c:218-226
        X[:, 0] = X[:, 0].real
        X[:, N // 2] = X[:, N // 2].real
    plan = M.Plan1D("c2r", N)
    dy = empty_dev((batch, N), np.float32, gpu)
    plan.execute_batched(dy, to_dev(X, gpu), batch)
    torch.cuda.synchronize()
    y = from_dev(dy)
    ref = fp64.c2r(X, N)  # keeps the Im(X0), Im(X_{N/2}) terms (kernel.c:36-45)
    assert np.max(np.abs(y - ref)) < abs_tol(N, factor=2e-6), (N, hermitian)
    orc = R.execute_plan_1d(R.create_plan_1d_c2r(N), X)
    assert rel_rms(y, orc) <= rms_tol_vs_mufft(N)
    out = np.zeros(N, np.float32)
    plan.execute(out, np.ascontiguousarray(X[1]))
    np.testing.assert_array_equal(out.view(np.uint32), y[1].view(np.uint32))


@pytest.mark.parametrize("N", [2 ** 17, 2 ** 18])
def test_real_large(gpu, N):
    """r2c / c2r beyond the fused kernels (generic passes + standalone resolve)."""
    import mufft_amd as M
    import torch
    x = real(11 + N, N, 0.5)
    p = M.Plan1D("r2c", N)
    dX = empty_dev((N // 2 + 1,), np.complex64, gpu)
    p.execute_batched(dX, to_dev(x, gpu), 1)
    q = M.Plan1D("c2r", N)
    dy = empty_dev((N,), np.float32, gpu)
    q.execute_batched(dy, dX, 1)
    torch.cuda.synchronize()
    X = from_dev(dX)
    assert rel_rms(X, fp64.r2c(x)) < 2e-6
    assert rel_rms(from_dev(dy) / N, x) < 2e-6


def test_invalid_sizes(gpu):
    import mufft_amd as M
    for bad in (0, 1, 3, 6, 1000):
        with pytest.raises(M.MufftError):
            M.Plan1D("c2c", bad, -1)
    for bad in (0, 1, 2, 6):  # r2c/c2r N=2 is refused like the reference (fft.c:485)
        with pytest.raises(M.MufftError):
            M.Plan1D("r2c", bad)
        with pytest.raises(M.MufftError):
            M.Plan1D("c2r", bad)


def test_roundtrip_batched_4096(gpu):
    """Forward then inverse recovers N*x (fft.h:70-71) on a large batch."""
    import mufft_amd as M
    import torch
    N, B = 4096, 512
    x = cplx(0x2001, N * B).reshape(B, N)
    f = M.Plan1D("c2c", N, -1)
    i = M.Plan1D("c2c", N, 1)
    dx = to_dev(x, gpu)
    dX = torch.empty_like(dx)
    dz = torch.empty_like(dx)
    f.execute_batched(dX, dx, B)
    i.execute_batched(dz, dX, B)
    torch.cuda.synchronize()
    z = from_dev(dz) / N
    assert rel_rms(z, x) < 2e-6
    # per-transform equality with single-transform execution
    one = torch.empty((N,), dtype=torch.complex64, device=gpu)
    f.execute_batched(one, dx[17], 1)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(from_dev(one).view(np.uint32), from_dev(dX[17]).view(np.uint32))
    assert math.isfinite(float(np.abs(from_dev(dX)).max()))


@pytest.mark.parametrize("kind", ["r2c", "c2r"])
def test_65536_real_small_and_large_launches(gpu, kind):
    """A 65536-point real plan on a launch of 3 transforms and on one of more
    than two transforms per CU (several rounds of workgroups): both against
    FP64 on the same leading transforms (test.c:58 bound), and the two results
    within rounding of each other."""
    import mufft_amd as M
    import torch
    from mufft_amd import _lib
    L = _lib.load()
    N = 65536
    tk = 2 if kind == "r2c" else 3
    assert L.mufft_hip_get_variant(N, tk) == L.mufft_hip_default_variant(N, tk)   # nothing chosen
    big = 2 * torch.cuda.get_device_properties(0).multi_processor_count + 8
    p = M.Plan1D(kind, N)
    outs = []
    for B in (3, big):
        if kind == "r2c":
            x = real(0x6501, B * N, 0.5).reshape(B, N)
            y = empty_dev((B, N // 2 + 1), np.complex64, gpu)
            ref = fp64.r2c(x[:3])
        else:
            x = cplx(0x6502, B * (N // 2 + 1), 0.5).reshape(B, N // 2 + 1)
            y = empty_dev((B, N), np.float32, gpu)
            ref = fp64.c2r(x[:3], N)
        p.execute_batched(y, to_dev(x, gpu), B)
        torch.cuda.synchronize()
        got = from_dev(y)[:3]
        assert np.max(np.abs(got - ref)) < abs_tol(N), (kind, B)
        outs.append(got)
    assert rel_rms(outs[1], outs[0]) < 1e-6
    p.close()
