"""Calls that share one plan from several streams, and per-plan variants.

Plan-owned scratch (the 2D c2r intermediate image, the four-step buffers,
the adjoint's folded spectrum) is shared by every call on the plan.  Calls
issued back to back on different streams without a host synchronisation
must give the same bits as the same calls run one after another: the
library orders the scratch users in GPU order.  The reference leaves
concurrent use of one plan undefined (its tmp_buffer is shared, fft.c:815).
"""
import numpy as np
import pytest

from oracle import fp64
from tests.helpers import cplx, from_dev, real, rel_rms, to_dev

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def test_2d_c2r_batched_on_two_streams(gpu):
    import mufft_amd as M
    import torch
    Nx, Ny, B = 256, 128, 3
    p = M.Plan2D("c2r", Nx, Ny)
    X = [to_dev(cplx(0x7001 + i, B * Ny * Nx, 0.5).reshape(B, Ny, Nx), gpu) for i in range(2)]
    seq = []
    for i in range(2):
        y = torch.zeros((B, Ny, Nx), dtype=torch.float32, device=gpu)
        p.execute_batched(y, X[i], B)
        torch.cuda.synchronize()
        seq.append(from_dev(y))
    s = [torch.cuda.Stream(gpu), torch.cuda.Stream(gpu)]
    ys = [torch.zeros((B, Ny, Nx), dtype=torch.float32, device=gpu) for _ in range(2)]
    for rep in range(3):
        for i in range(2):
            p.execute_batched(ys[i], X[i], B, stream=s[i])
        # and a drop-in call on the plan's own stream right behind them
        out = np.zeros(Nx * Ny, np.float32)
        p.execute(out, X[0][1])
        torch.cuda.synchronize()
        for i in range(2):
            np.testing.assert_array_equal(_bits(from_dev(ys[i])), _bits(seq[i]))
        np.testing.assert_array_equal(_bits(out), _bits(seq[0][1].reshape(-1)))


def test_four_step_batched_on_two_streams(gpu):
    import mufft_amd as M
    import torch
    N, B = 1 << 18, 2
    p = M.Plan1D("c2c", N, M.FORWARD)
    assert p.launches >= 2
    X = [to_dev(cplx(0x7101 + i, B * N, 0.5).reshape(B, N), gpu) for i in range(2)]
    seq = []
    for i in range(2):
        y = torch.empty_like(X[i])
        p.execute_batched(y, X[i], B)
        torch.cuda.synchronize()
        seq.append(from_dev(y))
    assert rel_rms(seq[0], fp64.c2c(from_dev(X[0]), -1)) < 2e-6
    s = [torch.cuda.Stream(gpu), torch.cuda.Stream(gpu)]
    ys = [torch.empty_like(X[i]) for i in range(2)]
    for i in range(2):
        p.execute_batched(ys[i], X[i], B, stream=s[i])
    torch.cuda.synchronize()
    for i in range(2):
        np.testing.assert_array_equal(_bits(from_dev(ys[i])), _bits(seq[i]))


def test_adjoint_on_two_streams(gpu):
    import mufft_amd as M
    import torch
    N, B = 4096, 4
    p = M.Plan1D("r2c", N, flags=M.FLAG_FULL_R2C)   # the adjoint folds into plan scratch
    G = [to_dev(cplx(0x7201 + i, B * N, 0.5).reshape(B, N), gpu) for i in range(2)]
    seq = []
    for i in range(2):
        y = torch.empty((B, N), dtype=torch.float32, device=gpu)
        p.execute_adjoint_batched(y, G[i], B)
        torch.cuda.synchronize()
        seq.append(from_dev(y))
    s = [torch.cuda.Stream(gpu), torch.cuda.Stream(gpu)]
    ys = [torch.empty((B, N), dtype=torch.float32, device=gpu) for _ in range(2)]
    for i in range(2):
        p.execute_adjoint_batched(ys[i], G[i], B, stream=s[i])
    torch.cuda.synchronize()
    for i in range(2):
        np.testing.assert_array_equal(_bits(from_dev(ys[i])), _bits(seq[i]))


@pytest.mark.parametrize("N,kind", [(4096, "c2c"), (65536, "r2c"), (65536, "c2r")])
def test_variant_per_plan(gpu, N, kind):
    """A plan created with an explicit variant runs it and leaves the
    process-wide choice alone."""
    import ctypes

    import mufft_amd as M
    import torch
    from mufft_amd import _lib
    from mufft_amd import api as A
    L = _lib.load()
    code = {"c2c": 0, "r2c": 2, "c2r": 3}[kind]
    buf = (ctypes.c_int * 64)()
    nv = L.mufft_hip_list_variants(N, code, buf, 64)
    before = L.mufft_hip_get_variant(N, code)
    B = 4
    if kind == "c2c":
        x = cplx(0x7301, B * N, 0.5).reshape(B, N)
    elif kind == "r2c":
        x = real(0x7302, B * N, 0.5).reshape(B, N)
    else:
        x = fp64.r2c(real(0x7303, B * N, 0.5).reshape(B, N)).astype(np.complex64) / np.float32(N)
    outs = []
    for v in list(buf[:nv]):
        p = M.Plan1D(kind, N, M.FORWARD, variant=v)
        assert p.variant == v
        shape = (B, N) if kind != "r2c" else (B, N // 2 + 1)
        dt = torch.float32 if kind == "c2r" else torch.complex64
        y = torch.empty(shape, dtype=dt, device=gpu)
        p.execute_batched(y, to_dev(x, gpu), B)
        torch.cuda.synchronize()
        outs.append(from_dev(y))
        assert L.mufft_hip_get_variant(N, code) == before
    ref = {"c2c": lambda: fp64.c2c(x, -1), "r2c": lambda: fp64.r2c(x),
           "c2r": lambda: fp64.c2r(x, N)}[kind]()
    for v, y in zip(list(buf[:nv]), outs):
        assert rel_rms(y, ref) < 2e-6, hex(v)
    with pytest.raises(A.MufftError):
        M.Plan1D(kind, N, M.FORWARD, variant=0x7F7F)
