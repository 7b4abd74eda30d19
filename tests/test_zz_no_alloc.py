"""The allocation-free batched 1D entry point against the allocating one.

mufft_hip_execute_plan_1d_batched_noalloc runs the same kernels as
mufft_hip_execute_plan_1d_batched, with the scratch of the multi-launch
paths (four-step, standalone resolve) in a fixed per-GPU device array
instead of plan-owned allocations.  Its output must be bit-identical to the
allocating call's, on two calls in a row on the same plan, the second with a
smaller batch (the shared scratch keeps the first call's contents).  Sizes
beyond the array are refused before any device work.  Both calls go through
the C ABI (ctypes).
"""
import ctypes

import numpy as np
import pytest

from oracle import fp64
from tests.helpers import cplx, from_dev, real, rel_rms, to_dev

pytestmark = pytest.mark.gpu

LIMIT_POINTS = 1 << 22


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _run(fn, plan, out, inp, howmany):
    import mufft_amd as M
    import torch
    rc = fn(plan.handle, ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(inp.data_ptr()),
            howmany, 0, 0, None)
    assert rc == 0, M._lib.load().mufft_hip_get_last_error().decode()
    torch.cuda.synchronize()
    return from_dev(out)


# (kind, N as passed to the creator, batch of the first call, of the second)
CASES = [
    ("c2c", 1 << 17, 8, 3),     # four-step, two launches
    ("c2c", 1 << 20, 4, 1),     # four-step, larger split
    ("c2c", 1 << 22, 1, 1),     # four-step at the scratch limit
    ("r2c", 1 << 18, 16, 5),    # four-step + standalone resolve (both regions)
    ("c2r", 1 << 18, 16, 5),    # standalone resolve + four-step
    ("c2c", 4096, 64, 7),       # fused: no scratch at all
    ("r2c", 65536, 6, 2),       # fused two-workgroup split: no scratch
]


@pytest.mark.parametrize("kind,N,b1,b2", CASES)
def test_noalloc_matches_batched(gpu, kind, N, b1, b2):
    import mufft_amd as M
    import torch
    print(f"noalloc {kind} N={N} batches {b1},{b2}", flush=True)
    lib = M._lib.load()
    if kind == "c2c":
        plan = M.Plan1D("c2c", N, M.FORWARD)
        make_in = lambda b, s: to_dev(cplx(s, b * N, 0.5).reshape(b, N), gpu)
        out_shape, out_dt = (lambda b: (b, N)), torch.complex64
    elif kind == "r2c":
        plan = M.Plan1D("r2c", N)
        make_in = lambda b, s: to_dev(real(s, b * N, 0.5).reshape(b, N), gpu)
        out_shape, out_dt = (lambda b: (b, N // 2 + 1)), torch.complex64
    else:
        plan = M.Plan1D("c2r", N)
        make_in = lambda b, s: to_dev(cplx(s, b * (N // 2 + 1), 0.5).reshape(b, N // 2 + 1), gpu)
        out_shape, out_dt = (lambda b: (b, N)), torch.float32
    for b, seed in ((b1, 0x9001), (b2, 0x9002)):
        x = make_in(b, seed + N)
        ya = torch.full(out_shape(b), float("nan"), dtype=out_dt, device=gpu)
        yn = torch.full(out_shape(b), float("nan"), dtype=out_dt, device=gpu)
        ref = _run(lib.mufft_hip_execute_plan_1d_batched, plan, ya, x, b)
        got = _run(lib.mufft_hip_execute_plan_1d_batched_noalloc, plan, yn, x, b)
        np.testing.assert_array_equal(_bits(got), _bits(ref))
        if kind == "c2c":
            xh = from_dev(x)
            assert rel_rms(got[-1], fp64.c2c(xh[-1], -1)) < 2e-6
    plan.close()


def test_noalloc_refuses_beyond_scratch(gpu):
    import mufft_amd as M
    import torch
    lib = M._lib.load()
    N = 1 << 20
    plan = M.Plan1D("c2c", N, M.FORWARD)
    b = LIMIT_POINTS // N + 1
    x = torch.zeros((b, N), dtype=torch.complex64, device=gpu)
    y = torch.zeros_like(x)
    rc = lib.mufft_hip_execute_plan_1d_batched_noalloc(plan.handle, ctypes.c_void_p(y.data_ptr()),
                                                       ctypes.c_void_p(x.data_ptr()), b, 0, 0, None)
    assert rc != 0
    assert b"at most" in lib.mufft_hip_get_last_error()
    plan.close()


def test_noalloc_refuses_pageable_memory(gpu):
    import mufft_amd as M
    lib = M._lib.load()
    N = 1 << 17
    plan = M.Plan1D("c2c", N, M.FORWARD)
    x = np.zeros(N, np.complex64)
    y = np.zeros(N, np.complex64)
    rc = lib.mufft_hip_execute_plan_1d_batched_noalloc(plan.handle, y.ctypes.data_as(ctypes.c_void_p),
                                                       x.ctypes.data_as(ctypes.c_void_p), 1, 0, 0, None)
    assert rc != 0
    plan.close()
