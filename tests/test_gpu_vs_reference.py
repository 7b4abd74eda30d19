"""GPU output against muFFT's own CPU outputs on identical seeded inputs.

Two sources: the committed golden vectors (written by the compiled reference,
AVX path and scalar path, tests/golden/make_golden.py) and, where the compiled
reference library is present next to the oracle, live runs of it at the
benchmark sizes.  Bound: relative RMS <= 5e-7*log2(N) (2D: log2(Nx*Ny)).
"""
import math
from pathlib import Path

import numpy as np
import pytest

from oracle import fp64, inputs
from tests.helpers import empty_dev, from_dev, to_dev

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden" / "mufft_ref_vectors.npz"


def bound(n):
    return 5e-7 * max(1.0, math.log2(n))


@pytest.fixture(scope="module")
def gold():
    return dict(np.load(GOLD))


def _cases(g, prefix):
    return sorted({k[:-4] for k in g if k.startswith(prefix) and k.endswith("_out")})


def test_golden_1d(gpu, gold):
    import mufft_amd as M
    for key in _cases(gold, "c2c_"):
        _, isa, N, d = key.split("_")
        N, d = int(N), int(d)
        out = np.zeros(N, np.complex64)
        M.Plan1D("c2c", N, d).execute(out, gold[key + "_in"])
        assert fp64.rel_rms(out, gold[key + "_out"]) <= bound(N), key
    for key in _cases(gold, "r2c_"):
        _, isa, N, mode = key.split("_")
        N = int(N)
        fl = {"plain": 0, "full": M.FLAG_FULL_R2C, "zp": M.FLAG_ZERO_PAD_UPPER_HALF}[mode]
        ref = gold[key + "_out"]
        out = np.zeros(N + 2, np.complex64)
        M.Plan1D("r2c", N, flags=fl).execute(out, gold[key + "_in"])
        assert fp64.rel_rms(out[: len(ref)], ref) <= bound(N), key
    for key in _cases(gold, "c2r_"):
        _, isa, N = key.split("_")
        N = int(N)
        out = np.zeros(N, np.float32)
        M.Plan1D("c2r", N).execute(out, gold[key + "_in"])
        assert fp64.rel_rms(out, gold[key + "_out"]) <= bound(N), key


def test_golden_2d(gpu, gold):
    import mufft_amd as M
    for key in _cases(gold, "c2c2d_"):
        _, isa, shp, d = key.split("_")
        Nx, Ny = map(int, shp.split("x"))
        out = np.zeros((Ny, Nx), np.complex64)
        M.Plan2D("c2c", Nx, Ny, int(d)).execute(out, gold[key + "_in"])
        assert fp64.rel_rms(out, gold[key + "_out"]) <= bound(Nx * Ny), key
    for key in _cases(gold, "r2c2d_"):
        _, isa, shp = key.split("_")
        Nx, Ny = map(int, shp.split("x"))
        out = np.zeros((Ny, Nx), np.complex64)
        M.Plan2D("r2c", Nx, Ny).execute(out, gold[key + "_in"])
        assert fp64.rel_rms(out[:, : Nx // 2 + 1], gold[key + "_out"]) <= bound(Nx * Ny), key
    for key in _cases(gold, "c2r2d_"):
        _, isa, shp = key.split("_")
        Nx, Ny = map(int, shp.split("x"))
        out = np.zeros(2 * Nx * Ny, np.float32)
        M.Plan2D("c2r", Nx, Ny).execute(out, gold[key + "_in"])
        assert fp64.rel_rms(out[: Nx * Ny].reshape(Ny, Nx), gold[key + "_out"]) <= bound(Nx * Ny), key


def test_golden_conv(gpu, gold):
    import mufft_amd as M
    for key in _cases(gold, "conv_"):
        _, isa, N, stereo = key.split("_")
        N, stereo = int(N), int(stereo)
        p = M.PlanConv(N, 0, stereo | 4 | 8)
        b0 = np.zeros(p.block_size // 8, np.complex64)
        b1 = np.zeros(p.block_size // 8, np.complex64)
        p.execute_input(0, b0, gold[key + "_a"])
        p.execute_input(1, b1, gold[key + "_b"])
        out = np.zeros(N, np.complex64 if stereo else np.float32)
        p.execute_output(out, b0, b1)
        assert fp64.rel_rms(out, gold[key + "_out"]) <= 2 * bound(N), key


@pytest.mark.parametrize("N", [1024, 4096, 65536])
def test_live_reference_avx_path(gpu, ref_lib, N):
    """Benchmark-size transforms against the compiled reference on this host."""
    import mufft_amd as M
    import torch
    B = 4
    x = inputs.complex_uniform(0x2001, B * N).reshape(B, N)
    for d in (-1, 1):
        p = M.Plan1D("c2c", N, d)
        dy = empty_dev((B, N), np.complex64, gpu)
        p.execute_batched(dy, to_dev(x, gpu), B)
        torch.cuda.synchronize()
        y = from_dev(dy)
        for b in range(B):
            assert fp64.rel_rms(y[b], ref_lib.fft_1d("c2c", N, x[b], d, 0)) <= bound(N)
    xr = inputs.uniform(0x2002, 2 * N)
    X = np.zeros(N + 1, np.complex64)
    M.Plan1D("r2c", 2 * N).execute(X, xr)
    assert fp64.rel_rms(X, ref_lib.fft_1d("r2c", 2 * N, xr, -1, 0)) <= bound(2 * N)
    z = np.zeros(2 * N, np.float32)
    M.Plan1D("c2r", 2 * N).execute(z, X)
    assert fp64.rel_rms(z, ref_lib.fft_1d("c2r", 2 * N, X, 1, 0)) <= bound(2 * N)
