"""GPU parity of the long fused column kernels and of the benchmark 2D shapes.

The reference runs every vertical pass over the whole image
(kernel.c:455-738, driven by fft.c:852-989).  Here one fused column kernel
per length runs all passes of a tile of adjacent columns; the lengths 2048,
4096 and 8192 use 4-column tiles, the split LDS exchange and split twiddles.
Each shape is checked per element against FP64 with the reference bound
|delta| < 1e-6*sqrt(Nx*Ny) (test.c:58) on zero-mean inputs, and against the
muFFT restatement (relative RMS <= 5e-7*log2(Nx*Ny)).  The benchmark shapes
(2D c2c 1024^2 x 16, r2c/c2r 2048^2 x 4) run at full batch.
"""
import math

import numpy as np
import pytest

from oracle import fp64
from oracle import restate as R
from tests.helpers import abs_tol, cplx, empty_dev, from_dev, real, rel_rms, rms_tol_vs_mufft, to_dev

pytestmark = pytest.mark.gpu

# (Nx, Ny): columns of 2048, 4096 and 8192 points; narrow images keep the
# FP64 and restatement references fast.
COLUMN_SHAPES = [(64, 2048), (32, 4096), (16, 8192), (8, 8192)]


def _flush(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("Nx,Ny", COLUMN_SHAPES)
@pytest.mark.parametrize("d", [-1, 1])
def test_c2c_long_columns(gpu, Nx, Ny, d):
    import mufft_amd as M
    import torch
    _flush(f"c2c 2D Nx={Nx} Ny={Ny} dir={d}")
    B = 2
    x = cplx(0x3101 + Nx + Ny, B * Nx * Ny, 0.5).reshape(B, Ny, Nx)
    p = M.Plan2D("c2c", Nx, Ny, d)
    assert p.launches == 2
    dy = empty_dev((B, Ny, Nx), np.complex64, gpu)
    p.execute_batched(dy, to_dev(x, gpu), B)
    torch.cuda.synchronize()
    y = from_dev(dy)
    assert np.max(np.abs(y - fp64.c2c_2d(x, d))) < abs_tol(Nx * Ny), (Nx, Ny, d)
    orc = R.execute_plan_2d(R.create_plan_2d_c2c(Nx, Ny, d), x[1])
    assert rel_rms(y[1], orc) <= rms_tol_vs_mufft(Nx * Ny)


@pytest.mark.parametrize("Nx,Ny", COLUMN_SHAPES)
@pytest.mark.parametrize("full", [False, True])
def test_r2c_long_columns(gpu, Nx, Ny, full):
    import mufft_amd as M
    import torch
    _flush(f"r2c 2D Nx={Nx} Ny={Ny} full={full}")
    B = 2
    x = real(0x3102 + Nx + Ny, B * Nx * Ny, 0.5).reshape(B, Ny, Nx)
    p = M.Plan2D("r2c", Nx, Ny, flags=M.FLAG_FULL_R2C if full else 0)
    dy = torch.zeros((B, Ny, Nx), dtype=torch.complex64, device=gpu)
    p.execute_batched(dy, to_dev(x, gpu), B)
    torch.cuda.synchronize()
    y = from_dev(dy)
    ncol = Nx if full else Nx // 2 + 1
    ref = fp64.r2c_2d(x, full=full)[..., :ncol]
    assert np.max(np.abs(y[..., :ncol] - ref)) < abs_tol(Nx * Ny), (Nx, Ny, full)
    fl = R.FLAG_FULL_R2C if full else 0
    orc = R.execute_plan_2d(R.create_plan_2d_r2c(Nx, Ny, fl), x[0]).reshape(Ny, Nx)[:, :ncol]
    assert rel_rms(y[0][:, :ncol], orc) <= rms_tol_vs_mufft(Nx * Ny)


@pytest.mark.parametrize("Nx,Ny", COLUMN_SHAPES)
def test_c2r_long_columns(gpu, Nx, Ny):
    import mufft_amd as M
    import torch
    _flush(f"c2r 2D Nx={Nx} Ny={Ny}")
    B = 2
    img = real(0x3103 + Nx + Ny, B * Nx * Ny, 0.5).reshape(B, Ny, Nx).astype(np.float64)
    X = np.zeros((B, Ny, Nx), np.complex64)
    X[..., : Nx // 2 + 1] = (np.fft.fft2(img) / math.sqrt(Nx * Ny))[..., : Nx // 2 + 1]
    p = M.Plan2D("c2r", Nx, Ny)
    dy = torch.zeros((B, Ny, Nx), dtype=torch.float32, device=gpu)
    p.execute_batched(dy, to_dev(X, gpu), B)
    torch.cuda.synchronize()
    y = from_dev(dy)
    ref = fp64.c2r_2d(X, Nx)
    # bins scaled by 1/sqrt(Nx*Ny): outputs of amplitude sqrt(Nx*Ny)*0.5
    assert np.max(np.abs(y - ref)) < abs_tol(Nx * Ny, amplitude=0.5 * math.sqrt(Nx * Ny)), (Nx, Ny)
    orc = R.execute_plan_2d(R.create_plan_2d_c2r(Nx, Ny), X[1]).reshape(-1)[: Nx * Ny]
    assert rel_rms(y[1].reshape(-1), orc) <= rms_tol_vs_mufft(Nx * Ny)


def _pixels(seed, shape):
    """Benchmark pixel law: uniform[0, 1) plus Gaussian noise of sigma 0.1."""
    from oracle import inputs
    n = int(np.prod(shape))
    u = (inputs.uniform(seed, n) + 1.0) * 0.5
    g = inputs.gaussian(seed + 1, n, sigma=0.1)
    return (u + g).astype(np.float32).reshape(shape)


def test_r2c_c2r_2048_squared_batch_of_4(gpu):
    """The 2D real benchmark shape at its full batch: per element against FP64
    on zero-mean pixels, relative RMS on the benchmark's pixel law, the
    restatement on one image, and the r2c -> c2r round trip."""
    import mufft_amd as M
    import torch
    n, B = 2048, 4
    _flush("2D r2c/c2r 2048x2048 x 4")
    pr = M.Plan2D("r2c", n, n)
    pc = M.Plan2D("c2r", n, n)
    for law in ("zero_mean", "pixels"):
        x = (real(0x3002, B * n * n, 0.5).reshape(B, n, n) if law == "zero_mean"
             else _pixels(0x3002, (B, n, n)))
        dx = to_dev(x, gpu)
        dY = torch.zeros((B, n, n), dtype=torch.complex64, device=gpu)
        dz = torch.zeros((B, n, n), dtype=torch.float32, device=gpu)
        pr.execute_batched(dY, dx, B)
        pc.execute_batched(dz, dY, B)
        torch.cuda.synchronize()
        Y = from_dev(dY)[..., : n // 2 + 1]
        z = from_dev(dz)
        for b in range(B):
            ref = fp64.r2c_2d(x[b])
            if law == "zero_mean":
                assert np.max(np.abs(Y[b] - ref)) < abs_tol(n * n), b
            assert rel_rms(Y[b], ref) < 2e-6, (law, b)
            # c2r of the GPU spectrum against FP64 c2r of the same spectrum
            zr = fp64.c2r_2d(Y[b], n)
            assert rel_rms(z[b], zr) < 2e-6, (law, b)
            assert rel_rms(z[b] / (n * n), x[b]) < 2e-6, (law, b)
        if law == "zero_mean":
            orc = R.execute_plan_2d(R.create_plan_2d_r2c(n, n), x[3]).reshape(n, n)[:, : n // 2 + 1]
            assert rel_rms(Y[3], orc) <= rms_tol_vs_mufft(n * n)
        del dx, dY, dz


def test_c2c_1024_squared_batch_of_16(gpu):
    """The 2D complex benchmark shape at its full batch, both directions."""
    import mufft_amd as M
    import torch
    n, B = 1024, 16
    _flush("2D c2c 1024x1024 x 16")
    x = cplx(0x3001, B * n * n, 0.5).reshape(B, n, n)
    dx = to_dev(x, gpu)
    for d in (-1, 1):
        p = M.Plan2D("c2c", n, n, d)
        dy = empty_dev((B, n, n), np.complex64, gpu)
        p.execute_batched(dy, dx, B)
        torch.cuda.synchronize()
        y = from_dev(dy)
        for b in (0, 7, 15):
            assert np.max(np.abs(y[b] - fp64.c2c_2d(x[b], d))) < abs_tol(n * n), (d, b)
        orc = R.execute_plan_2d(R.create_plan_2d_c2c(n, n, d), x[15])
        assert rel_rms(y[15], orc) <= rms_tol_vs_mufft(n * n)
        del dy
