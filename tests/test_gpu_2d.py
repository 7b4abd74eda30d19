"""GPU parity of the 2D transforms (test.c:541-575 sweep shape).

c2c over Nx, Ny = 2..512 both directions; r2c / c2r over Nx, Ny = 4..512 in
the reference's padded layouts (fft.h:238-260); checked against FP64 (per
element, |delta| < 1e-6*sqrt(Nx*Ny), test.c:58) and against the muFFT
restatement (relative RMS).  Larger shapes exercise the generic column path.
"""
import numpy as np
import pytest

from oracle import fp64
from oracle import restate as R
from tests.helpers import abs_tol, cplx, empty_dev, from_dev, real, rel_rms, rms_tol_vs_mufft, to_dev

pytestmark = pytest.mark.gpu

AXES = [2, 4, 8, 32, 128, 512]


@pytest.mark.parametrize("Nx", AXES)
@pytest.mark.parametrize("Ny", AXES)
def test_c2c_2d(gpu, Nx, Ny):
    import mufft_amd as M
    import torch
    B = 2
    x = cplx(0x3001 + Nx * 7 + Ny, B * Nx * Ny, 0.5).reshape(B, Ny, Nx)
    for d in (-1, 1):
        p = M.Plan2D("c2c", Nx, Ny, d)
        dy = empty_dev((B, Ny, Nx), np.complex64, gpu)
        p.execute_batched(dy, to_dev(x, gpu), B)
        torch.cuda.synchronize()
        y = from_dev(dy)
        ref = fp64.c2c_2d(x, d)
        assert np.max(np.abs(y - ref)) < abs_tol(Nx * Ny), (Nx, Ny, d)
        orc = R.execute_plan_2d(R.create_plan_2d_c2c(Nx, Ny, d), x)
        assert rel_rms(y, orc) <= rms_tol_vs_mufft(Nx * Ny)
        out = np.zeros((Ny, Nx), np.complex64)
        p.execute(out, np.ascontiguousarray(x[1]))
        np.testing.assert_array_equal(out.view(np.uint32), y[1].view(np.uint32))


@pytest.mark.parametrize("Nx", [4, 8, 64, 512])
@pytest.mark.parametrize("Ny", [2, 4, 32, 512])
@pytest.mark.parametrize("full", [False, True])
def test_r2c_2d(gpu, Nx, Ny, full):
    import mufft_amd as M
    import torch
    B = 2
    x = real(0x3002 + Nx * 3 + Ny, B * Nx * Ny, 0.5).reshape(B, Ny, Nx)
    p = M.Plan2D("r2c", Nx, Ny, flags=M.FLAG_FULL_R2C if full else 0)
    dy = torch.zeros((B, Ny, Nx), dtype=torch.complex64, device=gpu)
    p.execute_batched(dy, to_dev(x, gpu), B)
    torch.cuda.synchronize()
    y = from_dev(dy)
    ncol = Nx if full else Nx // 2 + 1
    ref = fp64.r2c_2d(x, full=full)[..., :ncol]
    assert np.max(np.abs(y[..., :ncol] - ref)) < abs_tol(Nx * Ny), (Nx, Ny, full)
    out = np.zeros((Ny, Nx), np.complex64)
    p.execute(out, np.ascontiguousarray(x[0]))
    np.testing.assert_array_equal(out[:, :ncol].view(np.uint32), y[0][:, :ncol].view(np.uint32))


@pytest.mark.parametrize("Nx", [4, 8, 64, 512])
@pytest.mark.parametrize("Ny", [2, 4, 32, 512])
def test_c2r_2d(gpu, Nx, Ny):
    import mufft_amd as M
    import torch
    B = 2
    # Hermitian-consistent spectrum of a real image (test.c:125-138 builds one by hand)
    img = real(0x3003 + Nx + Ny, B * Nx * Ny, 0.5).reshape(B, Ny, Nx).astype(np.float64)
    X = np.zeros((B, Ny, Nx), np.complex64)
    X[..., : Nx // 2 + 1] = (np.fft.fft2(img) / (Nx * Ny))[..., : Nx // 2 + 1]
    p = M.Plan2D("c2r", Nx, Ny)
    dy = torch.zeros((B, Ny, Nx), dtype=torch.float32, device=gpu)
    p.execute_batched(dy, to_dev(X, gpu), B)
    torch.cuda.synchronize()
    y = from_dev(dy)
    ref = fp64.c2r_2d(X, Nx)
    np.testing.assert_allclose(ref, img, atol=1e-5)  # the spectrum really is Hermitian
    assert np.max(np.abs(y - ref)) < abs_tol(Nx * Ny, amplitude=1.0), (Nx, Ny)
    orc = R.execute_plan_2d(R.create_plan_2d_c2r(Nx, Ny), X).reshape(B, -1)[:, : Nx * Ny]
    assert rel_rms(y.reshape(B, -1), orc) <= rms_tol_vs_mufft(Nx * Ny)
    out = np.zeros(2 * Nx * Ny, np.float32)  # reference contract: 2*Nx*Ny floats (fft.h:258-260)
    p.execute(out, np.ascontiguousarray(X[1]))
    np.testing.assert_array_equal(out[: Nx * Ny].view(np.uint32), y[1].reshape(-1).view(np.uint32))


def test_c2c_2d_non_hermitian_c2r_matches_restatement(gpu):
    """c2r on an arbitrary (non-Hermitian) spectrum follows the reference's arithmetic."""
    import mufft_amd as M
    import torch
    Nx, Ny = 32, 16
    X = cplx(99, Nx * Ny, 0.5).reshape(Ny, Nx)
    p = M.Plan2D("c2r", Nx, Ny)
    dy = torch.zeros((Ny, Nx), dtype=torch.float32, device=gpu)
    p.execute_batched(dy, to_dev(X, gpu), 1)
    torch.cuda.synchronize()
    orc = R.execute_plan_2d(R.create_plan_2d_c2r(Nx, Ny), X).reshape(-1)[: Nx * Ny]
    assert rel_rms(from_dev(dy).reshape(-1), orc) <= rms_tol_vs_mufft(Nx * Ny)


@pytest.mark.parametrize("Nx,Ny", [(1024, 1024), (4, 16384), (64, 65536), (65536, 4)])
def test_2d_large_and_generic(gpu, Nx, Ny):
    """Benchmark-size image plus shapes that take the generic (multi-launch) paths."""
    import mufft_amd as M
    import torch
    x = cplx(0x3004 + Nx, Nx * Ny, 0.5).reshape(Ny, Nx)
    for d in (-1, 1):
        p = M.Plan2D("c2c", Nx, Ny, d)
        dy = empty_dev((Ny, Nx), np.complex64, gpu)
        p.execute_batched(dy, to_dev(x, gpu), 1)
        torch.cuda.synchronize()
        y = from_dev(dy)
        assert rel_rms(y, fp64.c2c_2d(x, d)) < 2e-6, (Nx, Ny, d)
    xr = real(0x3005 + Nx, Nx * Ny, 0.5).reshape(Ny, Nx)
    pr = M.Plan2D("r2c", Nx, Ny)
    dY = torch.zeros((Ny, Nx), dtype=torch.complex64, device=gpu)
    pr.execute_batched(dY, to_dev(xr, gpu), 1)
    pc = M.Plan2D("c2r", Nx, Ny)
    dz = torch.zeros((Ny, Nx), dtype=torch.float32, device=gpu)
    pc.execute_batched(dz, dY, 1)
    torch.cuda.synchronize()
    Y = from_dev(dY)[:, : Nx // 2 + 1]
    assert rel_rms(Y, fp64.r2c_2d(xr)) < 2e-6
    assert rel_rms(from_dev(dz) / (Nx * Ny), xr) < 2e-6


@pytest.mark.parametrize("kind,n,batch", [("c2c", 1024, 40), ("r2c", 2048, 9), ("c2r", 2048, 9)])
def test_2d_batch_beyond_chunk_size(gpu, kind, n, batch):
    """A batch of more than 256 MiB of images (c2c: run in chunks, each
    chunk's row and column passes back to back): every image must equal the
    same image transformed alone, bit for bit, and the first must match FP64."""
    import torch

    import mufft_amd as M
    g = torch.Generator(device=gpu)
    g.manual_seed(0x3301)
    if kind == "c2c":
        x = torch.randn((batch, n, n), dtype=torch.complex64, device=gpu, generator=g)
        y = torch.empty_like(x)
    elif kind == "r2c":
        x = torch.rand((batch, n, n), device=gpu, generator=g)
        y = torch.empty((batch, n, n), dtype=torch.complex64, device=gpu)
    else:
        # c2r input: the padded layout, Nx/2+1 meaningful bins per row of Nx.
        x = torch.randn((batch, n, n), dtype=torch.complex64, device=gpu, generator=g)
        y = torch.empty((batch, n, n), device=gpu)
    p = M.Plan2D(kind, n, n, M.INVERSE if kind == "c2r" else M.FORWARD)
    assert max(x[0].numel() * x.element_size(), y[0].numel() * y.element_size()) * batch > 256 << 20
    p.execute_batched(y, x, batch)
    one = torch.empty_like(y[0:1])
    cols = n // 2 + 1 if kind == "r2c" else n   # r2c rows: Nx/2+1 bins of a padded row
    for b in (0, batch // 2, batch - 1):
        p.execute_batched(one, x[b:b + 1].clone(), 1)
        assert torch.equal(one[0, :, :cols], y[b, :, :cols]), (kind, b)
    x0 = x[0].cpu().numpy()
    if kind == "c2c":
        ref = np.fft.fft2(x0.astype(np.complex128))
        got = y[0].cpu().numpy()
    elif kind == "r2c":
        ref = np.fft.rfft2(x0.astype(np.float64))
        got = y[0, :, : n // 2 + 1].cpu().numpy()
    else:
        ref = fp64.c2r_2d(x0, n)   # columns first, then the rows' c2r (fft.c:863-916)
        got = y[0].cpu().numpy()
    assert rel_rms(got, ref) < 2e-6, kind
    p.close()
