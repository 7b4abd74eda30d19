"""Tuned schedules beyond the fused row variants: every runnable split of the
large transforms (two or three column launches) against FP64, the
autotuner's pick, and the alternative column tile widths, which must give
the same bits as the default width (a tile only selects which columns a
workgroup owns).  Choices are process-wide; each test restores the static
ones."""
import numpy as np
import pytest

from oracle import fp64
from tests.helpers import abs_tol, cplx, empty_dev, from_dev, real, rel_rms, to_dev

pytestmark = pytest.mark.gpu


def _codes(logn):
    out = []
    for n1 in range(6, 14):
        for a, b in [(0, 0)] + [(a, logn - n1 - a) for a in range(6, 11)]:
            code = (n1 << 8) | (a << 4) | b
            out.append(code)
    return out


@pytest.mark.parametrize("logn", [17, 20, 23, 24])
def test_every_runnable_split_matches_fp64(gpu, logn):
    import mufft_amd as M
    import torch
    from mufft_amd import _lib
    L = _lib.load()
    N = 1 << logn
    B = 2 if logn <= 20 else 1
    x = cplx(0x8100 + logn, B * N, 0.5).reshape(B, N)
    ref = fp64.c2c(x, -1)
    dx = to_dev(x, gpu)
    tried = 0
    try:
        for code in _codes(logn):
            if L.mufft_hip_set_large_split(N, 0, code) != 0:
                continue
            print(f"split 2^{logn} code {code:#x}", flush=True)
            p = M.Plan1D("c2c", N, M.FORWARD)
            assert p.launches == (3 if code & 0xFF else 2)
            dy = empty_dev((B, N), np.complex64, gpu)
            p.execute_batched(dy, dx, B)
            torch.cuda.synchronize()
            y = from_dev(dy)
            assert np.max(np.abs(y - ref)) < abs_tol(N), hex(code)
            assert rel_rms(y, ref) < 2e-6, hex(code)
            p.close()
            tried += 1
    finally:
        L.mufft_hip_set_large_split(N, 0, 0)
    assert tried >= 2


def test_autotuned_split_is_used_by_new_plans(gpu):
    import mufft_amd as M
    import torch
    from mufft_amd import api
    N, B = 1 << 18, 4
    try:
        code, us = api.autotune_large(N, "c2c_forward", B)
        assert us > 0 and api.get_large_split(N, "c2c_forward") == code
        for kind in ("r2c", "c2r"):
            # real transforms of 2^19 reals use the same 2^18-point split
            assert api.get_large_split(2 * N, kind) == code
        x = cplx(0x8200, B * N, 0.5).reshape(B, N)
        p = M.Plan1D("c2c", N, M.FORWARD)
        dy = empty_dev((B, N), np.complex64, gpu)
        p.execute_batched(dy, to_dev(x, gpu), B)
        torch.cuda.synchronize()
        assert rel_rms(from_dev(dy), fp64.c2c(x, -1)) < 2e-6
        xr = real(0x8201, 2 * N, 0.5)
        pr = M.Plan1D("r2c", 2 * N)
        dS = torch.zeros(N + 1, dtype=torch.complex64, device=gpu)
        pr.execute_batched(dS, to_dev(xr, gpu), 1)
        torch.cuda.synchronize()
        assert rel_rms(from_dev(dS), fp64.r2c(xr)) < 2e-6
    finally:
        api.set_large_split(N, "c2c_forward", 0)


@pytest.mark.parametrize("Ny", [128, 256, 512, 1024, 2048, 4096, 8192])
def test_column_tile_alternative_gives_identical_bits(gpu, Ny):
    import mufft_amd as M
    import torch
    from mufft_amd import api
    # 16 columns where the tiles are wider than the image (partial tile)
    Nx, B = (16 if Ny <= 256 or Ny >= 4096 else 64), 2
    x = cplx(0x8300 + Ny, B * Ny * Nx, 0.5).reshape(B, Ny, Nx)
    dx = to_dev(x, gpu)
    outs = []
    try:
        for tile in (0, -1):
            alt = {128: 64, 256: 32, 512: 16, 1024: 8, 2048: 8, 4096: 2, 8192: 2}[Ny]
            api.set_column_tile(Ny, 0 if tile == 0 else alt)
            print(f"2D columns {Ny} tile {api.get_column_tile(Ny)}", flush=True)
            p = M.Plan2D("c2c", Nx, Ny, M.FORWARD)
            dy = empty_dev((B, Ny, Nx), np.complex64, gpu)
            p.execute_batched(dy, dx, B)
            torch.cuda.synchronize()
            outs.append(from_dev(dy))
            p.close()
    finally:
        api.set_column_tile(Ny, 0)
    np.testing.assert_array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))
    assert rel_rms(outs[1], fp64.c2c_2d(x, -1)) < 2e-6


def test_column_tile_alternative_real_2d(gpu):
    """r2c and c2r 2048 x 2048 (the benchmark shape) with the alternative
    8-column tiles: the same bits as the 4-column default (1025 columns: the
    last tile is partial)."""
    import mufft_amd as M
    import torch
    from mufft_amd import api
    n, B = 2048, 1
    x = real(0x8400, B * n * n, 0.5).reshape(B, n, n)
    dx = to_dev(x, gpu)
    res = []
    try:
        for tile in (0, 8):
            api.set_column_tile(n, tile)
            pr, pc = M.Plan2D("r2c", n, n), M.Plan2D("c2r", n, n)
            dY = torch.zeros((B, n, n), dtype=torch.complex64, device=gpu)
            dz = torch.zeros((B, n, n), dtype=torch.float32, device=gpu)
            pr.execute_batched(dY, dx, B)
            pc.execute_batched(dz, dY, B)
            torch.cuda.synchronize()
            res.append((from_dev(dY)[..., : n // 2 + 1], from_dev(dz)))
            pr.close()
            pc.close()
    finally:
        api.set_column_tile(n, 0)
    for a, b in zip(res[0], res[1]):
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))


def test_autotune_2d_picks_a_compiled_width(gpu):
    from mufft_amd import api
    try:
        tile, us = api.autotune_2d(2048, 2048, "r2c", 2)
        assert tile in (4, 8) and us > 0
        assert api.get_column_tile(2048) == tile
        tile, us = api.autotune_2d(1024, 1024, "c2c_forward", 4)
        assert tile in (8, 16) and api.get_column_tile(1024) == tile
    finally:
        api.set_column_tile(2048, 0)
        api.set_column_tile(1024, 0)
