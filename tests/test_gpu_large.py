"""Transforms beyond the fused sizes: four-step (two fused column launches,
three from 2^24 on) for 2^17..2^26 points, and generic passes beyond that.

Every split N = N1 * N2 used by the planner is checked against an FP64 FFT
(numpy's pocketfft in double, itself checked against the direct DFT in
tests/test_oracle.py), in both directions, with a batch of two at padded
distances and with the zero-padded flag, plus the real transforms that run on
top of it (resolve pass + four-step)."""
import numpy as np
import pytest

from oracle import fp64
from tests.helpers import abs_tol, cplx, empty_dev, from_dev, real, rel_rms, to_dev

pytestmark = pytest.mark.gpu

THREE_PASS_FROM = 23   # first size (log2) the planner splits into three launches


def c2c_launches(logn):
    return 3 if logn >= THREE_PASS_FROM else 2


@pytest.mark.parametrize("logn", list(range(17, 27)))
@pytest.mark.parametrize("direction", [-1, 1])
def test_four_step_c2c(gpu, logn, direction):
    import torch
    import mufft_amd as M
    N = 1 << logn
    batch, pad = 2, 24
    p = M.Plan1D("c2c", N, direction)
    assert p.launches == c2c_launches(logn)
    x = cplx(0x9000 + logn, batch * (N + pad), 0.5).reshape(batch, N + pad)
    dy = torch.full((batch, N + pad), 3.0 + 0j, dtype=torch.complex64, device=gpu)
    p.execute_batched(dy, to_dev(x, gpu), batch, N + pad, N + pad)
    torch.cuda.synchronize()
    y = from_dev(dy)
    assert np.all(y[:, N:] == 3.0)   # nothing written past each transform
    for b in range(batch):
        ref = np.fft.fft(x[b, :N].astype(np.complex128)) if direction < 0 else \
            np.fft.ifft(x[b, :N].astype(np.complex128)) * N
        assert rel_rms(y[b, :N], ref) < 2e-6, (logn, direction, b)
        if logn <= 20:
            assert np.max(np.abs(y[b, :N] - ref)) < abs_tol(N), (logn, direction, b)


@pytest.mark.parametrize("logn", [17, 20, 24])
def test_four_step_zero_pad(gpu, logn):
    import torch
    import mufft_amd as M
    N = 1 << logn
    p = M.Plan1D("c2c", N, -1, M.FLAG_ZERO_PAD_UPPER_HALF)
    x = cplx(0x9100 + logn, N // 2, 0.5)
    dy = empty_dev((N,), np.complex64, gpu)
    p.execute_batched(dy, to_dev(x, gpu), 1)
    torch.cuda.synchronize()
    xf = np.zeros(N, np.complex128)
    xf[: N // 2] = x
    assert rel_rms(from_dev(dy), np.fft.fft(xf)) < 2e-6


@pytest.mark.parametrize("logn", [18, 21, 24, 25])
def test_four_step_real_roundtrip(gpu, logn):
    import torch
    import mufft_amd as M
    N = 1 << logn
    x = real(0x9200 + logn, 2 * N, 0.5).reshape(2, N)
    p = M.Plan1D("r2c", N)
    q = M.Plan1D("c2r", N)
    # an N-point real transform is an N/2-point complex one plus the resolve pass
    assert p.launches == q.launches == c2c_launches(logn - 1) + 1
    dX = empty_dev((2, N // 2 + 1), np.complex64, gpu)
    p.execute_batched(dX, to_dev(x, gpu), 2)
    dy = empty_dev((2, N), np.float32, gpu)
    q.execute_batched(dy, dX, 2)
    torch.cuda.synchronize()
    X = from_dev(dX)
    for b in range(2):
        assert rel_rms(X[b], np.fft.rfft(x[b].astype(np.float64))) < 2e-6
    assert rel_rms(from_dev(dy) / N, x) < 2e-6


def test_generic_beyond_four_step(gpu):
    """2^27 points: past the four-step range, generic passes."""
    import torch
    import mufft_amd as M
    N = 1 << 27
    p = M.Plan1D("c2c", N, -1)
    assert p.launches > 2
    x = torch.zeros(N, dtype=torch.complex64, device=gpu)
    x[5] = 1.0
    y = torch.empty_like(x)
    p.execute_batched(y, x, 1)
    torch.cuda.synchronize()
    k = torch.arange(0, N, 1 << 14, device=gpu, dtype=torch.float64)
    ref = torch.exp(-2j * np.pi * 5 * k / N)
    got = y[:: 1 << 14].to(torch.complex128)
    assert float((got - ref).abs().max()) < 2e-6
