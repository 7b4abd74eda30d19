"""GPU parity of the FFT convolution (test.c:405-488 sweep: N = 4..16384, mono
and stereo, both inputs zero padded), through the drop-in API on host buffers
and the batched device API, against direct O(N^2) convolution in FP64
(tolerance 2e-6*sqrt(N), test.c:431, 474) and the muFFT restatement."""
import numpy as np
import pytest

from oracle import fp64
from oracle import restate as R
from tests.helpers import cplx, from_dev, real, rel_rms, rms_tol_vs_mufft, to_dev

pytestmark = pytest.mark.gpu

ZP_BOTH = 4 | 8


@pytest.mark.parametrize("N", [4 * 2 ** k for k in range(0, 13)])
@pytest.mark.parametrize("stereo", [0, 1])
def test_conv_dropin(gpu, N, stereo):
    import mufft_amd as M
    a = cplx(0x4004 + N, N // 2, 0.5) if stereo else real(0x4002 + N, N // 2, 0.5)
    b = real(0x4003 + N, N // 2, 0.5)
    plan = M.PlanConv(N, 0, stereo | ZP_BOTH)
    bs = plan.block_size
    assert bs == (N * 8 if stereo else (N // 2 + 4) * 8)  # fft.c:586, 594
    b0 = np.zeros(bs // 8, np.complex64)
    b1 = np.zeros(bs // 8, np.complex64)
    plan.execute_input(0, b0, a)
    plan.execute_input(1, b1, b)
    out = np.zeros(N, np.complex64 if stereo else np.float32)
    plan.execute_output(out, b0, b1)
    direct = fp64.conv_direct(a.astype(np.complex128 if stereo else np.float64), b.astype(np.float64))[:N]
    assert np.max(np.abs(out - direct)) < 2e-6 * np.sqrt(N), (N, stereo)
    pc = R.create_plan_conv(N, 0, stereo | ZP_BOTH)
    orc = R.execute_conv_output(pc, R.execute_conv_input(pc, 0, a), R.execute_conv_input(pc, 1, b))
    assert rel_rms(out, orc) <= rms_tol_vs_mufft(N) * 2


@pytest.mark.parametrize("N", [256, 16384, 65536, 262144])
@pytest.mark.parametrize("stereo", [0, 1])
def test_conv_batched_device(gpu, N, stereo):
    """Batched convolution halves on device memory (fused multiply in the inverse);
    262144 takes the generic multi-launch path."""
    import mufft_amd as M
    import torch
    B = 3
    a = (cplx(7 + N, B * N // 2, 0.5) if stereo else real(7 + N, B * N // 2, 0.5)).reshape(B, N // 2)
    h = real(9 + N, N // 2, 0.5)
    plan = M.PlanConv(N, 0, stereo | ZP_BOTH)
    bd = plan.block_size // 8
    blocks = torch.zeros((B, bd), dtype=torch.complex64, device=gpu)
    hblock = torch.zeros((1, bd), dtype=torch.complex64, device=gpu)
    plan.execute_input_batched(0, blocks, to_dev(a, gpu), B)
    plan.execute_input_batched(1, hblock, to_dev(h, gpu), 1)
    out = torch.zeros((B, N), dtype=torch.complex64 if stereo else torch.float32, device=gpu)
    hrep = hblock.repeat(B, 1).contiguous()  # one filter spectrum per transform
    plan.execute_output_batched(out, blocks, hrep, B)
    torch.cuda.synchronize()
    y = from_dev(out)
    ref = fp64.conv_fft(a.astype(np.complex128 if stereo else np.float64), h.astype(np.float64), N)
    assert rel_rms(y, ref) < 2e-6, (N, stereo)


def test_convolve_func_pointer(gpu):
    """mufft_get_convolve_func returns a working GPU complex multiply (fft.c:549-563)."""
    import mufft_amd as M
    f = M.get_convolve_func(0)
    n = 1000
    a = cplx(1, n)
    b = cplx(2, n)
    out = np.zeros(n, np.complex64)
    f(out, a, b, 0.25, n)
    ref = 0.25 * (a.astype(np.complex128) * b)
    np.testing.assert_allclose(out, ref, rtol=1e-6, atol=1e-6)


def test_conv_invalid(gpu):
    import mufft_amd as M
    for bad in (0, 1, 2, 3, 12):
        with pytest.raises(M.MufftError):
            M.PlanConv(bad, 0, 0)
