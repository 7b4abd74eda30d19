"""The uneven convolution batch (one persistent launch: per-XCD work queues,
each job group's inverse tasks waiting for its forward tasks; hipGraph
replay) against FP64 convolution, direct O(N^2) convolution at small N, and
the per-plan drop-in path on the same jobs.  Mono N = 65536 and stereo
N = 32768 jobs run their 32768-point transforms as two-workgroup split
tasks.  Replays must give identical bits: the queue and readiness counters
are cleared inside the kernel by the last workgroup out."""
import numpy as np
import pytest

from oracle import fp64, inputs
from tests.helpers import from_dev, rel_rms, to_dev

pytestmark = pytest.mark.gpu

MONO = 0 | 4 | 8
STEREO = 1 | 4 | 8


def make_jobs(gpu, sizes_mono, sizes_stereo, seed=0x4001):
    import torch
    jobs, host = [], []
    for i, N in enumerate(sizes_mono):
        a = inputs.uniform(seed + 2 * i, N // 2)
        # decaying Gaussian taps, decay length N/8
        h = (inputs.gaussian(seed + 2 * i + 1, N // 2) * np.exp(-np.arange(N // 2) / (N / 8))).astype(np.float32)
        out = torch.zeros(N, dtype=torch.float32, device=gpu)
        jobs.append((N, MONO, to_dev(a, gpu), to_dev(h, gpu), out))
        host.append((N, 0, a, h))
    for i, N in enumerate(sizes_stereo):
        a = inputs.complex_uniform(seed + 1000 + i, N // 2)
        h = (inputs.gaussian(seed + 2000 + i, N // 2) * np.exp(-np.arange(N // 2) / (N / 8))).astype(np.float32)
        out = torch.zeros(N, dtype=torch.complex64, device=gpu)
        jobs.append((N, STEREO, to_dev(a, gpu), to_dev(h, gpu), out))
        host.append((N, 1, a, h))
    return jobs, host


def check(jobs, host):
    for (N, method, _, _, out), (_, stereo, a, h) in zip(jobs, host):
        y = from_dev(out)
        ref = fp64.conv_fft(a.astype(np.complex128 if stereo else np.float64), h.astype(np.float64), N)
        assert rel_rms(y, ref) < 2e-6, (N, stereo)
        if N <= 1024:
            direct = fp64.conv_direct(a.astype(np.complex128 if stereo else np.float64), h.astype(np.float64))[:N]
            assert np.max(np.abs(y - direct)) < 2e-6 * np.sqrt(N) * max(1.0, float(np.abs(h).max())), (N, stereo)


def test_uneven_batch_graph_and_direct(gpu):
    import mufft_amd as M
    import torch
    rng = np.random.default_rng(1)
    mono = [int(2 ** k) for k in rng.integers(3, 17, size=60)] + [8, 65536, 256]
    stereo = [8, 16, 1024, 16384, 32768, 16384]
    jobs, host = make_jobs(gpu, mono, stereo)
    b = M.ConvBatch(jobs)
    ntasks, grid = b.info()
    assert ntasks >= 3 and grid >= 1
    b.execute()                     # graph replay, waits
    assert b.status() == 0
    check(jobs, host)
    first = [from_dev(j[4]).copy() for j in jobs]
    for j in jobs:
        j[4].zero_()
    torch.cuda.synchronize()
    b.execute()                     # second replay: the ticket queue needs no reset
    assert b.status() == 0
    for j, f in zip(jobs, first):
        np.testing.assert_array_equal(from_dev(j[4]).view(np.uint32), f.view(np.uint32))
    for j in jobs:
        j[4].zero_()
    s = torch.cuda.Stream(gpu)
    b.execute(stream=s, graph=False)  # same kernels without the graph
    s.synchronize()
    assert b.status() == 0
    for j, f in zip(jobs, first):
        np.testing.assert_array_equal(from_dev(j[4]).view(np.uint32), f.view(np.uint32))


def test_batch_matches_per_plan_dropin(gpu):
    import mufft_amd as M
    import torch
    mono = [64, 4096, 65536]
    stereo = [512, 16384]
    jobs, host = make_jobs(gpu, mono, stereo, seed=77)
    b = M.ConvBatch(jobs)
    b.execute()
    for (N, method, a, h, out), (_, stereo_, _, _) in zip(jobs, host):
        p = M.PlanConv(N, 0, method)
        bd = p.block_size // 8
        b0 = torch.zeros(bd, dtype=torch.complex64, device=gpu)
        b1 = torch.zeros(bd, dtype=torch.complex64, device=gpu)
        ref = torch.zeros_like(out)
        p.execute_input(0, b0, a)
        p.execute_input(1, b1, h)
        p.execute_output(ref, b0, b1)
        assert rel_rms(from_dev(out), from_dev(ref)) < 2e-6, (N, method)


def test_batch_rejects_bad_jobs(gpu):
    import mufft_amd as M
    import torch
    x = torch.zeros(64, device=gpu)
    for bad in ((6, MONO), (4, MONO), (131072, MONO), (65536, STEREO)):
        with pytest.raises(M.MufftError):
            M.ConvBatch([(bad[0], bad[1], x, x, x)])
