"""Streaming overlap-add filtering against FP64 convolution, and
streaming in several calls equals one call (the tail carries over)."""
import numpy as np
import pytest
from scipy.signal import fftconvolve

from oracle import inputs
from tests.helpers import from_dev, to_dev

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,ntaps", [(4, 2), (256, 100), (4096, 2048), (65536, 1000)])
def test_ola_matches_direct_convolution(gpu, N, ntaps):
    import torch
    from mufft_amd.api import OverlapAdd
    hop = N // 2
    nblocks = max(3, (1 << 20) // hop) if N < 65536 else 8
    x = inputs.uniform(0x7001 + N, nblocks * hop)
    h = inputs.gaussian(0x7002 + N, ntaps) / np.float32(np.sqrt(ntaps))
    ola = OverlapAdd(N, torch.from_numpy(h).to(gpu))
    dx = to_dev(x, gpu)
    out = torch.zeros(nblocks * hop, dtype=torch.float32, device=gpu)
    ola(out, dx, nblocks)
    torch.cuda.synchronize()
    ref = fftconvolve(x.astype(np.float64), h.astype(np.float64))[: nblocks * hop]
    y = from_dev(out)
    rel = np.sqrt(np.mean((y - ref) ** 2) / np.mean(ref ** 2))
    assert rel < 2e-6, rel
    # streaming: the same signal in three calls (tail carried) gives the same samples
    ola.reset()
    parts = [1, nblocks - 2, 1]
    outs, off = [], 0
    for nb in parts:
        o = torch.zeros(nb * hop, dtype=torch.float32, device=gpu)
        ola(o, dx[off * hop:(off + nb) * hop], nb)
        outs.append(o)
        off += nb
    torch.cuda.synchronize()
    y2 = np.concatenate([from_dev(o) for o in outs])
    np.testing.assert_array_equal(y2.view(np.uint32), y.view(np.uint32))


def test_ola_rejects_bad_arguments(gpu):
    import mufft_amd as M
    from mufft_amd.api import OverlapAdd
    h = np.ones(10, np.float32)
    for N, taps in ((6, h), (8, h)):  # non power of two; more taps than N/2
        with pytest.raises(M.MufftError):
            OverlapAdd(N, taps)
