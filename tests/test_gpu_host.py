"""Host-memory buffers: the drop-in reads and writes pinned buffers in place,
batched calls stream host batches through device staging in overlapped
chunks, and the stream-ordered extensions refuse pageable memory instead of
handing it to a kernel.  Every result must equal the device-memory run of the
same plan bit for bit."""
import numpy as np
import pytest

from tests.helpers import cplx, empty_dev, from_dev, real, to_dev

pytestmark = pytest.mark.gpu


def _pinned(shape, dtype):
    import torch
    t = {np.complex64: torch.complex64, np.float32: torch.float32}[dtype]
    return torch.empty(shape, dtype=t, pin_memory=True)


@pytest.mark.parametrize("kind,N,batch", [("c2c", 4096, 2500), ("r2c", 65536, 300),
                                          ("c2r", 65536, 300), ("c2c", 1 << 17, 70)])
@pytest.mark.parametrize("mem", ["pageable", "pinned", "in_host", "out_host"])
def test_batched_host_equals_device(gpu, kind, N, batch, mem):
    import torch
    import mufft_amd as M
    if kind == "c2c":
        p = M.Plan1D("c2c", N, -1)
        x = cplx(0x8101, batch * N, 0.5).reshape(batch, N)
        oshape, odt = (batch, N), np.complex64
    elif kind == "r2c":
        p = M.Plan1D("r2c", N)
        x = real(0x8102, batch * N, 0.5).reshape(batch, N)
        oshape, odt = (batch, N // 2 + 1), np.complex64
    else:
        p = M.Plan1D("c2r", N)
        x = cplx(0x8103, batch * (N // 2 + 1), 0.5).reshape(batch, N // 2 + 1)
        oshape, odt = (batch, N), np.float32
    dy = empty_dev(oshape, odt, gpu)
    p.execute_batched(dy, to_dev(x, gpu), batch)
    torch.cuda.synchronize()
    ref = from_dev(dy)

    if mem == "pageable":
        xin, yout = np.ascontiguousarray(x), np.zeros(oshape, odt)
    elif mem == "pinned":
        xin = _pinned(x.shape, x.dtype.type)
        xin.copy_(torch.from_numpy(x))
        yout = _pinned(oshape, odt)
    elif mem == "in_host":
        xin, yout = np.ascontiguousarray(x), empty_dev(oshape, odt, gpu)
    else:
        xin, yout = to_dev(x, gpu), np.zeros(oshape, odt)
    p.execute_batched(yout, xin, batch)
    torch.cuda.synchronize()
    got = yout.cpu().numpy() if hasattr(yout, "cpu") else yout
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("kind", ["c2c", "r2c", "c2r"])
def test_2d_batched_host_equals_device(gpu, kind):
    import torch
    import mufft_amd as M
    nx, ny, b = 512, 256, 40
    if kind == "c2c":
        p = M.Plan2D("c2c", nx, ny, -1)
        x = cplx(0x8201, b * ny * nx, 0.5).reshape(b, ny, nx)
        oshape, odt = (b, ny, nx), np.complex64
    elif kind == "r2c":
        p = M.Plan2D("r2c", nx, ny, -1)
        x = real(0x8202, b * ny * nx, 0.5).reshape(b, ny, nx)
        oshape, odt = (b, ny, nx), np.complex64
    else:
        p = M.Plan2D("c2r", nx, ny, 1)
        x = cplx(0x8203, b * ny * nx, 0.5).reshape(b, ny, nx)
        oshape, odt = (b, ny, nx), np.float32
    # r2c leaves the padding columns of each row untouched: start both from the same bytes.
    init = np.full(oshape, 7.0, odt)
    dy = to_dev(init, gpu)
    p.execute_batched(dy, to_dev(x, gpu), b)
    torch.cuda.synchronize()
    ref = from_dev(dy)
    yout = init.copy()
    p.execute_batched(yout, np.ascontiguousarray(x), b)
    np.testing.assert_array_equal(yout.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("kind,N", [("c2c", 1024), ("r2c", 4096), ("c2r", 65536), ("c2c", 1 << 17)])
def test_dropin_pinned_equals_device(gpu, kind, N):
    """Pinned buffers run in place (one launch, no staging); pageable ones are staged."""
    import torch
    import mufft_amd as M
    if kind == "c2c":
        p = M.Plan1D("c2c", N, 1)
        x = cplx(0x8301, N, 0.5)
        oshape, odt = (N,), np.complex64
    elif kind == "r2c":
        p = M.Plan1D("r2c", N)
        x = real(0x8302, N, 0.5)
        oshape, odt = (N // 2 + 1,), np.complex64
    else:
        p = M.Plan1D("c2r", N)
        x = cplx(0x8303, N // 2 + 1, 0.5)
        oshape, odt = (N,), np.float32
    dy = empty_dev(oshape, odt, gpu)
    p.execute(dy, to_dev(x, gpu))
    torch.cuda.synchronize()
    ref = from_dev(dy)
    xin = _pinned(x.shape, x.dtype.type)
    xin.copy_(torch.from_numpy(x))
    yp = _pinned(oshape, odt)
    for _ in range(2):
        p.execute(yp, xin)
        np.testing.assert_array_equal(yp.numpy().view(np.uint32), ref.view(np.uint32))
    ypg = np.zeros(oshape, odt)
    p.execute(ypg, np.ascontiguousarray(x))
    np.testing.assert_array_equal(ypg.view(np.uint32), ref.view(np.uint32))


def test_stream_ordered_extensions_refuse_pageable(gpu):
    import mufft_amd as M
    N = 256
    p = M.Plan1D("c2c", N, -1)
    x = cplx(0x8401, N, 0.5)
    y = np.zeros(N, np.complex64)
    with pytest.raises(M.MufftError):
        p.execute_adjoint_batched(y, x, 1)
    dx = to_dev(x, gpu)
    with pytest.raises(M.MufftError):
        p.execute_mul_batched(y, dx, dx, 1)


def test_stream_copy(gpu):
    """The streaming copy used as the bandwidth reference copies exactly and
    refuses sizes that are not a multiple of 8 bytes."""
    import torch
    from mufft_amd import _lib
    L = _lib.load()
    for n in (1, 4095, 4096, 1 << 20):
        x = torch.randn(n, dtype=torch.complex64, device=gpu)
        y = torch.zeros_like(x)
        assert L.mufft_hip_stream_copy(y.data_ptr(), x.data_ptr(), n * 8, None) == 0
        torch.cuda.synchronize()
        assert torch.equal(torch.view_as_real(x), torch.view_as_real(y))
    assert L.mufft_hip_stream_copy(y.data_ptr(), x.data_ptr(), 12, None) != 0
