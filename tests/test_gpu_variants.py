"""Every registered fused-kernel variant is correct, and the autotuner works.

The autotuner may select any registered variant (points per thread, LDS
exchange mode, twiddle scheme), so each one is checked here against FP64 for
every transform kind, batch > 1 and a partial last workgroup.
"""
import numpy as np
import pytest

from oracle import fp64
from tests.helpers import abs_tol, cplx, empty_dev, from_dev, real, rel_rms, to_dev

pytestmark = pytest.mark.gpu

KINDS = {0: ("c2c", -1), 1: ("c2c", 1), 2: ("r2c", -1), 3: ("c2r", 1)}


def _variants(N, kind):
    import ctypes
    from mufft_amd import _lib
    L = _lib.load()
    buf = (ctypes.c_int * 64)()
    n = L.mufft_hip_list_variants(N, kind, buf, 64)
    assert n > 0
    return list(buf[:n])


def _run(gpu, kind, N, x, batch):
    import mufft_amd as M
    import torch
    name, d = KINDS[kind]
    if name == "c2c":
        p = M.Plan1D("c2c", N, d)
        dy = empty_dev((batch, N), np.complex64, gpu)
    elif name == "r2c":
        p = M.Plan1D("r2c", N)
        dy = empty_dev((batch, N // 2 + 1), np.complex64, gpu)
    else:
        p = M.Plan1D("c2r", N)
        dy = empty_dev((batch, N), np.float32, gpu)
    p.execute_batched(dy, to_dev(x, gpu), batch)
    torch.cuda.synchronize()
    return from_dev(dy)


@pytest.mark.parametrize("logn", range(1, 17))
@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_every_variant(gpu, logn, kind):
    from mufft_amd import _lib
    L = _lib.load()
    name, d = KINDS[kind]
    N = 1 << logn if name == "c2c" else 1 << (logn + 1)   # real sizes for r2c / c2r
    batch = 3
    if name == "c2c":
        x = cplx(logn * 11 + kind, batch * N, 0.5).reshape(batch, N)
        ref = fp64.c2c(x, d)
    elif name == "r2c":
        x = real(logn * 13 + kind, batch * N, 0.5).reshape(batch, N)
        ref = fp64.r2c(x)
    else:
        x = cplx(logn * 17 + kind, batch * (N // 2 + 1), 0.5).reshape(batch, N // 2 + 1)
        ref = fp64.c2r(x, N)
    try:
        for var in _variants(N, kind):
            assert L.mufft_hip_set_variant(N, kind, var) == 0
            assert L.mufft_hip_get_variant(N, kind) == var
            y = _run(gpu, kind, N, x, batch)
            assert rel_rms(y, ref) < 2e-6, (N, kind, hex(var))
            assert np.max(np.abs(y - ref)) < abs_tol(N, factor=2e-6), (N, kind, hex(var))
    finally:
        L.mufft_hip_set_variant(N, kind, 0)


@pytest.mark.parametrize("N,kind", [(4096, 0), (1024, 1), (65536, 2), (65536, 3)])
def test_autotune_selects_valid_variant(gpu, tmp_path, N, kind):
    import ctypes
    from mufft_amd import _lib
    L = _lib.load()
    us = ctypes.c_double(0)
    try:
        best = L.mufft_hip_autotune_1d(N, kind, 64, ctypes.byref(us))
        assert best > 0, _lib.last_error()
        assert best in _variants(N, kind)
        assert L.mufft_hip_get_variant(N, kind) == best and us.value > 0
        f = tmp_path / "tune.txt"
        assert L.mufft_hip_tune_save(str(f).encode()) == 0
        assert L.mufft_hip_set_variant(N, kind, 0) == 0
        assert L.mufft_hip_tune_load(str(f).encode()) >= 1
        assert L.mufft_hip_get_variant(N, kind) == best
        # a plan created now runs the tuned variant and stays correct
        name, d = KINDS[kind]
        batch = 2
        if name == "c2c":
            x = cplx(5, batch * N, 0.5).reshape(batch, N)
            ref = fp64.c2c(x, d)
        elif name == "r2c":
            x = real(5, batch * N, 0.5).reshape(batch, N)
            ref = fp64.r2c(x)
        else:
            x = cplx(5, batch * (N // 2 + 1), 0.5).reshape(batch, N // 2 + 1)
            ref = fp64.c2r(x, N)
        assert rel_rms(_run(gpu, kind, N, x, batch), ref) < 2e-6
    finally:
        L.mufft_hip_set_variant(N, kind, 0)


def _inner(a, b):
    a = np.asarray(a, np.complex128 if np.iscomplexobj(a) else np.float64)
    b = np.asarray(b, np.complex128 if np.iscomplexobj(b) else np.float64)
    return float(np.real(np.sum(a * np.conj(b))))


@pytest.mark.parametrize("logn", range(9, 16))
def test_c2r_variants_with_fused_prologue_flags(gpu, logn):
    """The c2r prologue of every variant under its runtime flags: the fused
    multiply with a conjugated second operand and a norm (the convolution's
    inverse), and the adjoint weighting of an r2c plan (plain and zero-padded),
    checked by FP64 and by the dot-product identity.  The variants read the
    bins in different orders and lane layouts (chunked, mirrored lanes), so
    each one is checked here with every flag."""
    import mufft_amd as M
    import torch
    from mufft_amd import _lib
    L = _lib.load()
    N = 1 << (logn + 1)
    nb = N // 2 + 1
    B = 3
    X = cplx(logn * 19 + 1, B * nb, 0.5).reshape(B, nb)
    Y = cplx(logn * 19 + 2, B * nb, 0.5).reshape(B, nb)
    ref_mul = fp64.c2r(0.25 * X.astype(np.complex128) * np.conj(Y), N)
    try:
        for var in _variants(N, 3):
            assert L.mufft_hip_set_variant(N, 3, var) == 0
            p = M.Plan1D("c2r", N)
            out = empty_dev((B, N), np.float32, gpu)
            p.execute_mul_batched(out, to_dev(X, gpu), to_dev(Y, gpu), B, idist2=nb, norm=0.25, conj2=True)
            torch.cuda.synchronize()
            y = from_dev(out)
            assert rel_rms(y, ref_mul) < 2e-6, (N, hex(var))
            p.close()
            for fl, nin in ((0, N), (M.FLAG_ZERO_PAD_UPPER_HALF, N // 2)):
                q = M.Plan1D("r2c", N, M.FORWARD, fl)
                x = real(logn * 19 + 3, B * nin).reshape(B, nin)
                g = cplx(logn * 19 + 4, B * nb).reshape(B, nb)
                Tx = empty_dev((B, nb), np.complex64, gpu)
                Tg = empty_dev((B, N), np.float32, gpu)
                q.execute_batched(Tx, to_dev(x, gpu), B)
                q.execute_adjoint_batched(Tg, to_dev(g, gpu), B)
                torch.cuda.synchronize()
                tx, tg = from_dev(Tx), from_dev(Tg)[:, :nin]
                lhs, rhs = _inner(tx, g), _inner(x, tg)
                assert abs(lhs - rhs) <= 2e-6 * np.sqrt(_inner(tx, tx) * _inner(g, g)), (N, hex(var), fl)
                q.close()
    finally:
        L.mufft_hip_set_variant(N, 3, 0)


@pytest.mark.parametrize("logn", range(9, 16))
def test_r2c_variants_with_epilogue_flags(gpu, logn):
    """The r2c epilogue of every variant under its runtime flags: the full
    spectrum (bins N..2N-1 too), zero-padded input, and the adjoint of a c2r
    plan (checked by the dot-product identity).  The variants pair bins k and
    N - k differently (through LDS, or in registers between mirrored lanes)."""
    import mufft_amd as M
    import torch
    from mufft_amd import _lib
    L = _lib.load()
    N = 1 << (logn + 1)
    nb = N // 2 + 1
    B = 3
    x = real(logn * 23 + 1, B * N, 0.5).reshape(B, N)
    full_ref = np.fft.fft(x.astype(np.float64), axis=-1)
    xz = x[:, : N // 2]
    zp_ref = fp64.r2c(np.concatenate([xz, np.zeros_like(xz)], axis=1))
    try:
        for var in _variants(N, 2):
            assert L.mufft_hip_set_variant(N, 2, var) == 0
            p = M.Plan1D("r2c", N, M.FORWARD, M.FLAG_FULL_R2C)
            out = empty_dev((B, N), np.complex64, gpu)
            p.execute_batched(out, to_dev(x, gpu), B)
            torch.cuda.synchronize()
            assert rel_rms(from_dev(out), full_ref) < 2e-6, (N, hex(var), "full")
            p.close()
            p = M.Plan1D("r2c", N, M.FORWARD, M.FLAG_ZERO_PAD_UPPER_HALF)
            out = empty_dev((B, nb), np.complex64, gpu)
            p.execute_batched(out, to_dev(np.ascontiguousarray(xz), gpu), B)
            torch.cuda.synchronize()
            assert rel_rms(from_dev(out), zp_ref) < 2e-6, (N, hex(var), "zero pad")
            p.close()
            q = M.Plan1D("c2r", N)
            X = cplx(logn * 23 + 2, B * nb).reshape(B, nb)
            g = real(logn * 23 + 3, B * N).reshape(B, N)
            Tx = empty_dev((B, N), np.float32, gpu)
            Tg = empty_dev((B, nb), np.complex64, gpu)
            q.execute_batched(Tx, to_dev(X, gpu), B)
            q.execute_adjoint_batched(Tg, to_dev(g, gpu), B)
            torch.cuda.synchronize()
            tx, tg = from_dev(Tx), from_dev(Tg)
            lhs, rhs = _inner(tx, g), _inner(X, tg)
            assert abs(lhs - rhs) <= 2e-6 * np.sqrt(_inner(tx, tx) * _inner(g, g)), (N, hex(var), "adjoint")
            q.close()
    finally:
        L.mufft_hip_set_variant(N, 2, 0)
