"""Host-side logic of the library that runs without a GPU: the registry of
fused-kernel variants and its static defaults, the tune cache, argument
validation, and plan creation failing cleanly where no device is present.
Nothing here launches a kernel."""
import ctypes

import pytest

from mufft_amd import _lib


def test_loading_the_library_first_leaves_one_hip_runtime():
    # The package binds to torch's bundled HIP runtime; a second copy from
    # /opt/rocm in the same process would leave torch without a device.
    import subprocess
    import sys
    code = ("import mufft_amd, torch\n"
            "libs = {l.split()[-1] for l in open('/proc/self/maps') if 'libamdhip64' in l}\n"
            "print(len(libs))\n")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.strip().splitlines()[-1] == "1"


@pytest.fixture(scope="module")
def lib():
    return _lib.load()


def _variants(L, n, kind):
    buf = (ctypes.c_int * 64)()
    cnt = L.mufft_hip_list_variants(n, kind, buf, 64)
    return list(buf[:max(cnt, 0)]), cnt


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_every_size_has_variants_and_a_registered_default(lib, kind):
    real = kind >= 2
    for logn in range(1, 17):
        n = 1 << (logn + 1 if real else logn)
        vs, cnt = _variants(lib, n, kind)
        assert cnt > 0 and len(vs) == cnt, (n, kind)
        assert len(set(vs)) == len(vs), "a variant is listed twice"
        d = lib.mufft_hip_default_variant(n, kind)
        assert d in vs, (n, kind, hex(d))
        assert lib.mufft_hip_get_variant(n, kind) == d   # untuned: the static default


def test_defaults_follow_the_measured_policy(lib):
    # 16 points per thread up to 8192 complex points (c2r: 8 below 1024, mirrored
    # lanes from 1024); 16384 with registers for two workgroups per CU (c2r with
    # chunked loads); 32768 complex points in one workgroup of 32 points per
    # thread (c2r with mirrored lanes), r2c split over two workgroups budgeted
    # to share a CU; 65536 split over two workgroups.
    M = 1 << 29
    assert lib.mufft_hip_default_variant(1024, 0) & 0xFF == 16
    assert lib.mufft_hip_default_variant(8192, 0) & 0xFF == 16
    assert lib.mufft_hip_default_variant(2048, 2) & 0xFF == 16
    assert lib.mufft_hip_default_variant(1024, 3) & 0xFF == 8
    assert lib.mufft_hip_default_variant(2048, 3) & M
    assert lib.mufft_hip_default_variant(16384, 3) & M
    assert not lib.mufft_hip_default_variant(16384, 2) & M
    assert (lib.mufft_hip_default_variant(16384, 0) >> 20) & 0xF == 2
    assert (lib.mufft_hip_default_variant(32768, 3) >> 20) & 0xF == 2
    assert (lib.mufft_hip_default_variant(32768, 0) >> 16) & 0xF == 0
    v = lib.mufft_hip_default_variant(65536, 2)
    assert v & 0xFF == 32 and (v >> 16) & 0xF == 1 and (v >> 20) & 0xF == 2
    v = lib.mufft_hip_default_variant(65536, 3)
    assert (v >> 16) & 0xF == 0 and v & M
    assert (lib.mufft_hip_default_variant(65536, 0) >> 16) & 0xF == 1


def test_invalid_sizes_and_variants_are_refused(lib):
    for n, kind in ((3, 0), (0, 0), (1 << 17, 0), (2, 2), (1 << 18, 3), (4096, 7)):
        assert lib.mufft_hip_list_variants(n, kind, None, 0) == -1
        assert _lib.last_error()
    assert lib.mufft_hip_set_variant(4096, 0, 0x7F7F) == -1
    assert lib.mufft_hip_default_variant(4096, 9) == -1


def test_tune_cache_roundtrip(lib, tmp_path):
    picks = {(4096, 0): None, (1024, 1): None, (65536, 2): None}
    try:
        for (n, kind) in picks:
            vs, _ = _variants(lib, n, kind)
            picks[(n, kind)] = vs[-1]
            assert lib.mufft_hip_set_variant(n, kind, vs[-1]) == 0
            assert lib.mufft_hip_get_variant(n, kind) == vs[-1]
        f = tmp_path / "tune.txt"
        assert lib.mufft_hip_tune_save(str(f).encode()) == 0
        text = f.read_text().splitlines()
        assert text[0].startswith("#") and len(text) == 1 + len(picks)
        for (n, kind) in picks:
            assert lib.mufft_hip_set_variant(n, kind, 0) == 0
        # lines that are comments, malformed, out of range or unregistered are skipped
        with open(f, "a") as fh:
            fh.write("garbage line\n12 0 999999 1.0\n40 0 16 1.0\n")
        assert lib.mufft_hip_tune_load(str(f).encode()) == len(picks)
        for (n, kind), v in picks.items():
            assert lib.mufft_hip_get_variant(n, kind) == v
        assert lib.mufft_hip_tune_load(str(tmp_path / "missing.txt").encode()) == -1
        assert "missing.txt" in _lib.last_error()
    finally:
        for (n, kind) in picks:
            lib.mufft_hip_set_variant(n, kind, 0)


def test_plan_creation_without_a_device_fails_cleanly(lib):
    import mufft_amd as M
    if M.device_count() > 0:
        pytest.skip("a GPU is present; the failure path needs a host without one")
    assert not lib.mufft_create_plan_1d_c2c(1024, -1, 0)
    assert "device" in _lib.last_error()
    assert not lib.mufft_create_plan_2d_c2c(64, 64, -1, 0)
    assert not lib.mufft_create_plan_conv(256, 0, 0)
    # invalid sizes are refused before any device call
    assert not lib.mufft_create_plan_1d_c2c(1000, -1, 0)
    assert "size" in _lib.last_error()
    assert not lib.mufft_create_plan_1d_r2c(2, 0)


# ---------------------------------------------------------------- argument checks
def test_buffer_layouts_of_the_reference_plans():
    from mufft_amd import api as A
    # c2c N, zero-padded input N/2 (fft.h:63-65); r2c N reals -> N/2+1 bins (N with
    # FULL_R2C); c2r N/2+1 bins -> N reals; the adjoint swaps the two sides.
    assert A.layout_1d("c2c", 64, 0) == (64, 8, 64, 8)
    assert A.layout_1d("c2c", 64, A.FLAG_ZERO_PAD_UPPER_HALF) == (32, 8, 64, 8)
    assert A.layout_1d("r2c", 64, 0) == (64, 4, 33, 8)
    assert A.layout_1d("r2c", 64, A.FLAG_FULL_R2C | A.FLAG_ZERO_PAD_UPPER_HALF) == (32, 4, 64, 8)
    assert A.layout_1d("c2r", 64, 0) == (33, 8, 64, 4)
    assert A.layout_1d("r2c", 64, 0, adjoint=True) == (33, 8, 64, 4)
    assert A.layout_1d("c2r", 64, 0, adjoint=True) == (64, 4, 33, 8)
    assert A.span_bytes(3, 0, 33, 8) == 3 * 33 * 8
    assert A.span_bytes(3, 40, 33, 8) == (2 * 40 + 33) * 8
    assert A.span_bytes(0, 40, 33, 8) == 0


def _plan_stub(cls, **attrs):
    # A plan object without a device handle: only its Python-side checks run.
    p = cls.__new__(cls)
    p.__dict__.update(attrs)
    p._h = None
    return p


def test_wrappers_refuse_bad_buffers_before_the_c_call():
    import numpy as np
    import torch

    from mufft_amd import api as A
    p = _plan_stub(A.Plan1D, kind="c2c", N=64, flags=0, direction=-1)
    good = np.zeros((4, 64), np.complex64)
    p._validate(good, good, 4, 0, 0)                       # fits
    with pytest.raises(ValueError, match="bytes"):
        p._validate(np.zeros((3, 64), np.complex64), good, 4, 0, 0)   # output one transform short
    with pytest.raises(ValueError, match="bytes"):
        p._validate(good, good, 4, 0, 65)                  # distance makes the output span longer
    with pytest.raises(ValueError, match="contiguous"):
        p._validate(np.zeros((4, 128), np.complex64)[:, ::2], good, 4, 0, 0)
    with pytest.raises(ValueError, match="dtype"):
        p._validate(np.zeros((4, 64), np.complex128), good, 4, 0, 0)
    t = torch.zeros((4, 128), dtype=torch.complex64)
    with pytest.raises(ValueError, match="contiguous"):
        p._validate(t[:, ::2], good, 4, 0, 0)
    with pytest.raises(ValueError, match="dtype"):
        p._validate(torch.zeros((4, 64), dtype=torch.float64), good, 4, 0, 0)
    p._validate(torch.zeros((4, 64), dtype=torch.complex64), good, 4, 0, 0)
    # a float32 buffer of the same byte size is accepted for complex data
    p._validate(torch.zeros((4, 128), dtype=torch.float32), good, 4, 0, 0)
    # the fused multiply's second operand: one block (idist2 = 0) or one per transform
    p._validate(good, good, 4, 0, 0, inp2=np.zeros(64, np.complex64), idist2=0)
    with pytest.raises(ValueError, match="second input"):
        p._validate(good, good, 4, 0, 0, inp2=np.zeros(64, np.complex64), idist2=64)
    r = _plan_stub(A.Plan1D, kind="r2c", N=64, flags=0, direction=-1)
    r._validate(np.zeros((2, 33), np.complex64), np.zeros((2, 64), np.float32), 2, 0, 0)
    with pytest.raises(ValueError, match="output"):
        r._validate(np.zeros((2, 32), np.complex64), np.zeros((2, 64), np.float32), 2, 0, 0)
    q = _plan_stub(A.Plan2D, kind="c2r", Nx=16, Ny=8)
    q._validate(np.zeros(16 * 8, np.float32), np.zeros((8, 16), np.complex64), 1)
    with pytest.raises(ValueError, match="input"):
        q._validate(np.zeros(16 * 8, np.float32), np.zeros((8, 9), np.complex64), 1)
    # raw addresses are passed through unchecked
    p._validate(0x1000, 0x2000, 4, 0, 0)


def _code(n1, a=0, b=0):
    return (n1 << 8) | (a << 4) | b


def test_large_splits_static_choice_and_validation(lib):
    # The static splits of the four-step sizes (a sweep on MI355X): 2^23..2^26
    # in three launches, 2^17..2^22 in two.
    assert lib.mufft_hip_get_large_split(1 << 23, 0) == _code(9, 8, 6)
    assert lib.mufft_hip_get_large_split(1 << 24, 0) == _code(10, 8, 6)
    assert lib.mufft_hip_get_large_split(1 << 26, 1) == _code(10, 10, 6)
    assert lib.mufft_hip_get_large_split(1 << 22, 0) == _code(12)
    assert lib.mufft_hip_get_large_split(1 << 17, 0) == _code(9)
    assert lib.mufft_hip_get_large_split(1 << 18, 2) == _code(9)   # r2c of 2^18 reals: 2^17 complex
    # fused sizes and sizes beyond the four-step range have no split
    for n, kind in ((1 << 16, 0), (1 << 27, 0), (1 << 17, 2), (3 << 17, 0)):
        assert lib.mufft_hip_get_large_split(n, kind) == -1
    # unrunnable codes: N1 beyond the four-step column kernel, N2 above 8192,
    # A or B beyond 1024, factors that do not multiply to N
    for n, code in ((1 << 20, _code(14)), (1 << 22, _code(8)), (1 << 24, _code(6, 11, 7)),
                    (1 << 24, _code(10, 8, 7)), (1 << 18, _code(5))):
        assert lib.mufft_hip_set_large_split(n, 0, code) == -1, (n, hex(code))
    try:
        assert lib.mufft_hip_set_large_split(1 << 20, 0, _code(12)) == 0
        assert lib.mufft_hip_get_large_split(1 << 20, 1) == _code(12)   # keyed by size, both kinds
    finally:
        assert lib.mufft_hip_set_large_split(1 << 20, 0, 0) == 0
    assert lib.mufft_hip_get_large_split(1 << 20, 0) == _code(10)


def test_column_tiles_static_choice_and_validation(lib):
    assert lib.mufft_hip_get_column_tile(1024) == 16
    assert lib.mufft_hip_get_column_tile(2048) == 4
    assert lib.mufft_hip_get_column_tile(8192) == 4
    assert lib.mufft_hip_get_column_tile(16384) == -1
    assert lib.mufft_hip_set_column_tile(2048, 5) == -1
    assert lib.mufft_hip_set_column_tile(64, 128) == -1      # no alternative below 128 points
    try:
        assert lib.mufft_hip_set_column_tile(2048, 8) == 0
        assert lib.mufft_hip_get_column_tile(2048) == 8
    finally:
        assert lib.mufft_hip_set_column_tile(2048, 0) == 0
    assert lib.mufft_hip_get_column_tile(2048) == 4


def test_tune_cache_carries_splits_and_tiles(lib, tmp_path):
    f = tmp_path / "tune.txt"
    f.write_text("# hand-written cache\n"
                 f"22 4 {_code(12)} 100.0\n"      # 2^22 in two launches, N1 = 4096
                 "11 5 8 20.5\n"                   # 2048-point columns, 8-column tiles
                 f"22 4 {_code(14)} 1.0\n"         # not runnable: ignored
                 "11 5 7 1.0\n")                   # not a compiled width: ignored
    try:
        assert lib.mufft_hip_tune_load(str(f).encode()) == 2
        assert lib.mufft_hip_get_large_split(1 << 22, 0) == _code(12)
        assert lib.mufft_hip_get_column_tile(2048) == 8
        g = tmp_path / "again.txt"
        assert lib.mufft_hip_tune_save(str(g).encode()) == 0
        lines = [l.split() for l in g.read_text().splitlines() if not l.startswith("#")]
        assert ["22", "4", str(_code(12))] in [l[:3] for l in lines]
        assert ["11", "5", "8"] in [l[:3] for l in lines]
    finally:
        lib.mufft_hip_set_large_split(1 << 22, 0, 0)
        lib.mufft_hip_set_column_tile(2048, 0)
