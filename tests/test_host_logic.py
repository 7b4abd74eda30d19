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
    # 16 points per thread for the real transforms at 1024 complex points, 32 for c2c;
    # c2r 8 points below 1024; c2c 16384 with registers for two workgroups per CU;
    # 32768 complex points in one workgroup (r2c 64 points per thread); 65536 split
    # over two workgroups.
    assert lib.mufft_hip_default_variant(1024, 0) & 0xFF == 32
    assert lib.mufft_hip_default_variant(8192, 0) & 0xFF == 32
    assert lib.mufft_hip_default_variant(2048, 2) & 0xFF == 16
    assert lib.mufft_hip_default_variant(1024, 3) & 0xFF == 8
    assert (lib.mufft_hip_default_variant(16384, 0) >> 20) & 0xF == 2
    assert (lib.mufft_hip_default_variant(32768, 0) >> 16) & 0xF == 0
    assert lib.mufft_hip_default_variant(65536, 2) & 0xFF == 64
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
