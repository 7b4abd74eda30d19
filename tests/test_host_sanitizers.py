"""The host side of the library under AddressSanitizer and
UndefinedBehaviorSanitizer (CPU only).

The reference relies on -Wall -Wextra and its internal asserts
(fft_internal.h:194-200).  Here the host units (planner, executor, tuner,
tune cache, C ABI) are rebuilt with the host-side sanitizers and linked with
the regular kernel units (mufft_amd/build.py: build_sanitized); a child
process with the ASan runtime preloaded drives every entry point that runs
without a GPU: the variant, split and tile registries, the tune cache
(including malformed and overlong lines), argument validation, NULL-safe
frees, and the creators' failure paths where no device is present; a
second child runs the CPU tests of the ABI and the host logic against the
sanitized build (MUFFT_AMD_LIBRARY).  Any memory error or undefined
behaviour aborts the child.
"""
import os
import subprocess
import sys
import textwrap
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent

CHILD = textwrap.dedent(r'''
    import ctypes, os, sys, tempfile
    L = ctypes.CDLL(sys.argv[1])
    vp, u, i, sz = ctypes.c_void_p, ctypes.c_uint, ctypes.c_int, ctypes.c_size_t
    def sig(name, res, args):
        f = getattr(L, name); f.restype = res; f.argtypes = args; return f
    list_variants = sig("mufft_hip_list_variants", i, [u, i, vp, i])
    default_variant = sig("mufft_hip_default_variant", i, [u, i])
    get_variant = sig("mufft_hip_get_variant", i, [u, i])
    set_variant = sig("mufft_hip_set_variant", i, [u, i, i])
    get_split = sig("mufft_hip_get_large_split", i, [u, i])
    set_split = sig("mufft_hip_set_large_split", i, [u, i, i])
    get_tile = sig("mufft_hip_get_column_tile", i, [u])
    set_tile = sig("mufft_hip_set_column_tile", i, [u, i])
    save = sig("mufft_hip_tune_save", i, [ctypes.c_char_p])
    load = sig("mufft_hip_tune_load", i, [ctypes.c_char_p])
    err = sig("mufft_hip_get_last_error", ctypes.c_char_p, [])
    buf = (ctypes.c_int * 64)()
    for kind in range(-1, 5):
        for logn in range(0, 30):
            for n in (1 << logn, (1 << logn) + 1, 3 << logn):
                cnt = list_variants(n, kind, buf, 64)
                list_variants(n, kind, None, 0)
                default_variant(n, kind)
                get_variant(n, kind)
                get_split(n, kind)
                if cnt > 0:
                    for v in list(buf[:min(cnt, 64)]):
                        assert set_variant(n, kind, v) == 0
                    set_variant(n, kind, 0)
                set_variant(n, kind, 0x7fffffff)
                for code in (0, 0x900, 0xa86, 0xfff, -1, 0x7fffffff):
                    set_split(n, kind, code)
                set_split(n, kind, 0)
    for L_ in [0, 1, 2, 3] + [1 << k for k in range(0, 20)]:
        get_tile(L_)
        for t in (-1, 0, 2, 4, 8, 16, 64, 1 << 30):
            set_tile(L_, t)
        set_tile(L_, 0)
    err()
    d = tempfile.mkdtemp()
    p = os.path.join(d, "cache.txt").encode()
    with open(p, "w") as f:
        f.write("# comment\n12 0 4112 1.0\n" + "9" * 5000 + "\n\n-5 2 3 x\n22 4 3072 5.0\n11 5 8 2.0\n"
                "99999999999999999999 4 1 1\n16 3 -7 1.0\n")
    assert load(p) >= 3, err()
    assert save(p) == 0
    assert load(p) >= 3
    assert load(os.path.join(d, "missing").encode()) == -1
    assert save(b"/nonexistent-dir/x") == -1
    for n in range(1, 30):
        set_variant(1 << n, 0, 0)
    set_split(1 << 22, 0, 0)
    set_tile(2048, 0)
    # creators and frees without a device: NULL, no crash
    c2c = sig("mufft_create_plan_1d_c2c", vp, [u, i, u])
    r2c = sig("mufft_create_plan_1d_r2c", vp, [u, u])
    c2r = sig("mufft_create_plan_1d_c2r", vp, [u, u])
    p2 = sig("mufft_create_plan_2d_c2c", vp, [u, u, i, u])
    conv = sig("mufft_create_plan_conv", vp, [u, u, u])
    for n in (0, 1, 2, 3, 4, 1024, 1 << 27, 1 << 28, 0xffffffff):
        for dr in (-1, 1, 0, 7):
            for pl in (c2c(n, dr, 0), p2(n, n, dr, 0)):
                assert pl is None or pl
        r2c(n, 0); c2r(n, 0); conv(n, 0, 0); conv(n, 0, 13)
    for name in ("mufft_free_plan_1d", "mufft_free_plan_2d", "mufft_free_plan_conv", "mufft_free",
                 "mufft_hip_free_conv_batch", "mufft_hip_free_ola", "mufft_hip_free"):
        sig(name, None, [vp])(None)
    sig("mufft_get_convolve_func", vp, [u])(0)
    sig("mufft_execute_plan_1d", None, [vp, vp, vp])(None, None, None)
    sig("mufft_execute_plan_2d", None, [vp, vp, vp])(None, None, None)
    assert sig("mufft_hip_execute_plan_1d_batched", i, [vp, vp, vp, u, sz, sz, vp])(None, None, None, 1, 0, 0, None) == -1
    assert sig("mufft_hip_execute_plan_1d_batched_noalloc", i, [vp, vp, vp, u, sz, sz, vp])(None, None, None, 1, 0, 0, None) == -1
    assert sig("mufft_hip_create_conv_batch", vp, [vp, u, u])(None, 0, 0) is None
    class Job(ctypes.Structure):
        _fields_ = [("N", u), ("method", u), ("first", vp), ("second", vp), ("output", vp)]
    jobs = (Job * 3)(Job(8, 0, 1, 1, 1), Job(3, 0, 1, 1, 1), Job(1 << 20, 1, 1, 1, 1))
    assert sig("mufft_hip_create_conv_batch", vp, [vp, u, u])(jobs, 3, 0) is None
    tm = sig("mufft_hip_twiddle_table_mode", i, [i, u, u, i, vp, u])
    for mode in (-1, 0, 1, 2, 3, 4):
        for n in (0, 1, 63, 64, 1 << 30):
            for prm in (0, 1, 64, n):
                tm(mode, n, prm, -1, None, 0)
    sig("mufft_hip_create_ola", vp, [u, vp, u])(0, None, 0)
    sig("mufft_hip_device_count", i, [])()
    print("SANITIZED_OK")
''')


@pytest.fixture(scope="module")
def sanitized_lib():
    sys.path.insert(0, str(ROOT / "mufft_amd"))
    try:
        import build
    finally:
        sys.path.pop(0)
    lib = build.SAN_LIB
    srcs = [build.CSRC / s for s in build.HOST_UNITS]
    newest = max(p.stat().st_mtime for p in srcs + list(build.CSRC.glob("*.h")))
    if not lib.exists() or lib.stat().st_mtime < newest:
        try:
            build.build_sanitized()
        except Exception as e:  # pragma: no cover - toolchain problems
            pytest.skip(f"sanitized build failed: {e}")
    rt = build.asan_runtime()
    if not rt.exists():
        pytest.skip("no AddressSanitizer runtime in this toolchain")
    return lib, rt


def _env(lib, rt):
    env = dict(os.environ)
    env["LD_PRELOAD"] = str(rt)
    env["MUFFT_AMD_LIBRARY"] = str(lib)
    # Leak checking would report the HIP runtime's and the interpreter's own
    # process-lifetime allocations; memory errors and UB abort regardless.
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    return env


def test_host_entry_points_under_asan_and_ubsan(sanitized_lib, tmp_path):
    lib, rt = sanitized_lib
    script = tmp_path / "child.py"
    script.write_text(CHILD)
    r = subprocess.run([sys.executable, str(script), str(lib)], capture_output=True, text=True,
                       env=_env(lib, rt), timeout=600)
    assert r.returncode == 0 and "SANITIZED_OK" in r.stdout, (r.stdout[-3000:] + r.stderr[-6000:])


def test_cpu_suite_of_the_host_logic_under_sanitizers(sanitized_lib):
    lib, rt = sanitized_lib
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", "-m", "not gpu",
                        "tests/test_abi.py", "tests/test_host_logic.py"],
                       cwd=ROOT, capture_output=True, text=True, env=_env(lib, rt), timeout=900)
    assert r.returncode == 0, (r.stdout[-4000:] + r.stderr[-4000:])
