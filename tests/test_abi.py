"""The C-ABI boundary, checked without a GPU.

The shared library loads, exports every function the public headers declare
(unmangled C symbols), keeps the reference's macro values, and its creators
refuse invalid sizes with NULL (fft.c:478-481, 628-631, 666-669) before any
device work.
"""
import os
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADERS = [ROOT / "include" / "mufft" / "fft.h", ROOT / "include" / "mufft" / "mufft_hip.h"]
LIB = ROOT / "mufft_amd" / "libmufft_hip.so"


def declared_functions():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"\b(mufft_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    names.discard("mufft_convolve_func")
    return names


def exported_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB)], capture_output=True, text=True, check=True)
    return {line.split()[-1] for line in out.stdout.splitlines() if " T " in line}


def test_library_loads():
    import mufft_amd
    assert mufft_amd.device_count() >= 0


def test_every_declared_function_is_exported():
    decl = declared_functions()
    assert len(decl) >= 31
    missing = decl - exported_symbols()
    assert not missing, f"declared but not exported: {sorted(missing)}"


def test_python_binding_covers_every_export():
    from mufft_amd import _lib
    bound = {name for name, _, _ in _lib.SIGNATURES}
    assert declared_functions() <= bound


def _macros(path):
    vals = {}
    for m in re.finditer(r"#define\s+(MUFFT_[A-Z0-9_]+)\s+(.+)", path.read_text()):
        vals[m.group(1)] = m.group(2).split("//")[0].strip()
    return vals


def test_macro_values_match_reference_api():
    mine = _macros(HEADERS[0])
    expected = {
        "MUFFT_FORWARD": -1, "MUFFT_INVERSE": 1, "MUFFT_FLAG_CPU_ANY": 0,
        "MUFFT_FLAG_CPU_NO_SIMD": (1 << 16) - 1, "MUFFT_FLAG_CPU_NO_AVX": 1,
        "MUFFT_FLAG_CPU_NO_SSE3": 2, "MUFFT_FLAG_CPU_NO_SSE": 4,
        "MUFFT_FLAG_FULL_R2C": 1 << 16, "MUFFT_FLAG_ZERO_PAD_UPPER_HALF": 1 << 17,
        "MUFFT_CONV_METHOD_FLAG_MONO_MONO": 0, "MUFFT_CONV_METHOD_FLAG_STEREO_MONO": 1,
        "MUFFT_CONV_METHOD_FLAG_ZERO_PAD_UPPER_HALF_FIRST": 4,
        "MUFFT_CONV_METHOD_FLAG_ZERO_PAD_UPPER_HALF_SECOND": 8,
        "MUFFT_CONV_BLOCK_FIRST": 0, "MUFFT_CONV_BLOCK_SECOND": 1,
    }
    for k, v in expected.items():
        assert eval(mine[k]) == v, k
    ref = Path("/root/reference/fft.h")
    if ref.exists():  # same values as the reference header where it is present
        theirs = _macros(ref)
        for k in expected:
            assert eval(theirs[k]) == eval(mine[k]), k


def test_invalid_sizes_return_null_without_gpu():
    from mufft_amd import _lib
    L = _lib.load()
    for n in (0, 1, 3, 12):
        assert not L.mufft_create_plan_1d_c2c(n, -1, 0)
    assert not L.mufft_create_plan_1d_c2c(16, 0, 0)  # invalid direction
    for n in (0, 2, 6):
        assert not L.mufft_create_plan_1d_r2c(n, 0)
        assert not L.mufft_create_plan_1d_c2r(n, 0)
        assert not L.mufft_create_plan_conv(n, 0, 0)
    assert not L.mufft_create_plan_2d_c2c(3, 4, -1, 0)
    assert not L.mufft_create_plan_2d_r2c(2, 4, 0)
    assert "invalid" in _lib.last_error()
    # free functions accept NULL (fft.c:993-996)
    L.mufft_free_plan_1d(None)
    L.mufft_free_plan_2d(None)
    L.mufft_free_plan_conv(None)
    L.mufft_free(None)


def test_package_refuses_to_import_without_library(tmp_path):
    """No CPU fallback: a copy of the package without the .so fails to import."""
    import shutil
    import sys
    pkg = tmp_path / "mufft_amd"
    shutil.copytree(ROOT / "mufft_amd", pkg, ignore=shutil.ignore_patterns("*.so", "_build", "csrc"))
    env = {k: v for k, v in os.environ.items() if k != "MUFFT_AMD_LIBRARY"}
    r = subprocess.run([sys.executable, "-c", "import mufft_amd"], cwd=tmp_path, capture_output=True, text=True,
                       env=env)
    assert r.returncode != 0 and "libmufft_hip.so" in r.stderr
