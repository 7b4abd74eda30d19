"""The drop-in at the C level: a C99 program (tests/c/dropin_check.c) that
includes "mufft/fft.h" and links libmufft_hip.so with the build line
INTEGRATION.md gives, the way the reference's callers include fft.h and link
its library (CMakeLists.txt:23-25).

Without a GPU: it compiles warning-free as strict C99, links, and reports the
C layout of mufft_hip_conv_job and the macro values, which must match the
Python binding.  With a GPU: it runs all 19 reference entry points on the
golden vectors the compiled reference wrote (tests/golden), and its outputs
must match them like the Python path's (relative RMS <= 5e-7*log2 N).
"""
import ctypes
import math
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
SRC = ROOT / "tests" / "c" / "dropin_check.c"
LIBDIR = ROOT / "mufft_amd"
GOLD = ROOT / "tests" / "golden" / "mufft_ref_vectors.npz"


def build_dropin(dest: Path) -> Path:
    """cc app.c -I$MUFFT_AMD/include -L$MUFFT_AMD/mufft_amd -lmufft_hip
    -Wl,-rpath,$MUFFT_AMD/mufft_amd -L/opt/rocm/lib -lamdhip64."""
    cc = shutil.which("cc") or shutil.which("gcc")
    if not cc:
        pytest.skip("no C compiler")
    exe = dest / "dropin_check"
    cmd = [cc, "-std=c99", "-pedantic", "-Wall", "-Wextra", "-Werror", "-O2", str(SRC),
           f"-I{ROOT / 'include'}", f"-L{LIBDIR}", "-lmufft_hip", f"-Wl,-rpath,{LIBDIR}",
           "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib", "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_c_caller_compiles_and_sees_the_same_layout(tmp_path):
    from mufft_amd import _lib
    from mufft_amd import api as A
    exe = build_dropin(tmp_path)
    r = subprocess.run([str(exe), "layout"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = dict(l.split(" ", 1) for l in r.stdout.strip().splitlines())
    size, *offs = map(int, lines["conv_job"].split())
    J = _lib.ConvJob
    assert size == ctypes.sizeof(J)
    assert offs == [getattr(J, f).offset for f in ("N", "method", "first", "second", "output")]
    assert list(map(int, lines["macros"].split())) == [
        A.FORWARD, A.INVERSE, A.FLAG_CPU_NO_SIMD, A.FLAG_FULL_R2C, A.FLAG_ZERO_PAD_UPPER_HALF,
        A.CONV_METHOD_FLAG_STEREO_MONO, A.CONV_METHOD_FLAG_ZERO_PAD_UPPER_HALF_FIRST,
        A.CONV_METHOD_FLAG_ZERO_PAD_UPPER_HALF_SECOND, A.CONV_BLOCK_FIRST, A.CONV_BLOCK_SECOND]


def _bound(n):
    return 5e-7 * max(1.0, math.log2(n))


def _rel_rms(y, ref):
    y = np.asarray(y, np.complex128)
    ref = np.asarray(ref, np.complex128)
    return float(np.sqrt(np.mean(np.abs(y - ref) ** 2) / np.mean(np.abs(ref) ** 2)))


@pytest.mark.gpu
def test_c_caller_runs_every_entry_point_on_the_golden_vectors(gpu, tmp_path):
    exe = build_dropin(tmp_path)
    g = dict(np.load(GOLD))
    lines, checks = [], []

    def put(name, arr):
        p = tmp_path / name
        np.ascontiguousarray(arr).tofile(p)
        return p

    def case(key, line_fmt, ref, dtype, n_cmp, bound):
        out = tmp_path / (key + ".out")
        lines.append(line_fmt.format(out=out))
        checks.append((key, out, ref, dtype, n_cmp, bound))

    for key in sorted(k[:-4] for k in g if k.endswith("_out") and "_avx_" in k):
        op = key.split("_")[0]
        if op == "c2c":
            _, _, N, d = key.split("_")
            fin = put(key + ".in", g[key + "_in"])
            case(key, f"c2c {N} {d} {fin} {{out}}", g[key + "_out"], np.complex64, int(N), _bound(int(N)))
        elif op == "r2c":
            _, _, N, mode = key.split("_")
            fl = {"plain": 0, "full": 1 << 16, "zp": 1 << 17}[mode]
            fin = put(key + ".in", g[key + "_in"])
            ref = g[key + "_out"]
            case(key, f"r2c {N} {fl} {fin} {{out}}", ref, np.complex64, len(ref), _bound(int(N)))
        elif op == "c2r":
            _, _, N = key.split("_")
            fin = put(key + ".in", g[key + "_in"])
            case(key, f"c2r {N} {fin} {{out}}", g[key + "_out"], np.float32, int(N), _bound(int(N)))
        elif op in ("c2c2d", "r2c2d", "c2r2d"):
            parts = key.split("_")
            Nx, Ny = map(int, parts[2].split("x"))
            fin = put(key + ".in", g[key + "_in"])
            ref = g[key + "_out"]
            if op == "c2c2d":
                case(key, f"c2c2d {Nx} {Ny} {parts[3]} {fin} {{out}}", ref, np.complex64, Nx * Ny,
                     _bound(Nx * Ny))
            elif op == "r2c2d":
                # compared on the Nx/2+1 meaningful columns of each row
                case(key, f"r2c2d {Nx} {Ny} {fin} {{out}}", ref, ("rows", Nx), Nx * Ny, _bound(Nx * Ny))
            else:
                case(key, f"c2r2d {Nx} {Ny} {fin} {{out}}", ref, np.float32, Nx * Ny, _bound(Nx * Ny))
    for key in sorted(k[:-4] for k in g if k.startswith("conv_avx_") and k.endswith("_out")):
        _, _, N, stereo = key.split("_")
        fa, fb = put(key + ".a", g[key + "_a"]), put(key + ".b", g[key + "_b"])
        method = int(stereo) | 4 | 8
        case(key, f"conv {N} {method} {fa} {fb} {{out}}", g[key + "_out"],
             np.complex64 if int(stereo) else np.float32, int(N), 2 * _bound(int(N)))
    # mufft_get_convolve_func: out = 0.25 * a * b (kernel.c:22-34)
    a = (np.arange(64) * 0.01 - 0.3).astype(np.float32).view(np.complex64)
    b = (np.cos(np.arange(64)) * 0.5).astype(np.float32).view(np.complex64)
    fa, fb = put("mul.a", a), put("mul.b", b)
    case("convolve", f"convolve {a.size} {fa} {fb} {{out}}", 0.25 * a.astype(np.complex128) * b,
         np.complex64, a.size, 1e-6)

    manifest = tmp_path / "manifest.txt"
    manifest.write_text("\n".join(lines) + "\n")
    print(f"C drop-in: {len(lines)} cases", flush=True)
    r = subprocess.run([str(exe), "run", str(manifest)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"ran {len(lines)} cases" in r.stdout
    for key, out, ref, dtype, n_cmp, bound in checks:
        if isinstance(dtype, tuple):   # padded r2c 2D rows
            Nx = dtype[1]
            y = np.fromfile(out, np.complex64).reshape(-1, Nx)[:, : Nx // 2 + 1]
        else:
            y = np.fromfile(out, dtype)[:n_cmp]
        assert _rel_rms(y.reshape(np.shape(ref)), ref) <= bound, key
