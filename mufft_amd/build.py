"""Builds libmufft_hip.so (HIP kernels for gfx950 + host planner + C ABI) in-tree."""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
BUILD = PKG / "_build"
LIB = PKG / "libmufft_hip.so"

# Kernel units first (the row-kernel instantiations take the longest), then
# the host units (planner, executor and C ABI).
SOURCES = [
    "rows_c2c_fwd_mid.hip", "rows_c2c_inv_mid.hip", "rows_c2r_mid.hip", "rows_r2c_mid.hip",
    "rows_c2r_large.hip", "rows_r2c_large.hip", "rows_c2c_fwd_large.hip", "rows_c2c_inv_large.hip",
    "rows_c2c_fwd.hip", "rows_c2c_inv.hip", "rows_r2c.hip", "rows_c2r.hip", "conv_batch.hip",
    "cols.hip", "misc.hip",
    "host_common.hip", "exec.hip", "plan_1d.hip", "plan_2d.hip", "plan_conv.hip", "runtime.hip",
    "autotune.hip", "conv_batch_host.hip", "ola.hip", "noalloc.hip",
]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed-FP32 instructions are disabled for all device code.
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I", str(CSRC), "-I", str(INCLUDE),
]


def _compile(src: str) -> Path:
    obj = BUILD / (Path(src).stem + ".o")
    cmd = [HIPCC, *FLAGS, "-c", str(CSRC / src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {src}")
    return obj


def build(jobs: int | None = None) -> Path:
    BUILD.mkdir(exist_ok=True)
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 2)))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    return LIB


# Host-only units (planner, executor, C ABI): rebuilt with AddressSanitizer
# and UndefinedBehaviorSanitizer on the host side only (GPU code is compiled
# unchanged; GPU sanitizers are not used) and linked with the kernel units of
# the regular build into a separate library that the CPU tests load.
HOST_UNITS = [s for s in SOURCES if not s.startswith("rows_") and s not in ("conv_batch.hip", "cols.hip", "misc.hip")]
SAN_LIB = BUILD / "libmufft_hip_san.so"
SAN_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
             "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer",
             "-Xarch_host", "-g"]


def asan_runtime() -> Path:
    """The shared AddressSanitizer runtime the sanitized library needs preloaded."""
    r = subprocess.run([HIPCC, "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True, text=True)
    return Path(r.stdout.strip())


def build_sanitized(jobs: int | None = None) -> Path:
    BUILD.mkdir(exist_ok=True)
    kernel_objs = [BUILD / (Path(s).stem + ".o") for s in SOURCES if s not in HOST_UNITS]
    if not all(o.exists() for o in kernel_objs):
        build(jobs)

    def compile_san(src: str) -> Path:
        obj = BUILD / (Path(src).stem + ".san.o")
        cmd = [HIPCC, *FLAGS, *SAN_FLAGS, "-c", str(CSRC / src), "-o", str(obj)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"hipcc (sanitized host) failed on {src}")
        return obj

    jobs = jobs or min(len(HOST_UNITS), max(1, (os.cpu_count() or 2)))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(compile_san, HOST_UNITS))
    tmp = SAN_LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-shared-libasan",
           "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-o", str(tmp), *map(str, objs), *map(str, kernel_objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("sanitized link failed")
    os.replace(tmp, SAN_LIB)
    return SAN_LIB


if __name__ == "__main__":
    print(build())
    if "--sanitized" in sys.argv:
        print(build_sanitized())
