"""ctypes binding to the compiled reference library (test infrastructure only).

``oracle/Makefile`` compiles the reference's own C99 sources into
``oracle/_ref/libmufft_ref.so``.  This module calls its public API (fft.h) on
host numpy buffers; it is used by the tests as a checker and by the CPU
baseline leg of the benchmark.  It never touches the GPU library.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_ref" / "libmufft_ref.so"
REFERENCE_DIR = Path("/root/reference")

_lib = None


def build(quiet: bool = True) -> bool:
    """Compile oracle/_ref from the reference sources when they are present."""
    if not (REFERENCE_DIR / "fft.c").exists():
        return LIB_PATH.exists()
    cmd = ["make", "-s", "-C", str(HERE), f"REF={REFERENCE_DIR}"]
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL if quiet else None)
    return LIB_PATH.exists()


def available() -> bool:
    return LIB_PATH.exists()


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(f"{LIB_PATH} is not built (make -C oracle)")
        L = ctypes.CDLL(str(LIB_PATH))  # RTLD_LOCAL: no clash with the GPU library
        vp, u, i, f, sz = ctypes.c_void_p, ctypes.c_uint, ctypes.c_int, ctypes.c_float, ctypes.c_size_t
        for name, res, args in [
            ("mufft_create_plan_1d_c2c", vp, [u, i, u]),
            ("mufft_create_plan_1d_r2c", vp, [u, u]),
            ("mufft_create_plan_1d_c2r", vp, [u, u]),
            ("mufft_execute_plan_1d", None, [vp, vp, vp]),
            ("mufft_free_plan_1d", None, [vp]),
            ("mufft_create_plan_2d_c2c", vp, [u, u, i, u]),
            ("mufft_create_plan_2d_r2c", vp, [u, u, u]),
            ("mufft_create_plan_2d_c2r", vp, [u, u, u]),
            ("mufft_execute_plan_2d", None, [vp, vp, vp]),
            ("mufft_free_plan_2d", None, [vp]),
            ("mufft_create_plan_conv", vp, [u, u, u]),
            ("mufft_execute_conv_input", None, [vp, u, vp, vp]),
            ("mufft_execute_conv_output", None, [vp, vp, vp, vp]),
            ("mufft_conv_get_transformed_block_size", sz, [vp]),
            ("mufft_free_plan_conv", None, [vp]),
            ("mufft_get_convolve_func", vp, [u]),
            ("mufft_alloc", vp, [sz]),
            ("mufft_free", None, [vp]),
            # oracle/ref_bench.c: the CPU-baseline workload on native threads
            ("mufft_ref_bench_run", ctypes.c_double, [i, i, u, vp, i, u, vp, i]),
            ("mufft_ref_time_1d", ctypes.c_double, [i, u, i, vp, i]),
            ("mufft_ref_time_2d", ctypes.c_double, [i, u, u, i, vp, i, i, i]),
            ("mufft_ref_time_conv", ctypes.c_double, [i, i, i, vp, vp, vp, vp]),
        ]:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


def _aligned(nbytes: int, dtype) -> np.ndarray:
    """64-byte aligned host buffer (the SIMD kernels use aligned loads, fft.h:178)."""
    raw = np.zeros(nbytes + 64, np.uint8)
    off = (-raw.ctypes.data) % 64
    return raw[off:off + nbytes].view(dtype)


def _copy_in(x, dtype) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype)
    buf = _aligned(x.nbytes, dtype)
    buf[...] = x.reshape(-1)
    return buf


class Plan1D:
    """One reference 1D plan (c2c / r2c / c2r)."""

    def __init__(self, kind: str, N: int, direction: int = -1, flags: int = 0):
        L = lib()
        self.kind, self.N, self.flags = kind, N, flags
        if kind == "c2c":
            self.h = L.mufft_create_plan_1d_c2c(N, direction, flags)
        elif kind == "r2c":
            self.h = L.mufft_create_plan_1d_r2c(N, flags)
        elif kind == "c2r":
            self.h = L.mufft_create_plan_1d_c2r(N, flags)
        else:
            raise ValueError(kind)
        if not self.h:
            raise ValueError(f"reference refused plan {kind} N={N} flags={flags:#x}")

    def out_like(self):
        N = self.N
        if self.kind == "c2c":
            return _aligned(N * 8, np.complex64)
        if self.kind == "r2c":
            return _aligned(N * 8, np.complex64)  # N complex (room for FULL_R2C)
        return _aligned(N * 4, np.float32)

    def execute(self, x, out=None):
        dt = np.float32 if self.kind == "r2c" else np.complex64
        xin = x if (isinstance(x, np.ndarray) and x.dtype == dt and x.ctypes.data % 64 == 0
                    and x.flags.c_contiguous) else _copy_in(x, dt)
        if out is None:
            out = self.out_like()
        lib().mufft_execute_plan_1d(self.h, _ptr(out), _ptr(xin))
        return out

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.mufft_free_plan_1d(self.h)
            self.h = None


def fft_1d(kind, N, x, direction=-1, flags=0):
    """Run one reference 1D transform; returns the meaningful part of the output."""
    p = Plan1D(kind, N, direction, flags)
    out = p.execute(x)
    if kind == "r2c":
        return out[: (N if flags & (1 << 16) else N // 2 + 1)].copy()
    return out.copy()


def fft_2d(kind, Nx, Ny, x, direction=-1, flags=0):
    L = lib()
    if kind == "c2c":
        h = L.mufft_create_plan_2d_c2c(Nx, Ny, direction, flags)
        xin = _copy_in(x, np.complex64)
        out = _aligned(Nx * Ny * 8, np.complex64)
    elif kind == "r2c":
        h = L.mufft_create_plan_2d_r2c(Nx, Ny, flags)
        xin = _copy_in(x, np.float32)
        out = _aligned(Nx * Ny * 8, np.complex64)
    else:
        h = L.mufft_create_plan_2d_c2r(Nx, Ny, flags)
        xin = _copy_in(x, np.complex64)
        out = _aligned(2 * Nx * Ny * 4, np.float32)
    if not h:
        raise ValueError("reference refused 2D plan")
    try:
        L.mufft_execute_plan_2d(h, _ptr(out), _ptr(xin))
    finally:
        L.mufft_free_plan_2d(h)
    if kind == "c2r":
        return out[: Nx * Ny].reshape(Ny, Nx).copy()
    return out.reshape(Ny, Nx).copy()


def conv(N, a, b, method, flags=0):
    """Full reference convolution: input(first), input(second), output (fft.c:799-809)."""
    L = lib()
    h = L.mufft_create_plan_conv(N, flags, method)
    if not h:
        raise ValueError("reference refused conv plan")
    try:
        bs = L.mufft_conv_get_transformed_block_size(h)
        b0 = _aligned(bs, np.uint8)
        b1 = _aligned(bs, np.uint8)
        stereo = method & 1
        da = np.complex64 if stereo else np.float32
        ain = _copy_in(a, da)
        bin_ = _copy_in(b, np.float32)
        L.mufft_execute_conv_input(h, 0, _ptr(b0), _ptr(ain))
        L.mufft_execute_conv_input(h, 1, _ptr(b1), _ptr(bin_))
        out = _aligned(N * 8, np.complex64 if stereo else np.float32)
        L.mufft_execute_conv_output(h, _ptr(out), _ptr(b0), _ptr(b1))
    finally:
        L.mufft_free_plan_conv(h)
    return out[:N].copy()


def threads_env() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
