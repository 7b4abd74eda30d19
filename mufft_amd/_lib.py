"""Loads libmufft_hip.so and declares the C ABI (include/mufft/fft.h, mufft_hip.h).

There is no fallback: if the shared library is missing or cannot be loaded,
importing the package raises, so nothing can silently run on the CPU.
"""
from __future__ import annotations

import ctypes
import os
import sys
from pathlib import Path

# MUFFT_AMD_LIBRARY selects another build of the same library (the
# host-sanitized one of the CPU tests); the default is the in-tree build.
LIB_PATH = Path(os.environ.get("MUFFT_AMD_LIBRARY") or Path(__file__).resolve().parent / "libmufft_hip.so")

_vp, _u, _i, _f, _sz = ctypes.c_void_p, ctypes.c_uint, ctypes.c_int, ctypes.c_float, ctypes.c_size_t

CONVOLVE_FUNC = ctypes.CFUNCTYPE(None, _vp, _vp, _vp, _f, _u)

# (name, restype, argtypes): every symbol the headers declare.
SIGNATURES = [
    # fft.h — reference API
    ("mufft_create_plan_1d_c2c", _vp, [_u, _i, _u]),
    ("mufft_create_plan_1d_r2c", _vp, [_u, _u]),
    ("mufft_create_plan_1d_c2r", _vp, [_u, _u]),
    ("mufft_execute_plan_1d", None, [_vp, _vp, _vp]),
    ("mufft_free_plan_1d", None, [_vp]),
    ("mufft_create_plan_conv", _vp, [_u, _u, _u]),
    ("mufft_execute_conv_input", None, [_vp, _u, _vp, _vp]),
    ("mufft_conv_get_transformed_block_size", _sz, [_vp]),
    ("mufft_execute_conv_output", None, [_vp, _vp, _vp, _vp]),
    ("mufft_free_plan_conv", None, [_vp]),
    ("mufft_get_convolve_func", _vp, [_u]),
    ("mufft_create_plan_2d_c2c", _vp, [_u, _u, _i, _u]),
    ("mufft_create_plan_2d_r2c", _vp, [_u, _u, _u]),
    ("mufft_create_plan_2d_c2r", _vp, [_u, _u, _u]),
    ("mufft_execute_plan_2d", None, [_vp, _vp, _vp]),
    ("mufft_free_plan_2d", None, [_vp]),
    ("mufft_alloc", _vp, [_sz]),
    ("mufft_calloc", _vp, [_sz]),
    ("mufft_free", None, [_vp]),
    # mufft_hip.h — extensions
    ("mufft_hip_get_last_error", ctypes.c_char_p, []),
    ("mufft_hip_device_count", _i, []),
    ("mufft_hip_malloc", _vp, [_sz]),
    ("mufft_hip_free", None, [_vp]),
    ("mufft_hip_memcpy", _i, [_vp, _vp, _sz]),
    ("mufft_hip_synchronize", _i, [_vp]),
    ("mufft_hip_execute_plan_1d_batched", _i, [_vp, _vp, _vp, _u, _sz, _sz, _vp]),
    ("mufft_hip_execute_plan_1d_batched_noalloc", _i, [_vp, _vp, _vp, _u, _sz, _sz, _vp]),
    ("mufft_hip_execute_plan_2d_batched", _i, [_vp, _vp, _vp, _u, _vp]),
    ("mufft_hip_execute_conv_input_batched", _i, [_vp, _u, _vp, _vp, _u, _sz, _sz, _vp]),
    ("mufft_hip_execute_conv_output_batched", _i, [_vp, _vp, _vp, _vp, _u, _sz, _sz, _sz, _vp]),
    ("mufft_hip_create_plan_1d_variant", _vp, [_u, _i, _u, _i]),
    ("mufft_hip_plan_1d_variant", _i, [_vp]),
    ("mufft_hip_twiddle_table_mode", _i, [_i, _u, _u, _i, _vp, _u]),
    ("mufft_hip_plan_1d_launches", _i, [_vp]),
    ("mufft_hip_plan_2d_launches", _i, [_vp]),
    ("mufft_hip_twiddle_table", _i, [_u, _i, _vp, _u]),
    ("mufft_hip_list_variants", _i, [_u, _i, _vp, _i]),
    ("mufft_hip_default_variant", _i, [_u, _i]),
    ("mufft_hip_get_variant", _i, [_u, _i]),
    ("mufft_hip_set_variant", _i, [_u, _i, _i]),
    ("mufft_hip_autotune_1d", _i, [_u, _i, _u, ctypes.POINTER(ctypes.c_double)]),
    ("mufft_hip_autotune_large_1d", _i, [_u, _i, _u, ctypes.POINTER(ctypes.c_double)]),
    ("mufft_hip_get_large_split", _i, [_u, _i]),
    ("mufft_hip_set_large_split", _i, [_u, _i, _i]),
    ("mufft_hip_autotune_2d", _i, [_u, _u, _i, _u, ctypes.POINTER(ctypes.c_double)]),
    ("mufft_hip_get_column_tile", _i, [_u]),
    ("mufft_hip_set_column_tile", _i, [_u, _i]),
    ("mufft_hip_tune_save", _i, [ctypes.c_char_p]),
    ("mufft_hip_tune_load", _i, [ctypes.c_char_p]),
    ("mufft_hip_execute_adjoint_1d_batched", _i, [_vp, _vp, _vp, _u, _sz, _sz, _vp]),
    ("mufft_hip_execute_plan_1d_mul_batched", _i,
     [_vp, _vp, _vp, _vp, _u, _sz, _sz, _sz, _f, _i, _i, _vp]),
    ("mufft_hip_mul_reduce", _i, [_vp, _vp, _vp, _u, _u, _sz, _sz, _f, _i, _vp]),
    ("mufft_hip_stream_copy", _i, [_vp, _vp, ctypes.c_size_t, _vp]),
    ("mufft_hip_create_ola", _vp, [_u, _vp, _u]),
    ("mufft_hip_execute_ola", _i, [_vp, _vp, _vp, _u, _vp]),
    ("mufft_hip_reset_ola", _i, [_vp, _vp]),
    ("mufft_hip_free_ola", None, [_vp]),
    ("mufft_hip_create_conv_batch", _vp, [_vp, _u, _u]),
    ("mufft_hip_execute_conv_batch", _i, [_vp, _vp]),
    ("mufft_hip_execute_conv_batch_direct", _i, [_vp, _vp]),
    ("mufft_hip_conv_batch_status", _i, [_vp]),
    ("mufft_hip_conv_batch_info", _i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    ("mufft_hip_free_conv_batch", None, [_vp]),
]


class ConvJob(ctypes.Structure):
    """mufft_hip_conv_job (include/mufft/mufft_hip.h)."""
    _fields_ = [("N", _u), ("method", _u), ("first", _vp), ("second", _vp), ("output", _vp)]

_lib = None


def _share_torch_runtime() -> None:
    """PyTorch's ROCm wheels carry their own HIP and HSA runtimes.  Loaded
    first, this library would bring /opt/rocm's copies into the process as
    well, and whichever pair initialises second finds no device (torch then
    reports no GPU).  Importing torch first, when it is installed, makes the
    library bind to torch's runtime, so both share one."""
    if "torch" in sys.modules:
        return
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
                " or `python mufft_amd/build.py`")
        _share_torch_runtime()
        L = ctypes.CDLL(str(LIB_PATH))
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    msg = load().mufft_hip_get_last_error()
    return msg.decode() if msg else ""
