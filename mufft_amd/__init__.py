"""muFFT for AMD MI355X (gfx950): the muFFT public API on hand-written HIP kernels.

The compute path is libmufft_hip.so (C ABI, include/mufft/*.h).  Importing
this package loads it and raises ImportError if it is missing; there is no
CPU fallback.
"""
from . import _lib

_lib.load()

from .api import (  # noqa: E402
    CONV_BLOCK_FIRST,
    ConvBatch,
    CONV_BLOCK_SECOND,
    CONV_METHOD_FLAG_MONO_MONO,
    CONV_METHOD_FLAG_STEREO_MONO,
    CONV_METHOD_FLAG_ZERO_PAD_UPPER_HALF_FIRST,
    CONV_METHOD_FLAG_ZERO_PAD_UPPER_HALF_SECOND,
    FLAG_CPU_ANY,
    FLAG_CPU_NO_SIMD,
    FLAG_FULL_R2C,
    FLAG_ZERO_PAD_UPPER_HALF,
    FORWARD,
    INVERSE,
    MufftError,
    OverlapAdd,
    Plan1D,
    Plan2D,
    PlanConv,
    device_count,
    get_convolve_func,
    twiddle_table,
)

__all__ = [
    "FORWARD", "INVERSE", "FLAG_CPU_ANY", "FLAG_CPU_NO_SIMD", "FLAG_FULL_R2C",
    "FLAG_ZERO_PAD_UPPER_HALF", "CONV_METHOD_FLAG_MONO_MONO", "CONV_METHOD_FLAG_STEREO_MONO",
    "CONV_METHOD_FLAG_ZERO_PAD_UPPER_HALF_FIRST", "CONV_METHOD_FLAG_ZERO_PAD_UPPER_HALF_SECOND",
    "CONV_BLOCK_FIRST", "CONV_BLOCK_SECOND", "MufftError", "Plan1D", "Plan2D", "PlanConv",
    "device_count", "get_convolve_func", "twiddle_table", "ConvBatch", "OverlapAdd",
]
