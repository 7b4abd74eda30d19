"""Python mirror of the muFFT plan interface over the C ABI of libmufft_hip.so.

Names, arguments and error behaviour follow the reference API (fft.h):
creators raise ``MufftError`` where the reference returns NULL (invalid size,
allocation failure), ``execute`` is synchronous, and the transforms are
unnormalised.  Buffers can be numpy arrays (host memory, staged by the
library), torch CUDA tensors (device memory, used in place) or raw integer
addresses.  ``execute_batched`` runs many transforms of one plan in a single
stream-ordered call on device memory (library extension).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

FORWARD = -1
INVERSE = 1
FLAG_CPU_ANY = 0
FLAG_CPU_NO_SIMD = (1 << 16) - 1
FLAG_FULL_R2C = 1 << 16
FLAG_ZERO_PAD_UPPER_HALF = 1 << 17
CONV_METHOD_FLAG_MONO_MONO = 0
CONV_METHOD_FLAG_STEREO_MONO = 1
CONV_METHOD_FLAG_ZERO_PAD_UPPER_HALF_FIRST = 1 << 2
CONV_METHOD_FLAG_ZERO_PAD_UPPER_HALF_SECOND = 1 << 3
CONV_BLOCK_FIRST = 0
CONV_BLOCK_SECOND = 1


class MufftError(RuntimeError):
    pass


def ptr(x) -> int | None:
    """Address of a buffer: torch tensor, numpy array or int."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        return int(x.data_ptr())
    if isinstance(x, np.ndarray):
        if not x.flags.c_contiguous:
            raise ValueError("buffer must be C-contiguous")
        return int(x.ctypes.data)
    raise TypeError(f"unsupported buffer type {type(x)!r}")


# Element types the kernels read and write: complex64 for complex data,
# float32 for real data (an operand may also be given in the other of the two,
# its size then counts in bytes).
_REAL, _CPLX = 4, 8


def check_buffer(x, need_bytes: int, what: str) -> None:
    """Checks one operand before its address goes to the C ABI: contiguous,
    float32 / complex64, on the GPU or in host memory, and at least
    ``need_bytes`` long.  Raw integer addresses are taken as they are."""
    if x is None or isinstance(x, int):
        return
    if hasattr(x, "data_ptr"):  # torch tensor
        import torch
        if x.dtype not in (torch.float32, torch.complex64):
            raise ValueError(f"{what}: dtype {x.dtype} (float32 or complex64 expected)")
        if not x.is_contiguous():
            raise ValueError(f"{what}: tensor must be contiguous")
        if x.device.type not in ("cuda", "cpu"):
            raise ValueError(f"{what}: tensor on unsupported device {x.device}")
        have = x.numel() * x.element_size()
    elif isinstance(x, np.ndarray):
        if x.dtype not in (np.float32, np.complex64):
            raise ValueError(f"{what}: dtype {x.dtype} (float32 or complex64 expected)")
        if not x.flags.c_contiguous:
            raise ValueError(f"{what}: array must be C-contiguous")
        have = x.nbytes
    else:
        raise TypeError(f"{what}: unsupported buffer type {type(x)!r}")
    if have < need_bytes:
        raise ValueError(f"{what}: {have} bytes, the call needs {need_bytes}")


def span_bytes(howmany: int, dist: int, packed: int, elem: int) -> int:
    """Bytes spanned by howmany blocks of `packed` elements at distance `dist`
    (0 = packed) in elements of `elem` bytes."""
    if howmany <= 0:
        return 0
    d = dist if dist else packed
    return ((howmany - 1) * d + packed) * elem


def layout_1d(kind: str, N: int, flags: int, adjoint: bool = False):
    """(input elements, input element bytes, output elements, output element
    bytes) of one transform of a 1D plan, or of its adjoint."""
    zp = bool(flags & FLAG_ZERO_PAD_UPPER_HALF)
    full = bool(flags & FLAG_FULL_R2C)
    if kind == "c2c":
        fwd = ((N // 2 if zp else N), _CPLX, N, _CPLX)
        return (N, _CPLX, N, _CPLX) if adjoint else fwd
    if kind == "r2c":
        spec = N if full else N // 2 + 1
        return (spec, _CPLX, N, _REAL) if adjoint else ((N // 2 if zp else N), _REAL, spec, _CPLX)
    if kind == "c2r":
        return (N, _REAL, N // 2 + 1, _CPLX) if adjoint else (N // 2 + 1, _CPLX, N, _REAL)
    raise ValueError(kind)


def stream_handle(s) -> int | None:
    if s is None:
        return None
    if isinstance(s, int):
        return s
    h = getattr(s, "cuda_stream", None)
    if h is not None:
        return int(h)
    raise TypeError(f"unsupported stream {type(s)!r}")


def _check(rc: int, what: str):
    if rc != 0:
        raise MufftError(f"{what}: {_lib.last_error()}")


class Plan1D:
    """mufft_plan_1d (fft.h:78-121).  kind: 'c2c', 'r2c' or 'c2r'."""

    def __init__(self, kind: str, N: int, direction: int = FORWARD, flags: int = 0,
                 variant: int = 0):
        """``variant`` fixes the fused-kernel variant of this plan (library
        extension; 0 = the process-wide choice)."""
        L = _lib.load()
        self.kind, self.N, self.flags = kind, N, flags
        self.direction = direction if kind == "c2c" else (FORWARD if kind == "r2c" else INVERSE)
        if variant:
            code = {"c2c": 0 if direction == FORWARD else 1, "r2c": 2, "c2r": 3}.get(kind)
            if code is None:
                raise ValueError(kind)
            h = L.mufft_hip_create_plan_1d_variant(N, code, flags, variant)
        elif kind == "c2c":
            h = L.mufft_create_plan_1d_c2c(N, direction, flags)
        elif kind == "r2c":
            h = L.mufft_create_plan_1d_r2c(N, flags)
        elif kind == "c2r":
            h = L.mufft_create_plan_1d_c2r(N, flags)
        else:
            raise ValueError(kind)
        if not h:
            raise MufftError(f"mufft_create_plan_1d_{kind}({N}) failed: {_lib.last_error()}")
        self._h = h

    @property
    def handle(self):
        return self._h

    @property
    def launches(self) -> int:
        return _lib.load().mufft_hip_plan_1d_launches(self._h)

    @property
    def variant(self) -> int:
        """The fused-kernel variant this plan runs (0: no fused row kernel)."""
        return int(_lib.load().mufft_hip_plan_1d_variant(self._h))

    def _validate(self, out, inp, howmany, idist, odist, adjoint=False, inp2=None, idist2=0):
        ni, ei, no, eo = layout_1d(self.kind, self.N, self.flags, adjoint)
        if inp2 is not None and self.kind == "c2c":
            ni = self.N  # the fused multiply reads every input point, zero padding or not
        check_buffer(inp, span_bytes(howmany, idist, ni, ei), "input")
        check_buffer(out, span_bytes(howmany, odist, no, eo), "output")
        if inp2 is not None:
            check_buffer(inp2, span_bytes(howmany if idist2 else 1, idist2, ni, ei), "second input")

    def execute(self, out, inp):
        self._validate(out, inp, 1, 0, 0)
        _lib.load().mufft_execute_plan_1d(self._h, ptr(out), ptr(inp))

    def execute_batched(self, out, inp, howmany: int, idist: int = 0, odist: int = 0, stream=None):
        self._validate(out, inp, howmany, idist, odist)
        rc = _lib.load().mufft_hip_execute_plan_1d_batched(
            self._h, ptr(out), ptr(inp), howmany, idist, odist, stream_handle(stream))
        _check(rc, "mufft_hip_execute_plan_1d_batched")

    def execute_adjoint_batched(self, out, inp, howmany: int, idist: int = 0, odist: int = 0,
                                stream=None):
        """Adjoint (conjugate transpose) of the plan's transform; see mufft_hip.h."""
        self._validate(out, inp, howmany, idist, odist, adjoint=True)
        rc = _lib.load().mufft_hip_execute_adjoint_1d_batched(
            self._h, ptr(out), ptr(inp), howmany, idist, odist, stream_handle(stream))
        _check(rc, "mufft_hip_execute_adjoint_1d_batched")

    def execute_mul_batched(self, out, inp, inp2, howmany: int, idist: int = 0, idist2: int = 0,
                            odist: int = 0, norm: float = 1.0, conj2: bool = False,
                            adjoint: bool = False, stream=None):
        """out = T(norm * inp * op(inp2)) with the multiply fused into the transform;
        idist2 = 0 applies one inp2 block to every transform."""
        self._validate(out, inp, howmany, idist, odist, adjoint=adjoint, inp2=inp2, idist2=idist2)
        rc = _lib.load().mufft_hip_execute_plan_1d_mul_batched(
            self._h, ptr(out), ptr(inp), ptr(inp2), howmany, idist, idist2, odist,
            ctypes.c_float(norm), int(conj2), int(adjoint), stream_handle(stream))
        _check(rc, "mufft_hip_execute_plan_1d_mul_batched")

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().mufft_free_plan_1d(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Plan2D:
    """mufft_plan_2d (fft.h:221-276).  kind: 'c2c', 'r2c' or 'c2r'."""

    def __init__(self, kind: str, Nx: int, Ny: int, direction: int = FORWARD, flags: int = 0):
        L = _lib.load()
        self.kind, self.Nx, self.Ny = kind, Nx, Ny
        if kind == "c2c":
            h = L.mufft_create_plan_2d_c2c(Nx, Ny, direction, flags)
        elif kind == "r2c":
            h = L.mufft_create_plan_2d_r2c(Nx, Ny, flags)
        elif kind == "c2r":
            h = L.mufft_create_plan_2d_c2r(Nx, Ny, flags)
        else:
            raise ValueError(kind)
        if not h:
            raise MufftError(f"mufft_create_plan_2d_{kind}({Nx}, {Ny}) failed: {_lib.last_error()}")
        self._h = h

    @property
    def launches(self) -> int:
        return _lib.load().mufft_hip_plan_2d_launches(self._h)

    def _validate(self, out, inp, howmany):
        px = self.Nx * self.Ny
        # r2c output / c2r input rows hold Nx complex values, Nx/2+1 of them
        # meaningful (fft.h:238-260); real images are Nx * Ny floats.
        ei = _REAL if self.kind == "r2c" else _CPLX
        eo = _REAL if self.kind == "c2r" else _CPLX
        check_buffer(inp, howmany * px * ei, "input")
        check_buffer(out, howmany * px * eo, "output")

    def execute(self, out, inp):
        self._validate(out, inp, 1)
        _lib.load().mufft_execute_plan_2d(self._h, ptr(out), ptr(inp))

    def execute_batched(self, out, inp, howmany: int, stream=None):
        self._validate(out, inp, howmany)
        rc = _lib.load().mufft_hip_execute_plan_2d_batched(
            self._h, ptr(out), ptr(inp), howmany, stream_handle(stream))
        _check(rc, "mufft_hip_execute_plan_2d_batched")

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().mufft_free_plan_2d(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PlanConv:
    """mufft_plan_conv (fft.h:131-207)."""

    def __init__(self, N: int, flags: int = 0, method: int = 0):
        L = _lib.load()
        h = L.mufft_create_plan_conv(N, flags, method)
        if not h:
            raise MufftError(f"mufft_create_plan_conv({N}) failed: {_lib.last_error()}")
        self._h, self.N, self.method = h, N, method

    @property
    def block_size(self) -> int:
        return int(_lib.load().mufft_conv_get_transformed_block_size(self._h))

    def execute_input(self, block: int, out, inp):
        _lib.load().mufft_execute_conv_input(self._h, block, ptr(out), ptr(inp))

    def execute_output(self, out, first, second):
        _lib.load().mufft_execute_conv_output(self._h, ptr(out), ptr(first), ptr(second))

    def execute_input_batched(self, block, out, inp, howmany, idist=0, odist=0, stream=None):
        rc = _lib.load().mufft_hip_execute_conv_input_batched(
            self._h, block, ptr(out), ptr(inp), howmany, idist, odist, stream_handle(stream))
        _check(rc, "mufft_hip_execute_conv_input_batched")

    def execute_output_batched(self, out, first, second, howmany, odist=0, dist_first=0,
                               dist_second=0, stream=None):
        rc = _lib.load().mufft_hip_execute_conv_output_batched(
            self._h, ptr(out), ptr(first), ptr(second), howmany, odist, dist_first, dist_second,
            stream_handle(stream))
        _check(rc, "mufft_hip_execute_conv_output_batched")

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().mufft_free_plan_conv(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def get_convolve_func(flags: int = 0):
    """mufft_get_convolve_func (fft.h:207) as a Python callable."""
    addr = _lib.load().mufft_get_convolve_func(flags)
    if not addr:
        raise MufftError("mufft_get_convolve_func returned NULL")
    f = _lib.CONVOLVE_FUNC(addr)

    def call(out, a, b, normalization: float, samples: int):
        f(ptr(out), ptr(a), ptr(b), ctypes.c_float(normalization), samples)

    return call


def twiddle_table(N: int, direction: int) -> np.ndarray:
    """The FP32 Stockham table of the c2c plan of size N, generated in FP64 on the GPU."""
    L = _lib.load()
    size = L.mufft_hip_twiddle_table(N, direction, None, 0)
    if size < 0:
        raise MufftError(_lib.last_error())
    out = np.zeros(max(size, 1), np.complex64)
    rc = L.mufft_hip_twiddle_table(N, direction, out.ctypes.data, size)
    if rc < 0:
        raise MufftError(_lib.last_error())
    return out[:size]


class ConvBatch:
    """mufft_hip_conv_batch: uneven convolution jobs in one persistent launch.

    ``jobs`` is a list of (N, method, first, second, output) with device
    buffers (torch tensors or addresses) laid out as in mufft_plan_conv."""

    def __init__(self, jobs, flags: int = 0):
        L = _lib.load()
        arr = (_lib.ConvJob * len(jobs))()
        for i, (N, method, first, second, out) in enumerate(jobs):
            arr[i] = _lib.ConvJob(N, method, ptr(first), ptr(second), ptr(out))
        self._keep = jobs
        h = L.mufft_hip_create_conv_batch(arr, len(jobs), flags)
        if not h:
            raise MufftError(f"mufft_hip_create_conv_batch failed: {_lib.last_error()}")
        self._h = h

    def execute(self, stream=None, graph: bool = True):
        L = _lib.load()
        fn = L.mufft_hip_execute_conv_batch if graph else L.mufft_hip_execute_conv_batch_direct
        _check(fn(self._h, stream_handle(stream)), "mufft_hip_execute_conv_batch")

    def status(self) -> int:
        return int(_lib.load().mufft_hip_conv_batch_status(self._h))

    def info(self) -> tuple[int, int]:
        nt, g = ctypes.c_int(0), ctypes.c_int(0)
        _lib.load().mufft_hip_conv_batch_info(self._h, ctypes.byref(nt), ctypes.byref(g))
        return nt.value, g.value

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().mufft_hip_free_conv_batch(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mul_reduce(out, a, b, n: int, howmany: int, adist: int = 0, bdist: int = 0,
               norm: float = 1.0, conj_b: bool = True, stream=None):
    """out[i] = norm * sum_b a_b[i] * op(b_b[i]) on the device (filter-gradient reduction)."""
    rc = _lib.load().mufft_hip_mul_reduce(ptr(out), ptr(a), ptr(b), n, howmany, adist, bdist,
                                          ctypes.c_float(norm), int(conj_b), stream_handle(stream))
    _check(rc, "mufft_hip_mul_reduce")


class OverlapAdd:
    """mufft_hip_ola: streaming FIR filtering by overlap-add (hop N/2)."""

    def __init__(self, N: int, taps):
        L = _lib.load()
        n = int(taps.numel() if hasattr(taps, "numel") else len(taps))
        h = L.mufft_hip_create_ola(N, ptr(taps), n)
        if not h:
            raise MufftError(f"mufft_hip_create_ola({N}) failed: {_lib.last_error()}")
        self._h, self.N = h, N

    def __call__(self, out, inp, nblocks: int, stream=None):
        rc = _lib.load().mufft_hip_execute_ola(self._h, ptr(out), ptr(inp), nblocks,
                                               stream_handle(stream))
        _check(rc, "mufft_hip_execute_ola")

    def reset(self, stream=None):
        _check(_lib.load().mufft_hip_reset_ola(self._h, stream_handle(stream)), "mufft_hip_reset_ola")

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().mufft_hip_free_ola(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


KIND_CODES = {"c2c_forward": 0, "c2c_inverse": 1, "r2c": 2, "c2r": 3}


def autotune(N: int, kind: str, howmany: int = 64) -> tuple[int, float]:
    """Times every fused-kernel variant of (N, kind) on the device, keeps the
    fastest one that passes the FP64 check for plans created afterwards, and
    returns (variant, microseconds)."""
    L = _lib.load()
    us = ctypes.c_double(0)
    v = L.mufft_hip_autotune_1d(N, KIND_CODES[kind], howmany, ctypes.byref(us))
    if v < 0:
        raise MufftError(f"autotune({N}, {kind}): {_lib.last_error()}")
    return v, us.value


def set_variant(N: int, kind: str, variant: int) -> None:
    """Selects a registered variant (0 = static default) for plans created afterwards."""
    _check(_lib.load().mufft_hip_set_variant(N, KIND_CODES[kind], variant), "mufft_hip_set_variant")


def get_variant(N: int, kind: str) -> int:
    return int(_lib.load().mufft_hip_get_variant(N, KIND_CODES[kind]))


def autotune_large(N: int, kind: str, howmany: int = 1) -> tuple[int, float]:
    """Times every runnable split of a transform beyond the fused sizes
    (2^17..2^26 complex points), keeps the fastest that passes the FP64 check
    for plans created afterwards, and returns (split code, microseconds)."""
    us = ctypes.c_double(0)
    c = _lib.load().mufft_hip_autotune_large_1d(N, KIND_CODES[kind], howmany, ctypes.byref(us))
    if c < 0:
        raise MufftError(f"autotune_large({N}, {kind}): {_lib.last_error()}")
    return c, us.value


def set_large_split(N: int, kind: str, code: int) -> None:
    """Selects a split code (0 = static) for plans created afterwards."""
    _check(_lib.load().mufft_hip_set_large_split(N, KIND_CODES[kind], code), "mufft_hip_set_large_split")


def get_large_split(N: int, kind: str) -> int:
    return int(_lib.load().mufft_hip_get_large_split(N, KIND_CODES[kind]))


def autotune_2d(Nx: int, Ny: int, kind: str, howmany: int = 1) -> tuple[int, float]:
    """Times the column tile widths of the Ny-point column kernel on an
    Nx x Ny image batch, keeps the faster for plans created afterwards and
    returns (tile width, microseconds)."""
    us = ctypes.c_double(0)
    t = _lib.load().mufft_hip_autotune_2d(Nx, Ny, KIND_CODES[kind], howmany, ctypes.byref(us))
    if t < 0:
        raise MufftError(f"autotune_2d({Nx}, {Ny}, {kind}): {_lib.last_error()}")
    return t, us.value


def set_column_tile(L: int, tile: int) -> None:
    _check(_lib.load().mufft_hip_set_column_tile(L, tile), "mufft_hip_set_column_tile")


def get_column_tile(L: int) -> int:
    return int(_lib.load().mufft_hip_get_column_tile(L))


def device_count() -> int:
    return int(_lib.load().mufft_hip_device_count())
