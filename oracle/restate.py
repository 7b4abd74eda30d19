"""float32 restatement of muFFT's scalar C path (test infrastructure only).

Every function names the reference lines it follows.  Arithmetic is done on
separate float32 real/imaginary arrays, one IEEE operation per numpy ufunc
call, in the reference's grouping, so the result is bit-identical to the
compiled reference built without SIMD (plan flags = MUFFT_FLAG_CPU_NO_SIMD).
That equality is what pins this restatement (tests/test_oracle.py); the FP64
oracle in ``oracle/fp64.py`` pins it a second time, within tolerance.

Arrays carry leading batch axes; the transform axis is the last one (1D and
row passes) or the second-to-last one (vertical passes).  Buffers are
functional (each pass returns a new array) instead of the reference's
ping-pong between ``output`` and ``tmp_buffer`` (fft.c:811-850); with
``restrict`` inputs the two are equivalent.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# Public constants, fft.h:44-65 and fft.h:131-150.
FORWARD = -1
INVERSE = 1
FLAG_CPU_ANY = 0
FLAG_CPU_NO_SIMD = (1 << 16) - 1
FLAG_CPU_NO_AVX = 1 << 0
FLAG_CPU_NO_SSE3 = 1 << 1
FLAG_CPU_NO_SSE = 1 << 2
FLAG_FULL_R2C = 1 << 16
FLAG_ZERO_PAD_UPPER_HALF = 1 << 17
CONV_METHOD_FLAG_MONO_MONO = 0
CONV_METHOD_FLAG_STEREO_MONO = 1
CONV_METHOD_FLAG_ZERO_PAD_UPPER_HALF_FIRST = 1 << 2
CONV_METHOD_FLAG_ZERO_PAD_UPPER_HALF_SECOND = 1 << 3
CONV_BLOCK_FIRST = 0
CONV_BLOCK_SECOND = 1

# Internal planner flags, fft_internal.h:169-192.
_CPU_AVX = FLAG_CPU_NO_AVX
_CPU_SSE3 = FLAG_CPU_NO_SSE3
_CPU_SSE = FLAG_CPU_NO_SSE
_DIR_INVERSE = 1 << 24
_DIR_FORWARD = 1 << 25
_DIR_ANY = 0
_R2C = 1 << 26
_C2R = 1 << 27
_NO_ZERO_PAD = 1 << 28
_PADDING_COMPLEX_SAMPLES = 4  # fft_internal.h:203

F32 = np.float32


# ---------------------------------------------------------------------------
# complex helpers, fft_internal.h:47-106 (cfloat_* in float32, no FMA)
# ---------------------------------------------------------------------------
def _mk(re, im) -> np.ndarray:
    re = np.asarray(re, dtype=F32)
    im = np.asarray(im, dtype=F32)
    out = np.empty(np.broadcast_shapes(re.shape, im.shape), np.complex64)
    out.real = re
    out.imag = im
    return out


def cadd(a, b):
    return _mk(a.real + b.real, a.imag + b.imag)


def csub(a, b):
    return _mk(a.real - b.real, a.imag - b.imag)


def cmul(a, b):
    """cfloat_mul, fft_internal.h:96-100: (ar*br - ai*bi, ar*bi + br*ai)."""
    return _mk(a.real * b.real - a.imag * b.imag, a.real * b.imag + b.real * a.imag)


def cmul_scalar(s, a):
    """cfloat_mul_scalar, fft_internal.h:102-106."""
    s = F32(s)
    return _mk(a.real * s, a.imag * s)


def cconj(a):
    return _mk(a.real, -a.imag)


# ---------------------------------------------------------------------------
# twiddles, fft.c:101-134 and fft.c:441-455
# ---------------------------------------------------------------------------
def twiddle(direction: int, k: int, p: int) -> np.complex64:
    """fft.c:102-106: phase = (M_PI*direction*k)/p in double, then (float)cos/sin."""
    phase = (math.pi * direction * k) / p
    return np.complex64(complex(F32(math.cos(phase)), F32(math.sin(phase))))


def build_twiddles(N: int, direction: int) -> np.ndarray:
    """fft.c:114-134.  Entry 3 is never initialised by the reference (never read); 0 here."""
    tw = np.zeros(N, np.complex64)
    off = 0
    p = 1
    while p < N:
        for k in range(p):
            tw[off + k] = twiddle(direction, k, p)
        off += 3 if p == 2 else p
        p <<= 1
    return tw


def build_r2c_twiddles(direction: int, N: int) -> np.ndarray:
    """fft.c:441-455: W[i] = (0, direction) * twiddle(direction, i, N), float32 complex multiply."""
    k = np.arange(N)
    phase = (math.pi * direction * k.astype(np.float64)) / N
    t = _mk(np.cos(phase).astype(F32), np.sin(phase).astype(F32))
    a = _mk(F32(0.0), F32(direction))
    return cmul(np.broadcast_to(a, t.shape), t)


# ---------------------------------------------------------------------------
# 1D kernels, kernel.c:84-453 (scalar "_c" variants)
# ---------------------------------------------------------------------------
def _zeros_like(x):
    return np.zeros(x.shape, np.complex64)


def radix2_p1(inp, tw, p, n):
    """kernel.c:84-102."""
    h = n >> 1
    a, b = inp[..., :h], inp[..., h:2 * h]
    out = np.empty(inp.shape[:-1] + (n,), np.complex64)
    out[..., 0::2] = cadd(a, b)
    out[..., 1::2] = csub(a, b)
    return out


def radix2_half_p1(inp, tw, p, n):
    """kernel.c:104-122 (upper half implicitly zero; the adds with 0 are kept)."""
    h = n >> 1
    a = inp[..., :h]
    b = _zeros_like(a)
    out = np.empty(inp.shape[:-1] + (n,), np.complex64)
    out[..., 0::2] = cadd(a, b)
    out[..., 1::2] = csub(a, b)
    return out


def radix2_generic(inp, tw, p, n):
    """kernel.c:151-168 (also kernel.c:124-149, the p == 2 special cases)."""
    h = n >> 1
    i = np.arange(h)
    k = i & (p - 1)
    a = inp[..., :h]
    b = cmul(tw[k], inp[..., h:2 * h])
    j = (i << 1) - k
    out = np.empty(inp.shape[:-1] + (n,), np.complex64)
    out[..., j] = cadd(a, b)
    out[..., j + p] = csub(a, b)
    return out


def forward_radix2_p2(inp, tw, p, n):
    """kernel.c:124-143 (inverse_radix2_p2, kernel.c:145-149, is the same code)."""
    return radix2_generic(inp, tw, 2, n)


def forward_radix4_p1(inp, tw, p, n):
    """kernel.c:170-198 (inverse_radix4_p1, kernel.c:228-232, is the same code)."""
    q = n >> 2
    a, b, c, d = (inp[..., m * q:(m + 1) * q] for m in range(4))
    r0, r1, r2, r3 = cadd(a, c), csub(a, c), cadd(b, d), csub(b, d)
    r3 = cmul(r3, tw[2])
    out = np.empty(inp.shape[:-1] + (n,), np.complex64)
    out[..., 0::4] = cadd(r0, r2)
    out[..., 1::4] = cadd(r1, r3)
    out[..., 2::4] = csub(r0, r2)
    out[..., 3::4] = csub(r1, r3)
    return out


def forward_half_radix4_p1(inp, tw, p, n):
    """kernel.c:200-226."""
    q = n >> 2
    a, b = inp[..., :q], inp[..., q:2 * q]
    r0, r1, r2 = a, a, b
    r3 = cmul(b, tw[2])
    out = np.empty(inp.shape[:-1] + (n,), np.complex64)
    out[..., 0::4] = cadd(r0, r2)
    out[..., 1::4] = cadd(r1, r3)
    out[..., 2::4] = csub(r0, r2)
    out[..., 3::4] = csub(r1, r3)
    return out


def radix4_generic(inp, tw, p, n):
    """kernel.c:234-271."""
    q = n >> 2
    i = np.arange(q)
    k = i & (p - 1)
    a = inp[..., 0:q]
    b = inp[..., q:2 * q]
    c = cmul(tw[k], inp[..., 2 * q:3 * q])
    d = cmul(tw[k], inp[..., 3 * q:4 * q])
    r0, r1, r2, r3 = cadd(a, c), csub(a, c), cadd(b, d), csub(b, d)
    r2 = cmul(r2, tw[p + k])
    r3 = cmul(r3, tw[p + k + p])
    o0, o1, o2, o3 = cadd(r0, r2), csub(r0, r2), cadd(r1, r3), csub(r1, r3)
    j = ((i - k) << 2) + k
    out = np.empty(inp.shape[:-1] + (n,), np.complex64)
    out[..., j] = o0
    out[..., j + p] = o2
    out[..., j + 2 * p] = o1
    out[..., j + 3 * p] = o3
    return out


def _radix8_tail(a, b, c, d, e, f, g, h, tw2, tw4, tw5, tw6, tw7):
    """Shared body of kernel.c:293-329 / kernel.c:349-385."""
    r0, r1, r2, r3 = cadd(a, e), csub(a, e), cadd(b, f), csub(b, f)
    r4, r5, r6, r7 = cadd(c, g), csub(c, g), cadd(d, h), csub(d, h)
    r5 = cmul(r5, tw2)
    r7 = cmul(r7, tw2)
    a, b, c, d = cadd(r0, r4), cadd(r1, r5), csub(r0, r4), csub(r1, r5)
    e, f, g, h = cadd(r2, r6), cadd(r3, r7), csub(r2, r6), csub(r3, r7)
    if tw4 is not None:
        e = cmul(e, tw4)
    f = cmul(f, tw5)
    g = cmul(g, tw6)
    h = cmul(h, tw7)
    return (cadd(a, e), cadd(b, f), cadd(c, g), cadd(d, h),
            csub(a, e), csub(b, f), csub(c, g), csub(d, h))


def forward_radix8_p1(inp, tw, p, n):
    """kernel.c:273-331 (inverse_radix8_p1, kernel.c:389-393, is the same code)."""
    o = n >> 3
    x = [inp[..., m * o:(m + 1) * o] for m in range(8)]
    ys = _radix8_tail(*x, tw[2], tw[4], tw[5], tw[6], tw[7])
    out = np.empty(inp.shape[:-1] + (n,), np.complex64)
    for m, y in enumerate(ys):
        out[..., m::8] = y
    return out


def forward_half_radix8_p1(inp, tw, p, n):
    """kernel.c:333-387: r0 = r1 = a, r2 = r3 = b, r4 = r5 = c, r6 = r7 = d."""
    o = n >> 3
    a, b, c, d = (inp[..., m * o:(m + 1) * o] for m in range(4))
    r0, r1, r2, r3, r4, r5, r6, r7 = a, a, b, b, c, c, d, d
    r5 = cmul(r5, tw[2])
    r7 = cmul(r7, tw[2])
    a2, b2, c2, d2 = cadd(r0, r4), cadd(r1, r5), csub(r0, r4), csub(r1, r5)
    e, f, g, h = cadd(r2, r6), cadd(r3, r7), csub(r2, r6), csub(r3, r7)
    e, f, g, h = cmul(e, tw[4]), cmul(f, tw[5]), cmul(g, tw[6]), cmul(h, tw[7])
    ys = (cadd(a2, e), cadd(b2, f), cadd(c2, g), cadd(d2, h),
          csub(a2, e), csub(b2, f), csub(c2, g), csub(d2, h))
    out = np.empty(inp.shape[:-1] + (n,), np.complex64)
    for m, y in enumerate(ys):
        out[..., m::8] = y
    return out


def radix8_generic(inp, tw, p, n):
    """kernel.c:395-453."""
    o = n >> 3
    i = np.arange(o)
    k = i & (p - 1)
    x = [inp[..., m * o:(m + 1) * o] for m in range(8)]
    a, b, c, d = x[0], x[1], x[2], x[3]
    e, f, g, h = (cmul(tw[k], x[m]) for m in range(4, 8))
    r0, r1, r2, r3 = cadd(a, e), csub(a, e), cadd(b, f), csub(b, f)
    r4, r5, r6, r7 = cadd(c, g), csub(c, g), cadd(d, h), csub(d, h)
    r4 = cmul(r4, tw[p + k])
    r5 = cmul(r5, tw[p + k + p])
    r6 = cmul(r6, tw[p + k])
    r7 = cmul(r7, tw[p + k + p])
    a, b, c, d = cadd(r0, r4), cadd(r1, r5), csub(r0, r4), csub(r1, r5)
    e, f, g, h = cadd(r2, r6), cadd(r3, r7), csub(r2, r6), csub(r3, r7)
    e = cmul(e, tw[3 * p + k])
    f = cmul(f, tw[3 * p + k + p])
    g = cmul(g, tw[3 * p + k + 2 * p])
    h = cmul(h, tw[3 * p + k + 3 * p])
    ys = (cadd(a, e), cadd(b, f), cadd(c, g), cadd(d, h),
          csub(a, e), csub(b, f), csub(c, g), csub(d, h))
    j = ((i - k) << 3) + k
    out = np.empty(inp.shape[:-1] + (n,), np.complex64)
    for m, y in enumerate(ys):
        out[..., j + m * p] = y
    return out


# ---------------------------------------------------------------------------
# vertical kernels, kernel.c:455-738, restated on (..., lines, columns) arrays
# ---------------------------------------------------------------------------
def _vout(inp):
    return np.empty(inp.shape, np.complex64)


def radix2_p1_vert(inp, tw, p):
    """kernel.c:455-481."""
    ny = inp.shape[-2]
    hl = ny >> 1
    a, b = inp[..., :hl, :], inp[..., hl:2 * hl, :]
    out = _vout(inp)
    out[..., 0::2, :] = cadd(a, b)
    out[..., 1::2, :] = csub(a, b)
    return out


def radix2_generic_vert(inp, tw, p):
    """kernel.c:483-511."""
    ny = inp.shape[-2]
    hl = ny >> 1
    line = np.arange(hl)
    k = line & (p - 1)
    j = (line << 1) - k
    a = inp[..., :hl, :]
    b = cmul(tw[k][:, None], inp[..., hl:2 * hl, :])
    out = _vout(inp)
    out[..., j, :] = cadd(a, b)
    out[..., j + p, :] = csub(a, b)
    return out


def forward_radix4_p1_vert(inp, tw, p):
    """kernel.c:513-553 (inverse variant kernel.c:555-559 is the same code)."""
    ny = inp.shape[-2]
    q = ny >> 2
    a, b, c, d = (inp[..., m * q:(m + 1) * q, :] for m in range(4))
    r0, r1, r2, r3 = cadd(a, c), csub(a, c), cadd(b, d), csub(b, d)
    r3 = cmul(r3, tw[2])
    out = _vout(inp)
    out[..., 0::4, :] = cadd(r0, r2)
    out[..., 1::4, :] = cadd(r1, r3)
    out[..., 2::4, :] = csub(r0, r2)
    out[..., 3::4, :] = csub(r1, r3)
    return out


def radix4_generic_vert(inp, tw, p):
    """kernel.c:561-602."""
    ny = inp.shape[-2]
    q = ny >> 2
    line = np.arange(q)
    k = line & (p - 1)
    t = tw[k][:, None]
    a = inp[..., 0:q, :]
    b = inp[..., q:2 * q, :]
    c = cmul(t, inp[..., 2 * q:3 * q, :])
    d = cmul(t, inp[..., 3 * q:4 * q, :])
    r0, r1, r2, r3 = cadd(a, c), csub(a, c), cadd(b, d), csub(b, d)
    r2 = cmul(r2, tw[p + k][:, None])
    r3 = cmul(r3, tw[p + k + p][:, None])
    a, b, c, d = cadd(r0, r2), cadd(r1, r3), csub(r0, r2), csub(r1, r3)
    j = ((line - k) << 2) + k
    out = _vout(inp)
    out[..., j, :] = a
    out[..., j + p, :] = b
    out[..., j + 2 * p, :] = c
    out[..., j + 3 * p, :] = d
    return out


def forward_radix8_p1_vert(inp, tw, p):
    """kernel.c:604-665 (no multiply by T[4], unlike the 1D kernel)."""
    ny = inp.shape[-2]
    o = ny >> 3
    x = [inp[..., m * o:(m + 1) * o, :] for m in range(8)]
    ys = _radix8_tail(*x, tw[2], None, tw[5], tw[6], tw[7])
    out = _vout(inp)
    for m, y in enumerate(ys):
        out[..., m::8, :] = y
    return out


def radix8_generic_vert(inp, tw, p):
    """kernel.c:673-738."""
    ny = inp.shape[-2]
    o = ny >> 3
    line = np.arange(o)
    k = line & (p - 1)
    x = [inp[..., m * o:(m + 1) * o, :] for m in range(8)]
    t = tw[k][:, None]
    a, b, c, d = x[0], x[1], x[2], x[3]
    e, f, g, h = (cmul(t, x[m]) for m in range(4, 8))
    r0, r1, r2, r3 = cadd(a, e), csub(a, e), cadd(b, f), csub(b, f)
    r4, r5, r6, r7 = cadd(c, g), csub(c, g), cadd(d, h), csub(d, h)
    r4 = cmul(r4, tw[p + k][:, None])
    r5 = cmul(r5, tw[p + k + p][:, None])
    r6 = cmul(r6, tw[p + k][:, None])
    r7 = cmul(r7, tw[p + k + p][:, None])
    a, b, c, d = cadd(r0, r4), cadd(r1, r5), csub(r0, r4), csub(r1, r5)
    e, f, g, h = cadd(r2, r6), cadd(r3, r7), csub(r2, r6), csub(r3, r7)
    e = cmul(e, tw[3 * p + k][:, None])
    f = cmul(f, tw[3 * p + k + p][:, None])
    g = cmul(g, tw[3 * p + k + 2 * p][:, None])
    h = cmul(h, tw[3 * p + k + 3 * p][:, None])
    ys = (cadd(a, e), cadd(b, f), cadd(c, g), cadd(d, h),
          csub(a, e), csub(b, f), csub(c, g), csub(d, h))
    j = ((line - k) << 3) + k
    out = _vout(inp)
    for m, y in enumerate(ys):
        out[..., j + m * p, :] = y
    return out


# ---------------------------------------------------------------------------
# resolve and convolve, kernel.c:22-82
# ---------------------------------------------------------------------------
def convolve(a, b, normalization, samples):
    """kernel.c:22-34: out[i] = normalization * (a[i] * b[i])."""
    return cmul_scalar(normalization, cmul(a[..., :samples], b[..., :samples]))


def resolve_c2r(inp, tw, samples):
    """kernel.c:36-45: out[i] = (a + b) + (a - b) * W[i], b = conj(in[samples - i])."""
    i = np.arange(samples)
    a = inp[..., i]
    b = cconj(inp[..., samples - i])
    return cadd(cadd(a, b), cmul(csub(a, b), tw[i]))


def _resolve_r2c(inp, tw, samples, full):
    """kernel.c:47-82."""
    n_out = 2 * samples if full else samples + 1
    out = np.zeros(inp.shape[:-1] + (n_out,), np.complex64)
    z0 = inp[..., 0]
    fe = _mk(z0.real, F32(0.0))
    fo = _mk(z0.imag, F32(0.0))
    out[..., 0] = cadd(fe, fo)
    out[..., samples] = csub(fe, fo)
    i = np.arange(1, samples)
    a = inp[..., i]
    b = cconj(inp[..., samples - i])
    fe = cadd(a, b)
    fo = cmul(tw[i], csub(a, b))
    out[..., i] = cmul_scalar(0.5, cadd(fe, fo))
    if full:
        out[..., i + samples] = cmul_scalar(0.5, csub(fe, fo))
    return out


def resolve_r2c(inp, tw, samples):
    return _resolve_r2c(inp, tw, samples, False)


def resolve_r2c_full(inp, tw, samples):
    return _resolve_r2c(inp, tw, samples, True)


# ---------------------------------------------------------------------------
# planner, fft.c:181-435 (registry tables, add_step, greedy first fit)
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class _Entry1D:
    name: str
    func: object
    radix: int
    min_mult: int
    fixed_p: int
    minimum_p: int
    flags: int


@dataclass(frozen=True)
class _Entry2D:
    name: str
    func: object
    radix: int
    min_y: int
    fixed_p: int
    minimum_p: int
    flags: int


_NEVER = 0xFFFFFFFF


def _table_1d():
    """fft.c:217-256 (STAMP_CPU_1D), arch bits in front, C last."""
    rows = [
        ("forward_radix8_p1", forward_radix8_p1, 8, 8, 1, _NEVER, _DIR_FORWARD | _NO_ZERO_PAD),
        ("forward_radix4_p1", forward_radix4_p1, 4, 4, 1, _NEVER, _DIR_FORWARD | _NO_ZERO_PAD),
        ("radix2_p1", radix2_p1, 2, 2, 1, _NEVER, _DIR_ANY | _NO_ZERO_PAD),
        ("forward_half_radix8_p1", forward_half_radix8_p1, 8, 8, 1, _NEVER,
         _DIR_FORWARD | FLAG_ZERO_PAD_UPPER_HALF),
        ("forward_half_radix4_p1", forward_half_radix4_p1, 4, 4, 1, _NEVER,
         _DIR_FORWARD | FLAG_ZERO_PAD_UPPER_HALF),
        ("radix2_half_p1", radix2_half_p1, 2, 2, 1, _NEVER, _DIR_ANY | FLAG_ZERO_PAD_UPPER_HALF),
        ("forward_radix2_p2", forward_radix2_p2, 2, 2, 2, _NEVER, _DIR_FORWARD),
        ("inverse_radix8_p1", forward_radix8_p1, 8, 8, 1, _NEVER, _DIR_INVERSE),
        ("inverse_radix4_p1", forward_radix4_p1, 4, 4, 1, _NEVER, _DIR_INVERSE),
        ("inverse_radix2_p2", forward_radix2_p2, 2, 2, 2, _NEVER, _DIR_INVERSE),
        ("radix8_generic", radix8_generic, 8, 8, 0, 8, _DIR_ANY),
        ("radix4_generic", radix4_generic, 4, 4, 0, 4, _DIR_ANY),
        ("radix2_generic", radix2_generic, 2, 2, 0, 4, _DIR_ANY),
    ]
    out = []
    for arch, min_x in ((_CPU_AVX, 4), (_CPU_SSE3, 2), (_CPU_SSE, 2), (0, 1)):
        for name, func, radix, mm, fixed_p, min_p, fl in rows:
            out.append((arch, min_x, _Entry1D(name, func, radix, mm * min_x, fixed_p, min_p, fl | arch)))
    return out


def _table_2d():
    """fft.c:258-287 (STAMP_CPU_2D)."""
    rows = [
        ("forward_radix8_p1_vert", forward_radix8_p1_vert, 8, 8, 1, _NEVER, _DIR_FORWARD),
        ("forward_radix4_p1_vert", forward_radix4_p1_vert, 4, 4, 1, _NEVER, _DIR_FORWARD),
        ("radix2_p1_vert", radix2_p1_vert, 2, 2, 1, _NEVER, _DIR_ANY),
        ("inverse_radix8_p1_vert", forward_radix8_p1_vert, 8, 8, 1, _NEVER, _DIR_INVERSE),
        ("inverse_radix4_p1_vert", forward_radix4_p1_vert, 4, 4, 1, _NEVER, _DIR_INVERSE),
        ("radix8_generic_vert", radix8_generic_vert, 8, 8, 0, 8, _DIR_ANY),
        ("radix4_generic_vert", radix4_generic_vert, 4, 4, 0, 4, _DIR_ANY),
        ("radix2_generic_vert", radix2_generic_vert, 2, 2, 0, 2, _DIR_ANY),
    ]
    out = []
    for arch, min_x in ((_CPU_AVX, 4), (_CPU_SSE3, 2), (_CPU_SSE, 2), (0, 1)):
        for name, func, radix, min_y, fixed_p, min_p, fl in rows:
            out.append((arch, min_x, _Entry2D(name, func, radix, min_y, fixed_p, min_p, fl | arch)))
    return out


_T1D = _table_1d()
_T2D = _table_2d()


@dataclass
class Step:
    name: str
    func: object
    radix: int
    p: int
    twiddle_offset: int


def _add_step(steps, entry, p):
    """fft.c:290-322."""
    off = 0
    if steps:
        prev = steps[-1]
        off = prev.twiddle_offset + (3 if prev.p == 2 else prev.p * (prev.radix - 1))
        if p >= 4 and prev.p == 1:
            off += 1
    steps.append(Step(entry.name, entry.func, entry.radix, p, off))


def _dir_flags(direction):
    if direction == FORWARD:
        return _DIR_FORWARD
    if direction == INVERSE:
        return _DIR_INVERSE
    return 0


def build_plan_1d(N, direction, flags, cpu_flags=0):
    """fft.c:325-379.  ``cpu_flags`` stands for mufft_get_cpu_flags() (cpu.c:78-114);
    the default 0 selects the scalar C table, which is the arithmetic restated here."""
    sf = _dir_flags(direction)
    sf |= cpu_flags & ~(FLAG_CPU_NO_SIMD & flags)
    sf |= FLAG_ZERO_PAD_UPPER_HALF if (flags & FLAG_ZERO_PAD_UPPER_HALF) else _NO_ZERO_PAD
    steps = []
    radix, p = N, 1
    while radix > 1:
        for _, _, e in _T1D:
            if (radix % e.radix == 0 and N >= e.min_mult and (sf & e.flags) == e.flags
                    and (p >= e.minimum_p or p == e.fixed_p)):
                _add_step(steps, e, p)
                radix //= e.radix
                p *= e.radix
                break
        else:
            return None
    return steps


def build_plan_2d(Nx, Ny, direction, flags, cpu_flags=0):
    """fft.c:382-435."""
    sf = _dir_flags(direction)
    sf |= cpu_flags & ~(FLAG_CPU_NO_SIMD & flags)
    steps = []
    radix, p = Ny, 1
    while radix > 1:
        for _, min_x, e in _T2D:
            if (radix % e.radix == 0 and Ny >= e.min_y and Nx >= min_x
                    and (sf & e.flags) == e.flags and (p >= e.minimum_p or p == e.fixed_p)):
                _add_step(steps, e, p)
                radix //= e.radix
                p *= e.radix
                break
        else:
            return None
    return steps


def _is_pow2(n):
    return n > 0 and (n & (n - 1)) == 0


# ---------------------------------------------------------------------------
# plans and executors, fft.c:476-989
# ---------------------------------------------------------------------------
@dataclass
class Plan1D:
    N: int  # complex transform size (fft.c:58)
    direction: int
    steps: list
    twiddles: np.ndarray
    kind: str = "c2c"  # c2c | r2c | r2c_full | c2r
    r2c_twiddles: np.ndarray = field(default=None)
    zero_pad: bool = False


def create_plan_1d_c2c(N, direction, flags=0):
    """fft.c:626-662."""
    if not _is_pow2(N) or N == 1:
        return None
    steps = build_plan_1d(N, direction, flags)
    if steps is None:
        return None
    return Plan1D(N, direction, steps, build_twiddles(N, direction),
                  zero_pad=bool(flags & FLAG_ZERO_PAD_UPPER_HALF))


def create_plan_1d_r2c(N, flags=0):
    """fft.c:476-513: an N/2 forward c2c plus the r2c resolve."""
    if not _is_pow2(N) or N == 1:
        return None
    plan = create_plan_1d_c2c(N // 2, FORWARD, flags)
    if plan is None:
        return None
    plan.r2c_twiddles = build_r2c_twiddles(FORWARD, N // 2)
    plan.kind = "r2c_full" if (flags & FLAG_FULL_R2C) else "r2c"
    return plan


def create_plan_1d_c2r(N, flags=0):
    """fft.c:515-547: the c2r resolve plus an N/2 inverse c2c."""
    if not _is_pow2(N) or N == 1:
        return None
    plan = create_plan_1d_c2c(N // 2, INVERSE, flags)
    if plan is None:
        return None
    plan.r2c_twiddles = build_r2c_twiddles(INVERSE, N // 2)
    plan.kind = "c2r"
    return plan


def _run_steps_1d(steps, tw, x, N):
    """Step loop of fft.c:827-843."""
    for s in steps:
        x = s.func(x, tw[s.twiddle_offset:], s.p, N)
    return x


def execute_plan_1d(plan: Plan1D, inp) -> np.ndarray:
    """fft.c:811-850.  Input/output types follow the plan kind:
    c2c complex64[..., N] (N/2 with zero pad) -> complex64[..., N];
    r2c float32[..., 2N] (N with zero pad) -> complex64[..., N+1] (2N when full);
    c2r complex64[..., N+1] -> float32[..., 2N]."""
    N = plan.N
    if plan.kind == "c2r":
        x = resolve_c2r(np.asarray(inp, np.complex64), plan.r2c_twiddles, N)
        y = _run_steps_1d(plan.steps, plan.twiddles, x, N)
        return y.view(np.float32)
    if plan.kind.startswith("r2c"):
        x = np.ascontiguousarray(inp, np.float32).view(np.complex64)
    else:
        x = np.asarray(inp, np.complex64)
    y = _run_steps_1d(plan.steps, plan.twiddles, x, N)
    if plan.kind == "r2c":
        return resolve_r2c(y, plan.r2c_twiddles, N)
    if plan.kind == "r2c_full":
        return resolve_r2c_full(y, plan.r2c_twiddles, N)
    return y


@dataclass
class Plan2D:
    Nx: int  # complex columns (Nx_real / 2 for r2c/c2r), fft.c:75
    Ny: int
    direction: int
    steps_x: list
    steps_y: list
    twiddles_x: np.ndarray
    twiddles_y: np.ndarray
    kind: str = "c2c"
    r2c_twiddles: np.ndarray = field(default=None)
    vertical_nx: int = 0


def create_plan_2d_c2c(Nx, Ny, direction, flags=0):
    """fft.c:664-716."""
    if not _is_pow2(Nx) or not _is_pow2(Ny) or Nx == 1 or Ny == 1:
        return None
    sx = build_plan_1d(Nx, direction, flags)
    sy = build_plan_2d(Nx, Ny, direction, flags)
    if sx is None or sy is None:
        return None
    return Plan2D(Nx, Ny, direction, sx, sy, build_twiddles(Nx, direction),
                  build_twiddles(Ny, direction), vertical_nx=Nx)


def create_plan_2d_r2c(Nx, Ny, flags=0):
    """fft.c:718-759."""
    if not _is_pow2(Nx) or not _is_pow2(Ny) or Nx == 1 or Ny == 1:
        return None
    plan = create_plan_2d_c2c(Nx // 2, Ny, FORWARD, flags | _R2C)
    if plan is None:
        return None
    plan.r2c_twiddles = build_r2c_twiddles(FORWARD, Nx // 2)
    if flags & FLAG_FULL_R2C:
        plan.kind, plan.vertical_nx = "r2c_full", Nx
    else:
        plan.kind, plan.vertical_nx = "r2c", Nx // 2 + 1
    return plan


def create_plan_2d_c2r(Nx, Ny, flags=0):
    """fft.c:761-792."""
    if not _is_pow2(Nx) or not _is_pow2(Ny) or Nx == 1 or Ny == 1:
        return None
    plan = create_plan_2d_c2c(Nx // 2, Ny, INVERSE, flags | _C2R)
    if plan is None:
        return None
    plan.r2c_twiddles = build_r2c_twiddles(INVERSE, Nx // 2)
    plan.kind = "c2r"
    plan.vertical_nx = Nx // 2 + 1
    return plan


def _run_steps_vert(steps, tw, x):
    """Vertical step loops of fft.c:882-889 and fft.c:978-987."""
    for s in steps:
        x = s.func(x, tw[s.twiddle_offset:], s.p)
    return x


def execute_plan_2d(plan: Plan2D, inp) -> np.ndarray:
    """fft.c:852-989.  Layouts (fft.h:225-260):
    c2c complex64[..., Ny, Nx] -> complex64[..., Ny, Nx];
    r2c float32[..., Ny, 2Nx] -> complex64[..., Ny, 2Nx], only the first Nx+1
        columns (all 2Nx when full) are meaningful, the rest are zero here;
    c2r complex64[..., Ny, 2Nx] (first Nx+1 columns read) -> float32[..., Ny, 2Nx]."""
    Nx, Ny = plan.Nx, plan.Ny
    if plan.kind == "c2r":
        x = np.asarray(inp, np.complex64)[..., :, :Nx + 1]
        x = _run_steps_vert(plan.steps_y, plan.twiddles_y, x)   # fft.c:882-889
        z = resolve_c2r(x, plan.r2c_twiddles, Nx)               # fft.c:894-897
        y = _run_steps_1d(plan.steps_x, plan.twiddles_x, z, Nx)  # fft.c:900-915
        return np.ascontiguousarray(y).view(np.float32)
    if plan.kind.startswith("r2c"):
        x = np.ascontiguousarray(inp, np.float32).view(np.complex64)
    else:
        x = np.asarray(inp, np.complex64)
    y = _run_steps_1d(plan.steps_x, plan.twiddles_x, x, Nx)     # fft.c:939-961
    if plan.kind == "r2c":
        r = resolve_r2c(y, plan.r2c_twiddles, Nx)               # fft.c:963-972
        full = np.zeros(r.shape[:-1] + (2 * Nx,), np.complex64)
        full[..., :Nx + 1] = r
        y = full
    elif plan.kind == "r2c_full":
        y = resolve_r2c_full(y, plan.r2c_twiddles, Nx)
    vnx = plan.vertical_nx
    cols = _run_steps_vert(plan.steps_y, plan.twiddles_y, y[..., :, :vnx])  # fft.c:978-987
    out = np.zeros(y.shape, np.complex64)
    out[..., :, :vnx] = cols
    return out


@dataclass
class PlanConv:
    N: int
    method: int
    plans: list
    output_plan: Plan1D
    block_size: int
    normalization: np.float32
    conv_multiply_n: int


def create_plan_conv(N, flags=0, method=0):
    """fft.c:565-624."""
    if not _is_pow2(N) or N == 1:
        return None
    ef = FLAG_ZERO_PAD_UPPER_HALF if method & CONV_METHOD_FLAG_ZERO_PAD_UPPER_HALF_FIRST else 0
    es = FLAG_ZERO_PAD_UPPER_HALF if method & CONV_METHOD_FLAG_ZERO_PAD_UPPER_HALF_SECOND else 0
    if (method & 1) == CONV_METHOD_FLAG_MONO_MONO:
        block_size = (N // 2 + _PADDING_COMPLEX_SAMPLES) * 8
        plans = [create_plan_1d_r2c(N, flags | ef), create_plan_1d_r2c(N, flags | es)]
        out_plan = create_plan_1d_c2r(N, flags)
        n = N // 2 + 1
    else:
        block_size = N * 8
        plans = [create_plan_1d_c2c(N, FORWARD, flags | ef),
                 create_plan_1d_r2c(N, flags | es | FLAG_FULL_R2C)]
        out_plan = create_plan_1d_c2c(N, INVERSE, flags)
        n = N
    if None in plans or out_plan is None:
        return None
    return PlanConv(N, method, plans, out_plan, block_size, F32(1.0) / F32(N), n)


def execute_conv_input(plan: PlanConv, block, inp):
    """fft.c:799-802."""
    return execute_plan_1d(plan.plans[block], inp)


def execute_conv_output(plan: PlanConv, first, second):
    """fft.c:804-809: multiply into conv_block, then the output plan."""
    block = convolve(first, second, plan.normalization, plan.conv_multiply_n)
    return execute_plan_1d(plan.output_plan, block)


def conv_transformed_block_size(plan: PlanConv) -> int:
    """fft.c:794-797."""
    return plan.block_size


def decomposition(N, direction, flags=0, cpu_flags=0):
    """Radix/p list the planner picks (documentation and planner tests)."""
    steps = build_plan_1d(N, direction, flags, cpu_flags)
    return None if steps is None else [(s.name, s.radix, s.p, s.twiddle_offset) for s in steps]
