"""Seeded synthetic inputs (test infrastructure and benchmark inputs).

SplitMix64 (Steele, Lea, Flood, "Fast splittable pseudorandom number
generators", OOPSLA 2014) replaces the reference tests' libc ``rand()``
(test.c:38-44), whose sequence depends on the C library.  Vectorised: value i
of stream ``seed`` is the SplitMix64 output for state seed + (i+1)*golden.
uniform[-1,1) = 2*((x >> 40) * 2^-24) - 1; Gaussian = Box-Muller in FP64,
rounded once to float32.
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, count: int, offset: int = 0) -> np.ndarray:
    i = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, count: int, lo: float = -1.0, hi: float = 1.0, offset: int = 0) -> np.ndarray:
    """float32 uniform in [lo, hi); [-1, 1) is exactly 2*u24 - 1."""
    u = (splitmix64(seed, count, offset) >> np.uint64(40)).astype(np.float64) * (2.0 ** -24)
    return (lo + (hi - lo) * u).astype(np.float32)


def gaussian(seed: int, count: int, sigma: float = 1.0) -> np.ndarray:
    m = (count + 1) // 2
    u1 = ((splitmix64(seed, m) >> np.uint64(11)).astype(np.float64) + 1.0) * (2.0 ** -53)
    u2 = (splitmix64(seed ^ 0x5DEECE66D, m) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    r = np.sqrt(-2.0 * np.log(u1))
    g = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])[:count]
    return (sigma * g).astype(np.float32)


def complex_uniform(seed: int, count: int) -> np.ndarray:
    """Interleaved re/im, both uniform[-1,1): complex64[count]."""
    return uniform(seed, 2 * count).view(np.complex64)
