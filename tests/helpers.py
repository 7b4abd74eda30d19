"""Shared helpers of the parity tests (inputs, tolerances, device transfer)."""
from __future__ import annotations

import math

import numpy as np

from oracle import fp64, inputs

# Reference per-element tolerance (test.c:58, 195): |delta| < 1e-6 * sqrt(N) for
# inputs uniform in [-0.5, 0.5]; scaled here by the input amplitude.
def abs_tol(n_total: int, amplitude: float = 0.5, factor: float = 1e-6) -> float:
    return factor * math.sqrt(n_total) * (amplitude / 0.5)


# Relative RMS bound against muFFT's own CPU output (north-star: 5e-7*log2 N).
def rms_tol_vs_mufft(n: int) -> float:
    return 5e-7 * max(1.0, math.log2(n))


def rel_rms(a, b) -> float:
    return fp64.rel_rms(a, b)


def cplx(seed: int, count: int, scale: float = 1.0) -> np.ndarray:
    return (inputs.complex_uniform(seed, count) * np.float32(scale)).astype(np.complex64)


def real(seed: int, count: int, scale: float = 1.0) -> np.ndarray:
    return (inputs.uniform(seed, count) * np.float32(scale)).astype(np.float32)


def to_dev(x: np.ndarray, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def empty_dev(shape, dtype, dev):
    import torch
    t = {np.complex64: torch.complex64, np.float32: torch.float32}[dtype]
    return torch.empty(shape, dtype=t, device=dev)


def from_dev(t) -> np.ndarray:
    return t.cpu().numpy()
