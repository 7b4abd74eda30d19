"""Writes tests/golden/mufft_ref_vectors.npz from the compiled reference library.

Inputs are SplitMix64 streams (oracle/inputs.py); outputs come from the
reference muFFT built from its own sources (oracle/Makefile -> oracle/_ref),
on its AVX path (flags = 0, what the reference runs on this host) and on its
scalar C path (flags = MUFFT_FLAG_CPU_NO_SIMD).  Re-run with
    make -C oracle && python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import inputs, ref  # noqa: E402

NO_SIMD = (1 << 16) - 1
FULL = 1 << 16
ZP = 1 << 17
OUT = Path(__file__).resolve().parent / "mufft_ref_vectors.npz"


def main():
    ref.build()
    vec = {}
    seed = 0x6000
    for isa, fl in (("avx", 0), ("c", NO_SIMD)):
        for N in (2, 8, 64, 1024):
            for d in (-1, 1):
                seed += 1
                x = inputs.complex_uniform(seed, N) * np.float32(0.5)
                vec[f"c2c_{isa}_{N}_{d}_in"] = x
                vec[f"c2c_{isa}_{N}_{d}_out"] = ref.fft_1d("c2c", N, x, d, fl)
        for N in (4, 16, 256, 4096):
            for mode, mf in (("plain", 0), ("full", FULL), ("zp", ZP)):
                seed += 1
                n_in = N // 2 if mode == "zp" else N
                x = inputs.uniform(seed, n_in) * np.float32(0.5)
                vec[f"r2c_{isa}_{N}_{mode}_in"] = x
                vec[f"r2c_{isa}_{N}_{mode}_out"] = ref.fft_1d("r2c", N, x, -1, fl | mf)
            seed += 1
            X = inputs.complex_uniform(seed, N // 2 + 1) * np.float32(0.5)
            vec[f"c2r_{isa}_{N}_in"] = X
            vec[f"c2r_{isa}_{N}_out"] = ref.fft_1d("c2r", N, X, 1, fl)
        for Nx, Ny in ((8, 4), (16, 32)):
            for d in (-1, 1):
                seed += 1
                x = (inputs.complex_uniform(seed, Nx * Ny) * np.float32(0.5)).reshape(Ny, Nx)
                vec[f"c2c2d_{isa}_{Nx}x{Ny}_{d}_in"] = x
                vec[f"c2c2d_{isa}_{Nx}x{Ny}_{d}_out"] = ref.fft_2d("c2c", Nx, Ny, x, d, fl)
            seed += 1
            xr = (inputs.uniform(seed, Nx * Ny) * np.float32(0.5)).reshape(Ny, Nx)
            vec[f"r2c2d_{isa}_{Nx}x{Ny}_in"] = xr
            vec[f"r2c2d_{isa}_{Nx}x{Ny}_out"] = ref.fft_2d("r2c", Nx, Ny, xr, -1, fl)[:, : Nx // 2 + 1]
            seed += 1
            Xc = (inputs.complex_uniform(seed, Nx * Ny) * np.float32(0.5)).reshape(Ny, Nx)
            vec[f"c2r2d_{isa}_{Nx}x{Ny}_in"] = Xc
            vec[f"c2r2d_{isa}_{Nx}x{Ny}_out"] = ref.fft_2d("c2r", Nx, Ny, Xc, 1, fl)
        for N in (64, 1024):
            for stereo in (0, 1):
                seed += 1
                m = stereo | 4 | 8
                a = (inputs.complex_uniform(seed, N // 2) if stereo else inputs.uniform(seed, N // 2))
                a = a * np.float32(0.5)
                h = inputs.uniform(seed ^ 0x55, N // 2) * np.float32(0.5)
                vec[f"conv_{isa}_{N}_{stereo}_a"] = a
                vec[f"conv_{isa}_{N}_{stereo}_b"] = h
                vec[f"conv_{isa}_{N}_{stereo}_out"] = ref.conv(N, a, h, m, fl)
    np.savez_compressed(OUT, **vec)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes, {len(vec)} arrays)")


if __name__ == "__main__":
    main()
