"""Pins the oracle (CPU only).

1. FP64 oracle: np.fft against brute-force DFT and closed forms.
2. float32 restatement of muFFT (oracle/restate.py):
   - bit-exact against the golden vectors the compiled reference wrote
     (scalar C path) and within tolerance of its AVX-path vectors;
   - bit-exact against the compiled reference itself when oracle/_ref is
     available (wider sweep);
   - against the FP64 oracle within the reference tolerances;
   - the reference's c2r quirk (Im of DC/Nyquist bins) by closed form.
3. Planner decompositions against the reference tables (fft.c:217-435).
4. SplitMix64 known answer.
A planted mistake (a dropped twiddle, a sign, a transposed index) in the
restatement breaks the bit-exact comparisons; the closed forms and the FP64
sweep catch it independently of the reference.
"""
import math
from pathlib import Path

import numpy as np
import pytest

from oracle import fp64, inputs
from oracle import restate as R

GOLD = Path(__file__).resolve().parent / "golden" / "mufft_ref_vectors.npz"
NO_SIMD = R.FLAG_CPU_NO_SIMD


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


# --------------------------------------------------------------------------- FP64
@pytest.mark.parametrize("N", [1, 2, 3, 8, 16, 64])
@pytest.mark.parametrize("d", [-1, 1])
def test_fp64_fft_matches_direct_dft(N, d):
    x = inputs.complex_uniform(N, N).astype(np.complex128)
    np.testing.assert_allclose(fp64.c2c(x, d), fp64.dft_direct(x, d), atol=1e-12 * N)


def test_fp64_closed_forms():
    N = 64
    delta = np.zeros(N, complex)
    delta[0] = 1
    np.testing.assert_allclose(fp64.c2c(delta, -1), np.ones(N), atol=1e-14)
    np.testing.assert_allclose(fp64.c2c(np.ones(N), -1), N * delta, atol=1e-12)
    k0 = 5
    n = np.arange(N)
    e = np.exp(2j * np.pi * k0 * n / N)
    X = fp64.c2c(e, -1)
    assert abs(X[k0] - N) < 1e-10 and np.max(np.abs(np.delete(X, k0))) < 1e-10
    # inverse(forward(x)) = N x  (unnormalised, fft.h:70-71)
    x = inputs.complex_uniform(3, N)
    np.testing.assert_allclose(fp64.c2c(fp64.c2c(x, -1), 1), N * x, atol=1e-10)


def test_fp64_direct_convolution_matches_fft_convolution():
    a = inputs.uniform(1, 32).astype(np.float64)
    b = inputs.uniform(2, 32).astype(np.float64)
    np.testing.assert_allclose(fp64.conv_direct(a, b), fp64.conv_fft(a, b, 64), atol=1e-12)
    assert abs(fp64.conv_direct(a, b)[5] - sum(a[5 - i] * b[i] for i in range(6))) < 1e-12


def _c2r_brute(X, N):
    """c2r by direct summation in FP64 with the reference's resolve semantics
    (kernel.c:36-45): the Hermitian part of bins 0..N/2 plus the imaginary
    parts of bins 0 and N/2 entering as -Im(X0)*(-1)^n - Im(X_{N/2})."""
    X = np.asarray(X, np.complex128)
    n = np.arange(N)
    y = np.full(X.shape[:-1] + (N,), 0.0)
    y += X[..., :1].real + X[..., N // 2: N // 2 + 1].real * np.where(n % 2 == 0, 1.0, -1.0)
    for k in range(1, N // 2):
        y += 2.0 * np.real(X[..., k: k + 1] * np.exp(2j * np.pi * k * n / N))
    y -= X[..., :1].imag * np.where(n % 2 == 0, 1.0, -1.0) + X[..., N // 2: N // 2 + 1].imag
    return y


@pytest.mark.parametrize("N", [4, 8, 32])
def test_fp64_c2r_matches_direct_sum(N):
    X = inputs.complex_uniform(0x77 + N, 3 * (N // 2 + 1)).reshape(3, N // 2 + 1)
    np.testing.assert_allclose(fp64.c2r(X, N), _c2r_brute(X, N), atol=1e-12 * N)
    np.testing.assert_allclose(fp64.r2c(inputs.uniform(N, N)), fp64.dft_direct(inputs.uniform(N, N))[: N // 2 + 1],
                               atol=1e-12 * N)


@pytest.mark.parametrize("Nx,Ny", [(4, 8), (8, 2), (16, 4)])
@pytest.mark.parametrize("d", [-1, 1])
def test_fp64_2d_matches_separable_direct_dft(Nx, Ny, d):
    """The 2D oracle against direct sums along each axis (rows, then columns)."""
    x = inputs.complex_uniform(Nx * 16 + Ny, 2 * Nx * Ny).reshape(2, Ny, Nx)
    rows = fp64.dft_direct(x, d)
    brute = np.swapaxes(fp64.dft_direct(np.swapaxes(rows, -1, -2), d), -1, -2)
    np.testing.assert_allclose(fp64.c2c_2d(x, d), brute, atol=1e-11 * Nx * Ny)
    # and against the definition X[ky, kx] = sum x[y, x] w^(d (kx x / Nx + ky y / Ny)) at one bin
    ky, kx = 1, Nx - 1
    yy, xx = np.meshgrid(np.arange(Ny), np.arange(Nx), indexing="ij")
    direct = np.sum(x[0] * np.exp(d * 2j * np.pi * (kx * xx / Nx + ky * yy / Ny)))
    assert abs(fp64.c2c_2d(x, d)[0, ky, kx] - direct) < 1e-11 * Nx * Ny


@pytest.mark.parametrize("Nx,Ny", [(4, 8), (16, 2)])
def test_fp64_real_2d_match_direct_sums(Nx, Ny):
    x = inputs.uniform(Nx + 3 * Ny, 2 * Nx * Ny).reshape(2, Ny, Nx)
    rows = fp64.dft_direct(x, -1)
    full = np.swapaxes(fp64.dft_direct(np.swapaxes(rows, -1, -2), -1), -1, -2)
    np.testing.assert_allclose(fp64.r2c_2d(x), full[..., : Nx // 2 + 1], atol=1e-11 * Nx * Ny)
    np.testing.assert_allclose(fp64.r2c_2d(x, full=True), full, atol=1e-11 * Nx * Ny)
    # c2r_2d: inverse column sums, then the per-row c2r with the reference's DC /
    # Nyquist imaginary-part terms, on an arbitrary (non-Hermitian) spectrum
    X = inputs.complex_uniform(Nx * Ny, 2 * Ny * (Nx // 2 + 1)).reshape(2, Ny, Nx // 2 + 1)
    cols = np.swapaxes(fp64.dft_direct(np.swapaxes(X, -1, -2), 1), -1, -2)
    np.testing.assert_allclose(fp64.c2r_2d(X, Nx), _c2r_brute(cols, Nx), atol=1e-11 * Nx * Ny)
    # and the round trip on the spectrum of a real image: c2r(r2c(x)) = Nx*Ny*x
    np.testing.assert_allclose(fp64.c2r_2d(fp64.r2c_2d(x), Nx), Nx * Ny * x, atol=1e-10 * Nx * Ny)


# --------------------------------------------------------------------------- golden
@pytest.fixture(scope="module")
def gold():
    return dict(np.load(GOLD))


def _golden_cases(g, prefix):
    return sorted({k[: -len("_out")] for k in g if k.startswith(prefix) and k.endswith("_out")})


def test_golden_c2c(gold):
    for key in _golden_cases(gold, "c2c_"):
        _, isa, N, d = key.split("_")
        N, d = int(N), int(d)
        x = gold[key + "_in"]
        y = R.execute_plan_1d(R.create_plan_1d_c2c(N, d, NO_SIMD), x)
        if isa == "c":
            np.testing.assert_array_equal(bits(y), bits(gold[key + "_out"]), err_msg=key)
        else:
            assert fp64.rel_rms(y, gold[key + "_out"]) < 5e-7 * max(1, math.log2(N)), key


def test_golden_real(gold):
    for key in _golden_cases(gold, "r2c_"):
        _, isa, N, mode = key.split("_")
        N = int(N)
        fl = {"plain": 0, "full": R.FLAG_FULL_R2C, "zp": R.FLAG_ZERO_PAD_UPPER_HALF}[mode]
        y = R.execute_plan_1d(R.create_plan_1d_r2c(N, NO_SIMD | fl), gold[key + "_in"])
        ref = gold[key + "_out"]
        if isa == "c":
            np.testing.assert_array_equal(bits(y), bits(ref), err_msg=key)
        else:
            assert fp64.rel_rms(y, ref) < 5e-7 * max(1, math.log2(N)), key
    for key in _golden_cases(gold, "c2r_"):
        _, isa, N = key.split("_")
        N = int(N)
        y = R.execute_plan_1d(R.create_plan_1d_c2r(N, NO_SIMD), gold[key + "_in"])
        ref = gold[key + "_out"]
        if isa == "c":
            np.testing.assert_array_equal(bits(y), bits(ref), err_msg=key)
        else:
            assert fp64.rel_rms(y, ref) < 5e-7 * max(1, math.log2(N)), key


def test_golden_2d(gold):
    for key in _golden_cases(gold, "c2c2d_"):
        _, isa, shp, d = key.split("_")
        Nx, Ny = map(int, shp.split("x"))
        y = R.execute_plan_2d(R.create_plan_2d_c2c(Nx, Ny, int(d), NO_SIMD), gold[key + "_in"])
        ref = gold[key + "_out"]
        if isa == "c":
            np.testing.assert_array_equal(bits(y), bits(ref), err_msg=key)
        else:
            assert fp64.rel_rms(y, ref) < 1e-6, key
    for key in _golden_cases(gold, "r2c2d_"):
        _, isa, shp = key.split("_")
        Nx, Ny = map(int, shp.split("x"))
        y = R.execute_plan_2d(R.create_plan_2d_r2c(Nx, Ny, NO_SIMD), gold[key + "_in"])[:, : Nx // 2 + 1]
        ref = gold[key + "_out"]
        if isa == "c":
            np.testing.assert_array_equal(bits(y), bits(ref), err_msg=key)
        else:
            assert fp64.rel_rms(y, ref) < 1e-6, key
    for key in _golden_cases(gold, "c2r2d_"):
        _, isa, shp = key.split("_")
        Nx, Ny = map(int, shp.split("x"))
        y = R.execute_plan_2d(R.create_plan_2d_c2r(Nx, Ny, NO_SIMD), gold[key + "_in"])
        y = y.reshape(-1)[: Nx * Ny].reshape(Ny, Nx)
        ref = gold[key + "_out"]
        if isa == "c":
            np.testing.assert_array_equal(bits(y), bits(ref), err_msg=key)
        else:
            assert fp64.rel_rms(y, ref) < 1e-6, key


def test_golden_conv(gold):
    for key in _golden_cases(gold, "conv_"):
        _, isa, N, stereo = key.split("_")
        N, stereo = int(N), int(stereo)
        pc = R.create_plan_conv(N, NO_SIMD, stereo | 4 | 8)
        y = R.execute_conv_output(pc, R.execute_conv_input(pc, 0, gold[key + "_a"]),
                                  R.execute_conv_input(pc, 1, gold[key + "_b"]))
        ref = gold[key + "_out"]
        if isa == "c":
            np.testing.assert_array_equal(bits(y), bits(ref), err_msg=key)
        else:
            assert fp64.rel_rms(y, ref) < 1e-6, key


# --------------------------------------------------------------------------- live reference
@pytest.mark.parametrize("N", [2, 4, 16, 128, 2048, 16384])
def test_restatement_bit_exact_vs_reference_c2c(ref_lib, N):
    for d in (-1, 1):
        for zp in (0, R.FLAG_ZERO_PAD_UPPER_HALF):
            x = inputs.complex_uniform(N + d + zp, N // 2 if zp else N)
            y = R.execute_plan_1d(R.create_plan_1d_c2c(N, d, NO_SIMD | zp), x)
            r = ref_lib.fft_1d("c2c", N, x, d, NO_SIMD | zp)
            np.testing.assert_array_equal(bits(y), bits(r))


@pytest.mark.parametrize("N", [4, 32, 512, 8192])
def test_restatement_bit_exact_vs_reference_real(ref_lib, N):
    for fl in (0, R.FLAG_FULL_R2C, R.FLAG_ZERO_PAD_UPPER_HALF):
        x = inputs.uniform(N + fl, N // 2 if fl & R.FLAG_ZERO_PAD_UPPER_HALF else N)
        y = R.execute_plan_1d(R.create_plan_1d_r2c(N, NO_SIMD | fl), x)
        np.testing.assert_array_equal(bits(y), bits(ref_lib.fft_1d("r2c", N, x, -1, NO_SIMD | fl)))
    X = inputs.complex_uniform(N * 3, N // 2 + 1)
    y = R.execute_plan_1d(R.create_plan_1d_c2r(N, NO_SIMD), X)
    np.testing.assert_array_equal(bits(y), bits(ref_lib.fft_1d("c2r", N, X, 1, NO_SIMD)))


@pytest.mark.parametrize("Nx,Ny", [(2, 2), (4, 64), (32, 8), (64, 64)])
def test_restatement_bit_exact_vs_reference_2d(ref_lib, Nx, Ny):
    for d in (-1, 1):
        x = inputs.complex_uniform(Nx * Ny + d, Nx * Ny).reshape(Ny, Nx)
        y = R.execute_plan_2d(R.create_plan_2d_c2c(Nx, Ny, d, NO_SIMD), x)
        np.testing.assert_array_equal(bits(y), bits(ref_lib.fft_2d("c2c", Nx, Ny, x, d, NO_SIMD)))
    if Nx >= 4:
        xr = inputs.uniform(Nx + Ny, Nx * Ny).reshape(Ny, Nx)
        y = R.execute_plan_2d(R.create_plan_2d_r2c(Nx, Ny, NO_SIMD), xr)[:, : Nx // 2 + 1]
        r = ref_lib.fft_2d("r2c", Nx, Ny, xr, -1, NO_SIMD)[:, : Nx // 2 + 1]
        np.testing.assert_array_equal(bits(y), bits(r))


def test_restatement_vs_reference_avx_path(ref_lib):
    for N in (8, 256, 4096, 65536):
        x = inputs.complex_uniform(N, N)
        for d in (-1, 1):
            y = R.execute_plan_1d(R.create_plan_1d_c2c(N, d), x)
            assert fp64.rel_rms(y, ref_lib.fft_1d("c2c", N, x, d, 0)) < 5e-7 * math.log2(N)


# --------------------------------------------------------------------------- vs FP64
@pytest.mark.parametrize("N", [2, 8, 64, 1024, 65536])
def test_restatement_vs_fp64(N):
    for d in (-1, 1):
        x = inputs.complex_uniform(7 * N, N) * np.float32(0.5)
        y = R.execute_plan_1d(R.create_plan_1d_c2c(N, d), x)
        assert np.max(np.abs(y - fp64.c2c(x, d))) < 1e-6 * math.sqrt(N)  # test.c:58
    if N >= 4:
        xr = inputs.uniform(9 * N, N) * np.float32(0.5)
        Y = R.execute_plan_1d(R.create_plan_1d_r2c(N), xr)
        assert np.max(np.abs(Y - fp64.r2c(xr))) < 1e-6 * math.sqrt(N)
        X = inputs.complex_uniform(11 * N, N // 2 + 1) * np.float32(0.5)
        z = R.execute_plan_1d(R.create_plan_1d_c2r(N), X)
        assert np.max(np.abs(z - fp64.c2r(X, N))) < 2e-6 * math.sqrt(N)


def test_c2r_quirk_closed_form():
    """Im of the DC and Nyquist bins reaches the output (kernel.c:36-45):
    c2r(i*e_0)[n] = -(-1)^n and c2r(i*e_{N/2})[n] = -1."""
    N = 32
    n = np.arange(N)
    X = np.zeros(N // 2 + 1, np.complex64)
    X[0] = 1j
    np.testing.assert_allclose(R.execute_plan_1d(R.create_plan_1d_c2r(N), X), -((-1.0) ** n), atol=1e-6)
    X = np.zeros(N // 2 + 1, np.complex64)
    X[N // 2] = 1j
    np.testing.assert_allclose(R.execute_plan_1d(R.create_plan_1d_c2r(N), X), -np.ones(N), atol=1e-6)
    np.testing.assert_allclose(fp64.c2r(X, N), -np.ones(N), atol=1e-12)


def test_restatement_conv_vs_direct():
    for N in (4, 64, 2048):
        a = inputs.uniform(N, N // 2) * np.float32(0.5)
        b = inputs.uniform(N + 1, N // 2) * np.float32(0.5)
        pc = R.create_plan_conv(N, 0, 0 | 4 | 8)
        y = R.execute_conv_output(pc, R.execute_conv_input(pc, 0, a), R.execute_conv_input(pc, 1, b))
        assert np.max(np.abs(y - fp64.conv_direct(a, b)[:N])) < 2e-6 * math.sqrt(N)  # test.c:431
        ac = inputs.complex_uniform(N + 2, N // 2) * np.float32(0.5)
        pc = R.create_plan_conv(N, 0, 1 | 4 | 8)
        y = R.execute_conv_output(pc, R.execute_conv_input(pc, 0, ac), R.execute_conv_input(pc, 1, b))
        assert np.max(np.abs(y - fp64.conv_direct(ac, b)[:N])) < 2e-6 * math.sqrt(N)


# --------------------------------------------------------------------------- planner
AVX = 7  # mufft_get_cpu_flags() on an AVX host (cpu.c:78-114)


def _short(steps):
    return [(s[0], s[1], s[2]) for s in steps]


def test_planner_decompositions_match_reference_tables():
    """Decompositions the reference picks on an AVX host (fft.c:217-287, 325-435)."""
    d = R.decomposition
    assert [(n, r, p) for n, r, p, _ in d(1024, -1, 0, AVX)] == [
        ("forward_radix8_p1", 8, 1), ("radix8_generic", 8, 8), ("radix8_generic", 8, 64),
        ("radix2_generic", 2, 512)]
    assert [(r, p) for _, r, p, _ in d(4096, -1, 0, AVX)] == [(8, 1), (8, 8), (8, 64), (8, 512)]
    # inverse plans start with direction-agnostic radix-2 (inverse radix-8/4 p1 never chosen)
    assert [(r, p) for _, r, p, _ in d(4096, 1, 0, AVX)] == [
        (2, 1), (2, 2), (4, 4), (8, 16), (8, 128), (4, 1024)]
    assert [n for n, *_ in d(2, -1, 0, AVX)] == ["radix2_p1"]
    assert [(r, p) for _, r, p, _ in d(8, -1, 0, AVX)] == [(2, 1), (2, 2), (2, 4)]
    assert d(65536, -1, R.FLAG_ZERO_PAD_UPPER_HALF, AVX)[0][0] == "forward_half_radix8_p1"
    assert R.create_plan_1d_c2c(12, -1) is None
    assert R.create_plan_1d_c2c(1, -1) is None
    assert R.create_plan_1d_r2c(2) is None  # fft.c:485 -> c2c(1) -> NULL


def test_twiddle_table_layout():
    """fft.c:114-134 and add_step offsets (fft.c:290-322): step at p reads T + off(p)."""
    for N, d in ((16, -1), (64, 1)):
        T = R.build_twiddles(N, d)
        assert T[0] == 1 and T[1] == 1
        assert abs(T[2] - np.exp(1j * np.pi * d / 2)) < 1e-7
        p = 4
        while p < N:
            k = np.arange(p)
            np.testing.assert_allclose(T[p:2 * p], np.exp(1j * np.pi * d * k / p), atol=1e-7)
            p *= 2
    for s in R.build_plan_1d(4096, -1, 0):
        assert s.twiddle_offset == (0 if s.p == 1 else (1 if s.p == 2 else s.p))


def test_splitmix64_known_answer():
    # SplitMix64 from state 0: first output 0xE220A8397B1DCDAF (Steele et al. 2014).
    assert int(inputs.splitmix64(0, 1)[0]) == 0xE220A8397B1DCDAF
    u = inputs.uniform(5, 100000)
    assert u.min() >= -1.0 and u.max() < 1.0 and abs(float(u.mean())) < 0.02
    np.testing.assert_array_equal(inputs.uniform(5, 10, offset=3), u[3:13])


def test_reference_baseline_driver(ref_lib):
    """oracle/ref_bench.c runs the benchmark workload on native threads: it
    reports a positive time for a valid run and -1 for invalid arguments or
    sizes the reference refuses (plan creation returns NULL)."""
    L = ref_lib.lib()
    xc = np.ascontiguousarray(inputs.complex_uniform(0x2001, 8 * 4096), dtype=np.complex64)
    xr = np.ascontiguousarray(inputs.uniform(0x2002, 2 * 65536), dtype=np.float32)
    t = L.mufft_ref_bench_run(4, 2, 4096, xc.ctypes.data, 8, 65536, xr.ctypes.data, 2)
    assert t > 0
    assert L.mufft_ref_bench_run(0, 1, 4096, xc.ctypes.data, 8, 65536, xr.ctypes.data, 2) == -1.0
    assert L.mufft_ref_bench_run(2, 1, 1000, xc.ctypes.data, 1, 65536, xr.ctypes.data, 1) == -1.0
