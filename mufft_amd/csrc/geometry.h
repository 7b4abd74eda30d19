// Shared host/device geometry and constants for the muFFT gfx950 kernels.
//
// Both the kernel templates and the host planner derive launch shapes, radix
// splits, LDS layouts and twiddle-table layouts from the constexpr functions
// below, so the two sides cannot disagree.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define MUFFT_HD __host__ __device__
#else
#define MUFFT_HD
#endif

namespace mufft {

enum Kind : int { KIND_C2C = 0, KIND_R2C = 1, KIND_C2R = 2 };

// Runtime flags of the fused kernels.
enum RowFlag : unsigned {
    RF_ZERO_PAD = 1u << 0,   // upper half of the (complex) input is zero, not read
    RF_FULL = 1u << 1,       // r2c writes all N bins (MUFFT_FLAG_FULL_R2C)
    RF_MUL = 1u << 2,        // input = norm * in * in2 (fused convolution multiply)
    RF_CONJ2 = 1u << 3,      // with RF_MUL: use conj(in2) (correlation / adjoint)
    // Adjoint weightings (gradient convention dL/dRe + i dL/dIm):
    RF_ADJ_R2C = 1u << 4,    // c2r prologue: interior bins x 1/2, Im of DC and Nyquist dropped
    RF_ADJ_C2R = 1u << 5,    // r2c epilogue: interior bins x 2, DC/Nyquist cross terms
    // The next kernel reads the output: store it cacheable (default: streaming
    // stores, for outputs that leave the device's caches anyway).
    RF_KEEP_OUT = 1u << 6,
};
// Every flag: a launch with none of them runs the kernels' plain
// instantiation, where the flags are compile-time zero.
constexpr unsigned RF_ALL = (1u << 7) - 1u;
// Template argument of the fused transforms: flags read at run time (any
// other value fixes them at compile time).
constexpr unsigned RF_DYNAMIC = ~0u;

MUFFT_HD constexpr int ilog2(long long n) { return n <= 1 ? 0 : 1 + ilog2(n >> 1); }

// ---------------------------------------------------------------------------
// Variants of the fused row kernel, one int of bit fields:
//   bits 0-7   E, points per thread (clamped to the transform length);
//   bits 8-11  LDS exchange mode (XMode in stockham.h: 0 padded complex,
//              1 split real/imaginary, 2 unpadded complex);
//   bits 12-15 twiddle mode (TwMode: 0 every twiddle from the table, 1 split
//              product);
//   bits 16-19 log2 N1: one transform split over N1 workgroups
//              (split_impl.h); E, T, the pass count and the LDS size then
//              refer to the N/N1-point sub-transform;
//   bits 20-23 workgroups per CU to budget registers for (0: no request);
//   bits 24-27 log2 of the radix of every pass but the remainder pass (0:
//              radix E); a radix below E gives each thread E/radix
//              butterflies per pass;
//   bit  28    c2r: the prologue reads its bins in chunks of 4 points per
//              thread, each chunk issued after the previous one's data has
//              arrived, in the order 0-3, E-4..E-1, 4-7, E-8..E-5, ...  Every
//              bin is read twice (as X[k] by one thread, as X[N - k] by its
//              mirror thread).  Issued all at once, both reads miss the L2
//              together and each costs an HBM read; chunk pairs make the
//              second read follow the first by one load round, so it hits
//              the line the first brought in (a whole transform's input, 256
//              KB per CU at 32768 points, does not stay in a 4 MB L2 shared
//              by 32 CUs for much longer).
//   bit  29    c2r with T >= 64 threads per transform: every bin is read
//              once.  Threads are numbered by mirror_lt (fft_device.h) so
//              the owners of bins k and N - k are lanes i and i ^ 32 of one
//              wave, which exchange the loaded bins in registers; the first
//              Stockham pass runs under that numbering (bit 28 is then
//              ignored).
// The planner's static choice is rv_default(); the autotuner may pick any
// registered variant.
// ---------------------------------------------------------------------------
MUFFT_HD constexpr int rv_make(int E, int xm, int twm) { return E | (xm << 8) | (twm << 12); }
MUFFT_HD constexpr int rv_make_split(int log2n1, int var) { return var | (log2n1 << 16); }
MUFFT_HD constexpr int rv_with_occ(int var, int wgs_per_cu) { return var | (wgs_per_cu << 20); }
MUFFT_HD constexpr int rv_with_radix(int var, int log2r) { return var | (log2r << 24); }
MUFFT_HD constexpr int rv_with_chunked_loads(int var) { return var | (1 << 28); }
MUFFT_HD constexpr bool rv_chunked_loads(int var) { return ((var >> 28) & 1) != 0; }
MUFFT_HD constexpr int rv_with_mirror_lanes(int var) { return var | (1 << 29); }
MUFFT_HD constexpr bool rv_mirror_lanes(int var) { return ((var >> 29) & 1) != 0; }
// Bits 28-29 change only the c2r kernels.
MUFFT_HD constexpr bool rv_c2r_only(int var) { return rv_chunked_loads(var) || rv_mirror_lanes(var); }
MUFFT_HD constexpr int rv_split(int var) { return (var >> 16) & 0xf; }   // log2 N1
MUFFT_HD constexpr int rv_occ(int var) { return (var >> 20) & 0xf; }
MUFFT_HD constexpr int rv_sub(int var) { return var & ~(0xf << 16); }    // variant of the sub-transform
MUFFT_HD constexpr int rv_n(int N, int var) { return N >> rv_split(var); }
MUFFT_HD constexpr int rv_E(int N, int var) { return (var & 0xff) >= rv_n(N, var) ? rv_n(N, var) : (var & 0xff); }
MUFFT_HD constexpr int rv_xm(int var) { return (var >> 8) & 0xf; }
MUFFT_HD constexpr int rv_twm(int var) { return (var >> 12) & 0xf; }
MUFFT_HD constexpr int rv_T(int N, int var) { return rv_n(N, var) / rv_E(N, var); }
MUFFT_HD constexpr int rv_B(int N, int var) { return rv_T(N, var) >= 256 ? 1 : 256 / rv_T(N, var); }
MUFFT_HD constexpr int rv_WG(int N, int var) { return rv_T(N, var) * rv_B(N, var); }
// Minimum waves per SIMD for the register budget (4 SIMDs per CU, 64-wide waves).
MUFFT_HD constexpr int rv_waves_per_simd(int N, int var) {
    return rv_occ(var) == 0 ? 1 : (rv_occ(var) * rv_WG(N, var) / 256 > 1 ? rv_occ(var) * rv_WG(N, var) / 256 : 1);
}
MUFFT_HD constexpr int rv_R(int N, int var) {
    return ((var >> 24) & 0xf) == 0 ? rv_E(N, var)
         : ((1 << ((var >> 24) & 0xf)) < rv_E(N, var) ? (1 << ((var >> 24) & 0xf)) : rv_E(N, var));
}
MUFFT_HD constexpr int rv_NP(int N, int var) {
    return rv_R(N, var) >= rv_n(N, var)
               ? 1
               : (ilog2(rv_n(N, var)) + ilog2(rv_R(N, var)) - 1) / ilog2(rv_R(N, var));
}
// LDS slots per transform (points for complex modes, floats for the split mode).
MUFFT_HD constexpr int rv_lds_per(int N, int var) {
    return rv_xm(var) == 2 ? rv_n(N, var) : rv_n(N, var) + rv_n(N, var) / rv_E(N, var);
}
// Static choice per size and transform kind (0/1 c2c, 2 r2c, 3 c2r; the real
// transforms' fused size N is half the real length), from sweeps of all
// registered variants on MI355X, cold caches, at 2^25-2^27 points per launch:
// split-product twiddles pay off from 512 points on; c2r below 1024 points
// runs 8 points per thread (-10% to -27%), from 1024 to 8192 points 16 with
// mirrored lanes (each bin read once, -5% to -10%); c2c and r2c run 16 points
// per thread up to 8192 (c2c 1024 and 8192: -7%, -3% against 32).  At 16384
// all kinds run 32 points per thread with the split exchange and registers
// budgeted for two workgroups per CU (c2c -20%, r2c -7%; c2r with chunked
// loads, -15%).  At 32768 c2c runs one workgroup of 32 points per thread
// per transform, c2r too with mirrored lanes (65.6 against 74.2 us for the
// chunked loads at 512 transforms, 509 against 580 at 4096); r2c splits each
// transform over two workgroups budgeted to share a CU (60 against 75 us at
// 512 transforms, 523 against 531 at 4096; the split reads its input twice,
// from the L2 the second time).  65536 runs on two workgroups per transform.
MUFFT_HD constexpr int rv_default_base(int N, int kind) {
    return N < 512 ? (kind == 3 && N == 256 ? rv_make(8, 0, 0) : rv_make(16, 0, 0))
         : N == 512 ? (kind == 3 ? rv_make(8, 0, 0) : rv_make(16, 0, 1))
         : N < 16384 ? (kind == 3 ? rv_with_mirror_lanes(rv_make(16, 0, 1)) : rv_make(16, 0, 1))
         : N == 16384 ? (kind == 3 ? rv_with_chunked_loads(rv_with_occ(rv_make(32, 1, 1), 2))
                                   : rv_with_occ(rv_make(32, 1, 1), 2))
         : N == 32768 ? (kind == 2 ? rv_with_occ(rv_make_split(1, rv_make(32, 1, 1)), 2)
                                   : kind == 3 ? rv_with_mirror_lanes(rv_make(32, 1, 1)) : rv_make(32, 1, 1))
         : rv_make_split(1, rv_make(32, 1, 1));
}
MUFFT_HD constexpr int rv_default(int N, int kind = 0) { return rv_default_base(N, kind); }

// Radix sequence of NP passes: NP - 1 passes of radix R and one of the
// remainder N / R^(NP-1), placed first (REM_FIRST) or last.  The twiddle
// reads of a pass cover (r - 1) * p table entries for stride p, so with the
// remainder first the row kernels keep the later passes' table regions small
// enough for the L1 cache (8192 points: passes 8, 32, 32 read 20 twiddles
// per thread from 20 KB instead of 32, 32, 8 reading 38 from 56 KB; 4-6%
// faster on MI355X, 2048 and 4096 real 3-5%).  The column kernels, whose
// lanes of one column thread share each twiddle, keep the remainder last:
// the other order made 1024-point columns 10-13% slower.
MUFFT_HD constexpr int seq_rem(int N, int R, int NP) { return N >> (ilog2(R) * (NP - 1)); }
MUFFT_HD constexpr int seq_radix(int N, int R, int NP, int s, bool rem_first) {
    return rem_first ? (s == 0 ? seq_rem(N, R, NP) : R) : (s < NP - 1 ? R : seq_rem(N, R, NP));
}
MUFFT_HD constexpr int seq_p(int N, int R, int NP, int s, bool rem_first) {
    return rem_first ? (s == 0 ? 1 : seq_rem(N, R, NP) << (ilog2(R) * (s - 1))) : (1 << (ilog2(R) * s));
}
// Radix of the first pass (the twiddle table holds N minus it entries).
MUFFT_HD constexpr int seq_first(int N, int R, int NP, bool rem_first) {
    return seq_radix(N, R, NP, 0, rem_first);
}

// ---------------------------------------------------------------------------
// Fused column (strided-axis) transform of length L over a tile of TX
// adjacent columns.  Lanes run along the columns (unit stride) so every row
// segment of a tile is one coalesced access.
// ---------------------------------------------------------------------------
// Points per thread: all of a tiny column, 16 up to 4096 points, 32 beyond.
MUFFT_HD constexpr int col_E(int L) { return L <= 16 ? L : (L <= 4096 ? 16 : 32); }
MUFFT_HD constexpr int col_T(int L) { return L / col_E(L); }
// Tile width of the four-step column kernel (whose transposed store needs the
// full complex tile in LDS).  Two columns from 4096 points on, so that two
// tiles share a CU: 2^22 x 4 -7%, the same time at 2^22 x 16.
MUFFT_HD constexpr int fs_TX(int L) {
    return col_T(L) <= 32 ? 256 / col_T(L) : (col_T(L) <= 128 ? 8 : (L >= 4096 ? 2 : 4));
}
MUFFT_HD constexpr int fs_WG(int L) { return col_T(L) * fs_TX(L); }
// The plain column kernel, from sweeps on MI355X.  From L = 1024 on it uses
// the split LDS exchange and split twiddles (half the LDS of a tile).  At
// L = 1024 a 16-column tile (one 128-byte line per row) is 15% faster than 8
// columns; at L = 2048 a 4-column tile (more, smaller workgroups per CU) is
// 12% faster than 8; at L = 4096 the split exchange lets two 4-column tiles
// share a CU (2D 4096^2: -5%); at L = 8192 it allows 4 columns instead of 2
// (2D 8192^2: -6%, four-step 2^26: -19%).
MUFFT_HD constexpr int col_TX(int L) { return L == 1024 ? 16 : (L == 2048 || L == 4096 || L == 8192 ? 4 : fs_TX(L)); }
// The one alternative tile width compiled per column length (0: none): twice
// the default below 1024 points, half or twice it from 1024 on where the LDS
// and the 1024-thread workgroup allow.  The tile width changes only which
// columns a workgroup owns, never the arithmetic, so every choice gives the
// same bits; the 2D autotuner times both on the caller's image shape.
MUFFT_HD constexpr int col_TX_alt(int L) {
    return L < 128 ? 0 : (L <= 512 ? 2 * col_TX(L) : (L == 1024 ? 8 : (L == 2048 ? 8 : 2)));
}
MUFFT_HD constexpr int col_xm(int L) { return L >= 1024 ? 1 : 0; }    // XMode
MUFFT_HD constexpr int col_twm(int L) { return L >= 1024 ? 1 : 0; }   // TwMode
MUFFT_HD constexpr int col_WG(int L) { return col_T(L) * col_TX(L); }
MUFFT_HD constexpr int col_lds_rows(int L) { return L + L / col_E(L); }
constexpr int COL_MAX_LOG2 = 13;   // fused column kernel covers L <= 8192
// Four-step split N = N1 * N2 for 2^17 <= N <= 2^26: log2 N1, the length of
// the first launch's columns (the second launch's are N2 = N / N1).  From a
// sweep of every split on MI355X at 512 MiB per launch; against the balanced
// split it saves 14% at 2^21, 21% at 2^22 and 23% at 2^25, where the chosen
// N2 is a 1024- or 4096-point column that tiles well.
// Three launches for the largest sizes, N = N1 * A * B: the four-step column
// kernel on N1-point columns, then A-point and B-point column passes (at most
// 1024 points, tiles of 64-128-byte row segments) over N2 = A * B.  Returns
// log2 N1 << 8 | log2 A << 4 | log2 B, or 0 where two launches are used.
// From a sweep of every split on MI355X at 512 MiB per launch, with the
// per-column twiddle steps: against the best two-launch split, -6% at 2^23,
// -8% at 2^24, -16% at 2^25, -25% at 2^26 (2^22 and below: no gain).
MUFFT_HD constexpr int fs3_split(int logn) {
    return logn == 23 ? (9 << 8) | (8 << 4) | 6
         : logn == 24 ? (10 << 8) | (8 << 4) | 6
         : logn == 25 ? (9 << 8) | (8 << 4) | 8
         : logn == 26 ? (10 << 8) | (10 << 4) | 6
         : 0;
}
// Split codes of a large transform (the planner's static choice below, or
// one the autotuner measured): log2 N1 << 8 | log2 A << 4 | log2 B, with
// A = B = 1 (code low bits 0) for two launches.
MUFFT_HD constexpr int fs_code_n1(int code) { return code >> 8; }
MUFFT_HD constexpr int fs_code_a(int code) { return (code >> 4) & 15; }
MUFFT_HD constexpr int fs_code_b(int code) { return code & 15; }
// Whether a split code is runnable for 2^logn points: the first launch's
// N1-point columns (64..8192) come in whole tiles of the four-step column
// kernel; two launches need N2 <= 8192 points per column, three launches
// A and B of 64..1024 points.
MUFFT_HD constexpr bool fs_code_valid(int logn, int code) {
    return fs_code_n1(code) >= 6 && fs_code_n1(code) <= 13 && fs_code_n1(code) < logn &&
           ((1LL << (logn - fs_code_n1(code))) % fs_TX(1 << fs_code_n1(code))) == 0 &&
           (((code & 0xff) == 0 && logn - fs_code_n1(code) >= 1 && logn - fs_code_n1(code) <= 13) ||
            ((code & 0xff) != 0 && fs_code_a(code) >= 6 && fs_code_a(code) <= 10 && fs_code_b(code) >= 6 &&
             fs_code_b(code) <= 10 && fs_code_n1(code) + fs_code_a(code) + fs_code_b(code) == logn));
}
MUFFT_HD constexpr int fs_log_n1(int logn) {
    return logn == 17 ? 9 : logn == 18 ? 8 : logn == 21 ? 11 : logn == 22 ? 12 : logn == 23 ? 11
         : logn == 25 ? 13 : logn / 2;
}
MUFFT_HD constexpr int fs_code_default(int logn) {
    return fs3_split(logn) ? fs3_split(logn) : (fs_log_n1(logn) << 8);
}
constexpr int ROW_MAX_LOG2 = 16;   // fused row kernels (one or two workgroups per transform) cover N <= 65536
constexpr int ROW_WG_MAX_LOG2 = 15;  // one workgroup per transform up to 32768 points

// Twiddle table layout for a Stockham radix sequence r_0, r_1, ...: pass s >= 1
// (stride p_s) owns (r_s - 1) * p_s entries starting at p_s - r_0; entry
// (m - 1) * p_s + k holds exp(dir * 2 pi i * k * m / (r_s * p_s)).  The table
// has N - r_0 entries in total.
MUFFT_HD constexpr int tw_offset(int r0, int p) { return p - r0; }

// cos and sin of 2*pi*j/r for any r > 0, in double, at compile time: the
// quadrant is split off in integers and the remaining angle, in [0, pi/2),
// goes through a Taylor series whose truncation error is below 1e-30.
MUFFT_HD constexpr double taylor_sin(double x) {
    double t = x, s = x;
    for (int n = 1; n < 20; ++n) {
        t *= -x * x / (double)((2 * n) * (2 * n + 1));
        s += t;
    }
    return s;
}
MUFFT_HD constexpr double taylor_cos(double x) {
    double t = 1.0, s = 1.0;
    for (int n = 1; n < 20; ++n) {
        t *= -x * x / (double)((2 * n - 1) * (2 * n));
        s += t;
    }
    return s;
}
MUFFT_HD constexpr double turn_part(long long j, long long r, bool want_sin) {
    const long long jj = ((j % r) + r) % r;
    const long long q = (4 * jj) / r;            // quadrant
    const long long rem = 4 * jj - q * r;        // angle within it, in units of (pi/2)/r
    const double phi = 1.5707963267948966 * (double)rem / (double)r;
    const double c = taylor_cos(phi), s = taylor_sin(phi);
    if (want_sin) return q == 0 ? s : (q == 1 ? c : (q == 2 ? -s : -c));
    return q == 0 ? c : (q == 1 ? -s : (q == 2 ? -c : s));
}
MUFFT_HD constexpr double cos_turn(long long j, long long r) { return turn_part(j, r, false); }
MUFFT_HD constexpr double sin_turn(long long j, long long r) { return turn_part(j, r, true); }

// cos(2 pi j / 64) and sin(2 pi j / 64): constant twiddles of the in-register
// radix-2..64 butterflies (exact at the quarter turns).
MUFFT_HD constexpr double cos64(int j) { return cos_turn(j, 64); }
MUFFT_HD constexpr double sin64(int j) { return sin_turn(j, 64); }
static_assert(cos64(16) == 0.0 && sin64(16) == 1.0 && cos64(32) == -1.0 && sin64(48) == -1.0,
              "quarter turns are exact");
static_assert(cos64(8) - 0.70710678118654752 < 4e-16 && 0.70710678118654752 - cos64(8) < 4e-16 &&
                  sin64(4) - 0.38268343236508977 < 4e-16 && 0.38268343236508977 - sin64(4) < 4e-16,
              "eighth and sixteenth turns within a few ulp");

}  // namespace mufft
