// Executor: the launch sequences of one transform axis (fused row kernel,
// fused column kernel, four-step column launches, generic passes).
#include "host.h"

namespace mufft {

namespace {

// Layout of one operand of a generic pass: transform (img, x) element i at
// base + img*bs + x + i*sy.
struct Lay {
    float2 *p;
    long long bs, sy;
};

// Runs the generic pass list from src; the last pass writes `fin`, earlier
// passes alternate between `fin` and `alt` (src must differ from both).
int run_generic(const Engine &e, const float2 *src, Lay s_lay, Lay fin, Lay alt, int ncols,
                int batch, bool zp, hipStream_t st) {
    const int P = (int)e.gen_radix.size();
    long long p = 1;
    const float2 *cur = src;
    long long cur_bs = s_lay.bs, cur_sy = s_lay.sy;
    for (int s = 0; s < P; ++s) {
        const Lay &d = ((P - 1 - s) % 2 == 0) ? fin : alt;
        PassArgs a{};
        a.in = cur;
        a.out = d.p;
        a.tw = (s == 0) ? nullptr : e.tw_gen + (p - e.gen_radix[0]);
        a.n = e.n;
        a.p = p;
        a.sy_in = cur_sy;
        a.sy_out = d.sy;
        a.bs_in = cur_bs;
        a.bs_out = d.bs;
        a.ncols = ncols;
        a.batch = batch;
        a.flags = (s == 0 && zp) ? RF_ZERO_PAD : 0u;
        HIP_TRY(launch_pass(e.gen_radix[s], e.direction, a, st));
        cur = d.p;
        cur_bs = d.bs;
        cur_sy = d.sy;
        p *= e.gen_radix[s];
    }
    return 0;
}

// c2c transforms beyond the fused row sizes, packed scratch `alt` of
// batch * n points: four-step (two fused column launches, three for the
// sizes of fs3_split) up to 2^(2 * COL_MAX_LOG2) points, generic passes
// beyond.  src, dst and alt must be distinct.
int run_large(const Engine &e, const float2 *src, long long sd, float2 *dst, long long dd,
              float2 *alt, int batch, bool zp, hipStream_t st) {
    if (!e.four_step())
        return run_generic(e, src, Lay{const_cast<float2 *>(src), sd, 1}, Lay{dst, dd, 1},
                           Lay{alt, e.n, 1}, 1, batch, zp, st);
    const int n1 = 1 << e.log_n1, n2 = 1 << e.log_n2;
    FourStepArgs f{};
    f.in = src;
    f.out = alt;
    f.tw = e.tw_fs1;
    f.coarse = e.tw_coarse;
    f.fine = e.tw_fine;
    f.in_bs = sd;
    f.out_bs = e.n;
    f.n2 = n2;
    f.ntiles = n2 / fs_TX(n1);
    f.batch = batch;
    f.log2s = e.log_s;
    f.flags = zp ? RF_ZERO_PAD : 0u;
    HIP_TRY(launch_four_step_cols(e.log_n1, e.direction, f, st));
    if (e.three_pass()) {
        // alt holds y[n2][k1], n2 = B*a + b.  Middle pass, in place: for each
        // (b, k1) the A-point transform over a (rows of B*N1 points), output
        // k_a times w_N2^(b*k_a).
        const int na = 1 << e.log_a, nb = 1 << e.log_b;
        ColArgs c{};
        c.in = alt;
        c.out = alt;
        c.tw = e.tw_pa;
        c.in_rs = c.out_rs = (long long)nb * n1;
        c.in_bs = c.out_bs = e.n;
        c.ncols = nb * n1;
        c.ntiles = (c.ncols + col_TX(na) - 1) / col_TX(na);
        c.batch = batch;
        c.coarse = e.tw_coarse2;
        c.fine = e.tw_fine2;
        c.log2s = e.log_s2;
        c.log2div = e.log_n1;
        HIP_TRY(launch_cols(e.log_a, e.direction, c, st, COL_TWIDDLE));
        // Last pass: for each (k_a, k1) the B-point transform over b; output
        // k_b is X[k1 + N1*(k_a + A*k_b)].  Image g = (transform, k_a).
        ColArgs d{};
        d.in = alt;
        d.out = dst;
        d.tw = e.tw_pb;
        d.in_rs = n1;
        d.out_rs = (long long)na * n1;
        d.in_bs = (long long)nb * n1;
        d.out_bs = n1;
        d.img_div = na;
        d.in_bs2 = e.n;
        d.out_bs2 = dd;
        d.ncols = n1;
        d.ntiles = (n1 + col_TX(nb) - 1) / col_TX(nb);
        d.batch = batch * na;
        HIP_TRY(launch_cols(e.log_b, e.direction, d, st, COL_TWO_LEVEL));
        return 0;
    }
    // Second half: the N2-point column transforms of the n2 x n1 matrix.
    ColArgs c{};
    c.in = alt;
    c.out = dst;
    c.tw = e.tw_fs2;
    c.in_rs = n1;
    c.out_rs = n1;
    c.in_bs = e.n;
    c.out_bs = dd;
    c.ncols = n1;
    c.ntiles = (n1 + col_TX(n2) - 1) / col_TX(n2);
    c.batch = batch;
    // Columns of 64-1024 points (tile rows of whole cache lines) stream:
    // -3% to -7% at 2^17-2^22 points on MI355X.  With narrower tile rows
    // (2048 points and up) neighbouring tiles share lines and streaming
    // accesses double the time.
    HIP_TRY(launch_cols(e.log_n2, e.direction, c, st, e.log_n2 >= 6 && e.log_n2 <= 10 ? COL_STREAM : COL_PLAIN));
    return 0;
}

}  // namespace

// Contiguous transforms of kind `kind` (rows / 1D): transform g reads
// in + g*idist (and in2 + g*idist2 with RF_MUL) and writes out + g*odist.
// Distances in complex elements.  s0/s1 are grown as scratch when needed.
int run_rows(const Engine &e, int kind, const float2 *rtw, const float2 *in, long long idist,
             const float2 *in2, long long idist2, float2 *out, long long odist, int batch,
             unsigned flags, float norm, DevBuf &s0, DevBuf &s1, hipStream_t st) {
    if (batch <= 0) return 0;
    if (e.row_fused()) {
        RowArgs a{};
        a.in = in;
        a.in2 = in2;
        a.out = out;
        a.tw = e.tw_row;
        a.rtw = rtw;
        a.stw = e.tw_split;
        a.idist = idist;
        a.idist2 = idist2;
        a.odist = odist;
        a.batch = batch;
        a.flags = flags;
        a.norm = norm;
        HIP_TRY(launch_rows(e.logn, e.direction, kind, e.row_var, a, st));
        return 0;
    }
    // Generic path: packed scratch of batch*n points each.
    const long long n = e.n;
    const size_t bytes = sizeof(float2) * (size_t)n * (size_t)batch;
    if (s0.ensure(bytes) || s1.ensure(bytes)) return -1;
    const bool zp = (flags & RF_ZERO_PAD) != 0;
    if (kind == KIND_C2C) {
        const float2 *src = in;
        long long sd = idist;
        bool zp_pass = zp;
        if (flags & RF_MUL) {
            MulArgs m{};
            m.a = in;
            m.b = in2;
            m.out = s0.f2();
            m.n = n;
            m.adist = idist;
            m.bdist = idist2;
            m.odist = n;
            m.batch = batch;
            m.norm = norm;
            m.flags = flags & RF_CONJ2;
            HIP_TRY(launch_mul(m, st));
            src = s0.f2();
            sd = n;
            zp_pass = false;
        }
        return run_large(e, src, sd, out, odist, s1.f2(), batch, zp_pass, st);
    }
    if (kind == KIND_R2C) {
        if (run_large(e, in, idist, s0.f2(), n, s1.f2(), batch, zp, st)) return -1;
        ResolveArgs r{};
        r.in = s0.f2();
        r.out = out;
        r.rtw = rtw;
        r.n = n;
        r.idist = n;
        r.odist = odist;
        r.batch = batch;
        r.flags = flags & (RF_FULL | RF_ADJ_C2R);
        HIP_TRY(launch_resolve_r2c(r, st));
        return 0;
    }
    // KIND_C2R: resolve (with the optional fused multiply) into s0, then passes.
    ResolveArgs r{};
    r.in = in;
    r.in2 = in2;
    r.out = s0.f2();
    r.rtw = rtw;
    r.n = n;
    r.idist = idist;
    r.idist2 = idist2;
    r.odist = n;
    r.batch = batch;
    // With RF_ZERO_PAD the resolve zeroes the upper half of its output: a
    // zero-padded c2r plan reads only the lower half of the resolved sequence
    // (radix2_half_p1 after resolve_c2r, fft.c:828-832).
    r.flags = flags & (RF_MUL | RF_CONJ2 | RF_ZERO_PAD | RF_ADJ_R2C);
    r.norm = norm;
    HIP_TRY(launch_resolve_c2r(r, st));
    return run_large(e, s0.f2(), n, out, odist, s1.f2(), batch, false, st);
}

// Column transforms of length e.n over `ncols` columns, `batch` images.
// Fused path works in place; the generic path needs src != dst and uses alt.
int run_cols(const Engine &e, const float2 *src, long long rs_src, long long bs_src, float2 *dst,
             long long rs_dst, long long bs_dst, int ncols, int batch, DevBuf &alt,
             hipStream_t st) {
    if (batch <= 0 || ncols <= 0) return 0;
    if (e.col_fused()) {
        ColArgs a{};
        a.in = src;
        a.out = dst;
        a.tw = e.tw_col;
        a.in_rs = rs_src;
        a.out_rs = rs_dst;
        a.in_bs = bs_src;
        a.out_bs = bs_dst;
        const int tx = e.col_tx ? e.col_tx : col_TX(e.n);
        a.ncols = ncols;
        a.ntiles = (ncols + tx - 1) / tx;
        a.batch = batch;
        HIP_TRY(launch_cols(e.logn, e.direction, a, st, COL_PLAIN, tx));
        return 0;
    }
    if (src == dst) {
        set_err("internal: generic column pass cannot run in place");
        return -1;
    }
    if (alt.ensure(sizeof(float2) * (size_t)bs_dst * (size_t)batch)) return -1;
    return run_generic(e, src, Lay{const_cast<float2 *>(src), bs_src, rs_src},
                       Lay{dst, bs_dst, rs_dst}, Lay{alt.f2(), bs_dst, rs_dst}, ncols, batch, false,
                       st);
}

}  // namespace mufft
