// Autotuner of the fused row kernels and the tune cache.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <complex>

#include "host.h"

using namespace mufft;

// ===========================================================================
// Autotuner: times every registered fused-kernel variant of a size and keeps
// the fastest one whose output passes an FP64 check.  The static choice
// rv_default() stays in force until a size is tuned or set explicitly.
// ===========================================================================
namespace {

using cd = std::complex<double>;

// Iterative radix-2 FFT in FP64, X[k] = sum x[n] exp(dir 2 pi i n k / n).
void fft64(std::vector<cd> &a, int dir) {
    const size_t n = a.size();
    for (size_t i = 1, j = 0; i < n; ++i) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        const double ang = dir * 2.0 * M_PI / (double)len;
        for (size_t i = 0; i < n; i += len)
            for (size_t k = 0; k < len / 2; ++k) {
                const cd w(std::cos(ang * (double)k), std::sin(ang * (double)k));
                const cd u = a[i + k], v = a[i + k + len / 2] * w;
                a[i + k] = u + v;
                a[i + k + len / 2] = u - v;
            }
    }
}

// FP64 reference of one transform of kind tk on complex length nc, using the
// reference semantics (r2c N/2+1 bins; c2r keeps Im of DC and Nyquist).
std::vector<cd> reference64(int tk, int nc, const float2 *in) {
    if (tk == TK_C2C_FWD || tk == TK_C2C_INV) {
        std::vector<cd> a(nc);
        for (int i = 0; i < nc; ++i) a[i] = cd(in[i].x, in[i].y);
        fft64(a, tk == TK_C2C_FWD ? -1 : 1);
        return a;
    }
    const int m = 2 * nc;
    if (tk == TK_R2C) {
        const float *x = reinterpret_cast<const float *>(in);
        std::vector<cd> a(m);
        for (int i = 0; i < m; ++i) a[i] = cd(x[i], 0.0);
        fft64(a, -1);
        a.resize(nc + 1);
        return a;
    }
    std::vector<cd> f(m);
    f[0] = cd(in[0].x, 0.0);
    f[nc] = cd(in[nc].x, 0.0);
    for (int k = 1; k < nc; ++k) {
        f[k] = cd(in[k].x, in[k].y);
        f[m - k] = std::conj(f[k]);
    }
    fft64(f, 1);
    std::vector<cd> y(nc);  // output reals packed as complex pairs
    for (int i = 0; i < m; ++i) {
        const double q2 = -(double)in[0].y * ((i & 1) ? -1.0 : 1.0) - (double)in[nc].y;
        const double val = f[i].real() + q2;
        if (i & 1) y[i / 2].imag(val);
        else y[i / 2].real(val);
    }
    return y;
}

// tk: TuneKind of a row variant, or one of these for the other tuned choices.
constexpr int TK_SPLIT = 4;     // split of a transform beyond the fused sizes (fs_code)
constexpr int TK_COLTILE = 5;   // tile width of the column kernel
struct TuneRecord {
    int logn, tk, variant;
    double us;
};
std::mutex g_tune_lock;
std::vector<TuneRecord> g_tune_records;

int tk_from_api(int kind) { return (kind >= 0 && kind <= 3) ? kind : -1; }

// Complex log2 length of an API size (real sizes for r2c / c2r).
int logn_from_api(unsigned N, int tk) {
    if (!is_pow2(N)) return -1;
    const int l = ilog2(N) - ((tk == TK_R2C || tk == TK_C2R) ? 1 : 0);
    return (l >= 1 && l <= ROW_MAX_LOG2) ? l : -1;
}

}  // namespace

extern "C" {

int mufft_hip_list_variants(unsigned N, int kind, int *out, int max) {
    const int tk = tk_from_api(kind);
    const int l = tk < 0 ? -1 : logn_from_api(N, tk);
    if (l < 0) {
        set_err("no fused kernel for this size");
        return -1;
    }
    return row_variant_list(l, out, max, tk == TK_C2R);
}

int mufft_hip_default_variant(unsigned N, int kind) {
    const int tk = tk_from_api(kind);
    const int l = tk < 0 ? -1 : logn_from_api(N, tk);
    if (l < 0) {
        set_err("no fused kernel for this size");
        return -1;
    }
    return rv_default(1 << l, tk);
}

int mufft_hip_get_variant(unsigned N, int kind) {
    const int tk = tk_from_api(kind);
    const int l = tk < 0 ? -1 : logn_from_api(N, tk);
    if (l < 0) {
        set_err("no fused kernel for this size");
        return -1;
    }
    return chosen_variant(l, tk);
}

int mufft_hip_set_variant(unsigned N, int kind, int variant) {
    const int tk = tk_from_api(kind);
    const int l = tk < 0 ? -1 : logn_from_api(N, tk);
    if (l < 0 || (variant != 0 && !row_variant_registered(l, variant))) {
        set_err("unknown size or variant");
        return -1;
    }
    set_chosen_variant(l, tk, variant);
    return 0;
}

static int mufft_hip_autotune_1d_impl(unsigned N, int kind, unsigned howmany, double *best_us) {
    const int tk = tk_from_api(kind);
    const int l = tk < 0 ? -1 : logn_from_api(N, tk);
    if (l < 0 || howmany == 0) {
        set_err("no fused kernel for this size");
        return -1;
    }
    const int nc = 1 << l;
    const int fkind = tk == TK_R2C ? KIND_R2C : (tk == TK_C2R ? KIND_C2R : KIND_C2C);
    const int dir = (tk == TK_C2C_FWD || tk == TK_R2C) ? -1 : 1;
    const long long idist = tk == TK_C2R ? nc + 1 : nc;
    const long long odist = tk == TK_R2C ? nc + 1 : nc;
    std::vector<float2> h_in((size_t)idist * howmany);
    uint64_t st = 0x9E3779B97F4A7C15ull ^ ((uint64_t)nc << 8) ^ (uint64_t)tk;
    for (auto &z : h_in) {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        z.x = (float)((double)(st >> 40) / 16777216.0 - 0.5);
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        z.y = (float)((double)(st >> 40) / 16777216.0 - 0.5);
    }
    const int nval = (int)std::min<unsigned>(howmany, 2);
    std::vector<std::vector<cd>> ref(nval);
    for (int b = 0; b < nval; ++b) ref[b] = reference64(tk, nc, h_in.data() + (size_t)b * idist);

    hipStream_t s = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    DevBuf din, dout;
    float2 *rtw = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int best = -1;
    double best_t = 1e30;
    int rc = 0;
    // Each timed launch starts with cold caches: streaming 256 MiB through
    // this buffer evicts the 256 MB Infinity Cache and the L2s, as a batch
    // streamed from HBM meets them.  Repeating a launch on the same buffers
    // would instead find its input cached and favour the variants that read
    // it twice (the two-workgroup split), which lose on cold data.
    constexpr long long kFlushPts = 1LL << 25;   // 256 MiB of float2 per half
    float2 *flush = nullptr;
    if (hipMalloc(&flush, 2 * kFlushPts * sizeof(float2)) != hipSuccess) {
        (void)hipGetLastError();   // no flush: warm timing
        flush = nullptr;
    }
    if (din.ensure(h_in.size() * sizeof(float2)) || dout.ensure((size_t)odist * howmany * sizeof(float2)) ||
        hipMemcpy(din.p, h_in.data(), h_in.size() * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess ||
        (fkind != KIND_C2C && build_resolve_table(nc, dir, &rtw, s)) ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        rc = -1;
    }
    int vars[64];
    const int nv = rc ? 0 : std::min(row_variant_list(l, vars, 64, tk == TK_C2R), 64);
    std::vector<float2> h_out((size_t)odist * nval);
    for (int i = 0; i < nv && !rc; ++i) {
        Engine e;
        if (e.init(nc, dir, fkind, true, false, s, vars[i])) {
            e.release();
            continue;
        }
        RowArgs a{};
        a.in = din.f2();
        a.out = dout.f2();
        a.tw = e.tw_row;
        a.rtw = rtw;
        a.stw = e.tw_split;
        a.idist = idist;
        a.odist = odist;
        a.batch = (int)howmany;
        bool ok = launch_rows(l, dir, fkind, vars[i], a, s) == hipSuccess &&
                  hipMemcpyAsync(h_out.data(), dout.p, h_out.size() * sizeof(float2),
                                 hipMemcpyDeviceToHost, s) == hipSuccess &&
                  hipStreamSynchronize(s) == hipSuccess;
        double err = 0;
        for (int b = 0; ok && b < nval; ++b) {
            double num = 0, den = 0;
            const size_t cnt = ref[b].size();
            for (size_t k = 0; k < cnt; ++k) {
                const float2 g = h_out[(size_t)b * odist + k];
                num += std::norm(cd(g.x, g.y) - ref[b][k]);
                den += std::norm(ref[b][k]);
            }
            err = std::max(err, std::sqrt(num / (den > 0 ? den : 1.0)));
        }
        ok = ok && err <= 3e-6;
        if (ok) {
            const int reps = 5;
            (void)launch_rows(l, dir, fkind, vars[i], a, s);   // warm-up
            double total_ms = 0;
            for (int r = 0; r < reps && ok; ++r) {
                if (flush) (void)launch_stream_copy(flush, flush + kFlushPts, kFlushPts, s);
                (void)hipEventRecord(e0, s);
                (void)launch_rows(l, dir, fkind, vars[i], a, s);
                (void)hipEventRecord(e1, s);
                float ms = 0;
                ok = hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
                total_ms += ms;
            }
            const double us = 1e3 * total_ms / reps;
            if (ok && us < best_t) {
                best_t = us;
                best = vars[i];
            }
        }
        (void)hipStreamSynchronize(s);
        e.release();
    }
    if (rtw) (void)hipFree(rtw);
    if (flush) (void)hipFree(flush);
    din.release();
    dout.release();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
    if (rc || best < 0) {
        if (!rc) set_err("no variant passed validation");
        return -1;
    }
    {
        set_chosen_variant(l, tk, best);
        std::lock_guard<std::mutex> g(g_tune_lock);
        g_tune_records.erase(std::remove_if(g_tune_records.begin(), g_tune_records.end(),
                                            [&](const TuneRecord &t) { return t.logn == l && t.tk == tk; }),
                             g_tune_records.end());
        g_tune_records.push_back({l, tk, best, best_t});
    }
    if (best_us) *best_us = best_t;
    return best;
}


int mufft_hip_autotune_1d(unsigned N, int kind, unsigned howmany, double *best_us) {
    return guarded<int>(-1, [&] { return mufft_hip_autotune_1d_impl(N, kind, howmany, best_us); });
}

// ===========================================================================
// Large transforms: the split N = N1 x N2 (two column launches) or
// N1 x A x B (three) is chosen by timing every runnable split at the
// caller's batch, each checked against an FP64 FFT (relative RMS <= 3e-6)
// first.  The static choice (fs_code_default) stays in force until then.
// ===========================================================================
static void record_tune(int l, int tk, int choice, double us) {
    std::lock_guard<std::mutex> g(g_tune_lock);
    g_tune_records.erase(std::remove_if(g_tune_records.begin(), g_tune_records.end(),
                                        [&](const TuneRecord &t) { return t.logn == l && t.tk == tk; }),
                         g_tune_records.end());
    g_tune_records.push_back({l, tk, choice, us});
}

// Complex log2 length of an API size beyond the fused row sizes.
static int large_logn(unsigned N, int kind) {
    const int tk = tk_from_api(kind);
    if (tk < 0 || !is_pow2(N)) return -1;
    const int l = ilog2(N) - ((tk == TK_R2C || tk == TK_C2R) ? 1 : 0);
    return (l > ROW_MAX_LOG2 && l <= 2 * COL_MAX_LOG2) ? l : -1;
}

static int mufft_hip_autotune_large_1d_impl(unsigned N, int kind, unsigned howmany, double *best_us) {
    const int l = large_logn(N, kind);
    if (l < 0 || howmany == 0) {
        set_err("no split to tune for this size (2^17..2^26 complex points)");
        return -1;
    }
    const int tk = tk_from_api(kind);
    const int dir = (tk == TK_C2C_FWD || tk == TK_R2C) ? -1 : 1;
    const long long nc = 1LL << l;
    std::vector<int> codes;
    for (int n1 = 6; n1 <= 13; ++n1) {
        if (fs_code_valid(l, n1 << 8)) codes.push_back(n1 << 8);
        for (int a = 6; a <= 10; ++a) {
            const int b = l - n1 - a;
            const int c = (n1 << 8) | (a << 4) | (b > 0 ? b : 0);
            if (b >= 6 && b <= 10 && fs_code_valid(l, c)) codes.push_back(c);
        }
    }
    // One transform for the FP64 check, the batch for the timing.
    std::vector<float2> h_in((size_t)nc);
    uint64_t st = 0x2545F4914F6CDD1Dull ^ (uint64_t)l;
    for (auto &z : h_in) {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        z.x = (float)((double)(st >> 40) / 16777216.0 - 0.5);
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        z.y = (float)((double)(st >> 40) / 16777216.0 - 0.5);
    }
    std::vector<cd> ref(h_in.size());
    for (size_t i = 0; i < h_in.size(); ++i) ref[i] = cd(h_in[i].x, h_in[i].y);
    fft64(ref, dir);
    double ref_norm = 0;
    for (const cd &z : ref) ref_norm += std::norm(z);

    hipStream_t s = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    DevBuf din, dout, s0, s1;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const size_t bytes = (size_t)nc * howmany * sizeof(float2);
    int rc = 0;
    if (din.ensure(bytes) || dout.ensure(bytes) || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        rc = -1;
    }
    for (unsigned b = 0; !rc && b < howmany; ++b)
        if (hipMemcpy(din.f2() + (size_t)b * nc, h_in.data(), nc * sizeof(float2), hipMemcpyHostToDevice) !=
            hipSuccess)
            rc = -1;
    int best = -1;
    double best_t = 1e30;
    std::vector<float2> h_out((size_t)nc);
    for (size_t i = 0; i < codes.size() && !rc; ++i) {
        Engine e;
        if (e.init((int)nc, dir, KIND_C2C, true, false, s, 0, codes[i])) {
            e.release();
            continue;
        }
        auto run = [&] {
            return run_rows(e, KIND_C2C, nullptr, din.f2(), nc, nullptr, 0, dout.f2(), nc, (int)howmany, 0u, 1.0f,
                            s0, s1, s);
        };
        bool ok = run() == 0 &&
                  hipMemcpyAsync(h_out.data(), dout.p, nc * sizeof(float2), hipMemcpyDeviceToHost, s) ==
                      hipSuccess &&
                  hipStreamSynchronize(s) == hipSuccess;
        if (ok) {
            double num = 0;
            for (size_t k = 0; k < h_out.size(); ++k) num += std::norm(cd(h_out[k].x, h_out[k].y) - ref[k]);
            ok = std::sqrt(num / (ref_norm > 0 ? ref_norm : 1.0)) <= 3e-6;
        }
        if (ok) {
            const int reps = 3;
            double total_ms = 0;
            for (int r = 0; r < reps && ok; ++r) {
                (void)hipEventRecord(e0, s);
                ok = run() == 0;
                (void)hipEventRecord(e1, s);
                float ms = 0;
                ok = ok && hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
                total_ms += ms;
            }
            const double us = 1e3 * total_ms / reps;
            if (ok && us < best_t) {
                best_t = us;
                best = codes[i];
            }
        }
        (void)hipStreamSynchronize(s);
        e.release();
    }
    din.release();
    dout.release();
    s0.release();
    s1.release();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
    if (rc || best < 0) {
        if (!rc) set_err("no split passed validation");
        return -1;
    }
    set_fs_code(l, best);
    record_tune(l, TK_SPLIT, best, best_t);
    if (best_us) *best_us = best_t;
    return best;
}

int mufft_hip_autotune_large_1d(unsigned N, int kind, unsigned howmany, double *best_us) {
    return guarded<int>(-1, [&] { return mufft_hip_autotune_large_1d_impl(N, kind, howmany, best_us); });
}

int mufft_hip_get_large_split(unsigned N, int kind) {
    const int l = large_logn(N, kind);
    if (l < 0) {
        set_err("no split for this size");
        return -1;
    }
    return chosen_fs_code(l);
}

int mufft_hip_set_large_split(unsigned N, int kind, int code) {
    const int l = large_logn(N, kind);
    if (l < 0 || (code != 0 && !fs_code_valid(l, code))) {
        set_err("unknown size or split");
        return -1;
    }
    set_fs_code(l, code);
    return 0;
}

// ===========================================================================
// 2D column passes: the tile width of the column kernel (col_TX or its
// alternative col_TX_alt) is chosen by timing both on the caller's image
// shape.  The width changes only which columns a workgroup owns, so both
// must give identical bits; a width whose output differs is refused.
// ===========================================================================
static int mufft_hip_autotune_2d_impl(unsigned Nx, unsigned Ny, int kind, unsigned howmany, double *best_us) {
    const int tk = tk_from_api(kind);
    if (tk < 0 || !is_pow2(Nx) || !is_pow2(Ny) || Ny < 2 || howmany == 0 ||
        ilog2(Ny) > COL_MAX_LOG2 || ((tk == TK_R2C || tk == TK_C2R) && Nx < 4)) {
        set_err("no column kernel to tune for this shape");
        return -1;
    }
    const int ly = ilog2(Ny);
    const int alt = col_TX_alt((int)Ny);
    const int dir = (tk == TK_C2C_FWD || tk == TK_R2C) ? -1 : 1;
    // Columns and row stride as the 2D plans lay them out (fft.c:664-792):
    // Nx complex columns per row, of which Nx/2 + 1 are transformed for r2c / c2r.
    const long long rs = (tk == TK_R2C || tk == TK_C2R) ? Nx / 2 : Nx;
    const int ncols = (int)((tk == TK_R2C || tk == TK_C2R) ? Nx / 2 + 1 : Nx);
    const long long img = (tk == TK_R2C || tk == TK_C2R) ? (long long)Nx * Ny : rs * Ny;
    const long long rstride = (tk == TK_R2C || tk == TK_C2R) ? (long long)Nx : rs;
    hipStream_t s = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    DevBuf din, dout, alt_buf;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const size_t bytes = (size_t)img * howmany * sizeof(float2);
    int rc = 0;
    std::vector<float2> h_in((size_t)img);
    uint64_t st = 0x9E3779B97F4A7C15ull ^ (uint64_t)Ny;
    for (auto &z : h_in) {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        z.x = (float)((double)(st >> 40) / 16777216.0 - 0.5);
        z.y = (float)((double)(st & 0xffffff) / 16777216.0 - 0.5);
    }
    if (din.ensure(bytes) || dout.ensure(bytes) || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess)
        rc = -1;
    for (unsigned b = 0; !rc && b < howmany; ++b)
        if (hipMemcpy(din.f2() + (size_t)b * img, h_in.data(), img * sizeof(float2), hipMemcpyHostToDevice) !=
            hipSuccess)
            rc = -1;
    const int cands[2] = {col_TX((int)Ny), alt};
    std::vector<float2> first, h_out((size_t)img);
    int best = -1;
    double best_t = 1e30;
    for (int ci = 0; ci < 2 && !rc; ++ci) {
        const int tx = cands[ci];
        if (tx <= 0) continue;
        Engine e;
        if (e.init((int)Ny, dir, KIND_C2C, false, true, s, 0, 0, tx)) {
            e.release();
            continue;
        }
        auto run = [&] {
            return run_cols(e, din.f2(), rstride, img, dout.f2(), rstride, img, ncols, (int)howmany, alt_buf, s);
        };
        bool ok = run() == 0 &&
                  hipMemcpyAsync(h_out.data(), dout.p, img * sizeof(float2), hipMemcpyDeviceToHost, s) ==
                      hipSuccess &&
                  hipStreamSynchronize(s) == hipSuccess;
        if (ok && first.empty()) first = h_out;
        else if (ok) {
            for (long long r = 0; r < Ny && ok; ++r)
                ok = memcmp(&h_out[(size_t)(r * rstride)], &first[(size_t)(r * rstride)],
                                 (size_t)ncols * sizeof(float2)) == 0;
        }
        if (ok) {
            const int reps = 5;
            double total_ms = 0;
            (void)run();
            for (int r = 0; r < reps && ok; ++r) {
                (void)hipEventRecord(e0, s);
                ok = run() == 0;
                (void)hipEventRecord(e1, s);
                float ms = 0;
                ok = ok && hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
                total_ms += ms;
            }
            const double us = 1e3 * total_ms / reps;
            if (ok && us < best_t) {
                best_t = us;
                best = tx;
            }
        }
        (void)hipStreamSynchronize(s);
        e.release();
    }
    din.release();
    dout.release();
    alt_buf.release();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
    if (rc || best < 0) {
        if (!rc) set_err("no tile width passed");
        return -1;
    }
    set_col_tx(ly, best);
    record_tune(ly, TK_COLTILE, best, best_t);
    if (best_us) *best_us = best_t;
    return best;
}

int mufft_hip_autotune_2d(unsigned Nx, unsigned Ny, int kind, unsigned howmany, double *best_us) {
    return guarded<int>(-1, [&] { return mufft_hip_autotune_2d_impl(Nx, Ny, kind, howmany, best_us); });
}

int mufft_hip_get_column_tile(unsigned L) {
    if (!is_pow2(L) || L < 2 || ilog2(L) > COL_MAX_LOG2) {
        set_err("no fused column kernel of this length");
        return -1;
    }
    return chosen_col_tx(ilog2(L));
}

int mufft_hip_set_column_tile(unsigned L, int tx) {
    if (!is_pow2(L) || L < 2 || ilog2(L) > COL_MAX_LOG2 ||
        (tx != 0 && tx != col_TX((int)L) && tx != col_TX_alt((int)L))) {
        set_err("unknown column length or tile width");
        return -1;
    }
    set_col_tx(ilog2(L), tx);
    return 0;
}

// Tune cache: one line per entry, "log2n kind variant microseconds".
static int mufft_hip_tune_save_impl(const char *path) {
    FILE *f = fopen(path, "w");
    if (!f) {
        set_err("cannot write %s", path);
        return -1;
    }
    std::lock_guard<std::mutex> g(g_tune_lock);
    fprintf(f, "# mufft_hip tune cache: log2(complex length) kind(0 c2c fwd,1 c2c inv,2 r2c,3 c2r) variant us\n");
    for (int l = 0; l <= ROW_MAX_LOG2; ++l)
        for (int tk = 0; tk < 4; ++tk)
            if (variant_is_set(l, tk)) {
                double us = 0;
                for (const auto &t : g_tune_records)
                    if (t.logn == l && t.tk == tk) us = t.us;
                fprintf(f, "%d %d %d %.3f\n", l, tk, chosen_variant(l, tk), us);
            }
    // Tuned splits (kind 4, a fs_code) and column tile widths (kind 5).
    for (const auto &t : g_tune_records)
        if (t.tk == TK_SPLIT || t.tk == TK_COLTILE) fprintf(f, "%d %d %d %.3f\n", t.logn, t.tk, t.variant, t.us);
    fclose(f);
    return 0;
}


int mufft_hip_tune_save(const char *path) {
    return guarded<int>(-1, [&] { return mufft_hip_tune_save_impl(path); });
}

static int mufft_hip_tune_load_impl(const char *path) {
    FILE *f = fopen(path, "r");
    if (!f) {
        set_err("cannot read %s", path);
        return -1;
    }
    char line[256];
    int n = 0;
    while (fgets(line, sizeof line, f)) {
        int l, tk, var;
        double us;
        if (line[0] == '#' || sscanf(line, "%d %d %d %lf", &l, &tk, &var, &us) != 4) continue;
        if (tk == TK_SPLIT) {
            if (l <= ROW_MAX_LOG2 || l > 2 * COL_MAX_LOG2 || !fs_code_valid(l, var)) continue;
            set_fs_code(l, var);
            record_tune(l, tk, var, us);
            ++n;
            continue;
        }
        if (tk == TK_COLTILE) {
            if (l < 1 || l > COL_MAX_LOG2 || (var != col_TX(1 << l) && var != col_TX_alt(1 << l))) continue;
            set_col_tx(l, var);
            record_tune(l, tk, var, us);
            ++n;
            continue;
        }
        if (l < 1 || l > ROW_MAX_LOG2 || tk < 0 || tk > 3 || !row_variant_registered(l, var)) continue;
        set_chosen_variant(l, tk, var);
        ++n;
    }
    fclose(f);
    return n;
}


int mufft_hip_tune_load(const char *path) {
    return guarded<int>(-1, [&] { return mufft_hip_tune_load_impl(path); });
}

}  // extern "C"

