// Planner, executor and C ABI of the MI355X muFFT drop-in.
//
// The reference builds a list of SIMD function pointers per plan and calls
// them one pass at a time on host memory (fft.c:325-435, fft.c:811-989).
// Here a plan owns device twiddle tables (generated in FP64 on the GPU), a
// private HIP stream and scratch space, and executing it enqueues one fused
// kernel per transform axis whenever the axis fits one or two workgroups
// (1D/rows up to 65536 complex points, columns up to 8192), or one generic
// Stockham pass per launch beyond that.  Host buffers are staged through device memory so
// the reference's host-memory calling convention keeps working.
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "conv_batch.h"
#include "kernels.h"
#include "mufft/fft.h"
#include "mufft/mufft_hip.h"

using namespace mufft;

namespace {

thread_local char g_err[512] = "";

void set_err(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

// C entry points that build host containers run their body through this, so
// an exception (std::bad_alloc) becomes an error return instead of crossing
// the C ABI.
template <typename R, typename F>
R guarded(R fail, F &&body) {
    try {
        return body();
    } catch (const std::exception &e) {
        set_err("host exception: %s", e.what());
    } catch (...) {
        set_err("host exception");
    }
    return fail;
}

#define HIP_TRY(expr)                                                              \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess) {                                                    \
            set_err("%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
            return -1;                                                             \
        }                                                                          \
    } while (0)

bool is_pow2(unsigned n) { return n != 0 && (n & (n - 1)) == 0; }

// Grow-only device buffer.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t n) {
        if (n <= bytes) return 0;
        release();
        if (hipMalloc(&p, n) != hipSuccess) {
            (void)hipGetLastError();
            set_err("hipMalloc(%zu bytes) failed", n);
            p = nullptr;
            return -1;
        }
        bytes = n;
        return 0;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    float2 *f2() const { return static_cast<float2 *>(p); }
};

// True when kernels cannot dereference the pointer directly (pageable or
// pinned host memory); such buffers are staged through device memory.
bool is_host_ptr(const void *ptr) {
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, ptr) != hipSuccess) {
        (void)hipGetLastError();
        return true;
    }
    return !(attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged ||
             attr.type == hipMemoryTypeArray);
}

// Where a buffer lives: device memory (used in place), pinned host memory
// (kernels may address it over PCIe through `dev`) or pageable host memory
// (only the copy engines can reach it).
enum PtrKind { PTR_DEVICE, PTR_PINNED, PTR_PAGEABLE };
struct PtrInfo {
    PtrKind kind;
    void *dev;
};
PtrInfo classify_ptr(const void *ptr) {
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, ptr) != hipSuccess) {
        (void)hipGetLastError();
        return {PTR_PAGEABLE, nullptr};
    }
    if (attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged ||
        attr.type == hipMemoryTypeArray)
        return {PTR_DEVICE, const_cast<void *>(ptr)};
    if (attr.type == hipMemoryTypeHost && attr.devicePointer) {
        // devicePointer is the mapping of the allocation's base address.
        const char *host_base = static_cast<const char *>(attr.hostPointer ? attr.hostPointer : ptr);
        return {PTR_PINNED, static_cast<char *>(attr.devicePointer) +
                                (static_cast<const char *>(ptr) - host_base)};
    }
    return {PTR_PAGEABLE, nullptr};
}

// Kernel-addressable form of a buffer for the stream-ordered extensions that
// do no staging: device pointers as they are, pinned host memory through its
// device mapping; pageable host memory is refused (a kernel cannot reach it).
int kernel_ptr(const void *ptr, const void **out) {
    const PtrInfo pi = classify_ptr(ptr);
    if (pi.kind == PTR_PAGEABLE) {
        set_err("buffer must be device memory or pinned host memory (mufft_alloc)");
        return -1;
    }
    *out = pi.dev;
    return 0;
}
template <typename T>
int kernel_ptrs(T **p) {
    const void *q = nullptr;
    if (kernel_ptr(*p, &q)) return -1;
    *p = static_cast<T *>(const_cast<void *>(q));
    return 0;
}
template <typename T, typename... R>
int kernel_ptrs(T **p, R... rest) {
    return kernel_ptrs(p) ? -1 : kernel_ptrs(rest...);
}

// Host-memory batches: the batch runs in chunks through two device staging
// slots on three streams, so the upload of chunk i+1 and the download of
// chunk i-1 overlap the transform of chunk i.  Device layouts mirror the host
// layouts (same distances), so the transform runs unchanged on each chunk.
struct HostPipe {
    static constexpr size_t kChunkBytes = 32u << 20;
    hipStream_t up = nullptr, down = nullptr;
    hipEvent_t ev_up[2] = {}, ev_comp[2] = {}, ev_down[2] = {};
    DevBuf in[2], out[2];
    int init() {
        if (up) return 0;
        HIP_TRY(hipStreamCreateWithFlags(&up, hipStreamNonBlocking));
        HIP_TRY(hipStreamCreateWithFlags(&down, hipStreamNonBlocking));
        for (int i = 0; i < 2; ++i) {
            HIP_TRY(hipEventCreateWithFlags(&ev_up[i], hipEventDisableTiming));
            HIP_TRY(hipEventCreateWithFlags(&ev_comp[i], hipEventDisableTiming));
            HIP_TRY(hipEventCreateWithFlags(&ev_down[i], hipEventDisableTiming));
        }
        return 0;
    }
    void release() {
        if (up) (void)hipStreamSynchronize(up);
        if (down) (void)hipStreamSynchronize(down);
        for (int i = 0; i < 2; ++i) {
            if (ev_up[i]) (void)hipEventDestroy(ev_up[i]);
            if (ev_comp[i]) (void)hipEventDestroy(ev_comp[i]);
            if (ev_down[i]) (void)hipEventDestroy(ev_down[i]);
            ev_up[i] = ev_comp[i] = ev_down[i] = nullptr;
            in[i].release();
            out[i].release();
        }
        if (up) (void)hipStreamDestroy(up);
        if (down) (void)hipStreamDestroy(down);
        up = down = nullptr;
    }
};

// One operand of a staged batch: unit u occupies [u*stride, u*stride + unit)
// bytes from `host`; `preload` uploads an output operand before the
// transform (for transforms that leave part of their output untouched).
struct StagedOperand {
    void *host;
    size_t unit, stride;
    bool staged;
};

// Runs run(dev_out, dev_in, count) over `howmany` units in chunks; returns
// when the last chunk's output is in host memory.
template <typename Run>
int run_staged_chunks(HostPipe &pp, hipStream_t comp, StagedOperand o, StagedOperand i,
                      bool preload_out, unsigned howmany, Run run) {
    const size_t big = std::max(o.stride, i.stride);
    const unsigned per = (unsigned)std::max<size_t>(1, HostPipe::kChunkBytes / std::max<size_t>(big, 1));
    const unsigned nchunks = (howmany + per - 1) / per;
    for (int sl = 0; sl < 2; ++sl) {
        const size_t cnt = std::min<size_t>(per, howmany);
        if (i.staged && pp.in[sl].ensure((cnt - 1) * i.stride + i.unit)) return -1;
        if (o.staged && pp.out[sl].ensure((cnt - 1) * o.stride + o.unit)) return -1;
    }
    // Work queued on `comp` before this call comes first.
    HIP_TRY(hipEventRecord(pp.ev_comp[0], comp));
    HIP_TRY(hipStreamWaitEvent(pp.up, pp.ev_comp[0], 0));
    HIP_TRY(hipStreamWaitEvent(pp.down, pp.ev_comp[0], 0));
    for (unsigned c = 0; c < nchunks; ++c) {
        const int sl = c & 1;
        const unsigned u0 = c * per;
        const unsigned cnt = std::min(per, howmany - u0);
        const size_t in_bytes = (size_t)(cnt - 1) * i.stride + i.unit;
        const size_t out_bytes = (size_t)(cnt - 1) * o.stride + o.unit;
        char *hin = static_cast<char *>(i.host) + (size_t)u0 * i.stride;
        char *hout = static_cast<char *>(o.host) + (size_t)u0 * o.stride;
        const bool up_work = i.staged || (o.staged && preload_out);
        if (c >= 2) {
            // Slot reuse: chunk c-2 must have been transformed and downloaded.
            HIP_TRY(hipStreamWaitEvent(pp.up, pp.ev_comp[sl], 0));
            if (o.staged) HIP_TRY(hipStreamWaitEvent(pp.up, pp.ev_down[sl], 0));
        }
        if (i.staged)
            HIP_TRY(hipMemcpyAsync(pp.in[sl].p, hin, in_bytes, hipMemcpyHostToDevice, pp.up));
        if (o.staged && preload_out)
            HIP_TRY(hipMemcpyAsync(pp.out[sl].p, hout, out_bytes, hipMemcpyHostToDevice, pp.up));
        if (up_work) {
            HIP_TRY(hipEventRecord(pp.ev_up[sl], pp.up));
            HIP_TRY(hipStreamWaitEvent(comp, pp.ev_up[sl], 0));
        }
        if (c >= 2 && o.staged) HIP_TRY(hipStreamWaitEvent(comp, pp.ev_down[sl], 0));
        void *dout = o.staged ? pp.out[sl].p : static_cast<void *>(hout);
        const void *din = i.staged ? pp.in[sl].p : static_cast<const void *>(hin);
        if (run(dout, din, cnt)) return -1;
        HIP_TRY(hipEventRecord(pp.ev_comp[sl], comp));
        if (o.staged) {
            HIP_TRY(hipStreamWaitEvent(pp.down, pp.ev_comp[sl], 0));
            HIP_TRY(hipMemcpyAsync(hout, pp.out[sl].p, out_bytes, hipMemcpyDeviceToHost, pp.down));
            HIP_TRY(hipEventRecord(pp.ev_down[sl], pp.down));
        }
    }
    HIP_TRY(hipStreamSynchronize(pp.down));
    HIP_TRY(hipStreamSynchronize(pp.up));
    HIP_TRY(hipStreamSynchronize(comp));
    return 0;
}

// As run_staged_chunks; after a failure the three streams are drained, so no
// copy into the staging slots is still in flight when the call returns.
template <typename Run>
int run_staged(HostPipe &pp, hipStream_t comp, StagedOperand o, StagedOperand i, bool preload_out,
               unsigned howmany, Run run) {
    if (pp.init()) return -1;
    const int rc = run_staged_chunks(pp, comp, o, i, preload_out, howmany, run);
    if (rc) {
        (void)hipStreamSynchronize(pp.up);
        (void)hipStreamSynchronize(pp.down);
        (void)hipStreamSynchronize(comp);
    }
    return rc;
}

// Stockham twiddle table for a radix sequence, generated on the GPU in FP64.
int build_stockham_table(int n, int direction, const std::vector<int> &radix, float2 **out,
                         hipStream_t s) {
    *out = nullptr;
    if (radix.size() <= 1) return 0;
    if (radix.size() > 24) {
        set_err("radix sequence too long");
        return -1;
    }
    TwiddleArgs a{};
    a.direction = direction;
    a.mode = 0;
    a.r0 = radix[0];
    a.npass = (int)radix.size();
    long long p = 1;
    for (size_t i = 0; i < radix.size(); ++i) {
        a.radix[i] = radix[i];
        a.p[i] = p;
        p *= radix[i];
    }
    a.n = n;
    a.count = n - radix[0];
    HIP_TRY(hipMalloc(out, sizeof(float2) * (size_t)a.count));
    a.out = *out;
    HIP_TRY(launch_twiddles(a, s));
    return 0;
}

int build_resolve_table(int n, int direction, float2 **out, hipStream_t s) {
    TwiddleArgs a{};
    a.direction = direction;
    a.mode = 1;
    a.n = n;
    a.count = n;
    HIP_TRY(hipMalloc(out, sizeof(float2) * (size_t)n));
    a.out = *out;
    HIP_TRY(launch_twiddles(a, s));
    return 0;
}

// exp(dir * 2 pi i * j / n) for j < count.
int build_plain_table(int n, int count, int direction, float2 **out, hipStream_t s) {
    TwiddleArgs a{};
    a.direction = direction;
    a.mode = 2;
    a.n = n;
    a.count = count;
    HIP_TRY(hipMalloc(out, sizeof(float2) * (size_t)count));
    a.out = *out;
    HIP_TRY(launch_twiddles(a, s));
    return 0;
}

// Twiddles of a transform of n points split into n / n2 sub-transforms:
// entry (k1 - 1) * n2 + j = exp(dir * 2 pi i * j * k1 / n).
int build_split_table(int n, int n2, int direction, float2 **out, hipStream_t s) {
    TwiddleArgs a{};
    a.direction = direction;
    a.mode = 3;
    a.n = n;
    a.p[0] = n2;
    a.count = n - n2;
    HIP_TRY(hipMalloc(out, sizeof(float2) * (size_t)a.count));
    a.out = *out;
    HIP_TRY(launch_twiddles(a, s));
    return 0;
}

// Radix sequence of the fused kernels with pass radix R: the remainder pass
// first for the row kernels, last for the column kernels (seq_radix()).
std::vector<int> fused_sequence(int n, int E, bool rem_first) {
    if (E > n) E = n;
    const int NP = E >= n ? 1 : (ilog2(n) + ilog2(E) - 1) / ilog2(E);
    std::vector<int> r;
    for (int s = 0; s < NP; ++s) r.push_back(seq_radix(n, E, NP, s, rem_first));
    return r;
}

std::vector<int> generic_sequence(int n) {
    std::vector<int> r;
    int left = n;
    while (left >= 16) {
        r.push_back(16);
        left /= 16;
    }
    if (left > 1) r.push_back(left);
    return r;
}

// ---------------------------------------------------------------------------
// Process-wide choice of the fused row-kernel variant per (size, transform
// kind): 0 means the static default rv_default(); the autotuner and
// mufft_hip_set_variant() fill it.  Plans read it when they are created.
// ---------------------------------------------------------------------------
enum TuneKind { TK_C2C_FWD = 0, TK_C2C_INV = 1, TK_R2C = 2, TK_C2R = 3 };

std::mutex g_var_lock;
int g_variant[ROW_MAX_LOG2 + 1][4] = {};

int tune_kind(int kind, int dir) {
    if (kind == KIND_C2C) return dir < 0 ? TK_C2C_FWD : TK_C2C_INV;
    return kind == KIND_R2C ? TK_R2C : TK_C2R;
}

int chosen_variant(int logn, int tk) {
    std::lock_guard<std::mutex> g(g_var_lock);
    const int v = g_variant[logn][tk];
    return v ? v : rv_default(1 << logn, tk);
}

// A complex transform of length n along one axis: fused row kernel, fused
// column kernel or generic passes, each with its own twiddle table.
struct Engine {
    int n = 0, logn = 0, direction = 0;
    int row_var = 0;              // fused row-kernel variant
    float2 *tw_row = nullptr;     // sequence of the row variant
    float2 *tw_col = nullptr;     // sequence of the column kernel
    float2 *tw_gen = nullptr;     // sequence of the generic pass kernels
    float2 *tw_split = nullptr;   // split variants: w_n^(dir*j*k1), 0 < k1 < N1, j < n/N1
    // Four-step rows (n = n1 * n2 beyond the fused sizes): column tables of
    // both halves and the two-level twiddle tables.
    int log_n1 = 0, log_n2 = 0, log_s = 0;
    float2 *tw_fs1 = nullptr, *tw_fs2 = nullptr, *tw_coarse = nullptr, *tw_fine = nullptr;
    // Three launches (fs3_split): N2 = A * B column passes, the middle one with
    // the twiddle w_N2^(b*k_a) from its own two-level tables.
    int log_a = 0, log_b = 0, log_s2 = 0;
    float2 *tw_pa = nullptr, *tw_pb = nullptr, *tw_coarse2 = nullptr, *tw_fine2 = nullptr;
    bool three_pass() const { return log_a > 0; }
    std::vector<int> gen_radix;
    bool row_fused() const { return logn <= ROW_MAX_LOG2; }
    bool col_fused() const { return logn <= COL_MAX_LOG2; }
    bool four_step() const { return logn > ROW_MAX_LOG2 && logn <= 2 * COL_MAX_LOG2; }

    // rows / cols: which roles this engine serves; kind: KIND_* of the row role.
    int init(int n_, int dir, int kind, bool rows, bool cols, hipStream_t s, int force_var = 0) {
        n = n_;
        logn = ilog2(n_);
        direction = dir;
        if (rows && row_fused()) {
            row_var = force_var ? force_var : chosen_variant(logn, tune_kind(kind, dir));
            const int n2 = rv_n(n, row_var);
            if (build_stockham_table(n2, dir, fused_sequence(n2, rv_R(n, row_var), true), &tw_row, s)) return -1;
            if (rv_split(row_var) > 0 && build_split_table(n, n2, dir, &tw_split, s)) return -1;
        }
        if (cols && col_fused()) {
            if (build_stockham_table(n, dir, fused_sequence(n, col_E(n), false), &tw_col, s)) return -1;
        }
        if (rows && four_step()) {
            const int f3 = fs3_split(logn);
            log_n1 = f3 ? (f3 >> 8) : fs_log_n1(logn);
            log_n2 = logn - log_n1;
            log_s = log_n2;
            const int n1 = 1 << log_n1, n2 = 1 << log_n2;
            if (build_stockham_table(n1, dir, fused_sequence(n1, col_E(n1), false), &tw_fs1, s) ||
                build_plain_table(n >> log_s, n >> log_s, dir, &tw_coarse, s) ||
                build_plain_table(n, 1 << log_s, dir, &tw_fine, s))
                return -1;
            if (f3) {
                log_a = (f3 >> 4) & 15;
                log_b = f3 & 15;
                log_s2 = log_n2 / 2;
                const int na = 1 << log_a, nb = 1 << log_b;
                if (build_stockham_table(na, dir, fused_sequence(na, col_E(na), false), &tw_pa, s) ||
                    build_stockham_table(nb, dir, fused_sequence(nb, col_E(nb), false), &tw_pb, s) ||
                    build_plain_table(n2 >> log_s2, n2 >> log_s2, dir, &tw_coarse2, s) ||
                    build_plain_table(n2, 1 << log_s2, dir, &tw_fine2, s))
                    return -1;
            } else if (build_stockham_table(n2, dir, fused_sequence(n2, col_E(n2), false), &tw_fs2, s)) {
                return -1;
            }
        }
        if ((rows && !row_fused() && !four_step()) || (cols && !col_fused())) {
            gen_radix = generic_sequence(n);
            if (build_stockham_table(n, dir, gen_radix, &tw_gen, s)) return -1;
        }
        return 0;
    }
    void release() {
        if (tw_row) (void)hipFree(tw_row);
        if (tw_col) (void)hipFree(tw_col);
        if (tw_gen) (void)hipFree(tw_gen);
        if (tw_split) (void)hipFree(tw_split);
        for (float2 *t : {tw_fs1, tw_fs2, tw_coarse, tw_fine, tw_pa, tw_pb, tw_coarse2, tw_fine2})
            if (t) (void)hipFree(t);
        tw_row = tw_col = tw_gen = tw_split = nullptr;
        tw_fs1 = tw_fs2 = tw_coarse = tw_fine = nullptr;
        tw_pa = tw_pb = tw_coarse2 = tw_fine2 = nullptr;
    }
    int launches_generic() const { return (int)gen_radix.size(); }
};

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
    HIP_TRY(launch_cols(e.log_n2, e.direction, c, st));
    return 0;
}

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
        a.ncols = ncols;
        a.ntiles = (ncols + col_TX(e.n) - 1) / col_TX(e.n);
        a.batch = batch;
        HIP_TRY(launch_cols(e.logn, e.direction, a, st));
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

// Batched entry points run on the caller's stream; NULL is the null stream.
hipStream_t resolve_stream(void *s, hipStream_t) { return static_cast<hipStream_t>(s); }

}  // namespace

// ===========================================================================
// 1D plans
// ===========================================================================
struct mufft_plan_1d {
    unsigned N = 0;      // complex length (reference plan->N, fft.c:58)
    int direction = 0;
    int kind = KIND_C2C;
    unsigned flags = 0;
    bool full = false, zero_pad = false;
    int device = 0;
    hipStream_t stream = nullptr;
    Engine eng;
    float2 *rtw = nullptr;
    DevBuf s0, s1, stage_in, stage_in2, stage_out;
    size_t in_bytes = 0, out_bytes = 0;   // per transform
    std::mutex lock;                       // serialises drop-in staging and host batches
    HostPipe pipe;                         // host-memory batches
    // Adjoint transform, built on first use.
    Engine adj;
    float2 *adj_rtw = nullptr;
    bool adj_ready = false;
    DevBuf fold;
};

static void destroy_1d(mufft_plan_1d *p) {
    if (!p) return;
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    p->eng.release();
    if (p->rtw) (void)hipFree(p->rtw);
    p->adj.release();
    if (p->adj_rtw) (void)hipFree(p->adj_rtw);
    p->fold.release();
    p->s0.release();
    p->s1.release();
    p->stage_in.release();
    p->stage_in2.release();
    p->stage_out.release();
    p->pipe.release();
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

static mufft_plan_1d *create_1d_impl(int kind, unsigned Nc, int direction, unsigned flags) {
    if (!is_pow2(Nc) || Nc < 2 || Nc > (1u << 27)) {
        set_err("invalid transform size");
        return nullptr;
    }
    mufft_plan_1d *p = new (std::nothrow) mufft_plan_1d;
    if (!p) return nullptr;
    p->N = Nc;
    p->kind = kind;
    p->direction = direction;
    p->flags = flags;
    p->full = (kind == KIND_R2C) && (flags & MUFFT_FLAG_FULL_R2C);
    p->zero_pad = (flags & MUFFT_FLAG_ZERO_PAD_UPPER_HALF) != 0;
    if (hipGetDevice(&p->device) != hipSuccess ||
        hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        set_err("no usable HIP device");
        p->stream = nullptr;
        destroy_1d(p);
        return nullptr;
    }
    if (p->eng.init((int)Nc, direction, kind, true, false, p->stream)) {
        destroy_1d(p);
        return nullptr;
    }
    if (kind != KIND_C2C && build_resolve_table((int)Nc, direction, &p->rtw, p->stream)) {
        destroy_1d(p);
        return nullptr;
    }
    if (hipStreamSynchronize(p->stream) != hipSuccess) {
        set_err("twiddle generation failed: %s", hipGetErrorString(hipGetLastError()));
        destroy_1d(p);
        return nullptr;
    }
    const size_t n = Nc;
    switch (kind) {
    case KIND_C2C:
        p->in_bytes = (p->zero_pad ? n / 2 : n) * 8;
        p->out_bytes = n * 8;
        break;
    case KIND_R2C:
        p->in_bytes = (p->zero_pad ? n : 2 * n) * 4;
        p->out_bytes = (p->full ? 2 * n : n + 1) * 8;
        break;
    default:
        p->in_bytes = (n + 1) * 8;
        p->out_bytes = 2 * n * 4;
        break;
    }
    return p;
}

static mufft_plan_1d *create_1d(int kind, unsigned Nc, int direction, unsigned flags) {
    return guarded<mufft_plan_1d *>(nullptr, [&] { return create_1d_impl(kind, Nc, direction, flags); });
}

// Distances in complex elements; extra = RF_MUL / RF_CONJ2 for fused multiplies.
static int exec_1d(mufft_plan_1d *p, void *out, const void *in, const void *in2,
                   unsigned howmany, long long idist, long long idist2, long long odist,
                   unsigned extra, float norm, hipStream_t st) {
    unsigned flags = extra;
    if (p->zero_pad) flags |= RF_ZERO_PAD;
    if (p->full) flags |= RF_FULL;
    return run_rows(p->eng, p->kind, p->rtw, static_cast<const float2 *>(in), idist,
                    static_cast<const float2 *>(in2), idist2, static_cast<float2 *>(out), odist,
                    (int)howmany, flags, norm, p->s0, p->s1, st);
}

// Kind / direction / weighting of the adjoint of a plan's transform:
// c2c -> opposite-direction c2c; r2c -> c2r with RF_ADJ_R2C; c2r -> r2c with RF_ADJ_C2R.
static void adjoint_shape(const mufft_plan_1d *p, int *kind, int *dir, unsigned *flags) {
    if (p->kind == KIND_C2C) {
        *kind = KIND_C2C;
        *dir = -p->direction;
        *flags = 0;
    } else if (p->kind == KIND_R2C) {
        *kind = KIND_C2R;
        *dir = MUFFT_INVERSE;
        *flags = RF_ADJ_R2C;
    } else {
        *kind = KIND_R2C;
        *dir = MUFFT_FORWARD;
        *flags = RF_ADJ_C2R;
    }
}

static int ensure_adjoint(mufft_plan_1d *p) {
    if (p->adj_ready) return 0;
    if (p->kind == KIND_C2R && p->zero_pad) {
        set_err("no adjoint for zero-padded c2r plans");
        return -1;
    }
    int kind, dir;
    unsigned fl;
    adjoint_shape(p, &kind, &dir, &fl);
    if (p->adj.init((int)p->N, dir, kind, true, false, p->stream)) return -1;
    if (kind != KIND_C2C && build_resolve_table((int)p->N, dir, &p->adj_rtw, p->stream)) return -1;
    HIP_TRY(hipStreamSynchronize(p->stream));
    p->adj_ready = true;
    return 0;
}

// Adjoint execution; distances in complex elements of the adjoint's input
// (the forward output) and output (the forward input).
static int exec_adjoint(mufft_plan_1d *p, void *out, const void *in, const void *in2,
                        unsigned howmany, long long idist, long long idist2, long long odist,
                        unsigned extra, float norm, hipStream_t st) {
    if (ensure_adjoint(p)) return -1;
    int kind, dir;
    unsigned fl;
    adjoint_shape(p, &kind, &dir, &fl);
    const float2 *src = static_cast<const float2 *>(in);
    long long sd = idist;
    if (p->kind == KIND_R2C && p->full) {
        if (extra & RF_MUL) {
            set_err("fused multiply is not available for the adjoint of a full r2c plan");
            return -1;
        }
        // Full spectrum (2N bins) folded onto bins 0..N first.
        const long long nb = (long long)p->N + 1;
        if (p->fold.ensure(sizeof(float2) * (size_t)nb * howmany)) return -1;
        MulArgs m{};
        m.a = src;
        m.out = p->fold.f2();
        m.n = 2LL * p->N;
        m.adist = idist;
        m.odist = nb;
        m.batch = (int)howmany;
        HIP_TRY(launch_fold(m, st));
        src = p->fold.f2();
        sd = nb;
    }
    const float2 *rtw = kind == KIND_C2C ? nullptr : p->adj_rtw;
    return run_rows(p->adj, kind, rtw, src, sd, static_cast<const float2 *>(in2), idist2,
                    static_cast<float2 *>(out), odist, (int)howmany, fl | extra, norm, p->s0, p->s1,
                    st);
}

// Packed distances (complex elements) of a plan's input/output.
static long long in_dist_c(const mufft_plan_1d *p) {
    if (p->kind == KIND_C2R) return (long long)p->N + 1;
    return p->zero_pad ? (long long)p->N / 2 : (long long)p->N;
}
static long long out_dist_c(const mufft_plan_1d *p) {
    if (p->kind == KIND_R2C) return p->full ? 2LL * p->N : (long long)p->N + 1;
    return p->N;
}

// Drop-in execution with staging of host buffers; synchronous.
static void exec_1d_dropin(mufft_plan_1d *p, void *out, const void *in, const void *in2,
                           size_t in2_bytes, unsigned extra, float norm) {
    std::lock_guard<std::mutex> g(p->lock);
    int dev_prev = 0;
    (void)hipGetDevice(&dev_prev);
    if (dev_prev != p->device) (void)hipSetDevice(p->device);
    const void *din = in, *din2 = in2;
    void *dout = out;
    const PtrInfo ci = classify_ptr(in), co = classify_ptr(out);
    const PtrInfo ci2 = in2 ? classify_ptr(in2) : PtrInfo{PTR_DEVICE, const_cast<void *>(in2)};
    // One fused launch touches each point once: pinned host buffers are then
    // read and written in place over PCIe, with no staging copies.
    const bool direct = p->eng.row_fused() && ci.kind != PTR_PAGEABLE && co.kind != PTR_PAGEABLE &&
                        ci2.kind != PTR_PAGEABLE;
    const bool hin = !direct && ci.kind != PTR_DEVICE, hout = !direct && co.kind != PTR_DEVICE;
    const bool hin2 = !direct && in2 && ci2.kind != PTR_DEVICE;
    if (direct) {
        din = ci.dev;
        din2 = ci2.dev;
        dout = co.dev;
    }
    int rc = 0;
    if (hin) {
        rc |= p->stage_in.ensure(p->in_bytes);
        if (!rc) rc |= hipMemcpyAsync(p->stage_in.p, in, p->in_bytes, hipMemcpyHostToDevice, p->stream) != hipSuccess;
        din = p->stage_in.p;
    }
    if (!rc && hin2) {
        rc |= p->stage_in2.ensure(in2_bytes);
        if (!rc) rc |= hipMemcpyAsync(p->stage_in2.p, in2, in2_bytes, hipMemcpyHostToDevice, p->stream) != hipSuccess;
        din2 = p->stage_in2.p;
    }
    if (!rc && hout) {
        rc |= p->stage_out.ensure(p->out_bytes);
        dout = p->stage_out.p;
    }
    if (!rc) rc = exec_1d(p, dout, din, din2, 1, in_dist_c(p), in_dist_c(p), out_dist_c(p), extra, norm, p->stream);
    if (!rc && hout)
        rc |= hipMemcpyAsync(out, p->stage_out.p, p->out_bytes, hipMemcpyDeviceToHost, p->stream) != hipSuccess;
    if (hipStreamSynchronize(p->stream) != hipSuccess || rc) {
        if (!rc) set_err("execution failed: %s", hipGetErrorString(hipGetLastError()));
        fprintf(stderr, "mufft_hip: %s\n", g_err);
    }
    if (dev_prev != p->device) (void)hipSetDevice(dev_prev);
}

extern "C" {

mufft_plan_1d *mufft_create_plan_1d_c2c(unsigned N, int direction, unsigned flags) {
    if (direction != MUFFT_FORWARD && direction != MUFFT_INVERSE) {
        set_err("invalid direction");
        return nullptr;
    }
    return create_1d(KIND_C2C, N, direction, flags);
}

mufft_plan_1d *mufft_create_plan_1d_r2c(unsigned N, unsigned flags) {
    if (!is_pow2(N) || N < 4) {  // N == 2 fails in the reference too (fft.c:485 -> c2c(1))
        set_err("invalid transform size");
        return nullptr;
    }
    return create_1d(KIND_R2C, N / 2, MUFFT_FORWARD, flags);
}

mufft_plan_1d *mufft_create_plan_1d_c2r(unsigned N, unsigned flags) {
    if (!is_pow2(N) || N < 4) {
        set_err("invalid transform size");
        return nullptr;
    }
    return create_1d(KIND_C2R, N / 2, MUFFT_INVERSE, flags);
}

void mufft_execute_plan_1d(mufft_plan_1d *plan, void *MUFFT_RESTRICT output,
                           const void *MUFFT_RESTRICT input) {
    if (!plan) return;
    exec_1d_dropin(plan, output, input, nullptr, 0, 0u, 1.0f);
}

void mufft_free_plan_1d(mufft_plan_1d *plan) { destroy_1d(plan); }

int mufft_hip_execute_plan_1d_batched(mufft_plan_1d *p, void *output, const void *input,
                                      unsigned howmany, size_t idist, size_t odist, void *stream) {
    if (!p || !output || !input) {
        set_err("null argument");
        return -1;
    }
    // Real buffers count floats; two floats make one complex element.
    long long id, od;
    if (p->kind == KIND_R2C) {
        if (idist & 1) {
            set_err("real input distance must be even");
            return -1;
        }
        id = idist ? (long long)idist / 2 : in_dist_c(p);
    } else {
        id = idist ? (long long)idist : in_dist_c(p);
    }
    if (p->kind == KIND_C2R) {
        if (odist & 1) {
            set_err("real output distance must be even");
            return -1;
        }
        od = odist ? (long long)odist / 2 : (long long)p->N;
    } else {
        od = odist ? (long long)odist : out_dist_c(p);
    }
    hipStream_t st = resolve_stream(stream, p->stream);
    const PtrInfo ci = classify_ptr(input), co = classify_ptr(output);
    if (ci.kind == PTR_DEVICE && co.kind == PTR_DEVICE)
        return exec_1d(p, output, input, nullptr, howmany, id, 0, od, 0u, 1.0f, st);
    // Host memory: chunked through device staging, copies overlapping the transforms.
    std::lock_guard<std::mutex> g(p->lock);
    const StagedOperand so{output, (size_t)out_dist_c(p) * 8u, (size_t)od * 8u, co.kind != PTR_DEVICE};
    const StagedOperand si{const_cast<void *>(input), (size_t)in_dist_c(p) * 8u, (size_t)id * 8u,
                           ci.kind != PTR_DEVICE};
    return run_staged(p->pipe, st, so, si, false, howmany, [&](void *o, const void *i, unsigned n) {
        return exec_1d(p, o, i, nullptr, n, id, 0, od, 0u, 1.0f, st);
    });
}

// Converts API distances (elements of each buffer's type) to complex elements.
static int api_dists(const mufft_plan_1d *p, bool adjoint, size_t idist, size_t odist,
                     long long *id, long long *od) {
    // Buffer types: forward c2c C->C, r2c R->C, c2r C->R; the adjoint swaps them.
    const bool in_real = adjoint ? (p->kind == KIND_C2R) : (p->kind == KIND_R2C);
    const bool out_real = adjoint ? (p->kind == KIND_R2C) : (p->kind == KIND_C2R);
    const long long in_packed = adjoint ? (p->kind == KIND_C2R ? (long long)p->N : out_dist_c(p))
                                        : in_dist_c(p);
    const long long out_packed = adjoint ? (p->kind == KIND_R2C ? (long long)p->N
                                            : (p->kind == KIND_C2R ? (long long)p->N + 1 : (long long)p->N))
                                         : out_dist_c(p);
    if ((in_real && (idist & 1)) || (out_real && (odist & 1))) {
        set_err("real buffer distances must be even");
        return -1;
    }
    *id = idist ? (long long)(in_real ? idist / 2 : idist) : in_packed;
    *od = odist ? (long long)(out_real ? odist / 2 : odist) : out_packed;
    return 0;
}

int mufft_hip_execute_adjoint_1d_batched(mufft_plan_1d *p, void *output, const void *input,
                                         unsigned howmany, size_t idist, size_t odist,
                                         void *stream) {
    if (!p || !output || !input) {
        set_err("null argument");
        return -1;
    }
    long long id, od;
    if (api_dists(p, true, idist, odist, &id, &od) || kernel_ptrs(&output, &input)) return -1;
    return exec_adjoint(p, output, input, nullptr, howmany, id, 0, od, 0u, 1.0f,
                        resolve_stream(stream, p->stream));
}

int mufft_hip_execute_plan_1d_mul_batched(mufft_plan_1d *p, void *output, const void *input,
                                          const void *input2, unsigned howmany, size_t idist,
                                          size_t idist2, size_t odist, float norm, int conj2,
                                          int adjoint, void *stream) {
    if (!p || !output || !input || !input2) {
        set_err("null argument");
        return -1;
    }
    // The multiply fuses into the prologue of kernels with complex input.
    const int kind_run = adjoint ? (p->kind == KIND_C2C ? KIND_C2C : (p->kind == KIND_R2C ? KIND_C2R : KIND_R2C))
                                 : p->kind;
    if (kind_run == KIND_R2C) {
        set_err("no fused multiply for transforms with real input");
        return -1;
    }
    long long id, od;
    if (api_dists(p, adjoint != 0, idist, odist, &id, &od) || kernel_ptrs(&output, &input, &input2))
        return -1;
    const unsigned extra = RF_MUL | (conj2 ? RF_CONJ2 : 0u);
    hipStream_t st = resolve_stream(stream, p->stream);
    if (adjoint)
        return exec_adjoint(p, output, input, input2, howmany, id, (long long)idist2, od, extra, norm, st);
    return exec_1d(p, output, input, input2, howmany, id, (long long)idist2, od, extra, norm, st);
}

int mufft_hip_mul_reduce(void *output, const void *a, const void *b, unsigned n, unsigned howmany,
                         size_t adist, size_t bdist, float norm, int conj_b, void *stream) {
    if (!output || !a || !b) {
        set_err("null argument");
        return -1;
    }
    if (kernel_ptrs(&output, &a, &b)) return -1;
    MulArgs m{};
    m.a = static_cast<const float2 *>(a);
    m.b = static_cast<const float2 *>(b);
    m.out = static_cast<float2 *>(output);
    m.n = n;
    m.adist = adist ? (long long)adist : n;
    m.bdist = bdist ? (long long)bdist : n;
    m.batch = (int)howmany;
    m.norm = norm;
    m.flags = conj_b ? RF_CONJ2 : 0u;
    HIP_TRY(launch_mul_reduce(m, static_cast<hipStream_t>(stream)));
    return 0;
}

int mufft_hip_stream_copy(void *dst, const void *src, size_t bytes, void *stream) {
    if (!dst || !src || (bytes & 7)) {
        set_err("null buffer or size not a multiple of 8 bytes");
        return -1;
    }
    if (kernel_ptrs(&dst, &src)) return -1;
    HIP_TRY(launch_stream_copy(static_cast<const float2 *>(src), static_cast<float2 *>(dst),
                               (long long)(bytes / 8), static_cast<hipStream_t>(stream)));
    return 0;
}

int mufft_hip_plan_1d_launches(const mufft_plan_1d *p) {
    if (!p) return -1;
    if (p->eng.row_fused()) return 1;
    const int c2c = p->eng.four_step() ? (p->eng.three_pass() ? 3 : 2) : p->eng.launches_generic();
    return c2c + (p->kind == KIND_C2C ? 0 : 1);
}

}  // extern "C"

// ===========================================================================
// 2D plans
// ===========================================================================
struct mufft_plan_2d {
    unsigned Nx = 0, Ny = 0;   // Nx = complex columns (Nx_real / 2 for r2c / c2r), fft.c:75
    int direction = 0;
    int kind = KIND_C2C;
    bool full = false, zero_pad = false;
    unsigned vertical_nx = 0;  // fft.c:85
    int device = 0;
    hipStream_t stream = nullptr;
    Engine ex, ey;
    float2 *rtw = nullptr;
    DevBuf s0, s1, s2, alt, stage_in, stage_out;
    size_t in_bytes = 0, out_bytes = 0;
    std::mutex lock;
    HostPipe pipe;
};

static void destroy_2d(mufft_plan_2d *p) {
    if (!p) return;
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    p->ex.release();
    p->ey.release();
    if (p->rtw) (void)hipFree(p->rtw);
    p->s0.release();
    p->s1.release();
    p->s2.release();
    p->alt.release();
    p->stage_in.release();
    p->stage_out.release();
    p->pipe.release();
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

static mufft_plan_2d *create_2d_impl(int kind, unsigned Nx_c, unsigned Ny, int direction,
                                unsigned flags) {
    if (!is_pow2(Nx_c) || !is_pow2(Ny) || Nx_c < 2 || Ny < 2 || Nx_c > (1u << 20) ||
        Ny > (1u << 20)) {
        set_err("invalid transform size");
        return nullptr;
    }
    mufft_plan_2d *p = new (std::nothrow) mufft_plan_2d;
    if (!p) return nullptr;
    p->Nx = Nx_c;
    p->Ny = Ny;
    p->kind = kind;
    p->direction = direction;
    p->full = kind == KIND_R2C && (flags & MUFFT_FLAG_FULL_R2C);
    p->zero_pad = (flags & MUFFT_FLAG_ZERO_PAD_UPPER_HALF) != 0;
    if (kind == KIND_C2C) p->vertical_nx = Nx_c;
    else if (kind == KIND_R2C) p->vertical_nx = p->full ? 2 * Nx_c : Nx_c + 1;
    else p->vertical_nx = Nx_c + 1;
    if (hipGetDevice(&p->device) != hipSuccess ||
        hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        set_err("no usable HIP device");
        p->stream = nullptr;
        destroy_2d(p);
        return nullptr;
    }
    if (p->ex.init((int)Nx_c, direction, kind, true, false, p->stream) ||
        p->ey.init((int)Ny, direction, KIND_C2C, false, true, p->stream) ||
        (kind != KIND_C2C && build_resolve_table((int)Nx_c, direction, &p->rtw, p->stream))) {
        destroy_2d(p);
        return nullptr;
    }
    if (hipStreamSynchronize(p->stream) != hipSuccess) {
        set_err("twiddle generation failed");
        destroy_2d(p);
        return nullptr;
    }
    const size_t px = (size_t)Nx_c * Ny;
    if (kind == KIND_C2C) {
        p->in_bytes = px * 8;
        p->out_bytes = px * 8;
    } else if (kind == KIND_R2C) {
        p->in_bytes = px * 8;          // Nx_real * Ny floats
        p->out_bytes = 2 * px * 8;     // Nx_real complex per row
    } else {
        p->in_bytes = 2 * px * 8;
        p->out_bytes = px * 8;         // Nx_real * Ny floats
    }
    return p;
}

static mufft_plan_2d *create_2d(int kind, unsigned Nx_c, unsigned Ny, int direction,
                                unsigned flags) {
    return guarded<mufft_plan_2d *>(nullptr, [&] { return create_2d_impl(kind, Nx_c, Ny, direction, flags); });
}

// howmany packed images.
static int exec_2d(mufft_plan_2d *p, void *out_, const void *in_, unsigned howmany,
                   hipStream_t st) {
    const long long Nx = p->Nx, Ny = p->Ny;
    const int hm = (int)howmany;
    const float2 *in = static_cast<const float2 *>(in_);
    float2 *out = static_cast<float2 *>(out_);
    const unsigned zp = p->zero_pad ? RF_ZERO_PAD : 0u;
    if (p->kind == KIND_C2C || p->kind == KIND_R2C) {
        const long long rs = p->kind == KIND_C2C ? Nx : 2 * Nx;   // output row stride
        const long long img = rs * Ny;
        float2 *rows_dst = out;
        if (!p->ey.col_fused()) {
            if (p->s2.ensure(sizeof(float2) * (size_t)img * hm)) return -1;
            rows_dst = p->s2.f2();
        }
        // The column pass reads the rows' output next: keep it cacheable.
        const unsigned f = zp | (p->full ? RF_FULL : 0u) | RF_KEEP_OUT;
        if (run_rows(p->ex, p->kind, p->rtw, in, Nx, nullptr, 0, rows_dst, rs, (int)(Ny * hm), f,
                     1.0f, p->s0, p->s1, st))
            return -1;
        return run_cols(p->ey, rows_dst, rs, img, out, rs, img, (int)p->vertical_nx, hm, p->alt, st);
    }
    // c2r: columns first (fft.c:863-916), then the per-row resolve and rows.
    const long long in_rs = 2 * Nx, in_img = in_rs * Ny;
    const long long mid_rs = Nx + 1, mid_img = mid_rs * Ny;
    if (p->s2.ensure(sizeof(float2) * (size_t)mid_img * hm)) return -1;
    if (run_cols(p->ey, in, in_rs, in_img, p->s2.f2(), mid_rs, mid_img, (int)(Nx + 1), hm, p->alt,
                 st))
        return -1;
    return run_rows(p->ex, KIND_C2R, p->rtw, p->s2.f2(), mid_rs, nullptr, 0, out, Nx,
                    (int)(Ny * hm), zp, 1.0f, p->s0, p->s1, st);
}

extern "C" {

mufft_plan_2d *mufft_create_plan_2d_c2c(unsigned Nx, unsigned Ny, int direction, unsigned flags) {
    if (direction != MUFFT_FORWARD && direction != MUFFT_INVERSE) {
        set_err("invalid direction");
        return nullptr;
    }
    return create_2d(KIND_C2C, Nx, Ny, direction, flags);
}

mufft_plan_2d *mufft_create_plan_2d_r2c(unsigned Nx, unsigned Ny, unsigned flags) {
    if (!is_pow2(Nx) || Nx < 4) {
        set_err("invalid transform size");
        return nullptr;
    }
    return create_2d(KIND_R2C, Nx / 2, Ny, MUFFT_FORWARD, flags);
}

mufft_plan_2d *mufft_create_plan_2d_c2r(unsigned Nx, unsigned Ny, unsigned flags) {
    if (!is_pow2(Nx) || Nx < 4) {
        set_err("invalid transform size");
        return nullptr;
    }
    return create_2d(KIND_C2R, Nx / 2, Ny, MUFFT_INVERSE, flags);
}

void mufft_execute_plan_2d(mufft_plan_2d *p, void *MUFFT_RESTRICT output,
                           const void *MUFFT_RESTRICT input) {
    if (!p) return;
    std::lock_guard<std::mutex> g(p->lock);
    int dev_prev = 0;
    (void)hipGetDevice(&dev_prev);
    if (dev_prev != p->device) (void)hipSetDevice(p->device);
    const void *din = input;
    void *dout = output;
    const bool hin = is_host_ptr(input), hout = is_host_ptr(output);
    int rc = 0;
    if (hin) {
        rc |= p->stage_in.ensure(p->in_bytes);
        if (!rc) rc |= hipMemcpyAsync(p->stage_in.p, input, p->in_bytes, hipMemcpyHostToDevice, p->stream) != hipSuccess;
        din = p->stage_in.p;
    }
    if (!rc && hout) {
        rc |= p->stage_out.ensure(p->out_bytes);
        dout = p->stage_out.p;
        // Columns the transform leaves untouched come back as they went in.
        if (!rc) rc |= hipMemcpyAsync(p->stage_out.p, output, p->out_bytes, hipMemcpyHostToDevice, p->stream) != hipSuccess;
    }
    if (!rc) rc = exec_2d(p, dout, din, 1, p->stream);
    if (!rc && hout)
        rc |= hipMemcpyAsync(output, p->stage_out.p, p->out_bytes, hipMemcpyDeviceToHost, p->stream) != hipSuccess;
    if (hipStreamSynchronize(p->stream) != hipSuccess || rc) {
        if (!rc) set_err("execution failed: %s", hipGetErrorString(hipGetLastError()));
        fprintf(stderr, "mufft_hip: %s\n", g_err);
    }
    if (dev_prev != p->device) (void)hipSetDevice(dev_prev);
}

void mufft_free_plan_2d(mufft_plan_2d *plan) { destroy_2d(plan); }

int mufft_hip_execute_plan_2d_batched(mufft_plan_2d *p, void *output, const void *input,
                                      unsigned howmany, void *stream) {
    if (!p || !output || !input) {
        set_err("null argument");
        return -1;
    }
    hipStream_t st = resolve_stream(stream, p->stream);
    const PtrInfo ci = classify_ptr(input), co = classify_ptr(output);
    if (ci.kind == PTR_DEVICE && co.kind == PTR_DEVICE) return exec_2d(p, output, input, howmany, st);
    // Host memory: chunked through device staging, copies overlapping the
    // transforms.  A padded r2c output keeps the columns the transform does
    // not write, so those images are uploaded first.
    std::lock_guard<std::mutex> g(p->lock);
    const bool preload = p->kind == KIND_R2C && !p->full;
    const StagedOperand so{output, p->out_bytes, p->out_bytes, co.kind != PTR_DEVICE};
    const StagedOperand si{const_cast<void *>(input), p->in_bytes, p->in_bytes, ci.kind != PTR_DEVICE};
    return run_staged(p->pipe, st, so, si, preload, howmany, [&](void *o, const void *i, unsigned n) {
        return exec_2d(p, o, i, n, st);
    });
}

int mufft_hip_plan_2d_launches(const mufft_plan_2d *p) {
    if (!p) return -1;
    int rows = p->ex.row_fused() ? 1
             : (p->ex.four_step() ? (p->ex.three_pass() ? 3 : 2) : p->ex.launches_generic()) + (p->kind == KIND_C2C ? 0 : 1);
    int cols = p->ey.col_fused() ? 1 : p->ey.launches_generic();
    return rows + cols;
}

}  // extern "C"

// ===========================================================================
// Convolution plans (fft.c:565-624, 794-809)
// ===========================================================================
struct mufft_plan_conv {
    unsigned N = 0;
    unsigned method = 0;
    bool stereo = false;
    mufft_plan_1d *plans[2] = {nullptr, nullptr};
    mufft_plan_1d *output_plan = nullptr;
    size_t block_size = 0;
    float normalization = 0.f;
    unsigned conv_multiply_n = 0;
};

extern "C" {

void mufft_free_plan_conv(mufft_plan_conv *plan) {
    if (!plan) return;
    destroy_1d(plan->plans[0]);
    destroy_1d(plan->plans[1]);
    destroy_1d(plan->output_plan);
    delete plan;
}

static mufft_plan_conv *mufft_create_plan_conv_impl(unsigned N, unsigned flags, unsigned method) {
    if (!is_pow2(N) || N < 4) {
        set_err("invalid transform size");
        return nullptr;
    }
    mufft_plan_conv *c = new (std::nothrow) mufft_plan_conv;
    if (!c) return nullptr;
    const unsigned ef = (method & MUFFT_CONV_METHOD_FLAG_ZERO_PAD_UPPER_HALF_FIRST) ? MUFFT_FLAG_ZERO_PAD_UPPER_HALF : 0u;
    const unsigned es = (method & MUFFT_CONV_METHOD_FLAG_ZERO_PAD_UPPER_HALF_SECOND) ? MUFFT_FLAG_ZERO_PAD_UPPER_HALF : 0u;
    c->N = N;
    c->method = method;
    c->stereo = (method & 1) == MUFFT_CONV_METHOD_FLAG_STEREO_MONO;
    if (!c->stereo) {
        c->block_size = (N / 2 + 4) * 8;   // fft.c:586 (MUFFT_PADDING_COMPLEX_SAMPLES = 4)
        c->plans[0] = mufft_create_plan_1d_r2c(N, flags | ef);
        c->plans[1] = mufft_create_plan_1d_r2c(N, flags | es);
        c->output_plan = mufft_create_plan_1d_c2r(N, flags);
        c->conv_multiply_n = N / 2 + 1;
    } else {
        c->block_size = (size_t)N * 8;
        c->plans[0] = mufft_create_plan_1d_c2c(N, MUFFT_FORWARD, flags | ef);
        c->plans[1] = mufft_create_plan_1d_r2c(N, flags | es | MUFFT_FLAG_FULL_R2C);
        c->output_plan = mufft_create_plan_1d_c2c(N, MUFFT_INVERSE, flags);
        c->conv_multiply_n = N;
    }
    c->normalization = 1.0f / (float)N;
    if (!c->plans[0] || !c->plans[1] || !c->output_plan) {
        mufft_free_plan_conv(c);
        return nullptr;
    }
    return c;
}


mufft_plan_conv *mufft_create_plan_conv(unsigned N, unsigned flags, unsigned method) {
    return guarded<mufft_plan_conv *>(nullptr, [&] { return mufft_create_plan_conv_impl(N, flags, method); });
}

size_t mufft_conv_get_transformed_block_size(mufft_plan_conv *plan) {
    return plan ? plan->block_size : 0;
}

void mufft_execute_conv_input(mufft_plan_conv *plan, unsigned block, void *output, const void *input) {
    if (!plan || block > 1) {
        set_err("invalid convolution block");
        return;
    }
    mufft_execute_plan_1d(plan->plans[block], output, input);
}

void mufft_execute_conv_output(mufft_plan_conv *plan, void *output, const void *input_first,
                               const void *input_second) {
    if (!plan) return;
    // The reference multiplies into conv_block and then runs the inverse plan
    // (fft.c:804-809); here the multiply is the inverse kernel's prologue.
    mufft_plan_1d *op = plan->output_plan;
    const size_t nb = (size_t)plan->conv_multiply_n * 8;
    exec_1d_dropin(op, output, input_first, input_second, nb, RF_MUL, plan->normalization);
}

int mufft_hip_execute_conv_input_batched(mufft_plan_conv *plan, unsigned block, void *output,
                                         const void *input, unsigned howmany, size_t idist,
                                         size_t odist, void *stream) {
    if (!plan || block > 1) {
        set_err("invalid convolution block");
        return -1;
    }
    mufft_plan_1d *p = plan->plans[block];
    const size_t od = odist ? odist : plan->block_size / 8;
    // Input distances of real inputs are in floats (mufft_hip_execute_plan_1d_batched).
    return mufft_hip_execute_plan_1d_batched(p, output, input, howmany, idist, od, stream);
}

int mufft_hip_execute_conv_output_batched(mufft_plan_conv *plan, void *output,
                                          const void *first, const void *second,
                                          unsigned howmany, size_t odist, size_t dist_first,
                                          size_t dist_second, void *stream) {
    if (!plan || !output || !first || !second) {
        set_err("null argument");
        return -1;
    }
    mufft_plan_1d *op = plan->output_plan;
    const long long bd = (long long)(plan->block_size / 8);
    const long long d1 = dist_first ? (long long)dist_first : bd;
    const long long d2 = dist_second ? (long long)dist_second : bd;
    long long od;
    if (op->kind == KIND_C2R) {
        if (odist & 1) {
            set_err("real output distance must be even");
            return -1;
        }
        od = odist ? (long long)odist / 2 : (long long)op->N;
    } else {
        od = odist ? (long long)odist : (long long)op->N;
    }
    if (kernel_ptrs(&output, &first, &second)) return -1;
    return exec_1d(op, output, first, second, howmany, d1, d2, od, RF_MUL, plan->normalization,
                   resolve_stream(stream, op->stream));
}

}  // extern "C"

// ===========================================================================
// Complex multiply function pointer (fft.c:549-563) and memory (fft.c:1032-1084)
// ===========================================================================
namespace {

struct MulContext {
    std::mutex lock;
    hipStream_t stream = nullptr;
    DevBuf a, b, o;
};

MulContext &mul_ctx() {
    static MulContext ctx;
    return ctx;
}

void gpu_convolve(void *output, const void *a, const void *b, float normalization, unsigned samples) {
    MulContext &c = mul_ctx();
    std::lock_guard<std::mutex> g(c.lock);
    if (!c.stream && hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) != hipSuccess) {
        set_err("no usable HIP device");
        c.stream = nullptr;
        return;
    }
    const size_t bytes = (size_t)samples * 8;
    const void *da = a, *db = b;
    void *dout = output;
    int rc = 0;
    if (is_host_ptr(a)) {
        rc |= c.a.ensure(bytes);
        if (!rc) rc |= hipMemcpyAsync(c.a.p, a, bytes, hipMemcpyHostToDevice, c.stream) != hipSuccess;
        da = c.a.p;
    }
    if (!rc && is_host_ptr(b)) {
        rc |= c.b.ensure(bytes);
        if (!rc) rc |= hipMemcpyAsync(c.b.p, b, bytes, hipMemcpyHostToDevice, c.stream) != hipSuccess;
        db = c.b.p;
    }
    const bool hout = is_host_ptr(output);
    if (!rc && hout) {
        rc |= c.o.ensure(bytes);
        dout = c.o.p;
    }
    if (!rc) {
        MulArgs m{};
        m.a = static_cast<const float2 *>(da);
        m.b = static_cast<const float2 *>(db);
        m.out = static_cast<float2 *>(dout);
        m.n = samples;
        m.batch = 1;
        m.norm = normalization;
        rc |= launch_mul(m, c.stream) != hipSuccess;
    }
    if (!rc && hout)
        rc |= hipMemcpyAsync(output, c.o.p, bytes, hipMemcpyDeviceToHost, c.stream) != hipSuccess;
    if (hipStreamSynchronize(c.stream) != hipSuccess || rc) {
        if (!rc) set_err("convolve failed");
        fprintf(stderr, "mufft_hip: %s\n", g_err);
    }
}

}  // namespace

extern "C" {

mufft_convolve_func mufft_get_convolve_func(unsigned flags) {
    (void)flags;  // CPU ISA bits have no meaning on the GPU
    return &gpu_convolve;
}

void *mufft_alloc(size_t size) {
    void *p = nullptr;
    if (hipHostMalloc(&p, size ? size : 1, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        set_err("hipHostMalloc(%zu) failed", size);
        return nullptr;
    }
    return p;
}

void *mufft_calloc(size_t size) {
    void *p = mufft_alloc(size);
    if (p) memset(p, 0, size);
    return p;
}

void mufft_free(void *ptr) {
    if (ptr) (void)hipHostFree(ptr);
}

const char *mufft_hip_get_last_error(void) { return g_err; }

int mufft_hip_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

void *mufft_hip_malloc(size_t bytes) {
    void *p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
        (void)hipGetLastError();
        set_err("hipMalloc(%zu) failed", bytes);
        return nullptr;
    }
    return p;
}

void mufft_hip_free(void *ptr) {
    if (ptr) (void)hipFree(ptr);
}

int mufft_hip_memcpy(void *dst, const void *src, size_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDefault));
    return 0;
}

int mufft_hip_synchronize(void *stream) {
    if (stream) {
        HIP_TRY(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    } else {
        HIP_TRY(hipDeviceSynchronize());
    }
    return 0;
}

static int mufft_hip_twiddle_table_impl(unsigned N, int direction, float *host_out, unsigned count) {
    if (!is_pow2(N) || N < 2 || N > (1u << ROW_MAX_LOG2)) {
        set_err("invalid transform size");
        return -1;
    }
    hipStream_t s = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    std::vector<int> seq = fused_sequence((int)N, rv_R((int)N, rv_default((int)N)), true);
    float2 *t = nullptr;
    int size = seq.size() > 1 ? (int)N - seq[0] : 0;
    int rc = build_stockham_table((int)N, direction, seq, &t, s);
    if (!rc && t && host_out && count) {
        const size_t n = count < (unsigned)size ? count : (unsigned)size;
        rc = hipMemcpyAsync(host_out, t, n * sizeof(float2), hipMemcpyDeviceToHost, s) != hipSuccess;
    }
    rc |= hipStreamSynchronize(s) != hipSuccess;
    if (t) (void)hipFree(t);
    (void)hipStreamDestroy(s);
    if (rc) {
        set_err("twiddle table generation failed");
        return -1;
    }
    return size;
}


int mufft_hip_twiddle_table(unsigned N, int direction, float *host_out, unsigned count) {
    return guarded<int>(-1, [&] { return mufft_hip_twiddle_table_impl(N, direction, host_out, count); });
}

}  // extern "C"

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

struct TuneRecord {
    int logn, tk, variant;
    double us;
};
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
    return row_variant_list(l, out, max);
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
    std::lock_guard<std::mutex> g(g_var_lock);
    g_variant[l][tk] = variant;
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
    const int nv = rc ? 0 : std::min(row_variant_list(l, vars, 64), 64);
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
        std::lock_guard<std::mutex> g(g_var_lock);
        g_variant[l][tk] = best;
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

// Tune cache: one line per entry, "log2n kind variant microseconds".
static int mufft_hip_tune_save_impl(const char *path) {
    FILE *f = fopen(path, "w");
    if (!f) {
        set_err("cannot write %s", path);
        return -1;
    }
    std::lock_guard<std::mutex> g(g_var_lock);
    fprintf(f, "# mufft_hip tune cache: log2(complex length) kind(0 c2c fwd,1 c2c inv,2 r2c,3 c2r) variant us\n");
    for (int l = 0; l <= ROW_MAX_LOG2; ++l)
        for (int tk = 0; tk < 4; ++tk)
            if (g_variant[l][tk]) {
                double us = 0;
                for (const auto &t : g_tune_records)
                    if (t.logn == l && t.tk == tk) us = t.us;
                fprintf(f, "%d %d %d %.3f\n", l, tk, g_variant[l][tk], us);
            }
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
        if (l < 1 || l > ROW_MAX_LOG2 || tk < 0 || tk > 3 || !row_variant_registered(l, var)) continue;
        std::lock_guard<std::mutex> g(g_var_lock);
        g_variant[l][tk] = var;
        ++n;
    }
    fclose(f);
    return n;
}


int mufft_hip_tune_load(const char *path) {
    return guarded<int>(-1, [&] { return mufft_hip_tune_load_impl(path); });
}

}  // extern "C"

// ===========================================================================
// Uneven convolution batch: two persistent launches (forward tasks, then
// inverse tasks) for many convolutions of mixed sizes and methods, replayed
// from a captured hipGraph.
// ===========================================================================
struct mufft_hip_conv_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<CBTask> tasks;
    std::vector<CBJob> jobs;
    CBArgs args{};
    int grid = 0;
    DevBuf d_tasks, d_jobs, d_flags, d_spectra;
    std::vector<float2 *> tables;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

void destroy_batch(mufft_hip_conv_batch *b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (float2 *t : b->tables)
        if (t) (void)hipFree(t);
    b->d_tasks.release();
    b->d_jobs.release();
    b->d_flags.release();
    b->d_spectra.release();
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

}  // namespace

extern "C" {

void mufft_hip_free_conv_batch(mufft_hip_conv_batch *b) { destroy_batch(b); }

static mufft_hip_conv_batch *mufft_hip_create_conv_batch_impl(const mufft_hip_conv_job *jobs, unsigned count,
                                                  unsigned flags) {
    (void)flags;  // CPU ISA bits have no meaning here
    if (!jobs || count == 0) {
        set_err("empty job list");
        return nullptr;
    }
    // Validate: mono N in [8, 65536], stereo N in [8, 32768], powers of two.
    for (unsigned i = 0; i < count; ++i) {
        const unsigned N = jobs[i].N;
        const bool stereo = (jobs[i].method & 1) == MUFFT_CONV_METHOD_FLAG_STEREO_MONO;
        if (!is_pow2(N) || N < 8 || N > (stereo ? 32768u : 65536u) || !jobs[i].first ||
            !jobs[i].second || !jobs[i].output) {
            set_err("job %u: unsupported size or null buffer", i);
            return nullptr;
        }
    }
    mufft_hip_conv_batch *b = new (std::nothrow) mufft_hip_conv_batch;
    if (!b) return nullptr;
    if (hipGetDevice(&b->device) != hipSuccess ||
        hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        set_err("no usable HIP device");
        b->stream = nullptr;
        destroy_batch(b);
        return nullptr;
    }
    // Order jobs so equal shapes are adjacent: (method, N) descending.
    std::vector<unsigned> order(count);
    for (unsigned i = 0; i < count; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](unsigned x, unsigned y) {
        const unsigned sx = jobs[x].method & 1, sy = jobs[y].method & 1;
        if (jobs[x].N != jobs[y].N) return jobs[x].N > jobs[y].N;
        return sx < sy;
    });
    // Spectra: mono (N/2 + 1) bins each; stereo N bins each.
    std::vector<size_t> spec_off(count);
    size_t spec_points = 0;
    for (unsigned r = 0; r < count; ++r) {
        const mufft_hip_conv_job &j = jobs[order[r]];
        const bool stereo = (j.method & 1) == MUFFT_CONV_METHOD_FLAG_STEREO_MONO;
        spec_off[r] = spec_points;
        const size_t n = stereo ? j.N : j.N / 2 + 1;
        spec_points += 2 * ((n + 7) & ~(size_t)7);
    }
    if (b->d_spectra.ensure(spec_points * sizeof(float2))) {
        destroy_batch(b);
        return nullptr;
    }
    b->jobs.resize(count);
    for (unsigned r = 0; r < count; ++r) {
        const mufft_hip_conv_job &j = jobs[order[r]];
        const bool stereo = (j.method & 1) == MUFFT_CONV_METHOD_FLAG_STEREO_MONO;
        const size_t n = stereo ? j.N : j.N / 2 + 1;
        CBJob &d = b->jobs[r];
        d.first = static_cast<const float2 *>(j.first);
        d.second = static_cast<const float2 *>(j.second);
        d.spec1 = b->d_spectra.f2() + spec_off[r];
        d.spec2 = d.spec1 + ((n + 7) & ~(size_t)7);
        d.out = static_cast<float2 *>(j.output);
    }
    // Twiddle tables per (log2 n, kind, direction) of the persistent kernel's variants.
    // Transforms of 2^CB_SPLIT_LOG2 points use the 2^(CB_SPLIT_LOG2-1)-point
    // sub-transform's table plus the split twiddles.
    struct TabKey {
        int logn, kind, dir;
        float2 *tw, *rtw, *stw;
    };
    std::vector<TabKey> tabs;
    auto table_for = [&](int logn, int kind, int dir, float2 **tw, float2 **rtw, float2 **stw) -> int {
        for (const auto &t : tabs)
            if (t.logn == logn && t.kind == kind && t.dir == dir) {
                *tw = t.tw;
                *rtw = t.rtw;
                *stw = t.stw;
                return 0;
            }
        TabKey k{logn, kind, dir, nullptr, nullptr, nullptr};
        const int n = 1 << logn;
        const int n2 = logn >= CB_SPLIT_LOG2 ? n / 2 : n;
        if (build_stockham_table(n2, dir, fused_sequence(n2, cb_points_per_thread(logn), true), &k.tw, b->stream))
            return -1;
        b->tables.push_back(k.tw);
        if (n2 != n) {
            if (build_split_table(n, n2, dir, &k.stw, b->stream)) return -1;
            b->tables.push_back(k.stw);
        }
        if (kind != KIND_C2C) {
            if (build_resolve_table(n, dir, &k.rtw, b->stream)) return -1;
            b->tables.push_back(k.rtw);
        }
        tabs.push_back(k);
        *tw = k.tw;
        *rtw = k.rtw;
        *stw = k.stw;
        return 0;
    };
    // Tasks: runs of equal (method, N) jobs, split by workgroup capacity.
    std::vector<CBTask> fwd, inv;
    for (unsigned r = 0; r < count;) {
        const mufft_hip_conv_job &j0 = jobs[order[r]];
        const bool stereo = (j0.method & 1) == MUFFT_CONV_METHOD_FLAG_STEREO_MONO;
        unsigned e = r;
        while (e < count && jobs[order[e]].N == j0.N && (jobs[order[e]].method & 1) == (j0.method & 1)) ++e;
        const int lN = ilog2(j0.N);
        // (type, kind, dir, logn, zero-pad, full, mul)
        struct Shape {
            int type, kind, dir, logn;
            unsigned flags;
        };
        std::vector<Shape> shapes;
        if (!stereo) {
            shapes.push_back({CB_FWD_FIRST, KIND_R2C, -1, lN - 1, 0u});
            shapes.push_back({CB_FWD_SECOND, KIND_R2C, -1, lN - 1, 0u});
            shapes.push_back({CB_INV, KIND_C2R, 1, lN - 1, (unsigned)RF_MUL});
        } else {
            shapes.push_back({CB_FWD_FIRST, KIND_C2C, -1, lN, 0u});
            shapes.push_back({CB_FWD_SECOND, KIND_R2C, -1, lN - 1, (unsigned)RF_FULL});
            shapes.push_back({CB_INV, KIND_C2C, 1, lN, (unsigned)RF_MUL});
        }
        for (Shape sh : shapes) {
            // zero padding per operand follows each job's method bits; a task
            // mixes only jobs with equal bits, so split runs on them too.
            const int cap = cb_jobs_per_task(sh.logn);
            for (unsigned q = r; q < e;) {
                const unsigned m0 = jobs[order[q]].method;
                unsigned q2 = q;
                while (q2 < e && q2 - q < (unsigned)cap && jobs[order[q2]].method == m0) ++q2;
                CBTask t{};
                t.type = sh.type;
                t.kind = sh.kind;
                t.dir = sh.dir;
                t.logn = sh.logn;
                t.job0 = (int)q;
                t.njobs = (int)(q2 - q);
                t.flags = sh.flags;
                if (sh.type != CB_INV) t.flags |= RF_KEEP_OUT;   // spectra are read by the inverse phase
                if (sh.type == CB_FWD_FIRST && (m0 & MUFFT_CONV_METHOD_FLAG_ZERO_PAD_UPPER_HALF_FIRST))
                    t.flags |= RF_ZERO_PAD;
                if (sh.type == CB_FWD_SECOND && (m0 & MUFFT_CONV_METHOD_FLAG_ZERO_PAD_UPPER_HALF_SECOND))
                    t.flags |= RF_ZERO_PAD;
                t.norm = 1.0f / (float)j0.N;
                float2 *tw = nullptr, *rtw = nullptr, *stw = nullptr;
                if (table_for(t.logn, t.kind, t.dir, &tw, &rtw, &stw)) {
                    destroy_batch(b);
                    return nullptr;
                }
                t.tw = tw;
                t.rtw = rtw;
                t.stw = stw;
                t.k1 = 0;
                (sh.type == CB_INV ? inv : fwd).push_back(t);
                if (t.logn >= CB_SPLIT_LOG2) {   // second half of the split transform
                    t.k1 = 1;
                    (sh.type == CB_INV ? inv : fwd).push_back(t);
                }
                q = q2;
            }
        }
        r = e;
    }
    // Longest-processing-time first inside each phase.
    auto cost = [](const CBTask &t) { return (double)t.njobs * (double)(1 << t.logn) * t.logn; };
    std::stable_sort(fwd.begin(), fwd.end(), [&](const CBTask &x, const CBTask &y) { return cost(x) > cost(y); });
    std::stable_sort(inv.begin(), inv.end(), [&](const CBTask &x, const CBTask &y) { return cost(x) > cost(y); });
    b->tasks = fwd;
    b->tasks.insert(b->tasks.end(), inv.begin(), inv.end());

    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, b->device) != hipSuccess) {
        set_err("hipGetDeviceProperties failed");
        destroy_batch(b);
        return nullptr;
    }
    b->grid = std::max(1, std::min<int>(prop.multiProcessorCount, (int)b->tasks.size()));
    if (b->d_tasks.ensure(b->tasks.size() * sizeof(CBTask)) ||
        b->d_jobs.ensure(b->jobs.size() * sizeof(CBJob)) ||
        b->d_flags.ensure(4 * sizeof(int)) ||
        hipMemcpyAsync(b->d_tasks.p, b->tasks.data(), b->tasks.size() * sizeof(CBTask),
                       hipMemcpyHostToDevice, b->stream) != hipSuccess ||
        hipMemcpyAsync(b->d_jobs.p, b->jobs.data(), b->jobs.size() * sizeof(CBJob),
                       hipMemcpyHostToDevice, b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess) {
        if (!g_err[0]) set_err("batch setup failed");
        destroy_batch(b);
        return nullptr;
    }
    CBArgs &a = b->args;
    a.tasks = static_cast<const CBTask *>(b->d_tasks.p);
    a.nfwd = (int)fwd.size();
    a.ntasks = (int)b->tasks.size();
    a.jobs = static_cast<const CBJob *>(b->d_jobs.p);
    a.njobs = (int)count;
    a.head = static_cast<int *>(b->d_flags.p);
    a.done = a.head + 2;
    // Capture reset + the two persistent launches once; executions replay the graph.
    if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess ||
        launch_conv_batch(a, b->grid, b->stream) != hipSuccess ||
        hipStreamEndCapture(b->stream, &b->graph) != hipSuccess ||
        hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGetLastError();
        set_err("graph capture of the batch failed");
        destroy_batch(b);
        return nullptr;
    }
    return b;
}


mufft_hip_conv_batch *mufft_hip_create_conv_batch(const mufft_hip_conv_job *jobs, unsigned count,
                                                  unsigned flags) {
    return guarded<mufft_hip_conv_batch *>(nullptr, [&] { return mufft_hip_create_conv_batch_impl(jobs, count, flags); });
}

int mufft_hip_execute_conv_batch(mufft_hip_conv_batch *b, void *stream) {
    if (!b) {
        set_err("null batch");
        return -1;
    }
    HIP_TRY(hipGraphLaunch(b->exec, stream ? static_cast<hipStream_t>(stream) : b->stream));
    if (!stream) HIP_TRY(hipStreamSynchronize(b->stream));
    return 0;
}

// Launches the batch without the graph (same kernels, same stream order).
int mufft_hip_execute_conv_batch_direct(mufft_hip_conv_batch *b, void *stream) {
    if (!b) {
        set_err("null batch");
        return -1;
    }
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : b->stream;
    HIP_TRY(launch_conv_batch(b->args, b->grid, s));
    if (!stream) HIP_TRY(hipStreamSynchronize(s));
    return 0;
}

// 0 when the last execution ran every task, 1 otherwise.  Synchronises the
// device (the batch may have run on a caller's stream).
int mufft_hip_conv_batch_status(mufft_hip_conv_batch *b) {
    if (!b) return -1;
    int done = 0;
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(&done, b->args.done, sizeof(int), hipMemcpyDeviceToHost));
    return done == b->args.ntasks ? 0 : 1;
}

int mufft_hip_conv_batch_info(const mufft_hip_conv_batch *b, int *ntasks, int *grid) {
    if (!b) return -1;
    if (ntasks) *ntasks = (int)b->tasks.size();
    if (grid) *grid = b->grid;
    return 0;
}

}  // extern "C"

// ===========================================================================
// Streaming overlap-add FIR filtering (the filter spectrum is computed once,
// as the reference's callers cache it, bench.c:251-257, fft.h:170-173).
// ===========================================================================
struct mufft_hip_ola {
    unsigned N = 0;
    mufft_plan_1d *fwd = nullptr;   // r2c, zero-padded input (hop N/2)
    mufft_plan_1d *inv = nullptr;   // c2r with the fused filter multiply
    DevBuf H, X, Y, tail[2];
    int cur = 0;
    size_t cap_blocks = 0;
};

extern "C" {

void mufft_hip_free_ola(mufft_hip_ola *o) {
    if (!o) return;
    if (o->fwd) (void)hipStreamSynchronize(o->fwd->stream);
    (void)hipDeviceSynchronize();
    destroy_1d(o->fwd);
    destroy_1d(o->inv);
    o->H.release();
    o->X.release();
    o->Y.release();
    o->tail[0].release();
    o->tail[1].release();
    delete o;
}

static mufft_hip_ola *mufft_hip_create_ola_impl(unsigned N, const float *taps, unsigned ntaps) {
    if (!is_pow2(N) || N < 4 || N > 65536 || !taps || ntaps == 0 || ntaps > N / 2) {
        set_err("invalid size or taps (N power of two in [4, 65536], 1 <= taps <= N/2)");
        return nullptr;
    }
    mufft_hip_ola *o = new (std::nothrow) mufft_hip_ola;
    if (!o) return nullptr;
    o->N = N;
    o->fwd = mufft_create_plan_1d_r2c(N, MUFFT_FLAG_ZERO_PAD_UPPER_HALF);
    o->inv = mufft_create_plan_1d_c2r(N, 0);
    const size_t h = N / 2;
    if (!o->fwd || !o->inv || o->H.ensure((h + 1) * sizeof(float2)) ||
        o->tail[0].ensure(h * sizeof(float)) || o->tail[1].ensure(h * sizeof(float))) {
        mufft_hip_free_ola(o);
        return nullptr;
    }
    std::vector<float> hp(h, 0.0f);
    const PtrInfo tp = classify_ptr(taps);
    if (tp.kind == PTR_PAGEABLE) {
        memcpy(hp.data(), taps, ntaps * sizeof(float));
    } else if (hipMemcpy(hp.data(), taps, ntaps * sizeof(float), hipMemcpyDefault) != hipSuccess) {
        set_err("copy of the filter taps failed");
        mufft_hip_free_ola(o);
        return nullptr;
    }
    DevBuf dh;
    hipStream_t st = o->fwd->stream;
    if (dh.ensure(h * sizeof(float)) ||
        hipMemcpy(dh.p, hp.data(), h * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        exec_1d(o->fwd, o->H.p, dh.p, nullptr, 1, (long long)h / 2, 0, (long long)h + 1, 0u, 1.0f, st) ||
        hipMemsetAsync(o->tail[0].p, 0, h * sizeof(float), st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        dh.release();
        if (!g_err[0]) set_err("filter transform failed");
        mufft_hip_free_ola(o);
        return nullptr;
    }
    dh.release();
    return o;
}


mufft_hip_ola *mufft_hip_create_ola(unsigned N, const float *taps, unsigned ntaps) {
    return guarded<mufft_hip_ola *>(nullptr, [&] { return mufft_hip_create_ola_impl(N, taps, ntaps); });
}

int mufft_hip_reset_ola(mufft_hip_ola *o, void *stream) {
    if (!o) return -1;
    HIP_TRY(hipMemsetAsync(o->tail[o->cur].p, 0, o->N / 2 * sizeof(float),
                           static_cast<hipStream_t>(stream)));
    return 0;
}

int mufft_hip_execute_ola(mufft_hip_ola *o, float *output, const float *input, unsigned nblocks,
                          void *stream) {
    if (!o || !output || !input) {
        set_err("null argument");
        return -1;
    }
    if (nblocks == 0) return 0;
    if (kernel_ptrs(&output, &input)) return -1;
    const long long n = o->N, h = n / 2;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (o->X.ensure(sizeof(float2) * (size_t)(h + 1) * nblocks) ||
        o->Y.ensure(sizeof(float) * (size_t)n * nblocks))
        return -1;
    // 1. spectra of every hop (zero padded to N); block b starts at input + b*h.
    if (exec_1d(o->fwd, o->X.p, input, nullptr, nblocks, h / 2, 0, h + 1, RF_KEEP_OUT, 1.0f, st))
        return -1;
    // 2. filtered blocks: c2r of X * H / N (H broadcast to every block).
    if (exec_1d(o->inv, o->Y.p, o->X.p, o->H.p, nblocks, h + 1, 0, h, RF_MUL | RF_KEEP_OUT,
                1.0f / (float)n, st))
        return -1;
    // 3. overlap-add with the tail of the previous call.
    float *tail_in = static_cast<float *>(o->tail[o->cur].p);
    float *tail_out = static_cast<float *>(o->tail[o->cur ^ 1].p);
    HIP_TRY(launch_ola(static_cast<const float *>(o->Y.p), output, tail_in, tail_out, n, nblocks, st));
    o->cur ^= 1;
    return 0;
}

}  // extern "C"
