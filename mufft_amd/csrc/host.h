// Host side of the library: error state, device buffers, pointer
// classification, twiddle-table builders, the per-axis Engine, the executor
// and the plan objects behind the C ABI.
//
// The reference builds a list of SIMD function pointers per plan and calls
// them one pass at a time on host memory (fft.c:325-435, fft.c:811-989).
// Here a plan owns device twiddle tables (generated in FP64 on the GPU), a
// private HIP stream and scratch space, and executing it enqueues one fused
// kernel per transform axis whenever the axis fits one or two workgroups
// (1D/rows up to 65536 complex points, columns up to 8192), two or three
// column launches (four-step) up to 2^26 points, or one generic Stockham pass
// per launch beyond that.
#pragma once

#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stddef.h>
#include <stdint.h>

#include <exception>
#include <mutex>
#include <vector>

#include "kernels.h"
#include "mufft/fft.h"
#include "mufft/mufft_hip.h"

namespace mufft {

// ---- error state -----------------------------------------------------------
// The message of the last failed call on this thread.
extern thread_local char g_err[512];
void set_err(const char *fmt, ...);

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

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            ::mufft::set_err("%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__,     \
                             __LINE__);                                                    \
            return -1;                                                                     \
        }                                                                                  \
    } while (0)

inline bool is_pow2(unsigned n) { return n != 0 && (n & (n - 1)) == 0; }

// ---- device memory ---------------------------------------------------------
// Grow-only device buffer.  Growing frees the old allocation; hipFree waits
// for the device, so no kernel still reads it.  A fixed buffer wraps memory
// the library does not own (a region of the allocation-free scratch): it
// never allocates or frees, and a request beyond its size is an error.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    bool fixed = false;
    int ensure(size_t n);
    void release();
    float2 *f2() const { return static_cast<float2 *>(p); }
    static DevBuf wrap(void *mem, size_t n) {
        DevBuf b;
        b.p = mem;
        b.bytes = n;
        b.fixed = true;
        return b;
    }
};

// Where a buffer lives: device memory (used in place), pinned host memory
// (kernels may address it over PCIe through `dev`) or pageable host memory
// (only the copy engines can reach it).
enum PtrKind { PTR_DEVICE, PTR_PINNED, PTR_PAGEABLE };
struct PtrInfo {
    PtrKind kind;
    void *dev;
};
PtrInfo classify_ptr(const void *ptr);
// True when kernels cannot use the pointer as a device address without
// staging (pinned or pageable host memory).
bool is_host_ptr(const void *ptr);

// Kernel-addressable form of a buffer for the stream-ordered extensions that
// do no staging: device pointers as they are, pinned host memory through its
// device mapping; pageable host memory is refused (a kernel cannot reach it).
int kernel_ptr(const void *ptr, const void **out);
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

// ---- twiddle tables (FP64 on the GPU, rounded once) ------------------------
int build_stockham_table(int n, int direction, const std::vector<int> &radix, float2 **out,
                         hipStream_t s);
int build_resolve_table(int n, int direction, float2 **out, hipStream_t s);
// exp(dir * 2 pi i * j / n) for j < count.
int build_plain_table(int n, int count, int direction, float2 **out, hipStream_t s);
// Twiddles of a transform of n points split into n / n2 sub-transforms:
// entry (k1 - 1) * n2 + j = exp(dir * 2 pi i * j * k1 / n).
int build_split_table(int n, int n2, int direction, float2 **out, hipStream_t s);

// Radix sequence of the fused kernels with pass radix E: the remainder pass
// first for the row kernels, last for the column kernels (seq_radix()).
std::vector<int> fused_sequence(int n, int E, bool rem_first);
// Radix-16 passes and one remainder pass (the generic one-pass-per-launch path).
std::vector<int> generic_sequence(int n);

// ---- fused-kernel variant choice -------------------------------------------
// Kind codes of the variant table and the autotuner.
enum TuneKind { TK_C2C_FWD = 0, TK_C2C_INV = 1, TK_R2C = 2, TK_C2R = 3 };
int tune_kind(int kind, int dir);
// Process-wide choice per (log2 size, kind): the autotuner's or
// mufft_hip_set_variant's, else the static default rv_default().
int chosen_variant(int logn, int tk);
void set_chosen_variant(int logn, int tk, int variant);
// Whether a size has a tuned or explicitly set choice.
bool variant_is_set(int logn, int tk);
// Process-wide split of transforms of 2^logn points beyond the fused sizes
// (a fs_code; 0 = the static fs_code_default) and tile width of the column
// kernel of 2^logL points (0 = col_TX): the autotuner's or an explicit choice.
int chosen_fs_code(int logn);
void set_fs_code(int logn, int code);
int chosen_col_tx(int logL);
void set_col_tx(int logL, int tx);

// ---- one transform axis ----------------------------------------------------
// A complex transform of length n along one axis: fused row kernel, fused
// column kernel, four-step column launches or generic passes, each with its
// own twiddle table.
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
    int fs_code = 0;              // the split in use (fs_code_*)
    int col_tx = 0;               // column tile width (0 = col_TX(n))
    float2 *tw_pa = nullptr, *tw_pb = nullptr, *tw_coarse2 = nullptr, *tw_fine2 = nullptr;
    std::vector<int> gen_radix;

    bool three_pass() const { return log_a > 0; }
    bool row_fused() const { return logn <= ROW_MAX_LOG2; }
    bool col_fused() const { return logn <= COL_MAX_LOG2; }
    bool four_step() const { return logn > ROW_MAX_LOG2 && logn <= 2 * COL_MAX_LOG2; }
    int launches_generic() const { return (int)gen_radix.size(); }

    // rows / cols: which roles this engine serves; kind: KIND_* of the row
    // role; var: the row variant (0 = the process-wide choice); fs: the split
    // beyond the fused sizes (0 = the process-wide choice); tx: the column
    // tile width (0 = the process-wide choice).
    int init(int n_, int dir, int kind, bool rows, bool cols, hipStream_t s, int var = 0, int fs = 0,
             int tx = 0);
    void release();
};

// ---- executor --------------------------------------------------------------
// Contiguous transforms of kind `kind` (rows / 1D): transform g reads
// in + g*idist (and in2 + g*idist2 with RF_MUL) and writes out + g*odist.
// Distances in complex elements.  Sizes beyond the fused row kernel use the
// scratch s0/s1 (grown on demand).
int run_rows(const Engine &e, int kind, const float2 *rtw, const float2 *in, long long idist,
             const float2 *in2, long long idist2, float2 *out, long long odist, int batch,
             unsigned flags, float norm, DevBuf &s0, DevBuf &s1, hipStream_t st);
// Column transforms of length e.n over `ncols` columns, `batch` images.
// The fused path works in place; the generic path needs src != dst and uses alt.
int run_cols(const Engine &e, const float2 *src, long long rs_src, long long bs_src, float2 *dst,
             long long rs_dst, long long bs_dst, int ncols, int batch, DevBuf &alt,
             hipStream_t st);
// Whether run_rows touches the scratch buffers for this engine.
inline bool rows_use_scratch(const Engine &e) { return !e.row_fused(); }

// ---- scratch ordering --------------------------------------------------------
// Plan-owned scratch is shared by every call on the plan, on any stream.  A
// call that uses it takes the mutex, makes its stream wait for the last use
// on another stream (an event), enqueues its work and records the event
// again: uses are serialised in GPU order, and the buffers never change size
// under in-flight work (growth frees through hipFree, which waits for the
// device).
struct ScratchOrder {
    std::mutex m;
    hipEvent_t ev = nullptr;
    hipStream_t last = nullptr;
    bool used = false;
    void release();
};
template <typename F>
int with_scratch(ScratchOrder &so, hipStream_t st, F &&body) {
    std::lock_guard<std::mutex> g(so.m);
    if (!so.ev) HIP_TRY(hipEventCreateWithFlags(&so.ev, hipEventDisableTiming));
    if (so.used && so.last != st) HIP_TRY(hipStreamWaitEvent(st, so.ev, 0));
    const int rc = body();
    // Record even after a failure: whatever was enqueued is ordered before the next use.
    HIP_TRY(hipEventRecord(so.ev, st));
    so.last = st;
    so.used = true;
    return rc;
}

}  // namespace mufft

// ---- plan objects (opaque in the public headers) -----------------------------
struct HostPipe;

struct mufft_plan_1d {
    unsigned N = 0;      // complex length (reference plan->N, fft.c:58)
    int direction = 0;
    int kind = mufft::KIND_C2C;
    unsigned flags = 0;
    bool full = false, zero_pad = false;
    int device = 0;
    hipStream_t stream = nullptr;
    mufft::Engine eng;
    float2 *rtw = nullptr;
    size_t in_bytes = 0, out_bytes = 0;   // per transform
    // Drop-in staging and host-memory batches, serialised by `lock`.
    std::mutex lock;
    mufft::DevBuf stage_in, stage_in2, stage_out;
    HostPipe *pipe = nullptr;
    // Scratch of the multi-launch paths and of the adjoint, ordered by `scratch`.
    mufft::ScratchOrder scratch;
    mufft::DevBuf s0, s1, fold;
    // Adjoint transform, built on first use (under scratch.m).
    mufft::Engine adj;
    float2 *adj_rtw = nullptr;
    bool adj_ready = false;
};

struct mufft_plan_2d {
    unsigned Nx = 0, Ny = 0;   // Nx = complex columns (Nx_real / 2 for r2c / c2r), fft.c:75
    int direction = 0;
    int kind = mufft::KIND_C2C;
    bool full = false, zero_pad = false;
    unsigned vertical_nx = 0;  // fft.c:85
    int device = 0;
    hipStream_t stream = nullptr;
    mufft::Engine ex, ey;
    float2 *rtw = nullptr;
    size_t in_bytes = 0, out_bytes = 0;
    std::mutex lock;
    mufft::DevBuf stage_in, stage_out;
    HostPipe *pipe = nullptr;
    mufft::ScratchOrder scratch;
    mufft::DevBuf s0, s1, s2, alt;
};

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

namespace mufft {

// 1D plan internals shared by the conv, overlap-add and allocation-free units.
mufft_plan_1d *create_1d(int kind, unsigned Nc, int direction, unsigned flags, int var);
void destroy_1d(mufft_plan_1d *p);
// Distances in complex elements; extra = RF_MUL / RF_CONJ2 / RF_KEEP_OUT.
int exec_1d(mufft_plan_1d *p, void *out, const void *in, const void *in2, unsigned howmany,
            long long idist, long long idist2, long long odist, unsigned extra, float norm,
            hipStream_t st);
// Synchronous drop-in execution with staging of pageable host buffers.
void exec_1d_dropin(mufft_plan_1d *p, void *out, const void *in, const void *in2,
                    size_t in2_bytes, unsigned extra, float norm);
// Packed distances (complex elements) of a plan's input / output.
long long in_dist_c(const mufft_plan_1d *p);
long long out_dist_c(const mufft_plan_1d *p);

// 2D: howmany packed images; s2 / alt are the intermediate image and the
// generic column scratch, s0 / s1 the generic row scratch.
int exec_2d_with(mufft_plan_2d *p, void *out, const void *in, unsigned howmany, DevBuf &s0,
                 DevBuf &s1, DevBuf &s2, DevBuf &alt, hipStream_t st);
// Complex elements of intermediate image scratch exec_2d_with needs per image (0: none).
size_t plan_2d_scratch_points(const mufft_plan_2d *p);

}  // namespace mufft
