// Persistent kernel for a batch of fast convolutions of uneven sizes.
//
// The reference runs one convolution as three calls (fft.c:799-809): forward
// transform of each input block, then multiply + inverse transform.  Here a
// whole batch of such jobs, of any mix of sizes and of both methods (mono:
// r2c / r2c / c2r, stereo: c2c / full r2c / c2c, fft.c:583-600), runs in two
// persistent launches: one over all forward tasks (both inputs of every job),
// one over all inverse tasks (with the spectrum multiply fused into their
// prologue).  Tasks of equal shape are grouped up to one workgroup's
// capacity and each phase is ordered longest-first (LPT); workgroups pull
// tasks from a device queue (an atomic counter).  The launch boundary orders
// the phases.  A device-wide wait inside one launch would need agent-scope
// release/acquire fences, which on gfx950 write back or invalidate a whole
// XCD's L2 (buffer_wbl2 / buffer_inv) at every task and slowed the batch
// five-fold; the queue atomics here are relaxed and no workgroup ever waits
// for another.
#include "conv_batch.h"
#include "rows_impl.h"

namespace mufft {

// Fused-kernel variant used inside the persistent kernel (1024 threads): 16
// points per thread with split twiddles; at 2^CB_SPLIT_LOG2 points the same
// variant runs each 16384-point half of the split transform.  It fits CB_LDS_BYTES.
MUFFT_HD constexpr int cb_variant(int N) {
    return rv_make(16, 0, 1);
}
MUFFT_HD constexpr int cb_sub_n(int logn) { return logn >= CB_SPLIT_LOG2 ? (1 << (logn - 1)) : (1 << logn); }

MUFFT_HD constexpr int cb_lds_bytes(int N) {
    return (CB_WG / rv_T(N, cb_variant(N))) * rv_lds_per(N, cb_variant(N)) *
           (rv_xm(cb_variant(N)) == 1 ? 4 : 8);
}

int cb_jobs_per_task(int logn) {
    const int N = cb_sub_n(logn);
    return CB_WG / rv_T(N, cb_variant(N));
}

int cb_points_per_thread(int logn) { return rv_E(cb_sub_n(logn), cb_variant(cb_sub_n(logn))); }

// Jobs [t.job0, t.job0 + t.njobs) of a task, one fused transform each.
template <int KIND, int DIR, int LOGN>
__device__ __forceinline__ void cb_run_rows(const CBTask &t, const CBJob *__restrict__ jobs, int tid,
                                            float2 *smem) {
    constexpr int N = 1 << LOGN;
    constexpr int VAR = cb_variant(N);
    constexpr int T = rv_T(N, VAR);
    constexpr int LDSN = rv_lds_per(N, VAR);
    static_assert(cb_lds_bytes(N) <= CB_LDS_BYTES, "LDS budget of the persistent kernel");
    const int lt = tid & (T - 1);
    const int sub = tid / T;
    const bool active = sub < t.njobs;
    const CBJob &j = jobs[t.job0 + (active ? sub : 0)];
    const float2 *X;
    const float2 *Y = nullptr;
    float2 *Xo;
    if (t.type == CB_FWD_FIRST) {
        X = j.first;
        Xo = j.spec1;
    } else if (t.type == CB_FWD_SECOND) {
        X = j.second;
        Xo = j.spec2;
    } else {
        X = j.spec1;
        Y = j.spec2;
        Xo = j.out;
    }
    row_transform<N, DIR, KIND, VAR>(X, Y, Xo, active, lt, t.tw, t.rtw, t.flags, t.norm,
                                      smem + sub * LDSN, reinterpret_cast<float *>(smem) + sub * LDSN);
}

// Half t.k1 of one job's transform over two workgroups (split_impl.h).
template <int KIND, int DIR, int LOGN>
__device__ __forceinline__ void cb_run_split(const CBTask &t, const CBJob *__restrict__ jobs, int tid,
                                             float2 *smem) {
    constexpr int N2 = 1 << (LOGN - 1);
    constexpr int VAR = cb_variant(N2);
    static_assert(cb_lds_bytes(N2) <= CB_LDS_BYTES, "LDS budget of the persistent kernel");
    const CBJob &j = jobs[t.job0];
    const float2 *X = t.type == CB_FWD_FIRST ? j.first : (t.type == CB_FWD_SECOND ? j.second : j.spec1);
    const float2 *Y = t.type == CB_INV ? j.spec2 : nullptr;
    float2 *Xo = t.type == CB_FWD_FIRST ? j.spec1 : (t.type == CB_FWD_SECOND ? j.spec2 : j.out);
    split_transform<N2, 2, DIR, KIND, VAR>(X, Y, Xo, t.k1, tid, t.tw, t.stw, t.rtw, t.flags, t.norm, smem,
                                           reinterpret_cast<float *>(smem));
}

template <int KIND, int DIR, int LOGN>
__device__ __noinline__ void cb_run(const CBTask &t, const CBJob *__restrict__ jobs, int tid,
                                       float2 *smem) {
    if constexpr (LOGN >= CB_SPLIT_LOG2) cb_run_split<KIND, DIR, LOGN>(t, jobs, tid, smem);
    else cb_run_rows<KIND, DIR, LOGN>(t, jobs, tid, smem);
}

template <int KIND, int DIR, int LO, int HI>
__device__ __forceinline__ void cb_dispatch_sizes(const CBTask &t, const CBJob *jobs, int tid,
                                                  float2 *smem) {
    static_for<LO, HI + 1>([&](auto L) {
        if (t.logn == decltype(L)::value) cb_run<KIND, DIR, decltype(L)::value>(t, jobs, tid, smem);
    });
}

__global__ void __launch_bounds__(CB_WG) conv_batch_kernel(CBArgs a, int phase) {
    __shared__ float2 smem[CB_LDS_BYTES / 8];
    __shared__ int s_task;
    const int tid = threadIdx.x;
    const int begin = phase == 0 ? 0 : a.nfwd;
    const int end = phase == 0 ? a.nfwd : a.ntasks;
    for (;;) {
        if (tid == 0) s_task = begin + atomicAdd(a.head + phase, 1);
        __syncthreads();
        const int ti = s_task;
        __syncthreads();
        if (ti >= end) break;
        const CBTask t = a.tasks[ti];
        if (t.kind == KIND_R2C) cb_dispatch_sizes<KIND_R2C, -1, CB_MIN_LOG2, 15>(t, a.jobs, tid, smem);
        else if (t.kind == KIND_C2R) cb_dispatch_sizes<KIND_C2R, 1, CB_MIN_LOG2, 15>(t, a.jobs, tid, smem);
        else if (t.dir < 0) cb_dispatch_sizes<KIND_C2C, -1, CB_MIN_LOG2, 15>(t, a.jobs, tid, smem);
        else cb_dispatch_sizes<KIND_C2C, 1, CB_MIN_LOG2, 15>(t, a.jobs, tid, smem);
        if (tid == 0) atomicAdd(a.done, 1);
    }
}

// Clears the queue heads and the completion count (a kernel, so the reset can
// live inside a captured graph).
__global__ void __launch_bounds__(64) conv_batch_reset_kernel(CBArgs a) {
    if (threadIdx.x < 2) a.head[threadIdx.x] = 0;
    if (threadIdx.x == 0) *a.done = 0;
}

hipError_t launch_conv_batch(const CBArgs &a, int grid, hipStream_t s) {
    hipLaunchKernelGGL(conv_batch_reset_kernel, dim3(1), dim3(64), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int g0 = a.nfwd < grid ? a.nfwd : grid;
    const int g1 = a.ntasks - a.nfwd < grid ? a.ntasks - a.nfwd : grid;
    if (g0 > 0) {
        hipLaunchKernelGGL(conv_batch_kernel, dim3(g0), dim3(CB_WG), 0, s, a, 0);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (g1 > 0) {
        hipLaunchKernelGGL(conv_batch_kernel, dim3(g1), dim3(CB_WG), 0, s, a, 1);
        e = hipGetLastError();
    }
    return e;
}

}  // namespace mufft
