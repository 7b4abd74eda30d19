// Persistent kernel for a batch of fast convolutions of uneven sizes.
//
// The reference runs one convolution as three calls (fft.c:799-809): forward
// transform of each input block, then multiply + inverse transform.  Here a
// whole batch of such jobs, of any mix of sizes and of both methods (mono:
// r2c / r2c / c2r, stereo: c2c / full r2c / c2c, fft.c:583-600), runs in one
// persistent launch: forward tasks (both inputs of every job) and inverse
// tasks (the spectrum product fused into their prologue), each inverse task
// starting as soon as its own job group's forward tasks are done.  Tasks of
// equal shape are grouped up to one workgroup's capacity, job groups are
// spread over the XCDs longest-first, and each XCD's workgroups pull its
// tasks from its own queue (a ticket counter); see CBArgs for the ordering.
#include "conv_batch.h"
#include "rows_impl.h"

namespace mufft {

// Fused-kernel variant used inside the persistent kernel (1024 threads): 16
// points per thread with split twiddles; at 2^CB_SPLIT_LOG2 points the same
// variant runs each 16384-point half of the split transform.  It fits
// CB_LDS_BYTES.  The c2r steps read each bin once through mirrored lanes
// where a transform has whole waves (N >= 1024), in chunks below.
MUFFT_HD constexpr int cb_variant(int N) {
    return N >= 1024 ? rv_with_mirror_lanes(rv_make(16, 0, 1)) : rv_with_chunked_loads(rv_make(16, 0, 1));
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

// Jobs [t.job0, t.job0 + t.njobs) of a task, one fused transform each; FL:
// the task's flags if fixed at compile time (RF_DYNAMIC: t.flags).
template <int KIND, int DIR, int LOGN, unsigned FL>
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
    row_transform<N, DIR, KIND, VAR, FL>(X, Y, Xo, active, lt, t.tw, t.rtw, t.flags, t.norm,
                                      smem + sub * LDSN, reinterpret_cast<float *>(smem) + sub * LDSN);
}

// Half t.k1 of one job's transform over two workgroups (split_impl.h).
template <int KIND, int DIR, int LOGN, unsigned FL>
__device__ __forceinline__ void cb_run_split(const CBTask &t, const CBJob *__restrict__ jobs, int tid,
                                             float2 *smem) {
    constexpr int N2 = 1 << (LOGN - 1);
    constexpr int VAR = cb_variant(N2);
    static_assert(cb_lds_bytes(N2) <= CB_LDS_BYTES, "LDS budget of the persistent kernel");
    const CBJob &j = jobs[t.job0];
    const float2 *X = t.type == CB_FWD_FIRST ? j.first : (t.type == CB_FWD_SECOND ? j.second : j.spec1);
    const float2 *Y = t.type == CB_INV ? j.spec2 : nullptr;
    float2 *Xo = t.type == CB_FWD_FIRST ? j.spec1 : (t.type == CB_FWD_SECOND ? j.spec2 : j.out);
    split_transform<N2, 2, DIR, KIND, VAR, FL>(X, Y, Xo, t.k1, tid, t.tw, t.stw, t.rtw, t.flags, t.norm, smem,
                                           reinterpret_cast<float *>(smem));
}

template <int KIND, int DIR, int LOGN, unsigned FL>
__device__ __forceinline__ void cb_run_fl(const CBTask &t, const CBJob *__restrict__ jobs, int tid,
                                          float2 *smem) {
    if constexpr (LOGN >= CB_SPLIT_LOG2) cb_run_split<KIND, DIR, LOGN, FL>(t, jobs, tid, smem);
    else cb_run_rows<KIND, DIR, LOGN, FL>(t, jobs, tid, smem);
}

// The flag sets of the zero-padded convolutions (both inputs padded, the
// usual block convolution) run with their flags fixed at compile time; any
// other set reads them at run time.  Mono: r2c, r2c, c2r with the multiply;
// stereo: c2c, full r2c, c2c with the multiply (conv_batch_host.hip).
template <int KIND, int DIR, int LOGN>
__device__ __noinline__ void cb_run(const CBTask &t, const CBJob *__restrict__ jobs, int tid,
                                    float2 *smem) {
    constexpr unsigned FWD = RF_ZERO_PAD | RF_KEEP_OUT;
    if constexpr (KIND == KIND_R2C) {
        if (t.flags == FWD) return cb_run_fl<KIND, DIR, LOGN, FWD>(t, jobs, tid, smem);
        if (t.flags == (FWD | RF_FULL)) return cb_run_fl<KIND, DIR, LOGN, FWD | RF_FULL>(t, jobs, tid, smem);
    } else if constexpr (KIND == KIND_C2C && DIR < 0) {
        if (t.flags == FWD) return cb_run_fl<KIND, DIR, LOGN, FWD>(t, jobs, tid, smem);
    } else {
        if (t.flags == RF_MUL) return cb_run_fl<KIND, DIR, LOGN, (unsigned)RF_MUL>(t, jobs, tid, smem);
    }
    cb_run_fl<KIND, DIR, LOGN, RF_DYNAMIC>(t, jobs, tid, smem);
}

template <int KIND, int DIR, int LO, int HI>
__device__ __forceinline__ void cb_dispatch_sizes(const CBTask &t, const CBJob *jobs, int tid,
                                                  float2 *smem) {
    static_for<LO, HI + 1>([&](auto L) {
        if (t.logn == decltype(L)::value) cb_run<KIND, DIR, decltype(L)::value>(t, jobs, tid, smem);
    });
}

template <int KIND_, int DIR_>
__device__ __forceinline__ void cb_dispatch_kind(const CBTask &t, const CBJob *jobs, int tid, float2 *smem) {
    cb_dispatch_sizes<KIND_, DIR_, CB_MIN_LOG2, 15>(t, jobs, tid, smem);
}

__device__ __forceinline__ void cb_run_task(const CBTask &t, const CBJob *jobs, int tid, float2 *smem) {
    if (t.kind == KIND_R2C) cb_dispatch_kind<KIND_R2C, -1>(t, jobs, tid, smem);
    else if (t.kind == KIND_C2R) cb_dispatch_kind<KIND_C2R, 1>(t, jobs, tid, smem);
    else if (t.dir < 0) cb_dispatch_kind<KIND_C2C, -1>(t, jobs, tid, smem);
    else cb_dispatch_kind<KIND_C2C, 1>(t, jobs, tid, smem);
}

// XCD (die) this workgroup runs on: the hardware XCC_ID register, bits 3:0.
__device__ __forceinline__ int cb_xcd() {
    return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & (CB_XCDS - 1);
}

// Each iteration keeps all of thread 0's work (publishing the previous
// task, the next ticket, the readiness wait) in one region at its top,
// followed by barriers, so the loop compiles to one uniform loop (a second
// thread-0 region at the end of the body let the compiler's control-flow
// structurization send the other lanes of wave 0 back to the barriers ahead
// of thread 0).
__global__ void __launch_bounds__(CB_WG) conv_batch_kernel(CBArgs a) {
    __shared__ float2 smem[CB_LDS_BYTES / 8];
    __shared__ int s_local;
    const int tid = threadIdx.x;
    const int x = cb_xcd();
    const int qb = a.qbegin[x], qn = a.qbegin[x + 1] - qb;
    int prev_group = -1;          // forward task whose completion is to be published
    bool finished_one = false;
    for (;;) {
        if (tid == 0) {
            if (prev_group >= 0)
                __hip_atomic_fetch_add(a.ready + prev_group, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (finished_one) atomicAdd(a.done, 1ull);
            const int local = (int)atomicAdd(a.head + x, 1u);
            s_local = local;
            if (local < qn) {
                const CBTask &t = a.tasks[qb + local];
                if (t.type == CB_INV) {
                    const unsigned need = (unsigned)t.nprod;
                    unsigned polls = 0;
                    while (__hip_atomic_load(a.ready + t.group, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
                        if (++polls > a.spin_limit) {
                            atomicAdd(a.timeouts, 1ull);
                            break;
                        }
                        __builtin_amdgcn_s_sleep(2);
                    }
                    // Drop stale lines of this CU's L1: the spectra are read from the L2.
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                }
            }
        }
        __syncthreads();
        const int local = __builtin_amdgcn_readfirstlane(s_local);
        __syncthreads();
        if (local >= qn) break;
        const CBTask t = a.tasks[qb + local];
        cb_run_task(t, a.jobs, tid, smem);
        // Forward task: every thread's spectrum stores are complete (in this
        // XCD's L2) before thread 0 publishes the task at the next iteration.
        __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0)
        __syncthreads();
        prev_group = t.type == CB_INV ? -1 : t.group;
        finished_one = true;
    }
    // The last workgroup out clears the queue and readiness counters for the
    // next launch (every other workgroup has made its last access to them).
    if (tid == 0) {
        const unsigned long long f = atomicAdd(a.finish, 1ull);
        if ((f + 1) % gridDim.x == 0) {
            for (int i = 0; i < CB_XCDS; ++i) __hip_atomic_store(a.head + i, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            for (int g = 0; g < a.ngroups; ++g)
                __hip_atomic_store(a.ready + g, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

int cb_resident_workgroups() {
    int dev = 0, ncu = 0, occ = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, conv_batch_kernel, CB_WG, 0) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return ncu * occ;
}

hipError_t launch_conv_batch(const CBArgs &a, int grid, hipStream_t s) {
    if (grid <= 0) return hipSuccess;
    hipLaunchKernelGGL(conv_batch_kernel, dim3(grid), dim3(CB_WG), 0, s, a);
    return hipGetLastError();
}

}  // namespace mufft
