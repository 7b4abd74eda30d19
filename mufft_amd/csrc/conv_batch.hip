// Persistent kernel for a batch of fast convolutions of uneven sizes.
//
// The reference runs one convolution as three calls (fft.c:799-809): forward
// transform of each input block, then multiply + inverse transform.  Here a
// whole batch of such jobs, of any mix of sizes and of both methods (mono:
// r2c / r2c / c2r, stereo: c2c / full r2c / c2c, fft.c:583-600), runs in one
// launch.  Each job is split into three tasks: forward(first), forward(second)
// and inverse (with the spectrum multiply fused into its prologue).  Tasks of
// equal shape are grouped up to one workgroup's capacity; the host orders all
// forward tasks before all inverse tasks, each part longest-first (LPT).
// Workgroups pull tasks from a device queue (an atomic counter).  A forward
// task publishes its spectra (release fence at agent scope) and increments
// its jobs' counters; an inverse task waits until both counters of its jobs
// reach 2 (acquire).  Forward tasks never wait and every task is pulled by a
// running workgroup, so the waits always end; a spin limit nevertheless bounds
// them and raises an error flag instead of hanging.
#include "conv_batch.h"
#include "rows_impl.h"

namespace mufft {

// Fused-kernel variant used inside the persistent kernel (1024 threads):
// 16 points per thread with split twiddles, 32 with the split exchange at
// 32768 points.  Both fit CB_LDS_BYTES.
MUFFT_HD constexpr int cb_variant(int N) {
    return N >= 32768 ? rv_make(32, 1, 1) : rv_make(16, 0, 1);
}

MUFFT_HD constexpr int cb_lds_bytes(int N) {
    return (CB_WG / rv_T(N, cb_variant(N))) * rv_lds_per(N, cb_variant(N)) *
           (rv_xm(cb_variant(N)) == 1 ? 4 : 8);
}

int cb_jobs_per_task(int logn) {
    const int N = 1 << logn;
    return CB_WG / rv_T(N, cb_variant(N));
}

int cb_points_per_thread(int logn) { return rv_E(1 << logn, cb_variant(1 << logn)); }

template <int KIND, int DIR, int LOGN>
__device__ __noinline__ void cb_run(const CBTask &t, const CBJob *__restrict__ jobs, int tid,
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

template <int KIND, int DIR, int LO, int HI>
__device__ __forceinline__ void cb_dispatch_sizes(const CBTask &t, const CBJob *jobs, int tid,
                                                  float2 *smem) {
    static_for<LO, HI + 1>([&](auto L) {
        if (t.logn == decltype(L)::value) cb_run<KIND, DIR, decltype(L)::value>(t, jobs, tid, smem);
    });
}

__global__ void __launch_bounds__(CB_WG) conv_batch_kernel(CBArgs a) {
    __shared__ float2 smem[CB_LDS_BYTES / 8];
    __shared__ int s_task;
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) s_task = atomicAdd(a.head, 1);
        __syncthreads();
        const int ti = s_task;
        __syncthreads();
        if (ti >= a.ntasks) break;
        const CBTask t = a.tasks[ti];
        if (t.type == CB_INV) {
            if (tid == 0) {
                long long spins = 0;  // one budget for the whole task
                for (int j = 0; j < t.njobs && spins <= a.spin_limit; ++j) {
                    int *c = a.counters + t.job0 + j;
                    while (__hip_atomic_load(c, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < 2) {
                        __builtin_amdgcn_s_sleep(2);
                        if (++spins > a.spin_limit) {
                            atomicExch(a.error, 1);
                            break;
                        }
                    }
                }
            }
            __syncthreads();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        if (t.kind == KIND_R2C) cb_dispatch_sizes<KIND_R2C, -1, CB_MIN_LOG2, 15>(t, a.jobs, tid, smem);
        else if (t.kind == KIND_C2R) cb_dispatch_sizes<KIND_C2R, 1, CB_MIN_LOG2, 15>(t, a.jobs, tid, smem);
        else if (t.dir < 0) cb_dispatch_sizes<KIND_C2C, -1, CB_MIN_LOG2, 15>(t, a.jobs, tid, smem);
        else cb_dispatch_sizes<KIND_C2C, 1, CB_MIN_LOG2, 15>(t, a.jobs, tid, smem);
        if (t.type != CB_INV) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __syncthreads();
            if (tid < t.njobs) {
                __hip_atomic_fetch_add(a.counters + t.job0 + tid, 1, __ATOMIC_RELEASE,
                                       __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
}

// Clears the queue head, the error flag and the per-job counters (a kernel,
// so the reset can live inside a captured graph).
__global__ void __launch_bounds__(256) conv_batch_reset_kernel(CBArgs a) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < a.njobs) a.counters[i] = 0;
    if (i == 0) {
        *a.head = 0;
        *a.error = 0;
    }
}

hipError_t launch_conv_batch(const CBArgs &a, int grid, hipStream_t s) {
    const int nreset = (a.njobs + 255) / 256;
    hipLaunchKernelGGL(conv_batch_reset_kernel, dim3(nreset > 0 ? nreset : 1), dim3(256), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(conv_batch_kernel, dim3(grid), dim3(CB_WG), 0, s, a);
    return hipGetLastError();
}

}  // namespace mufft
