// Allocation-free batched 1D execution.
//
// mufft_hip_execute_plan_1d_batched grows plan-owned scratch with hipMalloc
// on demand (the multi-launch paths beyond the fused sizes: four-step column
// launches, the standalone resolve, generic passes).  Callers that cannot
// allocate (code under HIP graph capture, serving loops, harnesses that forbid
// it) use this variant instead: the scratch lives in a fixed device array,
// one copy per GPU, and the call issues no allocation, no free and no event.
// The plan itself (twiddle tables) is built beforehand by the plan creator.
#include <string.h>

#include "host.h"

namespace mufft {

namespace {

// Two regions of kNoAllocPoints complex points each (64 MiB in all): the
// generic path's s0 (multiply / resolve output, r2c intermediate) and s1
// (four-step intermediate).  Every call on a GPU shares its copy of the array,
// and each use overwrites what it reads before reading it: nothing needs to
// start at zero.
constexpr size_t kNoAllocPoints = size_t(1) << 22;
static __device__ __align__(256) float2 scratch[2 * kNoAllocPoints];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_m;
float2 *g_scratch_ptr[kMaxDevices] = {};

// Device address of this GPU's copy of `scratch`, looked up once per device
// with that device current.  Host code must never pass `scratch` itself to a
// kernel: in host code the name is only a stand-in for the symbol.
int scratch_for_device(int dev, float2 **out) {
    if (dev < 0 || dev >= kMaxDevices) {
        set_err("device %d beyond the allocation-free scratch table", dev);
        return -1;
    }
    std::lock_guard<std::mutex> g(g_scratch_m);
    if (!g_scratch_ptr[dev]) {
        int cur = 0;
        HIP_TRY(hipGetDevice(&cur));
        if (cur != dev) HIP_TRY(hipSetDevice(dev));
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess || !ptr) {
            (void)hipGetLastError();
            set_err("hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
            return -1;
        }
        g_scratch_ptr[dev] = static_cast<float2 *>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

}  // namespace

}  // namespace mufft

using namespace mufft;

extern "C" {

int mufft_hip_execute_plan_1d_batched_noalloc(mufft_plan_1d *p, void *output, const void *input,
                                              unsigned howmany, size_t idist, size_t odist,
                                              void *stream) {
    if (!p || !output || !input) {
        set_err("null argument");
        return -1;
    }
    if (howmany == 0) return 0;
    // Distances in complex elements (real buffers count two floats per element).
    const bool in_real = p->kind == KIND_R2C, out_real = p->kind == KIND_C2R;
    if ((in_real && (idist & 1)) || (out_real && (odist & 1))) {
        set_err("real buffer distances must be even");
        return -1;
    }
    const long long id = idist ? (long long)(in_real ? idist / 2 : idist) : in_dist_c(p);
    const long long od = odist ? (long long)(out_real ? odist / 2 : odist) : out_dist_c(p);
    // Device memory, or pinned host memory through its device mapping; pageable
    // host memory would need staging buffers and is refused.
    if (kernel_ptrs(&output, &input)) return -1;
    DevBuf s0, s1;
    if (rows_use_scratch(p->eng)) {
        const size_t points = (size_t)p->eng.n * howmany;
        if (points > kNoAllocPoints) {
            set_err("allocation-free execution holds at most %zu complex points per call "
                    "(N x howmany = %zu)", kNoAllocPoints, points);
            return -1;
        }
        float2 *base = nullptr;
        if (scratch_for_device(p->device, &base)) return -1;
        s0 = DevBuf::wrap(base, kNoAllocPoints * sizeof(float2));
        s1 = DevBuf::wrap(base + kNoAllocPoints, kNoAllocPoints * sizeof(float2));
    }
    unsigned flags = 0;
    if (p->zero_pad) flags |= RF_ZERO_PAD;
    if (p->full) flags |= RF_FULL;
    return run_rows(p->eng, p->kind, p->rtw, static_cast<const float2 *>(input), id, nullptr, 0,
                    static_cast<float2 *>(output), od, (int)howmany, flags, 1.0f, s0, s1,
                    static_cast<hipStream_t>(stream));
}

}  // extern "C"
