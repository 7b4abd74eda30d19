// Streaming overlap-add FIR filtering on the fused real-transform kernels.
#include <string.h>

#include <new>
#include <vector>

#include "host.h"

using namespace mufft;

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
