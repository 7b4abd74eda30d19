// Runtime entry points: the convolve function pointer (fft.c:549-563),
// memory (fft.c:1032-1084), device helpers, the twiddle-table inspection
// hook and the device-side utility kernels behind the extensions.
#include <stdio.h>
#include <string.h>

#include "host.h"

using namespace mufft;

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

// The auxiliary FP64-built tables, for inspection:
//   mode 1: resolve twiddles W[k], k < n (the r2c/c2r table of complex length n);
//   mode 2: exp(dir * 2 pi i * j / n), j < param (the four-step fine/coarse tables);
//   mode 3: split twiddles, entry (k1 - 1) * param + j = exp(dir * 2 pi i * j * k1 / n).
static int mufft_hip_twiddle_table_mode_impl(int mode, unsigned n, unsigned param, int direction,
                                             float *host_out, unsigned count) {
    if (!is_pow2(n) || n < 2 || n > (1u << 27) || (direction != -1 && direction != 1)) {
        set_err("invalid table size or direction");
        return -1;
    }
    size_t size = 0;
    if (mode == 1) size = n;
    else if (mode == 2 && param >= 1 && param <= n) size = param;
    else if (mode == 3 && is_pow2(param) && param < n) size = n - param;
    else {
        set_err("invalid table mode or parameter");
        return -1;
    }
    if (!host_out) return (int)size;
    hipStream_t s = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    float2 *t = nullptr;
    int rc = mode == 1 ? build_resolve_table((int)n, direction, &t, s)
           : mode == 2 ? build_plain_table((int)n, (int)param, direction, &t, s)
                       : build_split_table((int)n, (int)param, direction, &t, s);
    if (!rc && t && count) {
        const size_t m = count < size ? count : size;
        rc = hipMemcpyAsync(host_out, t, m * sizeof(float2), hipMemcpyDeviceToHost, s) != hipSuccess;
    }
    rc |= hipStreamSynchronize(s) != hipSuccess;
    if (t) (void)hipFree(t);
    (void)hipStreamDestroy(s);
    if (rc) {
        set_err("twiddle table generation failed");
        return -1;
    }
    return (int)size;
}

int mufft_hip_twiddle_table_mode(int mode, unsigned n, unsigned param, int direction, float *host_out,
                                 unsigned count) {
    return guarded<int>(-1, [&] {
        return mufft_hip_twiddle_table_mode_impl(mode, n, param, direction, host_out, count);
    });
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

}  // extern "C"
