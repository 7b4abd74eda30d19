// Row-kernel dispatch, the generic one-pass-per-launch Stockham kernel (any
// size, any strides), standalone resolve/multiply kernels and the FP64
// twiddle generator.
#include "kernels.h"
#include "row_variants.h"
#include "stockham.h"

namespace mufft {

// Each fills its (direction, kind) table: entries [0, kRowMidIndex), the
// _mid ones [kRowMidIndex, kRowSplitIndex), the _large ones
// [kRowSplitIndex, kNumRowVariants).
void row_kernels_c2c_fwd(const void **out);
void row_kernels_c2c_inv(const void **out);
void row_kernels_r2c(const void **out);
void row_kernels_c2r(const void **out);
void row_kernels_c2c_fwd_mid(const void **out);
void row_kernels_c2c_inv_mid(const void **out);
void row_kernels_r2c_mid(const void **out);
void row_kernels_c2r_mid(const void **out);
void row_kernels_c2c_fwd_large(const void **out);
void row_kernels_c2c_inv_large(const void **out);
void row_kernels_r2c_large(const void **out);
void row_kernels_c2r_large(const void **out);

namespace {
struct RowTables {
    const void *t[4][kNumRowVariants];
    RowTables() {
        row_kernels_c2c_fwd(t[0]);
        row_kernels_c2c_inv(t[1]);
        row_kernels_r2c(t[2]);
        row_kernels_c2r(t[3]);
        row_kernels_c2c_fwd_mid(t[0]);
        row_kernels_c2c_inv_mid(t[1]);
        row_kernels_r2c_mid(t[2]);
        row_kernels_c2r_mid(t[3]);
        row_kernels_c2c_fwd_large(t[0]);
        row_kernels_c2c_inv_large(t[1]);
        row_kernels_r2c_large(t[2]);
        row_kernels_c2r_large(t[3]);
    }
};
const RowTables &row_tables() {
    static const RowTables tabs;
    return tabs;
}
int row_variant_index(int logN, int var) {
    for (int i = 0; i < kNumRowVariants; ++i)
        if (kRowVariants[i].logN == logN && kRowVariants[i].var == var) return i;
    return -1;
}
}  // namespace

int row_variant_list(int logN, int *vars, int max, bool c2r) {
    int n = 0;
    for (int i = 0; i < kNumRowVariants; ++i) {
        if (kRowVariants[i].logN != logN) continue;
        if (!c2r && rv_c2r_only(kRowVariants[i].var)) continue;
        if (vars && n < max) vars[n] = kRowVariants[i].var;
        ++n;
    }
    return n;
}

bool row_variant_registered(int logN, int var) { return row_variant_index(logN, var) >= 0; }

hipError_t launch_rows(int logN, int direction, int kind, int var, const RowArgs &a, hipStream_t s) {
    const int idx = row_variant_index(logN, var);
    if (idx < 0) return hipErrorInvalidValue;
    int table;
    if (kind == KIND_C2C) table = direction < 0 ? 0 : 1;
    else if (kind == KIND_R2C && direction < 0) table = 2;
    else if (kind == KIND_C2R && direction > 0) table = 3;
    else return hipErrorInvalidValue;
    const void *k = row_tables().t[table][idx];
    if (a.batch <= 0) return hipSuccess;
    const int N = 1 << logN;
    const long long nblocks = rv_split(var) > 0
                                  ? (long long)a.batch << rv_split(var)
                                  : ((long long)a.batch + rv_B(N, var) - 1) / rv_B(N, var);
    if (nblocks > 0x7fffffffLL) return hipErrorInvalidValue;
    RowArgs args = a;
    void *params[] = {&args};
    return hipLaunchKernel(k, dim3((unsigned)nblocks), dim3(rv_WG(N, var)), params, 0, s);
}

// ---------------------------------------------------------------------------
// One Stockham pass over global memory (reference pass semantics,
// kernel.c:151-168/234-271/395-453 and their vertical forms kernel.c:483-738,
// generalised to radix 2..16).  Used only where no fused kernel applies.
// ---------------------------------------------------------------------------
template <int R, int DIR>
__global__ void __launch_bounds__(256) fft_pass_kernel(PassArgs a) {
    const long long nb = a.n / R;
    const long long total = nb * (long long)a.ncols * a.batch;
    const bool zp = (a.flags & RF_ZERO_PAD) != 0;
    for (long long gid = (long long)blockIdx.x * 256 + threadIdx.x; gid < total;
         gid += (long long)gridDim.x * 256) {
        const long long x = gid % a.ncols;
        const long long rest = gid / a.ncols;
        const long long b = rest % nb;
        const long long img = rest / nb;
        const float2 *src = a.in + img * a.bs_in + x;
        float2 u[R];
        static_for<0, R>([&](auto M) {
            constexpr int m = decltype(M)::value;
            u[m] = (zp && m >= R / 2) ? make_float2(0.f, 0.f) : src[(b + m * nb) * a.sy_in];
        });
        const long long k = b & (a.p - 1);
        if (a.p > 1) {
            static_for<1, R>([&](auto M) {
                constexpr int m = decltype(M)::value;
                u[m] = cmul(u[m], a.tw[(m - 1) * a.p + k]);
            });
        }
        dft<R, DIR>(u);
        float2 *dst = a.out + img * a.bs_out + x;
        const long long o = (b - k) * R + k;
        static_for<0, R>([&](auto Q) {
            constexpr int q = decltype(Q)::value;
            dst[(o + q * a.p) * a.sy_out] = u[q];
        });
    }
}

template <int DIR>
static const void *pass_kernel_ptr(int radix) {
    switch (radix) {
    case 2: return (const void *)&fft_pass_kernel<2, DIR>;
    case 4: return (const void *)&fft_pass_kernel<4, DIR>;
    case 8: return (const void *)&fft_pass_kernel<8, DIR>;
    case 16: return (const void *)&fft_pass_kernel<16, DIR>;
    default: return nullptr;
    }
}

static unsigned grid_for(long long work, int per_block) {
    long long g = (work + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > 1048576) g = 1048576;  // grid-stride loops cover the rest
    return (unsigned)g;
}

hipError_t launch_pass(int radix, int direction, const PassArgs &a, hipStream_t s) {
    const void *k = direction < 0 ? pass_kernel_ptr<-1>(radix) : pass_kernel_ptr<1>(radix);
    if (!k) return hipErrorInvalidValue;
    const long long work = a.n / radix * (long long)a.ncols * a.batch;
    if (work <= 0) return hipSuccess;
    PassArgs args = a;
    void *params[] = {&args};
    return hipLaunchKernel(k, dim3(grid_for(work, 256)), dim3(256), params, 0, s);
}

// ---------------------------------------------------------------------------
// Standalone real-transform resolves (kernel.c:36-82) for sizes beyond the
// fused kernels, and the pointwise multiply (kernel.c:22-34).
// ---------------------------------------------------------------------------
// n is a power of two (the launchers check): the bin and transform of an
// element come from a mask and a shift, not a 64-bit division.
__global__ void __launch_bounds__(256) resolve_r2c_kernel(ResolveArgs a) {
    const long long total = a.n * a.batch;
    const bool full = (a.flags & RF_FULL) != 0;
    const int lg = __builtin_ctzll((unsigned long long)a.n);
    for (long long gid = (long long)blockIdx.x * 256 + threadIdx.x; gid < total;
         gid += (long long)gridDim.x * 256) {
        const long long k = gid & (a.n - 1);
        const long long g = gid >> lg;
        const float2 *Z = a.in + g * a.idist;
        float2 *X = a.out + g * a.odist;
        const float2 za = Z[k];
        const bool adj = (a.flags & RF_ADJ_C2R) != 0;
        if (k == 0) {
            const float r0 = za.x + za.y, rn = za.x - za.y;
            X[0] = make_float2(r0, adj ? -rn : 0.f);
            X[a.n] = make_float2(rn, adj ? -r0 : 0.f);
        } else {
            const float2 zb = cconj(Z[a.n - k]);
            const float2 fe = cadd(za, zb);
            const float2 fo = cmul(a.rtw[k], csub(za, zb));
            const float sc = adj ? 1.0f : 0.5f;
            X[k] = cscale(cadd(fe, fo), sc);
            if (full) X[k + a.n] = cscale(csub(fe, fo), sc);
        }
    }
}

__global__ void __launch_bounds__(256) resolve_c2r_kernel(ResolveArgs a) {
    const long long total = a.n * a.batch;
    const bool mul = (a.flags & RF_MUL) != 0;
    const bool cj = (a.flags & RF_CONJ2) != 0;
    const int lg = __builtin_ctzll((unsigned long long)a.n);
    for (long long gid = (long long)blockIdx.x * 256 + threadIdx.x; gid < total;
         gid += (long long)gridDim.x * 256) {
        const long long k = gid & (a.n - 1);
        const long long g = gid >> lg;
        if ((a.flags & RF_ZERO_PAD) && k >= a.n / 2) {
            a.out[g * a.odist + k] = make_float2(0.f, 0.f);
            continue;
        }
        const float2 *X = a.in + g * a.idist;
        float2 xa = X[k];
        float2 xb = X[a.n - k];
        if (mul) {
            const float2 *Y = a.in2 + g * a.idist2;
            float2 ya = Y[k], yb = Y[a.n - k];
            if (cj) {
                ya = cconj(ya);
                yb = cconj(yb);
            }
            xa = cscale(cmul(xa, ya), a.norm);
            xb = cscale(cmul(xb, yb), a.norm);
        }
        if (a.flags & RF_ADJ_R2C) {
            xa = k == 0 ? make_float2(xa.x, 0.f) : cscale(xa, 0.5f);
            xb = k == 0 ? make_float2(xb.x, 0.f) : cscale(xb, 0.5f);
        }
        const float2 bc = cconj(xb);
        a.out[g * a.odist + k] = cadd(cadd(xa, bc), cmul(csub(xa, bc), a.rtw[k]));
    }
}

__global__ void __launch_bounds__(256) mul_kernel(MulArgs a) {
    const long long total = a.n * a.batch;
    const bool cj = (a.flags & RF_CONJ2) != 0;
    for (long long gid = (long long)blockIdx.x * 256 + threadIdx.x; gid < total;
         gid += (long long)gridDim.x * 256) {
        const long long i = gid % a.n;
        const long long g = gid / a.n;
        float2 y = a.b[g * a.bdist + i];
        if (cj) y = cconj(y);
        a.out[g * a.odist + i] = cscale(cmul(a.a[g * a.adist + i], y), a.norm);
    }
}

// Folds a full n-bin spectrum onto bins 0..n/2 for the adjoint of a full r2c:
// G_k = H_k + conj(H_{n-k}) for 0 < k < n/2, G_0 = H_0, G_{n/2} = H_{n/2}.
__global__ void __launch_bounds__(256) fold_kernel(MulArgs a) {
    const long long half = a.n / 2;
    const long long total = (half + 1) * a.batch;
    for (long long gid = (long long)blockIdx.x * 256 + threadIdx.x; gid < total;
         gid += (long long)gridDim.x * 256) {
        const long long k = gid % (half + 1);
        const long long g = gid / (half + 1);
        const float2 *H = a.a + g * a.adist;
        float2 v = H[k];
        if (k > 0 && k < half) v = cadd(v, cconj(H[a.n - k]));
        a.out[g * a.odist + k] = v;
    }
}

// out[i] = norm * sum_b a_b[i] * op(b_b[i]) (op = conj with RF_CONJ2): the
// batch reduction of a filter gradient.  One thread per bin, the batch loop
// in a fixed order (deterministic).
__global__ void __launch_bounds__(256) mul_reduce_kernel(MulArgs a) {
    const bool cj = (a.flags & RF_CONJ2) != 0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < a.n;
         i += (long long)gridDim.x * 256) {
        float2 acc = make_float2(0.f, 0.f);
        for (int g = 0; g < a.batch; ++g) {
            float2 y = a.b[g * a.bdist + i];
            if (cj) y = cconj(y);
            acc = cadd(acc, cmul(a.a[g * a.adist + i], y));
        }
        a.out[i] = cscale(acc, a.norm);
    }
}

// Overlap-add of nblocks filtered blocks of n reals at hop n/2:
// out[b*h + i] = y_b[i] + (b ? y_{b-1}[h + i] : tail_in[i]); tail_out = y_last[h..n).
__global__ void __launch_bounds__(256) ola_kernel(const float *__restrict__ y, float *__restrict__ out,
                                                  const float *__restrict__ tail_in,
                                                  float *__restrict__ tail_out, long long n,
                                                  long long nblocks) {
    const long long h = n / 2;
    const long long total = h * nblocks;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < total + h;
         i += (long long)gridDim.x * 256) {
        if (i < total) {
            const long long b = i / h, r = i % h;
            out[i] = y[b * n + r] + (b > 0 ? y[(b - 1) * n + h + r] : tail_in[r]);
        } else {
            const long long r = i - total;
            tail_out[r] = y[(nblocks - 1) * n + h + r];
        }
    }
}

hipError_t launch_ola(const float *y, float *out, const float *tail_in, float *tail_out,
                      long long n, long long nblocks, hipStream_t s) {
    const long long work = n / 2 * (nblocks + 1);
    if (nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(ola_kernel, dim3(grid_for(work, 256)), dim3(256), 0, s, y, out, tail_in,
                       tail_out, n, nblocks);
    return hipGetLastError();
}

// Streaming copy with the row kernels' access width and cache policy (8 bytes
// per lane, non-temporal): the achievable-HBM reference the FFT kernels are
// compared with.  Each lane moves 16 elements per step.
__global__ void __launch_bounds__(256) stream_copy_kernel(const u32x2 *__restrict__ in,
                                                          u32x2 *__restrict__ out, long long n) {
    constexpr int PER = 16;
    const long long chunk = 256LL * PER;
    for (long long base = (long long)blockIdx.x * chunk; base < n; base += (long long)gridDim.x * chunk) {
        u32x2 v[PER];
        for (int e = 0; e < PER; ++e) {
            const long long i = base + threadIdx.x + 256LL * e;
            v[e] = i < n ? __builtin_nontemporal_load(in + i) : u32x2{0u, 0u};
        }
        for (int e = 0; e < PER; ++e) {
            const long long i = base + threadIdx.x + 256LL * e;
            if (i < n) __builtin_nontemporal_store(v[e], out + i);
        }
    }
}

hipError_t launch_stream_copy(const float2 *in, float2 *out, long long n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const long long blocks = (n + 4095) / 4096;
    hipLaunchKernelGGL(stream_copy_kernel, dim3((unsigned)(blocks < 16384 ? blocks : 16384)), dim3(256), 0,
                       s, reinterpret_cast<const u32x2 *>(in), reinterpret_cast<u32x2 *>(out), n);
    return hipGetLastError();
}

hipError_t launch_fold(const MulArgs &a, hipStream_t s) {
    const long long work = (a.n / 2 + 1) * a.batch;
    if (work <= 0) return hipSuccess;
    hipLaunchKernelGGL(fold_kernel, dim3(grid_for(work, 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_mul_reduce(const MulArgs &a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(mul_reduce_kernel, dim3(grid_for(a.n, 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_resolve_r2c(const ResolveArgs &a, hipStream_t s) {
    const long long work = a.n * a.batch;
    if (work <= 0) return hipSuccess;
    if ((a.n & (a.n - 1)) != 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(resolve_r2c_kernel, dim3(grid_for(work, 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_resolve_c2r(const ResolveArgs &a, hipStream_t s) {
    const long long work = a.n * a.batch;
    if (work <= 0) return hipSuccess;
    if ((a.n & (a.n - 1)) != 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(resolve_c2r_kernel, dim3(grid_for(work, 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_mul(const MulArgs &a, hipStream_t s) {
    const long long work = a.n * a.batch;
    if (work <= 0) return hipSuccess;
    hipLaunchKernelGGL(mul_kernel, dim3(grid_for(work, 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Twiddles in FP64, rounded once to FP32.  The phase expression is the
// reference's twiddle() (fft.c:102-106): (M_PI * direction * k) / p in double.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) twiddle_kernel(TwiddleArgs a) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= a.count) return;
    const double pi = 3.14159265358979323846;
    const double dir = (double)a.direction;
    if (a.mode == 0) {
        // Locate the pass owning entry e: pass s covers [p_s - r0, p_{s+1} - r0).
        int s = 1;
        while (s + 1 < a.npass && (long long)e >= a.p[s + 1] - a.r0) ++s;
        const long long local = (long long)e - (a.p[s] - a.r0);
        const long long m = local / a.p[s] + 1;
        const long long k = local % a.p[s];
        const long long half = (long long)a.radix[s] * a.p[s] / 2;
        const double phase = (pi * dir * (double)(k * m)) / (double)half;
        a.out[e] = make_float2((float)cos(phase), (float)sin(phase));
    } else if (a.mode == 1) {
        // fft.c:441-455: W = (0, dir) * twiddle(dir, e, n), complex multiply in float.
        const double phase = (pi * dir * (double)e) / (double)a.n;
        const float c = (float)cos(phase);
        const float sn = (float)sin(phase);
        const float d = (float)a.direction;
        const float zero = 0.0f;
        a.out[e] = make_float2(__fsub_rn(__fmul_rn(zero, c), __fmul_rn(d, sn)),
                               __fadd_rn(__fmul_rn(zero, sn), __fmul_rn(c, d)));
    } else if (a.mode == 2) {
        const double phase = (pi * dir * (double)e) / (double)(a.n / 2);
        a.out[e] = make_float2((float)cos(phase), (float)sin(phase));
    } else {
        const long long n = e % a.p[0];
        const long long k1 = e / a.p[0] + 1;
        const double phase = (pi * dir * (double)(n * k1)) / (double)(a.n / 2);
        a.out[e] = make_float2((float)cos(phase), (float)sin(phase));
    }
}

hipError_t launch_twiddles(const TwiddleArgs &a, hipStream_t s) {
    if (a.count <= 0) return hipSuccess;
    hipLaunchKernelGGL(twiddle_kernel, dim3((a.count + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace mufft
