// Fused single-workgroup transform of contiguous sequences: the hot kernel
// for 1D c2c/r2c/c2r and for 2D row passes.
//
// One HBM read and one HBM write per point: the load phase is the reference's
// first pass (kernel.c:84-122, 170-232, 273-393; zero padding as in the
// "half" kernels), the store phase is its last pass, and the real-transform
// resolve passes (kernel.c:36-82) are fused into the prologue (c2r) or the
// epilogue (r2c).  The convolution multiply (kernel.c:22-34) can be fused
// into the prologue so mufft_execute_conv_output is a single launch.
#pragma once

#include "kernels.h"
#include "stockham.h"

namespace mufft {

template <int N, int DIR, int KIND>
__global__ void __launch_bounds__(row_WG(N)) fft_rows_kernel(RowArgs a) {
    constexpr int E = row_E(N);
    constexpr int LOGE = ilog2(E);
    constexpr int T = row_T(N);
    constexpr int B = row_B(N);
    constexpr bool SPLIT = row_split(N);
    constexpr int LDSN = row_lds_per(N);
    constexpr int SMEM = SPLIT ? (LDSN + 1) / 2 : B * LDSN;
    __shared__ float2 smem[SMEM];

    const int tid = threadIdx.x;
    const int lt = tid & (T - 1);
    const int sub = tid / T;
    const long long g = (long long)blockIdx.x * B + sub;
    const bool active = g < a.batch;
    float2 *L = smem + (SPLIT ? 0 : sub * LDSN);
    float *Lf = reinterpret_cast<float *>(smem);
    auto slot = [](int o) { return pad_slot<LOGE>(o); };

    float2 v[E];
    const unsigned flags = a.flags;
    const bool zp = (flags & RF_ZERO_PAD) != 0;

    if constexpr (KIND == KIND_C2R) {
        // Pre-pass resolve (kernel.c:36-45), optionally after the fused multiply.
        const float2 *X = a.in + g * a.idist;
        const float2 *Y = a.in2 + g * a.idist2;
        const bool mul = (flags & RF_MUL) != 0;
        const bool cj = (flags & RF_CONJ2) != 0;
        static_for<0, E>([&](auto C) {
            constexpr int c = decltype(C)::value;
            const int k = lt + T * c;
            float2 z = make_float2(0.f, 0.f);
            if (active && !(zp && c >= E / 2)) {
                float2 xa = X[k];
                float2 xb = X[N - k];
                if (mul) {
                    float2 ya = Y[k];
                    float2 yb = Y[N - k];
                    if (cj) {
                        ya = cconj(ya);
                        yb = cconj(yb);
                    }
                    xa = cscale(cmul(xa, ya), a.norm);
                    xb = cscale(cmul(xb, yb), a.norm);
                }
                const float2 bc = cconj(xb);
                z = cadd(cadd(xa, bc), cmul(csub(xa, bc), a.rtw[k]));
            }
            v[c] = z;
        });
    } else {
        const float2 *X = a.in + g * a.idist;
        if (KIND == KIND_C2C && (flags & RF_MUL)) {
            const float2 *Y = a.in2 + g * a.idist2;
            const bool cj = (flags & RF_CONJ2) != 0;
            static_for<0, E>([&](auto C) {
                constexpr int c = decltype(C)::value;
                float2 z = make_float2(0.f, 0.f);
                if (active) {
                    float2 y = Y[lt + T * c];
                    if (cj) y = cconj(y);
                    z = cscale(cmul(X[lt + T * c], y), a.norm);
                }
                v[c] = z;
            });
        } else {
            static_for<0, E>([&](auto C) {
                constexpr int c = decltype(C)::value;
                v[c] = (active && !(zp && c >= E / 2)) ? X[lt + T * c] : make_float2(0.f, 0.f);
            });
        }
    }

    stockham_passes<N, DIR, SPLIT>(v, lt, a.tw, L, Lf, slot);

    float2 *Xo = a.out + g * a.odist;
    if constexpr (KIND == KIND_R2C) {
        // Post-pass resolve (kernel.c:47-82): bin k needs Z[k] and Z[N - k],
        // which another thread holds, so Z goes through LDS once more.
        const bool full = (flags & RF_FULL) != 0;
        auto emit = [&](int c, int k, float2 za, float2 zp_) {
            if (k == 0) {
                Xo[0] = make_float2(za.x + za.y, 0.f);
                Xo[N] = make_float2(za.x - za.y, 0.f);
            } else {
                const float2 zb = cconj(zp_);
                const float2 fe = cadd(za, zb);
                const float2 fo = cmul(a.rtw[k], csub(za, zb));
                Xo[k] = cscale(cadd(fe, fo), 0.5f);
                if (full) Xo[k + N] = cscale(csub(fe, fo), 0.5f);
            }
        };
        if constexpr (!SPLIT) {
            static_for<0, E>([&](auto C) {
                constexpr int c = decltype(C)::value;
                L[slot(lt + T * c)] = v[c];
            });
            __syncthreads();
            if (active) {
                static_for<0, E>([&](auto C) {
                    constexpr int c = decltype(C)::value;
                    const int k = lt + T * c;
                    emit(c, k, v[c], L[slot((N - k) & (N - 1))]);
                });
            }
        } else {
            float px[E];
            static_for<0, E>([&](auto C) {
                constexpr int c = decltype(C)::value;
                Lf[slot(lt + T * c)] = v[c].x;
            });
            __syncthreads();
            static_for<0, E>([&](auto C) {
                constexpr int c = decltype(C)::value;
                px[c] = Lf[slot((N - (lt + T * c)) & (N - 1))];
            });
            __syncthreads();
            static_for<0, E>([&](auto C) {
                constexpr int c = decltype(C)::value;
                Lf[slot(lt + T * c)] = v[c].y;
            });
            __syncthreads();
            if (active) {
                static_for<0, E>([&](auto C) {
                    constexpr int c = decltype(C)::value;
                    const int k = lt + T * c;
                    emit(c, k, v[c], make_float2(px[c], Lf[slot((N - k) & (N - 1))]));
                });
            }
        }
    } else {
        if (active) {
            static_for<0, E>([&](auto C) {
                constexpr int c = decltype(C)::value;
                Xo[lt + T * c] = v[c];
            });
        }
    }
}

template <int DIR, int KIND>
static const void *row_kernel_ptr(int logN) {
    switch (logN) {
    case 1: return (const void *)&fft_rows_kernel<2, DIR, KIND>;
    case 2: return (const void *)&fft_rows_kernel<4, DIR, KIND>;
    case 3: return (const void *)&fft_rows_kernel<8, DIR, KIND>;
    case 4: return (const void *)&fft_rows_kernel<16, DIR, KIND>;
    case 5: return (const void *)&fft_rows_kernel<32, DIR, KIND>;
    case 6: return (const void *)&fft_rows_kernel<64, DIR, KIND>;
    case 7: return (const void *)&fft_rows_kernel<128, DIR, KIND>;
    case 8: return (const void *)&fft_rows_kernel<256, DIR, KIND>;
    case 9: return (const void *)&fft_rows_kernel<512, DIR, KIND>;
    case 10: return (const void *)&fft_rows_kernel<1024, DIR, KIND>;
    case 11: return (const void *)&fft_rows_kernel<2048, DIR, KIND>;
    case 12: return (const void *)&fft_rows_kernel<4096, DIR, KIND>;
    case 13: return (const void *)&fft_rows_kernel<8192, DIR, KIND>;
    case 14: return (const void *)&fft_rows_kernel<16384, DIR, KIND>;
    case 15: return (const void *)&fft_rows_kernel<32768, DIR, KIND>;
    default: return nullptr;
    }
}

}  // namespace mufft
