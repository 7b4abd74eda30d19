// Fused multi-pass Stockham transform on register-resident points.
//
// Semantics follow the reference's pass definition (kernel.c:151-168,
// 234-271, 395-453 and the Stockham narrative in doxygen/fft.md:244-334):
// pass s has radix r and stride p; butterfly b (b < N/r) reads points
// b + m*N/r, multiplies point m by exp(dir*2*pi*i*k*m/(r*p)) with k = b mod p,
// takes an r-point DFT and writes output q to (b - k)*r + k + q*p.  Unlike the
// reference, which runs one radix-2/4/8 pass per sweep over memory, all passes
// of a transform run inside one workgroup: each thread keeps E points in
// registers, butterflies are in-register DFTs of radix up to 32, and points
// move between passes through LDS only.
#pragma once

#include "fft_device.h"

namespace mufft {

// v[c] holds point lt + T*c on entry (T = N/E) and holds output point lt + T*c
// on exit.  `tw` is the twiddle table of the sequence (E, ..., E, r_last).
// `Slot(o)` maps a point index to its LDS slot; `L` is the complex LDS view,
// `Lf` the float view used by the split exchange.
template <int N, int DIR, bool SPLIT, typename Slot>
__device__ __forceinline__ void stockham_passes(float2 (&v)[row_E(N)], int lt,
                                                const float2 *__restrict__ tw,
                                                float2 *L, float *Lf, Slot slot) {
    constexpr int E = row_E(N);
    constexpr int T = N / E;
    constexpr int NP = row_NP(N);
    static_for<0, NP>([&](auto S) {
        constexpr int s = decltype(S)::value;
        constexpr int r = seq_radix(N, E, NP, s);
        constexpr int p = seq_p(E, s);
        constexpr int J = E / r;   // butterflies per thread in this pass
        static_for<0, J>([&](auto JJ) {
            constexpr int j = decltype(JJ)::value;
            float2 u[r];
            static_for<0, r>([&](auto M) { u[decltype(M)::value] = v[j + decltype(M)::value * J]; });
            if constexpr (s > 0) {
                const int k = (lt + T * j) & (p - 1);
                const float2 *t = tw + tw_offset(E, p) + k;
                static_for<1, r>([&](auto M) {
                    constexpr int m = decltype(M)::value;
                    u[m] = cmul(u[m], t[(m - 1) * p]);
                });
            }
            dft<r, DIR>(u);
            static_for<0, r>([&](auto Q) { v[j + decltype(Q)::value * J] = u[decltype(Q)::value]; });
        });
        if constexpr (s < NP - 1) {
            // Not the last pass: r == E and J == 1.  Output q goes to point
            // (lt - k)*E + k + q*p; the next pass reads points lt + T*c.
            const int k = lt & (p - 1);
            const int base = (lt - k) * E + k;
            if constexpr (!SPLIT) {
                static_for<0, E>([&](auto Q) {
                    constexpr int q = decltype(Q)::value;
                    L[slot(base + q * p)] = v[q];
                });
                __syncthreads();
                static_for<0, E>([&](auto C) {
                    constexpr int c = decltype(C)::value;
                    v[c] = L[slot(lt + T * c)];
                });
                __syncthreads();
            } else {
                static_for<0, E>([&](auto Q) {
                    constexpr int q = decltype(Q)::value;
                    Lf[slot(base + q * p)] = v[q].x;
                });
                __syncthreads();
                static_for<0, E>([&](auto C) {
                    constexpr int c = decltype(C)::value;
                    v[c].x = Lf[slot(lt + T * c)];
                });
                __syncthreads();
                static_for<0, E>([&](auto Q) {
                    constexpr int q = decltype(Q)::value;
                    Lf[slot(base + q * p)] = v[q].y;
                });
                __syncthreads();
                static_for<0, E>([&](auto C) {
                    constexpr int c = decltype(C)::value;
                    v[c].y = Lf[slot(lt + T * c)];
                });
                __syncthreads();
            }
        }
    });
}

}  // namespace mufft
