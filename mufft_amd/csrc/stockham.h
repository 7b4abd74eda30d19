// Fused multi-pass Stockham transform on register-resident points.
//
// Semantics follow the reference's pass definition (kernel.c:151-168,
// 234-271, 395-453 and the Stockham narrative in doxygen/fft.md:244-334):
// pass s has radix r and stride p; butterfly b (b < N/r) reads points
// b + m*N/r, multiplies point m by exp(dir*2*pi*i*k*m/(r*p)) with k = b mod p,
// takes an r-point DFT and writes output q to (b - k)*r + k + q*p.  Unlike the
// reference, which runs one radix-2/4/8 pass per sweep over memory, all passes
// of a transform run inside one workgroup: each thread keeps E points in
// registers, butterflies are in-register DFTs of radix up to 64, and points
// move between passes through LDS only.
#pragma once

#include "fft_device.h"

namespace mufft {

// Exchange modes between passes.
enum XMode : int {
    XM_PAD = 0,    // complex points, one spare LDS slot every E points
    XM_SPLIT = 1,  // real parts then imaginary parts through a float array (half the LDS)
    XM_FLAT = 2,   // complex points, no padding (smallest footprint)
};

// Twiddle application: TW_FULL loads every exp(dir*2*pi*i*k*m/(r*p)) from the
// table; TW_SPLIT loads only m < M and multiples of M (M = 8 for r = 64/32,
// 4 for r = 16) and forms the rest as one product, cutting the loads per
// radix-32 butterfly from 31 to 10 at the cost of about one extra rounding.
enum TwMode : int { TW_FULL = 0, TW_SPLIT = 1 };

template <int r>
constexpr int tw_split_m() { return r >= 32 ? 8 : (r >= 16 ? 4 : r); }

// Multiplies u[m] by the pass twiddle exp(dir*2*pi*i*k*m/(r*p)) = table
// entry base + (m-1)*p + k, read through twl(index, k) = table[index + k].
template <int r, int p, int TWM, typename TwLoad>
__device__ __forceinline__ void apply_twiddles(float2 (&u)[r], int k, int base, TwLoad twl) {
    if constexpr (TWM == TW_SPLIT && r >= 16) {
        constexpr int ML = tw_split_m<r>();
        float2 lo[ML];
        static_for<1, ML>([&](auto M) {
            constexpr int m = decltype(M)::value;
            lo[m] = twl(base + (m - 1) * p, k);
            u[m] = cmul(u[m], lo[m]);
        });
        static_for<1, r / ML>([&](auto H) {
            constexpr int mh = decltype(H)::value;
            const float2 hi = twl(base + (mh * ML - 1) * p, k);
            u[mh * ML] = cmul(u[mh * ML], hi);
            static_for<1, ML>([&](auto M) {
                constexpr int m = decltype(M)::value;
                u[mh * ML + m] = cmul(u[mh * ML + m], cmul(lo[m], hi));
            });
        });
    } else {
        static_for<1, r>([&](auto M) {
            constexpr int m = decltype(M)::value;
            u[m] = cmul(u[m], twl(base + (m - 1) * p, k));
        });
    }
}

// Twiddle-table readers: plain global loads, or buffer loads with the table
// index as a scalar offset.
struct TwGlobal {
    const float2 *__restrict__ t;
    __device__ __forceinline__ float2 operator()(int idx, int k) const { return t[idx + k]; }
};
struct TwBuffer {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ float2 operator()(int idx, int k) const {
        return bload(r, (unsigned)k * 8u, (unsigned)idx * 8u);
    }
};

// Number of passes of a sequence whose passes have radix E but one (the
// remainder, radix N / E^(passes-1); see seq_radix()).
MUFFT_HD constexpr int np_for(int N, int E) {
    return E >= N ? 1 : (ilog2(N) + ilog2(E) - 1) / ilog2(E);
}

// Orders the LDS exchange between the threads of one transform.  When they
// all sit in one wavefront (WAVE), the wave's own LDS accesses are already
// executed in order, so only the compiler must not move LDS accesses across
// the point (wavefront-scope fences): no workgroup barrier, and the other
// waves of the workgroup run on.
template <bool WAVE>
__device__ __forceinline__ void exchange_sync() {
    if constexpr (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// v[c] holds point lt + T*c on entry (T = N/E) and holds output point lt + T*c
// on exit.  The passes follow seq_radix(): radix R (R <= E, default E) and
// one remainder pass, first or last (REM_FIRST); each thread runs J = E/r
// butterflies of a pass, butterfly j being b = lt + T*j over its points
// v[j + m*J].  `twl` reads the twiddle table of that sequence.  `Slot(o)`
// maps a point index to its LDS slot; `L` is the complex LDS view, `Lf` the
// float view used by the split exchange.  The first pass may hold its points
// under another thread numbering: with lt0 >= 0, v[c] holds point lt0 + T*c
// on entry (lt0 a permutation of the threads); the first pass has no
// twiddles and its results go through LDS, so the later passes and the
// output use lt.
template <int N, int E, int DIR, int XM, int TWM, int R = E, bool WAVE = false, bool REM_FIRST = true,
          typename TwLoad, typename Slot>
__device__ __forceinline__ void stockham_passes(float2 (&v)[E], int lt, TwLoad twl,
                                                float2 *L, float *Lf, Slot slot, int lt0 = -1) {
    constexpr int T = N / E;
    constexpr int NP = np_for(N, R);
    if (lt0 < 0) lt0 = lt;
    static_assert(R <= E && E % R == 0, "pass radix divides the points per thread");
    static_assert(!WAVE || T <= 64, "wave-local exchange needs the transform within one wave");
    // Slot of point lt + T*c for the exchange reads.  The slot maps used here
    // (padded: o + (o >> log2 E), scaled by a tile width and offset by a lane)
    // split exactly over a multiple of E, so when E divides T every read is
    // one per-thread base plus a compile-time offset; the compiler then keeps
    // one address live instead of one per point, which it would otherwise
    // hold (or spill) across all passes.
    const int rd0 = slot(lt), rz = slot(0);
    auto rd_slot = [&](auto C) {
        constexpr int c = decltype(C)::value;
        if constexpr (T % E == 0) return rd0 + (slot(T * c) - rz);
        else return slot(lt + T * c);
    };
    static_for<0, NP>([&](auto S) {
        constexpr int s = decltype(S)::value;
        constexpr int r = seq_radix(N, R, NP, s, REM_FIRST);
        constexpr int p = seq_p(N, R, NP, s, REM_FIRST);
        constexpr int J = E / r;   // butterflies per thread in this pass
        static_for<0, J>([&](auto JJ) {
            constexpr int j = decltype(JJ)::value;
            float2 u[r];
            static_for<0, r>([&](auto M) { u[decltype(M)::value] = v[j + decltype(M)::value * J]; });
            if constexpr (s > 0) {
                const int k = (lt + T * j) & (p - 1);
                apply_twiddles<r, p, TWM>(u, k, tw_offset(seq_first(N, R, NP, REM_FIRST), p), twl);
            }
            dft<r, DIR>(u);
            static_for<0, r>([&](auto Q) { v[j + decltype(Q)::value * J] = u[decltype(Q)::value]; });
        });
        if constexpr (s < NP - 1) {
            // Output q of butterfly b goes to point (b - k)*r + k + q*p with
            // k = b mod p; the next pass reads points lt + T*c.
            auto publish = [&](auto get, auto put) {
                static_for<0, J>([&](auto JJ) {
                    constexpr int j = decltype(JJ)::value;
                    const int b = (s == 0 ? lt0 : lt) + T * j;
                    const int k = b & (p - 1);
                    const int base = (b - k) * r + k;
                    static_for<0, r>([&](auto Q) {
                        constexpr int q = decltype(Q)::value;
                        put(slot(base + q * p), get(v[j + q * J]));
                    });
                });
            };
            if constexpr (XM != XM_SPLIT) {
                publish([](float2 z) { return z; }, [&](int o, float2 z) { L[o] = z; });
                exchange_sync<WAVE>();
                static_for<0, E>([&](auto C) { v[decltype(C)::value] = L[rd_slot(C)]; });
                exchange_sync<WAVE>();
            } else {
                publish([](float2 z) { return z.x; }, [&](int o, float x) { Lf[o] = x; });
                exchange_sync<WAVE>();
                static_for<0, E>([&](auto C) { v[decltype(C)::value].x = Lf[rd_slot(C)]; });
                exchange_sync<WAVE>();
                publish([](float2 z) { return z.y; }, [&](int o, float y) { Lf[o] = y; });
                exchange_sync<WAVE>();
                static_for<0, E>([&](auto C) { v[decltype(C)::value].y = Lf[rd_slot(C)]; });
                exchange_sync<WAVE>();
            }
        }
    });
}

}  // namespace mufft
