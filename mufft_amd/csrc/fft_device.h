// Device-side building blocks: complex arithmetic, constant twiddles and the
// in-register radix-R DFT used by every gfx950 kernel of this library.
#pragma once

#include <hip/hip_runtime.h>
#include <type_traits>

#include "geometry.h"

namespace mufft {

// Compile-time loop: f(std::integral_constant<int, i>) for i in [B, E).
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F &&f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// a * (dir * i)
template <int DIR>
__device__ __forceinline__ float2 mul_dir_i(float2 a) {
    if constexpr (DIR > 0) return make_float2(-a.y, a.x);
    else return make_float2(a.y, -a.x);
}

// a * exp(DIR * 2 pi i * j / 64).
template <int j, int DIR>
__device__ __forceinline__ float2 twc64(float2 a) {
    constexpr float h = 0.70710678118654752440f;
    constexpr float d = DIR > 0 ? 1.0f : -1.0f;
    if constexpr (j == 0) return a;
    else if constexpr (j == 16) return mul_dir_i<DIR>(a);
    else if constexpr (j == 32) return make_float2(-a.x, -a.y);
    else if constexpr (j == 48) return mul_dir_i<-DIR>(a);
    else if constexpr (j == 8) return make_float2(h * (a.x - d * a.y), h * (a.y + d * a.x));
    else if constexpr (j == 24) return make_float2(h * (-a.x - d * a.y), h * (d * a.x - a.y));
    else if constexpr (j == 40) return make_float2(h * (d * a.y - a.x), h * (-a.y - d * a.x));
    else if constexpr (j == 56) return make_float2(h * (a.x + d * a.y), h * (a.y - d * a.x));
    else {
        constexpr float c = (float)cos64(j);
        constexpr float s = (float)(d * sin64(j));
        return make_float2(a.x * c - a.y * s, a.x * s + a.y * c);
    }
}

// a * exp(DIR * 2 pi i * J / R) with a compile-time exponent; the eighth-turn
// cases use the exact structure of the constant instead of a full multiply.
// R is a power of two; above 64 the constant comes from cos_turn / sin_turn.
template <int J, int R, int DIR>
__device__ __forceinline__ float2 twc(float2 a) {
    static_assert(R > 0 && (R & (R - 1)) == 0, "R is a power of two");
    if constexpr (R > 64) {
        constexpr int jr = ((J % R) + R) % R;
        if constexpr (jr % (R / 64) == 0) {
            return twc<jr / (R / 64), 64, DIR>(a);
        } else {
            constexpr float d = DIR > 0 ? 1.0f : -1.0f;
            constexpr float c = (float)cos_turn(jr, R);
            constexpr float s = (float)(d * sin_turn(jr, R));
            return make_float2(a.x * c - a.y * s, a.x * s + a.y * c);
        }
    } else {
        return twc64<((((J % R) + R) % R) * (64 / R)) & 63, DIR>(a);
    }
}

__host__ __device__ constexpr int bitrev(int x, int bits) {
    return bits == 0 ? 0 : (((x & 1) << (bits - 1)) | bitrev(x >> 1, bits - 1));
}

// In-register DFT of R points, X[k] = sum_n v[n] exp(DIR * 2 pi i n k / R).
// Decimation in frequency in place; the bit-reversed result order is undone
// by a compile-time register renaming (no data movement after unrolling).
template <int R, int DIR>
__device__ __forceinline__ void dft(float2 (&v)[R]) {
    if constexpr (R > 1) {
        constexpr int L = ilog2(R);
        static_for<0, L>([&](auto st) {
            constexpr int span = R >> (decltype(st)::value + 1);
            static_for<0, R / (2 * span)>([&](auto blk) {
                static_for<0, span>([&](auto jj) {
                    constexpr int j = decltype(jj)::value;
                    constexpr int i0 = decltype(blk)::value * 2 * span + j;
                    const float2 a = v[i0];
                    const float2 b = v[i0 + span];
                    v[i0] = cadd(a, b);
                    v[i0 + span] = twc<j, 2 * span, DIR>(csub(a, b));
                });
            });
        });
        float2 t[R];
        static_for<0, R>([&](auto k) { t[decltype(k)::value] = v[bitrev(decltype(k)::value, L)]; });
        static_for<0, R>([&](auto k) { v[decltype(k)::value] = t[decltype(k)::value]; });
    }
}

// ---------------------------------------------------------------------------
// Buffer-resource access.  A strided access base + lt + T*c becomes one VGPR
// offset (lt) plus a scalar offset (T*c) per slot, instead of one 64-bit
// address per slot; the descriptor's byte limit also bounds every access
// (out-of-range loads read 0, out-of-range stores are dropped).
// ---------------------------------------------------------------------------
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// Word 3 of a raw (untyped) gfx9 buffer descriptor: 32-bit data format.
constexpr int kBufWord3 = 0x00020000;

// The base pointer and size must be uniform across the wave.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *p, unsigned bytes) {
    const unsigned long long u = reinterpret_cast<unsigned long long>(p);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
    void *q = reinterpret_cast<void *>(((unsigned long long)hi << 32) | lo);
    return __builtin_amdgcn_make_buffer_rsrc(q, (short)0, (int)__builtin_amdgcn_readfirstlane(bytes),
                                             kBufWord3);
}

__device__ __forceinline__ float2 bload(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}

__device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff,
                                       float2 v) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, (int)voff, (int)soff, 0);
}

// Streaming variants (cache policy "nt", aux bit 1): for data that one
// kernel reads or writes exactly once, so it does not displace reused lines
// (twiddle tables, operands read twice) from the caches.
__device__ __forceinline__ float2 bload_nt(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 2));
}

__device__ __forceinline__ void bstore_nt(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff,
                                          float2 v) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, (int)voff, (int)soff, 2);
}

// Thread numbering for the c2r pre-pass resolve, which pairs bin k with bin
// N - k.  When T >= 64 threads hold points lt + T*c, the owners of k and
// N - k are threads lt and T - lt; numbering thread t as mirror_lt<T>(t)
// puts every such pair in lanes i and i ^ 32 of one wave.  Lanes 0 and 32 of
// the first wave get 0 and T/2, which are their own partners.
template <int T>
__device__ __forceinline__ int mirror_lt(int t) {
    static_assert(T >= 64 && (T & (T - 1)) == 0, "whole waves of one transform");
    const int q = ((t >> 6) << 5) | (t & 31);
    return (t & 32) == 0 ? q : (q == 0 ? T / 2 : T - q);
}

// The value v of lane i ^ 32 (the other half of the wave).
__device__ __forceinline__ float2 swap_half_waves(float2 v) {
    return make_float2(__shfl_xor(v.x, 32), __shfl_xor(v.y, 32));
}

// Padded LDS slot of point o when every E points get one spare slot.
template <int LOGE>
__device__ __forceinline__ int pad_slot(int o) { return o + (o >> LOGE); }

// blockIdx remap so consecutive logical blocks land on the same XCD (gfx950
// dispatches workgroups round-robin over its 8 XCDs, each with its own L2):
// neighbouring column tiles share cache lines, so keeping them on one L2
// avoids fetching the same line into several L2s.
// The blocks beyond the last multiple of 8 keep their own index.
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb) {
    const unsigned full = nb & ~7u;
    if (b >= full) return b;
    return (b & 7u) * (full >> 3) + (b >> 3);
}

}  // namespace mufft
