// One transform of N = N1 * N2 complex points spread over N1 workgroups
// (decimation in frequency by N1, then a fused N2-point transform).
//
// Workgroup k1 of a transform produces the output points X[N1*k2 + k1]:
//   y[n]          = w_N^(dir*n*k1) * sum_{n1 < N1} x[n + N2*n1] * w_N1^(dir*n1*k1),   n < N2
//   X[N1*k2 + k1] = sum_{n < N2} y[n] * w_N2^(dir*n*k2)
// so each workgroup keeps the register and LDS footprint of an N2-point
// transform (several workgroups per CU) while the whole transform still costs
// one compulsory HBM read and write.  The N1 workgroups of a transform are
// consecutive logical blocks on one XCD (xcd_block): the input each of them
// reads in full is fetched from HBM once and served to the others from that
// XCD's L2, and their interleaved output points fill the same L2 lines.
//
// The real transforms use N1 = 2.  r2c: the reference's post-pass resolve
// (kernel.c:47-82) pairs Z[k] with Z[N - k]; with k = 2*k2 + k1 the partner
// is 2*(N2 - k2 - k1) + k1, in the same workgroup.  c2r: the pre-pass resolve
// (kernel.c:36-45) forms each Z[m] from X[m] and X[N - m] as the points are
// loaded.  The flags are those of the single-workgroup kernel (rows_impl.h):
// zero padding, fused multiply, full r2c spectrum and the adjoint weightings.
#pragma once

#include "kernels.h"
#include "stockham.h"

namespace mufft {

// Workgroup k1 of one transform: X / Y / Xo are the transform's input, second
// operand and output, L / Lf its LDS (rv_lds_per(N2, VAR) slots).  All T
// threads of the workgroup call it.
// FIXED: compile-time flags, or RF_DYNAMIC (as in row_transform).
template <int N2, int N1, int DIR, int KIND, int VAR, unsigned FIXED = RF_DYNAMIC>
__device__ __forceinline__ void split_transform(const float2 *X, const float2 *Y, float2 *Xo, int k1,
                                                int lt, const float2 *tw, const float2 *stw,
                                                const float2 *rtw, unsigned flags, float norm,
                                                float2 *L, float *Lf) {
    constexpr int E = rv_E(N2, VAR);
    constexpr int LOGE = ilog2(E);
    constexpr int T = rv_T(N2, VAR);
    constexpr int XM = rv_xm(VAR);
    constexpr int TWM = rv_twm(VAR);
    constexpr bool SPLIT = XM == XM_SPLIT;
    constexpr int NT = N1 * N2;
    static_assert(rv_B(N2, VAR) == 1, "one sub-transform per workgroup");
    static_assert(N1 == 2 || N1 == 4, "split factor");
    static_assert(KIND == KIND_C2C || N1 == 2, "real transforms split in two");
    auto slot = [](int o) { return XM == XM_FLAT ? o : pad_slot<LOGE>(o); };
    k1 = __builtin_amdgcn_readfirstlane(k1);   // uniform: it feeds scalar buffer offsets

    if constexpr (FIXED != RF_DYNAMIC) flags = FIXED;
    const bool zp = (flags & RF_ZERO_PAD) != 0;
    const bool mul = (flags & RF_MUL) != 0;
    const bool cj = (flags & RF_CONJ2) != 0;
    const bool full = (flags & RF_FULL) != 0;
    const bool adj = (flags & (RF_ADJ_R2C | RF_ADJ_C2R)) != 0;

    const unsigned nin = KIND == KIND_C2R ? NT + 1 : ((zp && !mul) ? NT / 2 : NT);
    const unsigned nout = KIND == KIND_R2C ? (full ? 2 * NT : NT + 1) : NT;
    const __amdgpu_buffer_rsrc_t rin = make_rsrc(X, nin * 8u);
    const __amdgpu_buffer_rsrc_t rin2 = make_rsrc(Y, mul ? (KIND == KIND_C2R ? NT + 1 : NT) * 8u : 0u);
    const __amdgpu_buffer_rsrc_t rout = make_rsrc(Xo, nout * 8u);
    const __amdgpu_buffer_rsrc_t rrtw = make_rsrc(rtw, KIND == KIND_C2C ? 0u : NT * 8u);
    const __amdgpu_buffer_rsrc_t rstw = make_rsrc(stw, (N1 - 1) * N2 * 8u);
    auto ld = [](__amdgpu_buffer_rsrc_t r, int vp, int sp) {
        return bload(r, (unsigned)vp * 8u, (unsigned)sp * 8u);
    };

    // w_N1^(dir*n1*k1) is one of 1, i, -1, -i: as an exact complex factor.
    auto rot = [&](int n1) {
        const int q = (n1 * k1 * (4 / N1)) & 3;  // quarter turns
        const float c = q == 0 ? 1.f : (q == 2 ? -1.f : 0.f);
        const float s = q == 1 ? (float)DIR : (q == 3 ? -(float)DIR : 0.f);
        return make_float2(c, s);
    };

    // Per-point twiddles from one table entry per thread and compile-time
    // rotations (the tables' direction is the kernel's):
    //   resolve  W[m] = i*DIR*exp(i*pi*DIR*m/NT), W[m + 2T*c] = W[m] * exp(i*pi*DIR*c/E),
    //            W[NT - m] = conj(W[m]);
    //   split    w_NT^(k1*n) for k1 = 1: w_NT^(lt + T*c) = w_NT^lt * exp(2*pi*i*DIR*c/(2E)).
    // With N1 = 4 the split twiddle's rotation depends on k1, so it is read per point.
    // (r2c reads its base W[2*lt + k1] after the passes, where it is used.)
    // c2r with mirrored lanes (rv_mirror_lanes): loads and the first pass
    // run under the numbering lm = mirror_lt(lt).
    constexpr bool MIRROR = KIND == KIND_C2R && rv_mirror_lanes(VAR) && T >= 64;
    int lm = lt;
    if constexpr (MIRROR) lm = mirror_lt<T>(lt);
    float2 wb = make_float2(0.f, 0.f);
    if constexpr (KIND == KIND_C2R) wb = ld(rrtw, lm, 0);
    float2 sb = make_float2(0.f, 0.f);
    if constexpr (N1 == 2) {
        if (k1 > 0) sb = ld(rstw, lm, 0);
    }

    float2 v[E];
    if constexpr (MIRROR) {
        // Each bin read once.  Point (c, n1) = lm + T*c + N2*n1; X[NT - m]
        // is point (E-1-c, 1-n1) of the other half-wave's lane.  Lane 32
        // (lm = T/2) holds its own partners; lane 0 (lm = 0) holds them at
        // (E-c, 1-n1), except X[NT] for (0, 0) and itself for (0, 1).
        auto bin = [&](int sp, bool edge) -> float2 {
            float2 x = ld(rin, lm, sp);
            if (mul) {
                float2 y = ld(rin2, lm, sp);
                if (cj) y = cconj(y);
                x = cscale(cmul(x, y), norm);
            }
            if (adj) x = edge ? make_float2(x.x, 0.f) : cscale(x, 0.5f);
            return x;
        };
        float2 xn = ld(rin, 0, NT);   // bin NT, for lane 0
        if (mul) {
            float2 y = ld(rin2, 0, NT);
            if (cj) y = cconj(y);
            xn = cscale(cmul(xn, y), norm);
        }
        if (adj) xn = make_float2(xn.x, 0.f);
        const bool l0 = lt == 0, lh = lt == 32;
        float2 x0[E], x1[E];
        // v[c] from its own points (c, 0), (c, 1) and their partners.
        auto finish = [&](auto C) {
            constexpr int c = decltype(C)::value;
            float2 p0 = swap_half_waves(x1[E - 1 - c]);
            float2 p1 = swap_half_waves(x0[E - 1 - c]);
            p0 = lh ? x1[E - 1 - c] : p0;
            p1 = lh ? x0[E - 1 - c] : p1;
            p0 = l0 ? (c == 0 ? xn : x1[(E - c) % E]) : p0;
            p1 = l0 ? (c == 0 ? x1[0] : x0[(E - c) % E]) : p1;
            // W[lm + T*c + N2*n1] = W[lm] * exp(2*pi*i*DIR*(c + E*n1)/(4E)).
            const float2 b0 = cconj(p0), b1 = cconj(p1);
            const float2 z0 = cadd(cadd(x0[c], b0), cmul(csub(x0[c], b0), twc<c, 4 * E, DIR>(wb)));
            const float2 z1 = cadd(cadd(x1[c], b1), cmul(csub(x1[c], b1), twc<c + E, 4 * E, DIR>(wb)));
            float2 acc = zp ? z0 : cadd(z0, cmul(z1, rot(1)));
            if (k1 > 0) acc = cmul(acc, twc<c, 2 * E, DIR>(sb));
            v[c] = acc;
        };
        // Points in groups {g, E-1-g}, closed under the partner map; lane 0
        // also needs group g-1's E-g for point g and group g+1's g+1 for point
        // E-1-g, so point E-1-g is finished one group later.
        static_for<0, E / 2 + 1>([&](auto G) {
            constexpr int g = decltype(G)::value;
            if constexpr (g < E / 2) {
                x0[g] = bin(T * g, g == 0 && lm == 0);
                x1[g] = bin(T * g + N2, false);
                x0[E - 1 - g] = bin(T * (E - 1 - g), false);
                x1[E - 1 - g] = bin(T * (E - 1 - g) + N2, false);
                finish(G);
            }
            if constexpr (g > 0) finish(std::integral_constant<int, E - g>{});
        });
    } else static_for<0, E>([&](auto J) {
        // c2r chunked loads (rv_chunked_loads): as in rows_impl.h, points
        // taken in chunks from the front and the back in turn, so the mirror
        // thread's second read of each bin follows the first by one chunk.
        constexpr int j = decltype(J)::value;
        constexpr int CH = 4;
        constexpr bool CHUNKED = KIND == KIND_C2R && rv_chunked_loads(VAR) && E >= 2 * CH;
        constexpr int q = j / CH, p = q / 2, r = j % CH;
        constexpr int c = !CHUNKED ? j : (q % 2 == 0 ? p * CH + r : E - (p + 1) * CH + r);
        if constexpr (CHUNKED && j > 0 && r == 0) __builtin_amdgcn_sched_barrier(0);
        float2 acc = make_float2(0.f, 0.f);
        static_for<0, N1>([&](auto M1) {
            constexpr int n1 = decltype(M1)::value;
            constexpr int sp = T * c + N2 * n1;  // point m = lt + sp
            if (zp && n1 >= N1 / 2) return;
            float2 z;
            if constexpr (KIND == KIND_C2R) {
                // X[NT - m] is element (T - lt) + (NT - T*(c+1) - N2*n1).
                constexpr int spb = NT - T * (c + 1) - N2 * n1;
                float2 xa = ld(rin, lt, sp);
                float2 xb = ld(rin, T - lt, spb);
                if (mul) {
                    float2 ya = ld(rin2, lt, sp);
                    float2 yb = ld(rin2, T - lt, spb);
                    if (cj) {
                        ya = cconj(ya);
                        yb = cconj(yb);
                    }
                    xa = cscale(cmul(xa, ya), norm);
                    xb = cscale(cmul(xb, yb), norm);
                }
                if (adj) {
                    const bool m0 = sp == 0 && lt == 0;
                    xa = m0 ? make_float2(xa.x, 0.f) : cscale(xa, 0.5f);
                    xb = m0 ? make_float2(xb.x, 0.f) : cscale(xb, 0.5f);
                }
                const float2 bc = cconj(xb);
                // W[lt + T*c + N2*n1] = W[lt] * exp(2*pi*i*DIR*(c + E*n1)/(4E)).
                z = cadd(cadd(xa, bc), cmul(csub(xa, bc), twc<c + E * n1, 4 * E, DIR>(wb)));
            } else if (KIND == KIND_C2C && mul) {
                float2 y = ld(rin2, lt, sp);
                if (cj) y = cconj(y);
                z = cscale(cmul(ld(rin, lt, sp), y), norm);
            } else {
                z = ld(rin, lt, sp);
            }
            acc = n1 == 0 ? z : cadd(acc, cmul(z, rot(n1)));
        });
        if constexpr (N1 == 2) {
            if (k1 > 0) acc = cmul(acc, twc<c, 2 * E, DIR>(sb));
        } else {
            if (k1 > 0) acc = cmul(acc, ld(rstw, lt, (k1 - 1) * N2 + T * c));
        }
        v[c] = acc;
    });

    constexpr int RAD = rv_R(N2, VAR);
    static_assert(!MIRROR || np_for(N2, RAD) > 1, "the first pass hands its points over through LDS");
    stockham_passes<N2, E, DIR, XM, TWM, RAD>(v, lt, TwBuffer{make_rsrc(tw, (N2 - seq_first(N2, RAD, np_for(N2, RAD), true)) * 8u)}, L, Lf,
                                              slot, lm);

    // v[c] now holds X[N1*k2 + k1] for k2 = lt + T*c.
    if constexpr (KIND != KIND_R2C) {
        static_for<0, E>([&](auto C) {
            constexpr int c = decltype(C)::value;
            bstore(rout, (unsigned)(N1 * lt) * 8u, (unsigned)(N1 * T * c + k1) * 8u, v[c]);
        });
    } else {
        // Bin k = 2*k2 + k1 needs Z[k] and Z[NT - k], which is v at
        // k2' = N2 - k2 - k1 (mod N2) of this workgroup.  As in rows_impl.h the
        // upper half (k2 >= N2/2) goes through LDS once and the owner of each
        // lower-half point emits both bins of its pair.
        const float scale = adj ? 1.0f : 0.5f;
        wb = bload(rrtw, (unsigned)(2 * lt) * 8u, (unsigned)k1 * 8u);   // W[2*lt + k1]
        // Bin 2*(vp + sp) + k1 from za = Z[k], zn = Z[NT - k] and w = W[k].
        auto emit = [&](int vp, int sp, float2 za, float2 zn, float2 w) {
            const unsigned vo = (unsigned)(2 * vp) * 8u;
            const unsigned so = (unsigned)(2 * sp + k1) * 8u;
            const float2 zb = cconj(zn);
            const float2 fe = cadd(za, zb);
            const float2 fo = cmul(w, csub(za, zb));
            bstore(rout, vo, so, cscale(cadd(fe, fo), scale));
            if (full) bstore(rout, vo, so + (unsigned)NT * 8u, cscale(csub(fe, fo), scale));
        };
        float2 *Lc = SPLIT ? reinterpret_cast<float2 *>(Lf) : L;
        static_for<E / 2, E>([&](auto C) {
            constexpr int c = decltype(C)::value;
            Lc[slot(lt + T * c - N2 / 2)] = v[c];
        });
        __syncthreads();
        static_for<0, E / 2>([&](auto C) {
            constexpr int c = decltype(C)::value;
            if (c == 0 && lt == 0 && k1 == 0) {
                // Bins 0 and NT from Z[0]; bin NT/2 (k2 = N2/2) is its own partner.
                const float2 z0 = v[0];
                const float r0 = z0.x + z0.y, rn = z0.x - z0.y;
                bstore(rout, 0u, 0u, make_float2(r0, adj ? -rn : 0.f));
                bstore(rout, 0u, (unsigned)NT * 8u, make_float2(rn, adj ? -r0 : 0.f));
                const float2 zh = Lc[slot(0)];
                emit(0, N2 / 2, zh, zh, twc<E / 2, 2 * E, DIR>(wb));   // W[N2], lt = k1 = 0
            } else {
                // Partner k2' = N2 - k2 - k1 = (T - lt - k1) + T*(E-1-c), upper half.
                const float2 zn = Lc[slot(N2 / 2 - lt - T * c - k1)];
                const float2 w = twc<c, 2 * E, DIR>(wb);   // W[2*(lt + T*c) + k1]
                emit(lt, T * c, v[c], zn, w);
                emit(T - lt - k1, T * (E - 1 - c), zn, v[c], cconj(w));
            }
        });
    }
}

template <int N2, int N1, int DIR, int KIND, int VAR>
__global__ void __launch_bounds__(rv_T(N2, VAR), rv_waves_per_simd(N2, VAR)) fft_split_kernel(RowArgs a) {
    constexpr int LDSN = rv_lds_per(N2, VAR);
    __shared__ float2 smem[rv_xm(VAR) == XM_SPLIT ? (LDSN + 1) / 2 : LDSN];
    const unsigned g = xcd_block(blockIdx.x, gridDim.x);
    const long long t = g / N1;
    const int k1 = (int)(g % N1);
    // The split's stores are cacheable anyway: RF_KEEP_OUT changes nothing.
    if ((a.flags & RF_ALL & ~(unsigned)RF_KEEP_OUT) == 0)
        split_transform<N2, N1, DIR, KIND, VAR, 0u>(a.in + t * a.idist, a.in2 + t * a.idist2, a.out + t * a.odist,
                                                      k1, threadIdx.x, a.tw, a.stw, a.rtw, 0u, a.norm, smem,
                                                      reinterpret_cast<float *>(smem));
    else
        split_transform<N2, N1, DIR, KIND, VAR>(a.in + t * a.idist, a.in2 + t * a.idist2, a.out + t * a.odist,
                                                k1, threadIdx.x, a.tw, a.stw, a.rtw, a.flags, a.norm, smem,
                                                reinterpret_cast<float *>(smem));
}

}  // namespace mufft
