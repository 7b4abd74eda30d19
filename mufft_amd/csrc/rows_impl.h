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
#include "row_variants.h"
#include "split_impl.h"
#include "stockham.h"

namespace mufft {

// One fused transform (N points, variant VAR) by the T = N/E threads of a
// workgroup slot.  X / Y / Xo are this transform's input, second operand and
// output; L / Lf its LDS region (rv_lds_per(N, VAR) slots).  Every thread of
// the workgroup must call it (the exchanges synchronise the workgroup);
// inactive slots pass active = false.  With T >= 64 a wave never spans two
// transforms, so the pointers are wave-uniform and the strided traffic goes
// through buffer descriptors (one VGPR offset per thread, scalar offsets per
// slot).
// FIXED: the flags, when the caller knows them at compile time (RF_DYNAMIC:
// read at run time).  Fixed flags fold to constants, so no branch on them
// splits the load or store sequences (such branches kept loads from issuing
// ahead of their use and raised register pressure).
template <int N, int DIR, int KIND, int VAR, unsigned FIXED = RF_DYNAMIC>
__device__ __forceinline__ void row_transform(const float2 *X, const float2 *Y, float2 *Xo,
                                              bool active, int lt, const float2 *tw,
                                              const float2 *rtw, unsigned flags, float norm,
                                              float2 *L, float *Lf) {
    constexpr int E = rv_E(N, VAR);
    constexpr int LOGE = ilog2(E);
    constexpr int T = rv_T(N, VAR);
    constexpr int XM = rv_xm(VAR);
    constexpr int TWM = rv_twm(VAR);
    constexpr int RAD = rv_R(N, VAR);
    constexpr bool SPLIT = XM == XM_SPLIT;
    constexpr bool BUF = T >= 64;
    // A transform of at most 64 threads lies in one wave (its threads are
    // consecutive): its exchanges need no workgroup barrier.
    constexpr bool WAVE = T <= 64;
    // c2r with mirrored lanes: each bin read once (rv_mirror_lanes).
    constexpr bool MIRROR = KIND == KIND_C2R && BUF && rv_mirror_lanes(VAR);
    auto slot = [](int o) { return XM == XM_FLAT ? o : pad_slot<LOGE>(o); };

    if constexpr (FIXED != RF_DYNAMIC) flags = FIXED;
    const bool zp = (flags & RF_ZERO_PAD) != 0;
    const bool mul = (flags & RF_MUL) != 0;
    const bool cj = (flags & RF_CONJ2) != 0;
    const bool full = (flags & RF_FULL) != 0;
    const bool adj = (flags & (RF_ADJ_R2C | RF_ADJ_C2R)) != 0;
    // Element vp + sp of an operand: vp per lane, sp a compile-time slot offset.
    __amdgpu_buffer_rsrc_t rin, rin2, rout, rrtw;  // set and used only when BUF
    if constexpr (BUF) {
        const unsigned nin = KIND == KIND_C2R ? N + 1 : ((zp && !mul) ? N / 2 : N);
        const unsigned nout = KIND == KIND_R2C ? (full ? 2 * N : N + 1) : N;
        rin = make_rsrc(X, active ? nin * 8u : 0u);
        rin2 = make_rsrc(Y, (active && mul) ? (KIND == KIND_C2R ? N + 1 : N) * 8u : 0u);
        rout = make_rsrc(Xo, active ? nout * 8u : 0u);
        rrtw = make_rsrc(rtw, KIND == KIND_C2C ? 0u : N * 8u);
    }
    // c2c / r2c inputs are read once (streaming); c2r reads each bin twice
    // (as X[k] and as X[N - k]), so its input stays cacheable.
    auto ld_in = [&](int vp, int sp) -> float2 {
        if constexpr (BUF && KIND != KIND_C2R) return bload_nt(rin, (unsigned)vp * 8u, (unsigned)sp * 8u);
        else if constexpr (BUF) return bload(rin, (unsigned)vp * 8u, (unsigned)sp * 8u);
        else return X[vp + sp];
    };
    auto ld_in2 = [&](int vp, int sp) -> float2 {
        if constexpr (BUF) return bload(rin2, (unsigned)vp * 8u, (unsigned)sp * 8u);
        else return Y[vp + sp];
    };
    auto ld_rtw = [&](int vp, int sp) -> float2 {
        if constexpr (BUF) return bload(rrtw, (unsigned)vp * 8u, (unsigned)sp * 8u);
        else return rtw[vp + sp];
    };
    // r2c rows of N + 1 bins start at any 8-byte offset, so the lines at
    // their ends are written in parts, by different store instructions and
    // often different workgroups; cacheable stores let the L2 merge them
    // (on MI355X, cold: r2c 65536 x 512 -12%, 4096-16384 real points -8% to
    // -16%, within +-3% elsewhere).  Other outputs stream unless kept.
    const bool keep = KIND == KIND_R2C || (flags & RF_KEEP_OUT) != 0;
    auto st_out = [&](int vp, int sp, float2 val) {
        if constexpr (BUF) {
            if (keep) bstore(rout, (unsigned)vp * 8u, (unsigned)sp * 8u, val);
            else bstore_nt(rout, (unsigned)vp * 8u, (unsigned)sp * 8u, val);
        } else {
            Xo[vp + sp] = val;
        }
    };

    // Resolve twiddles W[k] = i*DIR*exp(i*pi*DIR*k/N) (the table's direction is
    // the kernel's: r2c forward, c2r inverse).  A thread's points k = lt + T*c
    // take W[lt] * exp(i*pi*DIR*c/E), a compile-time rotation, so one table
    // entry per thread is read instead of one per point; W[N - k] = conj(W[k]).
    // (r2c reads W[lt] after the passes, where it is used.)
    float2 wb = make_float2(0.f, 0.f);
    int lm = lt;   // numbering of the loads and the first pass
    if constexpr (MIRROR) lm = mirror_lt<T>(lt);
    if constexpr (KIND == KIND_C2R) wb = ld_rtw(lm, 0);

    float2 v[E];
    if constexpr (MIRROR) {
        // Pre-pass resolve (kernel.c:36-45), each bin read once: this thread
        // holds bins lm + T*c, and X[N - k] for k = lm + T*c is bin
        // (T - lm) + T*(E-1-c) of the other half-wave's lane.  Lane 32
        // (lm = T/2) holds its own partners; lane 0 (lm = 0) too, at
        // X[N - T*c] = a[E - c], except bin N.  Selects, no branches: the
        // prologue stays one block, so the loads issue ahead of their use.
        auto bin = [&](int vp, int sp, bool edge) -> float2 {
            float2 x = ld_in(vp, sp);
            if (mul) {
                float2 y = ld_in2(vp, sp);
                if (cj) y = cconj(y);
                x = cscale(cmul(x, y), norm);
            }
            // Adjoint of r2c (as below): interior bins halved, bins 0 and N real.
            if (adj) x = edge ? make_float2(x.x, 0.f) : cscale(x, 0.5f);
            return x;
        };
        const float2 xn = bin(0, N, true);
        const bool l0 = lt == 0, lh = lt == 32;
        float2 a[E];
        auto finish = [&](auto C) {
            constexpr int c = decltype(C)::value;
            float2 b = swap_half_waves(a[E - 1 - c]);
            b = lh ? a[E - 1 - c] : b;
            b = l0 ? (c == 0 ? xn : a[(E - c) % E]) : b;
            const float2 bc = cconj(b);
            const float2 z = cadd(cadd(a[c], bc), cmul(csub(a[c], bc), twc<c, 2 * E, DIR>(wb)));
            v[c] = (zp && c >= E / 2) ? make_float2(0.f, 0.f) : z;
        };
        // Points in groups {g, E-1-g}, closed under the partner map; lane 0
        // also needs a[E-g] (group g-1) for point g and a[g+1] (group g+1)
        // for point E-1-g, so point E-1-g is finished one group later.
        static_for<0, E / 2 + 1>([&](auto G) {
            constexpr int g = decltype(G)::value;
            if constexpr (g < E / 2) {
                a[g] = bin(lm, T * g, g == 0 && lm == 0);
                a[E - 1 - g] = bin(lm, T * (E - 1 - g), false);
                finish(G);
            }
            if constexpr (g > 0) finish(std::integral_constant<int, E - g>{});
        });
    } else if constexpr (KIND == KIND_C2R) {
        // Pre-pass resolve (kernel.c:36-45), optionally after the fused multiply.
        // X[N - k] for k = lt + T*c is element (T - lt) + T*(E-1-c).
        static_for<0, E>([&](auto J) {
            // Chunked loads (rv_chunked_loads): point c of step j, chunks of
            // CH points taken from the front and the back in turn; nothing
            // of a chunk moves above its boundary, so its loads issue after
            // the previous chunk's data arrived.
            constexpr int j = decltype(J)::value;
            constexpr int CH = 4;
            constexpr bool CHUNKED = rv_chunked_loads(VAR) && E >= 2 * CH;
            constexpr int q = j / CH, p = q / 2, r = j % CH;
            constexpr int c = !CHUNKED ? j : (q % 2 == 0 ? p * CH + r : E - (p + 1) * CH + r);
            if constexpr (CHUNKED && j > 0 && r == 0) __builtin_amdgcn_sched_barrier(0);
            const int k = lt + T * c;
            float2 z = make_float2(0.f, 0.f);
            if ((BUF || active) && !(zp && c >= E / 2)) {
                float2 xa = ld_in(lt, T * c);
                float2 xb = ld_in(T - lt, T * (E - 1 - c));
                if (mul) {
                    float2 ya = ld_in2(lt, T * c);
                    float2 yb = ld_in2(T - lt, T * (E - 1 - c));
                    if (cj) {
                        ya = cconj(ya);
                        yb = cconj(yb);
                    }
                    xa = cscale(cmul(xa, ya), norm);
                    xb = cscale(cmul(xb, yb), norm);
                }
                if (adj) {
                    // Adjoint of r2c: x = Re(sum_{k<=N} G_k w^{-nk}) equals c2r of G with
                    // interior bins halved and the imaginary parts of bins 0 and N dropped.
                    xa = c == 0 && lt == 0 ? make_float2(xa.x, 0.f) : cscale(xa, 0.5f);
                    xb = c == 0 && lt == 0 ? make_float2(xb.x, 0.f) : cscale(xb, 0.5f);
                }
                const float2 bc = cconj(xb);
                z = cadd(cadd(xa, bc), cmul(csub(xa, bc), twc<c, 2 * E, DIR>(wb)));
            }
            (void)k;
            v[c] = z;
        });
    } else {
        if (KIND == KIND_C2C && mul) {
            static_for<0, E>([&](auto C) {
                constexpr int c = decltype(C)::value;
                float2 z = make_float2(0.f, 0.f);
                if (BUF || active) {
                    float2 y = ld_in2(lt, T * c);
                    if (cj) y = cconj(y);
                    z = cscale(cmul(ld_in(lt, T * c), y), norm);
                }
                v[c] = z;
            });
        } else {
            static_for<0, E>([&](auto C) {
                constexpr int c = decltype(C)::value;
                v[c] = ((BUF || active) && !(zp && c >= E / 2)) ? ld_in(lt, T * c) : make_float2(0.f, 0.f);
            });
        }
    }

    if constexpr (BUF) {
        const unsigned ntw = (unsigned)(N - seq_first(N, RAD, rv_NP(N, VAR), true));   // 0 for one pass
        static_assert(!MIRROR || rv_NP(N, VAR) > 1, "the first pass hands its points over through LDS");
        stockham_passes<N, E, DIR, XM, TWM, RAD, WAVE>(v, lt, TwBuffer{make_rsrc(tw, ntw * 8u)}, L, Lf, slot, lm);
    } else {
        stockham_passes<N, E, DIR, XM, TWM, RAD, WAVE>(v, lt, TwGlobal{tw}, L, Lf, slot);
    }

    if constexpr (KIND == KIND_R2C) {
        // Post-pass resolve (kernel.c:47-82).  Bins k and N - k both come from
        // the pair Z[k], Z[N - k]: the upper half of Z (k >= N/2) goes through
        // LDS once and the owner of each lower-half Z[k] emits both bins, so
        // the upper-half registers die at the LDS store.  With RF_ADJ_C2R the
        // epilogue emits the adjoint of c2r (muFFT semantics, kernel.c:36-45):
        // 2*R_k for interior bins, R_0 - i R_N at bin 0 and R_N - i R_0 at bin N,
        // where R = r2c of the incoming gradient.
        const float scale = adj ? 1.0f : 0.5f;
        wb = ld_rtw(lt, 0);   // W[lt]
        // Bin (vp + sp) from za = Z[k], zn = Z[N - k] and w = W[k].
        auto emit = [&](int vp, int sp, float2 za, float2 zn, float2 w) {
            const float2 zb = cconj(zn);
            const float2 fe = cadd(za, zb);
            const float2 fo = cmul(w, csub(za, zb));
            st_out(vp, sp, cscale(cadd(fe, fo), scale));
            if (full) st_out(vp, sp + N, cscale(csub(fe, fo), scale));
        };
        float2 *Lc = SPLIT ? reinterpret_cast<float2 *>(Lf) : L;
        static_for<E / 2, E>([&](auto C) {
            constexpr int c = decltype(C)::value;
            Lc[slot(lt + T * c - N / 2)] = v[c];
        });
        exchange_sync<WAVE>();
        if (BUF || active) {
            static_for<0, E / 2>([&](auto C) {
                constexpr int c = decltype(C)::value;
                if (c == 0 && lt == 0) {
                    // Bins 0 and N from Z[0]; bin N/2 from Z[N/2], its own partner.
                    const float2 z0 = v[0];
                    const float r0 = z0.x + z0.y, rn = z0.x - z0.y;
                    st_out(0, 0, make_float2(r0, adj ? -rn : 0.f));
                    st_out(0, N, make_float2(rn, adj ? -r0 : 0.f));
                    const float2 zh = Lc[slot(0)];
                    emit(0, N / 2, zh, zh, twc<E / 2, 2 * E, DIR>(wb));   // W[N/2], lt = 0
                } else {
                    // Z[N - k] for k = lt + T*c is upper-half point N/2 - k.
                    const float2 zn = Lc[slot(N / 2 - lt - T * c)];
                    const float2 w = twc<c, 2 * E, DIR>(wb);
                    emit(lt, T * c, v[c], zn, w);
                    emit(T - lt, T * (E - 1 - c), zn, v[c], cconj(w));
                }
            });
        }
    } else {
        if (BUF || active) {
            static_for<0, E>([&](auto C) {
                constexpr int c = decltype(C)::value;
                st_out(lt, T * c, v[c]);
            });
        }
    }
}


template <int N, int DIR, int KIND, int VAR>
__global__ void __launch_bounds__(rv_WG(N, VAR), rv_waves_per_simd(N, VAR)) fft_rows_kernel(RowArgs a) {
    constexpr int T = rv_T(N, VAR);
    constexpr int B = rv_B(N, VAR);
    constexpr int LDSN = rv_lds_per(N, VAR);
    constexpr int SMEM = rv_xm(VAR) == XM_SPLIT ? (B * LDSN + 1) / 2 : B * LDSN;
    __shared__ float2 smem[SMEM];
    const int tid = threadIdx.x;
    const int lt = tid & (T - 1);
    const int sub = tid / T;
    const long long g = (long long)blockIdx.x * B + sub;
    // r2c stores cacheable anyway: RF_KEEP_OUT changes nothing there.
    constexpr unsigned MOOT = KIND == KIND_R2C ? (unsigned)RF_KEEP_OUT : 0u;
    if ((a.flags & RF_ALL & ~MOOT) == 0) {
        row_transform<N, DIR, KIND, VAR, MOOT>(a.in + g * a.idist, a.in2 + g * a.idist2, a.out + g * a.odist,
                                               g < a.batch, lt, a.tw, a.rtw, 0u, a.norm, smem + sub * LDSN,
                                               reinterpret_cast<float *>(smem) + sub * LDSN);
        return;
    }
    if constexpr (KIND == KIND_C2C) {
        // The 2D row pass (its output is read by the column pass): 2D c2c
        // 1024^2 x 16 -3% to -6% with the flag fixed.
        if ((a.flags & RF_ALL) == RF_KEEP_OUT) {
            row_transform<N, DIR, KIND, VAR, RF_KEEP_OUT>(a.in + g * a.idist, a.in2 + g * a.idist2,
                                                          a.out + g * a.odist, g < a.batch, lt, a.tw, a.rtw,
                                                          a.flags, a.norm, smem + sub * LDSN,
                                                          reinterpret_cast<float *>(smem) + sub * LDSN);
            return;
        }
    }
    row_transform<N, DIR, KIND, VAR>(a.in + g * a.idist, a.in2 + g * a.idist2, a.out + g * a.odist, g < a.batch,
                                     lt, a.tw, a.rtw, a.flags, a.norm, smem + sub * LDSN,
                                     reinterpret_cast<float *>(smem) + sub * LDSN);
}

// Entries [I, END) of the table of the instantiated variants of one
// (direction, kind).
template <int DIR, int KIND, int I, int END>
void fill_row_kernels(const void **out) {
    if constexpr (I < END) {
        constexpr RowVariantEntry e = kRowVariants[I];
        if constexpr (rv_split(e.var) > 0) {
            constexpr int s = rv_split(e.var);
            out[I] = reinterpret_cast<const void *>(
                &fft_split_kernel<(1 << (e.logN - s)), (1 << s), DIR, KIND, rv_sub(e.var)>);
        } else {
            out[I] = reinterpret_cast<const void *>(&fft_rows_kernel<(1 << e.logN), DIR, KIND, e.var>);
        }
        fill_row_kernels<DIR, KIND, I + 1, END>(out);
    }
}

}  // namespace mufft
