// Fused column (strided-axis) transforms: the 2D vertical passes.
//
// The reference runs each vertical radix pass as a separate sweep over the
// whole image (kernel.c:455-738, driven by fft.c:978-987).  Here one
// workgroup owns a tile of TX adjacent columns and runs every pass of the
// column transforms in registers/LDS: one HBM read and one HBM write per
// point, lanes along the unit-stride axis so each row segment of the tile is
// a single coalesced access.  Tiles are assigned XCD-aware so neighbouring
// tiles, which share cache lines, run under the same L2.  In place is safe:
// a workgroup reads its whole tile before it writes any of it.
#include "kernels.h"
#include "stockham.h"

namespace mufft {

// TXP: tile width (columns per workgroup); 0 = the default col_TX(L).
template <int L_, int DIR, int MODE = COL_PLAIN, int TXP = 0>
__global__ void __launch_bounds__(col_T(L_) * (TXP ? TXP : col_TX(L_))) fft_cols_kernel(ColArgs a) {
    constexpr int E = col_E(L_);
    constexpr int LOGE = ilog2(E);
    constexpr int T = col_T(L_);
    constexpr int TX = TXP ? TXP : col_TX(L_);
    constexpr int ROWS = col_lds_rows(L_);
    constexpr int XM = col_xm(L_);
    __shared__ float2 smem[XM == XM_SPLIT ? (ROWS * TX + 1) / 2 : ROWS * TX];

    const int tid = threadIdx.x;
    const int x = tid & (TX - 1);
    const int t = tid / TX;
    const unsigned bid = xcd_block(blockIdx.x, gridDim.x);
    const int img = (int)(bid / (unsigned)a.ntiles);
    const int tile = (int)(bid % (unsigned)a.ntiles);
    const int col = tile * TX + x;
    const bool active = col < a.ncols && img < a.batch;
    auto slot = [x](int o) { return pad_slot<LOGE>(o) * TX + x; };

    long long ib = (long long)img * a.in_bs, ob = (long long)img * a.out_bs;
    if constexpr (MODE == COL_TWO_LEVEL) {
        ib = (long long)(img / a.img_div) * a.in_bs2 + (long long)(img % a.img_div) * a.in_bs;
        ob = (long long)(img / a.img_div) * a.out_bs2 + (long long)(img % a.img_div) * a.out_bs;
    }

    float2 v[E];
    if constexpr (MODE == COL_STREAM) {
        // Image bytes below 2^32 (launch_cols checks): one per-lane offset,
        // the row offset of each point in a scalar register.
        const unsigned span = (unsigned)(((long long)(L_ - 1) * a.in_rs + a.ncols) * 8);
        const __amdgpu_buffer_rsrc_t rin = make_rsrc(a.in + ib, span);
        const unsigned vo = (unsigned)((long long)t * a.in_rs + col) * 8u;
        static_for<0, E>([&](auto C) {
            constexpr int c = decltype(C)::value;
            const unsigned so = (unsigned)__builtin_amdgcn_readfirstlane((int)((long long)(T * c) * a.in_rs)) * 8u;
            v[c] = active ? bload_nt(rin, vo, so) : make_float2(0.f, 0.f);
        });
    } else {
        const float2 *src = a.in + ib + col;
        static_for<0, E>([&](auto C) {
            constexpr int c = decltype(C)::value;
            v[c] = active ? src[(long long)(t + T * c) * a.in_rs] : make_float2(0.f, 0.f);
        });
    }

    // COL_TWIDDLE: the twiddle of point t + T*c of a column with factor b is
    // w^(b*t) * w^(b*T*c), a factor per thread and E steps per column in LDS
    // (as in the four-step column kernel below), where a workgroup has a
    // thread per step.
    constexpr bool STEPS = MODE == COL_TWIDDLE && E <= T;
    __shared__ float2 steps[STEPS ? E * TX : 1];
    float2 w0 = make_float2(1.f, 0.f);
    if constexpr (STEPS) {
        const unsigned smask = (1u << a.log2s) - 1u;
        if (tid < E * TX) {   // step tid / TX of this thread's own column (tid % TX == x)
            const unsigned es = ((unsigned)col >> a.log2div) * (unsigned)(T * (tid / TX));
            steps[tid] = col < a.ncols ? cmul(a.coarse[es >> a.log2s], a.fine[es & smask]) : make_float2(1.f, 0.f);
        }
        if (active) {
            const unsigned e0 = ((unsigned)col >> a.log2div) * (unsigned)t;
            w0 = cmul(a.coarse[e0 >> a.log2s], a.fine[e0 & smask]);
        }
    }

    // Remainder pass last (seq_radix): the lanes of one column thread share its
    // twiddles.  Its exchanges synchronise the workgroup (the steps are visible
    // after them).
    stockham_passes<L_, E, DIR, XM, col_twm(L_), E, false, false>(v, t, TwGlobal{a.tw}, smem,
                                                                  reinterpret_cast<float *>(smem), slot);

    if (active) {
        if constexpr (STEPS) {
            static_for<0, E>([&](auto C) {
                constexpr int c = decltype(C)::value;
                v[c] = cmul(v[c], cmul(w0, steps[c * TX + x]));
            });
        } else if constexpr (MODE == COL_TWIDDLE) {
            const unsigned b = (unsigned)col >> a.log2div;
            const unsigned smask = (1u << a.log2s) - 1u;
            static_for<0, E>([&](auto C) {
                constexpr int c = decltype(C)::value;
                const unsigned e = b * (unsigned)(t + T * c);
                v[c] = cmul(v[c], cmul(a.coarse[e >> a.log2s], a.fine[e & smask]));
            });
        }
        if constexpr (MODE == COL_STREAM) {
            const unsigned span = (unsigned)(((long long)(L_ - 1) * a.out_rs + a.ncols) * 8);
            const __amdgpu_buffer_rsrc_t rout = make_rsrc(a.out + ob, span);
            const unsigned vo = (unsigned)((long long)t * a.out_rs + col) * 8u;
            static_for<0, E>([&](auto C) {
                constexpr int c = decltype(C)::value;
                const unsigned so = (unsigned)__builtin_amdgcn_readfirstlane((int)((long long)(T * c) * a.out_rs)) * 8u;
                bstore_nt(rout, vo, so, v[c]);
            });
        } else {
            float2 *dst = a.out + ob + col;
            static_for<0, E>([&](auto C) {
                constexpr int c = decltype(C)::value;
                dst[(long long)(t + T * c) * a.out_rs] = v[c];
            });
        }
    }
}

// Four-step first half (FourStepArgs): the column transforms above plus the
// w_N^(dir*n2*k1) twiddle, stored transposed through LDS so that each column
// of the tile leaves as one contiguous run of N1 points.
template <int L_, int DIR>
__global__ void __launch_bounds__(fs_WG(L_)) fft_four_step_cols_kernel(FourStepArgs a) {
    constexpr int E = col_E(L_);
    constexpr int LOGE = ilog2(E);
    constexpr int T = col_T(L_);
    constexpr int TX = fs_TX(L_);
    constexpr int WG = fs_WG(L_);
    constexpr int ROWS = col_lds_rows(L_);
    constexpr int LP = L_ + 1;   // row pitch of the transposed tile
    static_assert(LP <= ROWS, "transposed tile fits the exchange LDS");
    __shared__ float2 smem[ROWS * TX];

    const int tid = threadIdx.x;
    const int x = tid & (TX - 1);
    const int t = tid / TX;
    const unsigned bid = xcd_block(blockIdx.x, gridDim.x);
    const int img = (int)(bid / (unsigned)a.ntiles);
    const int tile = (int)(bid % (unsigned)a.ntiles);
    const int col = tile * TX + x;   // n2; the host launches whole tiles only
    const bool zp = (a.flags & RF_ZERO_PAD) != 0;
    auto slot = [x](int o) { return pad_slot<LOGE>(o) * TX + x; };

    float2 v[E];
    const float2 *src = a.in + (long long)img * a.in_bs + col;
    static_for<0, E>([&](auto C) {
        constexpr int c = decltype(C)::value;
        const int row = t + T * c;
        v[c] = (zp && row >= L_ / 2) ? make_float2(0.f, 0.f) : src[(long long)row * a.n2];
    });

    // The twiddle of point (n2, k1) = (col, t + T*c) is w^(col*t) * w^(col*T*c):
    // a factor per thread and E steps per column, the steps of the tile's
    // columns in LDS, one coarse*fine product each (every product is within
    // about 1.5 ulp; a point's twiddle within about 3.5).  Two table gathers
    // per point instead cost 10-16% of the whole transform.  Used where a
    // workgroup has a thread per step (E <= T, columns from 256 points).
    constexpr bool STEPS = E <= T;
    __shared__ float2 steps[STEPS ? E * TX : 1];
    const int smask = (1 << a.log2s) - 1;
    float2 w0 = make_float2(1.f, 0.f);
    if constexpr (STEPS) {
        if (tid < E * TX) {
            const unsigned es = (unsigned)(tile * TX + tid % TX) * (unsigned)(T * (tid / TX));   // < N
            steps[tid] = cmul(a.coarse[es >> a.log2s], a.fine[es & smask]);
        }
        const unsigned e0 = (unsigned)col * (unsigned)t;
        w0 = cmul(a.coarse[e0 >> a.log2s], a.fine[e0 & smask]);
    }

    // The passes' exchanges synchronise the workgroup: the steps are
    // visible to every thread after them.  Split pass twiddles from 1024
    // points on, as in the column kernel (2^20, 2^22: -2.5%).
    stockham_passes<L_, E, DIR, XM_PAD, col_twm(L_), E, false, false>(v, t, TwGlobal{a.tw}, smem, nullptr, slot);

    static_for<0, E>([&](auto C) {
        constexpr int c = decltype(C)::value;
        float2 w;
        if constexpr (STEPS) {
            w = cmul(w0, steps[c * TX + x]);
        } else {
            const unsigned e = (unsigned)col * (unsigned)(t + T * c);   // < N
            w = cmul(a.coarse[e >> a.log2s], a.fine[e & smask]);
        }
        smem[x * LP + t + T * c] = cmul(v[c], w);
    });
    __syncthreads();
    float2 *dst = a.out + (long long)img * a.out_bs + (long long)tile * TX * L_;
    static_for<0, E>([&](auto J) {
        constexpr int j = decltype(J)::value;
        const int i = tid + WG * j;   // (column, k1) = (i / L, i % L)
        dst[i] = smem[(i / L_) * LP + (i % L_)];
    });
}

template <int DIR>
static const void *four_step_kernel_ptr(int logL) {
    switch (logL) {
    case 6: return (const void *)&fft_four_step_cols_kernel<64, DIR>;
    case 7: return (const void *)&fft_four_step_cols_kernel<128, DIR>;
    case 8: return (const void *)&fft_four_step_cols_kernel<256, DIR>;
    case 9: return (const void *)&fft_four_step_cols_kernel<512, DIR>;
    case 10: return (const void *)&fft_four_step_cols_kernel<1024, DIR>;
    case 11: return (const void *)&fft_four_step_cols_kernel<2048, DIR>;
    case 12: return (const void *)&fft_four_step_cols_kernel<4096, DIR>;
    case 13: return (const void *)&fft_four_step_cols_kernel<8192, DIR>;
    default: return nullptr;
    }
}

hipError_t launch_four_step_cols(int logN1, int direction, const FourStepArgs &a, hipStream_t s) {
    const void *k = direction < 0 ? four_step_kernel_ptr<-1>(logN1) : four_step_kernel_ptr<1>(logN1);
    if (!k) return hipErrorInvalidValue;
    const int L = 1 << logN1;
    if (a.n2 % fs_TX(L) != 0 || a.ntiles != a.n2 / fs_TX(L)) return hipErrorInvalidValue;
    const long long nblocks = (long long)a.ntiles * a.batch;
    if (nblocks <= 0) return hipSuccess;
    if (nblocks > 0x7fffffffLL) return hipErrorInvalidValue;
    FourStepArgs args = a;
    void *params[] = {&args};
    return hipLaunchKernel(k, dim3((unsigned)nblocks), dim3(fs_WG(L)), params, 0, s);
}

// The alternative tile width of a column length (col_TX_alt), plain mode.
template <int DIR>
static const void *col_alt_kernel_ptr(int logL) {
    switch (logL) {
    case 7: return (const void *)&fft_cols_kernel<128, DIR, COL_PLAIN, col_TX_alt(128)>;
    case 8: return (const void *)&fft_cols_kernel<256, DIR, COL_PLAIN, col_TX_alt(256)>;
    case 9: return (const void *)&fft_cols_kernel<512, DIR, COL_PLAIN, col_TX_alt(512)>;
    case 10: return (const void *)&fft_cols_kernel<1024, DIR, COL_PLAIN, col_TX_alt(1024)>;
    case 11: return (const void *)&fft_cols_kernel<2048, DIR, COL_PLAIN, col_TX_alt(2048)>;
    case 12: return (const void *)&fft_cols_kernel<4096, DIR, COL_PLAIN, col_TX_alt(4096)>;
    case 13: return (const void *)&fft_cols_kernel<8192, DIR, COL_PLAIN, col_TX_alt(8192)>;
    default: return nullptr;
    }
}

template <int DIR>
static const void *col_kernel_ptr(int logL) {
    switch (logL) {
    case 1: return (const void *)&fft_cols_kernel<2, DIR>;
    case 2: return (const void *)&fft_cols_kernel<4, DIR>;
    case 3: return (const void *)&fft_cols_kernel<8, DIR>;
    case 4: return (const void *)&fft_cols_kernel<16, DIR>;
    case 5: return (const void *)&fft_cols_kernel<32, DIR>;
    case 6: return (const void *)&fft_cols_kernel<64, DIR>;
    case 7: return (const void *)&fft_cols_kernel<128, DIR>;
    case 8: return (const void *)&fft_cols_kernel<256, DIR>;
    case 9: return (const void *)&fft_cols_kernel<512, DIR>;
    case 10: return (const void *)&fft_cols_kernel<1024, DIR>;
    case 11: return (const void *)&fft_cols_kernel<2048, DIR>;
    case 12: return (const void *)&fft_cols_kernel<4096, DIR>;
    case 13: return (const void *)&fft_cols_kernel<8192, DIR>;
    default: return nullptr;
    }
}

// The middle and last passes of the three-pass transform: columns of 64 to
// 1024 points.
template <int DIR, int MODE>
static const void *col_mode_kernel_ptr(int logL) {
    switch (logL) {
    case 6: return (const void *)&fft_cols_kernel<64, DIR, MODE>;
    case 7: return (const void *)&fft_cols_kernel<128, DIR, MODE>;
    case 8: return (const void *)&fft_cols_kernel<256, DIR, MODE>;
    case 9: return (const void *)&fft_cols_kernel<512, DIR, MODE>;
    case 10: return (const void *)&fft_cols_kernel<1024, DIR, MODE>;
    default: return nullptr;
    }
}

hipError_t launch_cols(int logL, int direction, const ColArgs &a, hipStream_t s, int mode, int tx) {
    const void *k = nullptr;
    const int L = 1 << logL;
    if (tx == 0) tx = col_TX(L);
    if (tx != col_TX(L) && (mode != COL_PLAIN || logL < 1 || logL > COL_MAX_LOG2 || tx != col_TX_alt(L)))
        return hipErrorInvalidValue;
    if (a.ntiles <= 0 || (long long)a.ntiles * tx < a.ncols) return hipErrorInvalidValue;
    if (mode == COL_PLAIN && tx != col_TX(L))
        k = direction < 0 ? col_alt_kernel_ptr<-1>(logL) : col_alt_kernel_ptr<1>(logL);
    else if (mode == COL_PLAIN) k = direction < 0 ? col_kernel_ptr<-1>(logL) : col_kernel_ptr<1>(logL);
    else if (mode == COL_TWIDDLE)
        k = direction < 0 ? col_mode_kernel_ptr<-1, COL_TWIDDLE>(logL) : col_mode_kernel_ptr<1, COL_TWIDDLE>(logL);
    else if (mode == COL_TWO_LEVEL)
        k = direction < 0 ? col_mode_kernel_ptr<-1, COL_TWO_LEVEL>(logL) : col_mode_kernel_ptr<1, COL_TWO_LEVEL>(logL);
    else if (mode == COL_STREAM) {
        // Buffer offsets are 32-bit: the image must span less than 4 GiB.
        const long long rs = a.in_rs > a.out_rs ? a.in_rs : a.out_rs;
        if (((long long)(L - 1) * rs + a.ncols) * 8 >= (1LL << 32) || a.ncols > rs) return hipErrorInvalidValue;
        k = direction < 0 ? col_mode_kernel_ptr<-1, COL_STREAM>(logL) : col_mode_kernel_ptr<1, COL_STREAM>(logL);
    }
    if (!k) return hipErrorInvalidValue;
    if (mode == COL_TWIDDLE && (!a.coarse || !a.fine)) return hipErrorInvalidValue;
    if (mode == COL_TWO_LEVEL && a.img_div <= 0) return hipErrorInvalidValue;
    const long long nblocks = (long long)a.ntiles * a.batch;
    if (nblocks <= 0) return hipSuccess;
    if (nblocks > 0x7fffffffLL) return hipErrorInvalidValue;
    ColArgs args = a;
    void *params[] = {&args};
    return hipLaunchKernel(k, dim3((unsigned)nblocks), dim3(col_T(L) * tx), params, 0, s);
}

}  // namespace mufft
