// Host-visible launch interface of the gfx950 kernels (no HIP device code here).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "geometry.h"

namespace mufft {

// Fused single-workgroup transforms of contiguous sequences (1D, 2D rows).
struct RowArgs {
    const float2 *in;    // c2c/r2c: N points per transform (N/2 with zero pad); c2r: N+1 bins
    const float2 *in2;   // second operand of the fused multiply (RF_MUL)
    float2 *out;         // c2c: N points; r2c: N+1 bins (2N with RF_FULL); c2r: 2N reals
    const float2 *tw;    // Stockham twiddles, layout of tw_offset()
    const float2 *rtw;   // real-transform resolve twiddles W[k], k < N
    const float2 *stw;   // split variants: w_N^(dir*n*k1) at (k1-1)*(N/N1) + n, 0 < k1 < N1
    long long idist, idist2, odist;  // distance between transforms, in complex elements
    int batch;
    unsigned flags;      // RowFlag
    float norm;          // scale of the fused multiply
};

// Fused column transforms of length L over tiles of adjacent columns.
struct ColArgs {
    const float2 *in;
    float2 *out;
    const float2 *tw;
    long long in_rs, out_rs;   // row stride (complex elements)
    long long in_bs, out_bs;   // image stride (complex elements)
    int ncols;                 // columns to transform (<= row stride)
    int ntiles;                // column tiles per image
    int batch;                 // images
    // Mode COL_TWIDDLE: output row r of column c is multiplied by
    // w_M^(dir*(c >> log2div)*r), w_M^e = coarse[e >> log2s] * fine[e & (2^log2s - 1)].
    const float2 *coarse, *fine;
    int log2s, log2div;
    // Mode COL_TWO_LEVEL: image g sits at (g / img_div) * in_bs2 + (g % img_div) * in_bs
    // (and likewise for the output).
    int img_div;
    long long in_bs2, out_bs2;
};

// Column-kernel modes (the plain one serves the 2D and four-step passes;
// COL_TWIDDLE and COL_TWO_LEVEL are the middle and last passes of the
// three-pass transform; COL_STREAM is the plain transform with streaming
// (non-temporal) buffer accesses, for the two-pass four-step's second launch
// on columns of 64 to 1024 points, whose tile rows are whole cache lines).
enum ColMode : int { COL_PLAIN = 0, COL_TWIDDLE = 1, COL_TWO_LEVEL = 2, COL_STREAM = 3 };

// First half of a four-step transform of N = N1 * N2 points (n = n1*N2 + n2):
// the N1-point transforms down the columns of the N1 x N2 matrix, times
// w_N^(dir*n2*k1), written transposed (element (k1, n2) to n2*N1 + k1) so the
// second half is a plain column transform of length N2 over N1 columns.
// w_N^e = coarse[e >> log2S] * fine[e & (S - 1)].
struct FourStepArgs {
    const float2 *in;
    float2 *out;
    const float2 *tw;      // Stockham table of the N1-point column kernel
    const float2 *coarse;  // w_N^(dir*j*S), j < N/S
    const float2 *fine;    // w_N^(dir*j), j < S
    long long in_bs, out_bs;  // transform distance (complex elements)
    int n2;                // columns = input row stride
    int ntiles;            // column tiles per transform
    int batch;
    int log2s;
    unsigned flags;        // RF_ZERO_PAD: rows n1 >= N1/2 are zero
};

// One Stockham pass of radix R over global memory (fallback for any size).
struct PassArgs {
    const float2 *in;
    float2 *out;
    const float2 *tw;          // this pass's section: entry (m-1)*p + k
    long long n;               // transform length
    long long p;               // Stockham stride
    long long sy_in, sy_out;   // distance between consecutive points of one transform
    long long bs_in, bs_out;   // distance between transforms
    int ncols;                 // transforms interleaved with unit stride (columns)
    int batch;                 // transforms with stride bs
    unsigned flags;            // RF_ZERO_PAD (first pass only)
};

struct ResolveArgs {
    const float2 *in;
    const float2 *in2;         // RF_MUL second operand (c2r only)
    float2 *out;
    const float2 *rtw;
    long long n;               // complex length N (real length 2N)
    long long idist, idist2, odist;
    int batch;
    unsigned flags;            // RF_FULL (r2c), RF_MUL / RF_CONJ2 (c2r)
    float norm;
};

struct MulArgs {
    const float2 *a, *b;
    float2 *out;
    long long n;
    long long adist, bdist, odist;
    int batch;
    float norm;
    unsigned flags;            // RF_CONJ2
};

// Twiddle generation in FP64 on the GPU (rounded once to FP32).
struct TwiddleArgs {
    float2 *out;
    int direction;
    int count;
    // mode 0: Stockham table for a radix sequence (npass passes with radix[s], p[s]).
    // mode 1: real-transform resolve twiddles W[k] = (0, dir) * exp(i pi dir k / n).
    // mode 2: plain table exp(i pi dir k / (n/2)) for k < count.
    // mode 3: split twiddles exp(i pi dir (e mod p[0]) (e / p[0] + 1) / (n/2)), e < count.
    int mode;
    int r0;
    int npass;
    int radix[24];
    long long p[24];
    long long n;
};

// Fused row kernel of size 2^logN in variant `var` (geometry.h, row_variants.h).
hipError_t launch_rows(int logN, int direction, int kind, int var, const RowArgs &a, hipStream_t s);
// Registered variants of a size; returns the count (at most max are written).
// Without c2r, the entries that differ from another only in c2r-specific
// bits are left out (they run the same kernel for the other kinds).
int row_variant_list(int logN, int *vars, int max, bool c2r = true);
bool row_variant_registered(int logN, int var);
// tx: tile width, col_TX(L) (0) or col_TX_alt(L) (plain mode only); ntiles must cover ncols.
hipError_t launch_cols(int logL, int direction, const ColArgs &a, hipStream_t s, int mode = COL_PLAIN, int tx = 0);
hipError_t launch_pass(int radix, int direction, const PassArgs &a, hipStream_t s);
hipError_t launch_resolve_r2c(const ResolveArgs &a, hipStream_t s);
hipError_t launch_resolve_c2r(const ResolveArgs &a, hipStream_t s);
hipError_t launch_mul(const MulArgs &a, hipStream_t s);
hipError_t launch_fold(const MulArgs &a, hipStream_t s);
hipError_t launch_mul_reduce(const MulArgs &a, hipStream_t s);
hipError_t launch_four_step_cols(int logN1, int direction, const FourStepArgs &a, hipStream_t s);
hipError_t launch_stream_copy(const float2 *in, float2 *out, long long n, hipStream_t s);
hipError_t launch_ola(const float *y, float *out, const float *tail_in, float *tail_out,
                      long long n, long long nblocks, hipStream_t s);
hipError_t launch_twiddles(const TwiddleArgs &a, hipStream_t s);


}  // namespace mufft
