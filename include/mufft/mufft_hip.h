/* Extensions of the muFFT API for MI355X (gfx950): batched and stream-ordered
 * execution on device-resident buffers, device memory helpers and error
 * reporting.  None of these has a counterpart in the reference (which runs
 * exactly one transform per call on host memory, fft.h:117, fft.h:272); they
 * use a separate mufft_hip_ prefix so the reference names stay untouched.
 *
 * Conventions
 *  - Return value: 0 on success, a negative value on error; the message is
 *    available from mufft_hip_get_last_error() on the calling thread.
 *  - stream: a hipStream_t passed as void* (NULL = the null stream).
 *    Batched calls are stream-ordered and asynchronous with respect to the
 *    host; the drop-in fft.h calls run on a plan-private stream and return
 *    after the transform has finished.
 *  - Distances are in elements of the buffer's type (complex float for
 *    complex data, float for real data); 0 selects the packed layout.
 */
#ifndef MUFFT_HIP_EXT_H
#define MUFFT_HIP_EXT_H

#include <stddef.h>

#include "fft.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Error of the last failed call on this thread ("" when none). */
const char *mufft_hip_get_last_error(void);
/* Number of visible GPUs (0 when the HIP runtime finds none). */
int mufft_hip_device_count(void);

/* Device memory and synchronisation helpers. */
void *mufft_hip_malloc(size_t bytes);
void mufft_hip_free(void *ptr);
int mufft_hip_memcpy(void *dst, const void *src, size_t bytes);
int mufft_hip_synchronize(void *stream);

/* howmany 1D transforms of one plan.  Input element count per transform:
 * c2c N (N/2 with ZERO_PAD), r2c N reals (N/2 with ZERO_PAD), c2r N/2+1
 * complex; output: c2c N, r2c N/2+1 complex (N with FULL_R2C), c2r N reals. */
int mufft_hip_execute_plan_1d_batched(mufft_plan_1d *plan, void *output, const void *input,
                                      unsigned howmany, size_t idist, size_t odist, void *stream);

/* Allocation-free form of mufft_hip_execute_plan_1d_batched, with no
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, layouts and results (the same kernels run);
 * input and output must be device memory or pinned host memory (pageable
 * memory is refused).  Sizes that run as one fused kernel (up to 65536
 * complex points) need no scratch.  Larger sizes take their scratch from a
 * fixed 64 MiB device array, one per GPU, held until the process ends: the
 * call is refused when N x howmany exceeds 2^22 complex points (N = the
 * plan's complex length, N/2 of the real length for r2c / c2r).  All calls
 * on one GPU share that array, so such calls that overlap in time (other
 * streams, other threads, other plans) race and must be ordered by the
 * caller.  The plan is created beforehand, as usual. */
int mufft_hip_execute_plan_1d_batched_noalloc(mufft_plan_1d *plan, void *output, const void *input,
                                              unsigned howmany, size_t idist, size_t odist,
                                              void *stream);

/* howmany packed 2D images (layouts of mufft_execute_plan_2d, one after another). */
int mufft_hip_execute_plan_2d_batched(mufft_plan_2d *plan, void *output, const void *input,
                                      unsigned howmany, void *stream);

/* Batched convolution halves; block layouts as in mufft_execute_conv_input/output.
 * Distances between blocks are in complex elements (0 = block size / 8). */
int mufft_hip_execute_conv_input_batched(mufft_plan_conv *plan, unsigned block, void *output,
                                         const void *input, unsigned howmany, size_t idist,
                                         size_t odist, void *stream);
int mufft_hip_execute_conv_output_batched(mufft_plan_conv *plan, void *output,
                                          const void *input_first, const void *input_second,
                                          unsigned howmany, size_t odist, size_t dist_first,
                                          size_t dist_second, void *stream);

/* ---- Adjoints (for differentiation) ----
 * The adjoint (conjugate transpose) of a plan's transform, on device memory,
 * gradient convention dL/dRe + i dL/dIm:
 *   c2c: the opposite-direction unnormalised c2c;
 *   r2c: N/2+1 complex gradient bins (N with FULL_R2C) -> N real gradients;
 *   c2r: N real gradients -> N/2+1 complex gradient bins, including the
 *        reference's use of Im(X_0), Im(X_{N/2}) (kernel.c:36-45).
 * Zero-padded plans return gradients for all N inputs; the caller keeps the
 * first half.  Not available for zero-padded c2r plans.  Distances are in
 * elements of each buffer's type, 0 = packed. */
int mufft_hip_execute_adjoint_1d_batched(mufft_plan_1d *plan, void *output, const void *input,
                                         unsigned howmany, size_t idist, size_t odist,
                                         void *stream);
/* output = T(norm * input * op(input2)), op = conj when conj2 != 0, T the plan's
 * transform (adjoint = 0) or its adjoint (adjoint = 1); the multiply is fused
 * into the transform's first pass.  idist2 = 0 uses one input2 block for every
 * transform.  Needs a complex-input transform (c2c, c2r, adjoint of r2c). */
int mufft_hip_execute_plan_1d_mul_batched(mufft_plan_1d *plan, void *output, const void *input,
                                          const void *input2, unsigned howmany, size_t idist,
                                          size_t idist2, size_t odist, float norm, int conj2,
                                          int adjoint, void *stream);
/* output[i] = norm * sum_b a_b[i] * op(b_b[i]), i < n (complex), op = conj when
 * conj_b != 0: the batch reduction of a filter gradient.  Deterministic. */
int mufft_hip_mul_reduce(void *output, const void *a, const void *b, unsigned n, unsigned howmany,
                         size_t adist, size_t bdist, float norm, int conj_b, void *stream);

/* As mufft_create_plan_1d_c2c / _r2c / _c2r, with the fused-kernel variant
 * fixed for this plan.  kind: 0 c2c forward, 1 c2c inverse, 2 r2c, 3 c2r (the
 * codes of the variant calls below); N as for those creators; variant: a
 * registered variant of the size (mufft_hip_list_variants), 0 = the
 * process-wide choice.  No counterpart in the reference, whose plans always
 * take its static decomposition (fft.c:325-379). */
mufft_plan_1d *mufft_hip_create_plan_1d_variant(unsigned N, int kind, unsigned flags, int variant);
/* The fused-kernel variant a plan runs (0 when it runs no fused row kernel). */
int mufft_hip_plan_1d_variant(const mufft_plan_1d *plan);

/* Number of kernel launches one execution of the plan issues. */
int mufft_hip_plan_1d_launches(const mufft_plan_1d *plan);
int mufft_hip_plan_2d_launches(const mufft_plan_2d *plan);

/* The FP32 Stockham twiddle table the plan of a c2c transform of size N
 * uses, generated in FP64 on the GPU; copied to host_out (count complex
 * values, at most the table size).  Returns the table size. */
int mufft_hip_twiddle_table(unsigned N, int direction, float *host_out, unsigned count);

/* The auxiliary FP64-built tables, for inspection (count complex values
 * copied to host_out; returns the table size, or the size alone when
 * host_out is NULL):
 *   mode 1: the r2c / c2r resolve twiddles W[k] = (0, dir) * exp(i pi dir k / n),
 *           k < n, of a transform of complex length n (fft.c:441-455);
 *   mode 2: exp(dir * 2 pi i * j / n), j < param (four-step twiddles);
 *   mode 3: exp(dir * 2 pi i * j * k1 / n) at (k1 - 1) * param + j, 0 < k1 < n / param,
 *           j < param (twiddles of a transform split over n / param workgroups). */
int mufft_hip_twiddle_table_mode(int mode, unsigned n, unsigned param, int direction,
                                 float *host_out, unsigned count);

/* ---- Fused-kernel variants and the autotuner ----
 * A size of up to 32768 complex points runs as one fused kernel whose
 * variant fixes the points per thread, the LDS exchange mode and the twiddle
 * scheme.  kind: 0 c2c forward, 1 c2c inverse, 2 r2c, 3 c2r; N is the size
 * passed to the plan creator (real length for r2c / c2r).  A choice applies
 * to plans created afterwards, process-wide. */
int mufft_hip_list_variants(unsigned N, int kind, int *out, int max);
int mufft_hip_default_variant(unsigned N, int kind);
int mufft_hip_get_variant(unsigned N, int kind);
/* variant 0 restores the static default. */
int mufft_hip_set_variant(unsigned N, int kind, int variant);
/* Times every variant on howmany transforms, checks each against an FP64 FFT
 * (relative RMS <= 3e-6) and selects the fastest that passes; returns it.
 * Each timed launch starts with cold caches (a 256 MiB stream evicts the
 * Infinity Cache first), as a batch streamed from HBM would meet them. */
int mufft_hip_autotune_1d(unsigned N, int kind, unsigned howmany, double *best_us);
/* Transforms beyond the fused sizes (2^17..2^26 complex points; N as for
 * the plan creators, kind as above) run as two or three column launches over
 * a split N = N1 x N2 or N1 x A x B, coded log2 N1 << 8 | log2 A << 4 |
 * log2 B (A = B = 1 for two launches).  The autotuner times every runnable
 * split at howmany transforms, checks each against an FP64 FFT first
 * (relative RMS <= 3e-6) and selects the fastest; plans created afterwards
 * use it (process-wide, as the variants).  No counterpart in the reference,
 * whose large transforms run its fixed radix passes (README.md:49-51). */
int mufft_hip_autotune_large_1d(unsigned N, int kind, unsigned howmany, double *best_us);
int mufft_hip_get_large_split(unsigned N, int kind);
/* code 0 restores the static split. */
int mufft_hip_set_large_split(unsigned N, int kind, int code);
/* 2D column passes: the fused column kernel of length Ny owns tiles of
 * adjacent columns; two tile widths are compiled per length from 128 points
 * on.  The autotuner times both on an Nx x Ny image batch of the given kind
 * (real Nx for r2c / c2r, whose columns are Nx/2 + 1), requires identical
 * bits and selects the faster for plans created afterwards.  Returns the
 * width.  The reference picks its vertical passes by a fixed rule
 * (fft.c:382-435). */
int mufft_hip_autotune_2d(unsigned Nx, unsigned Ny, int kind, unsigned howmany, double *best_us);
int mufft_hip_get_column_tile(unsigned L);
/* tile 0 restores the static width. */
int mufft_hip_set_column_tile(unsigned L, int tile);
/* Tune cache as text, one "log2n kind choice us" line per tuned entry (kind
 * 0-3 fused variants, 4 large-transform splits, 5 column tile widths). */
int mufft_hip_tune_save(const char *path);
int mufft_hip_tune_load(const char *path);

/* ---- Uneven convolution batch ----
 * Many fast convolutions of mixed sizes and methods in one persistent kernel
 * launch fed by a device work queue, replayed from a captured hipGraph.  Each
 * job is one mufft_plan_conv round (both inputs transformed, multiplied,
 * inverse transformed, normalised by 1/N) on device buffers:
 *   first:  N/2 reals (mono) or N/2 complex (stereo) with the matching
 *           MUFFT_CONV_METHOD_FLAG_ZERO_PAD_UPPER_HALF_FIRST bit, else N;
 *   second: N/2 reals with ..._SECOND, else N reals;
 *   output: N reals (mono) or N complex (stereo).
 * Mono N = 8..65536, stereo N = 8..32768, powers of two.  A batch is a single
 * kernel launch: each job group is assigned to one XCD, whose workgroups run
 * its forward transforms and then its inverse transforms as soon as the
 * group's spectra are ready.  The batch owns the spectra scratch; the job
 * pointers are fixed at creation, and executions of one batch must not
 * overlap in time. */
typedef struct mufft_hip_conv_job {
    unsigned N;
    unsigned method;
    const void *first;
    const void *second;
    void *output;
} mufft_hip_conv_job;

typedef struct mufft_hip_conv_batch mufft_hip_conv_batch;

mufft_hip_conv_batch *mufft_hip_create_conv_batch(const mufft_hip_conv_job *jobs, unsigned count,
                                                  unsigned flags);
/* Replays the captured graph on stream (NULL: the batch's stream, and waits). */
int mufft_hip_execute_conv_batch(mufft_hip_conv_batch *batch, void *stream);
/* Same kernel without the graph. */
int mufft_hip_execute_conv_batch_direct(mufft_hip_conv_batch *batch, void *stream);
/* 0 = every execution so far ran every task, 1 = one did not (synchronises the device). */
int mufft_hip_conv_batch_status(mufft_hip_conv_batch *batch);
int mufft_hip_conv_batch_info(const mufft_hip_conv_batch *batch, int *ntasks, int *grid);
void mufft_hip_free_conv_batch(mufft_hip_conv_batch *batch);

/* Device-to-device copy with the FFT kernels' access width and streaming
 * cache policy (bytes a multiple of 8): the achievable-bandwidth reference for
 * their roofline figures.  No counterpart in the reference. */
int mufft_hip_stream_copy(void *dst, const void *src, size_t bytes, void *stream);

/* ---- Streaming overlap-add FIR filtering ----
 * y = x * h for an unbounded stream x, hop N/2, FFT size N (power of two,
 * 4..65536), real filter of ntaps <= N/2 taps (host or device pointer, copied).
 * Each call filters nblocks * N/2 input samples into as many output samples
 * (device buffers); the convolution tail carries over to the next call.
 * Calls on one object must be ordered on one stream. */
typedef struct mufft_hip_ola mufft_hip_ola;
mufft_hip_ola *mufft_hip_create_ola(unsigned N, const float *taps, unsigned ntaps);
int mufft_hip_execute_ola(mufft_hip_ola *ola, float *output, const float *input, unsigned nblocks,
                          void *stream);
/* Clears the carried tail (start of a new stream). */
int mufft_hip_reset_ola(mufft_hip_ola *ola, void *stream);
void mufft_hip_free_ola(mufft_hip_ola *ola);

#ifdef __cplusplus
}
#endif

#endif /* MUFFT_HIP_EXT_H */
