/* muFFT-compatible public C API, implemented on AMD MI355X (gfx950) GPUs.
 *
 * Drop-in for the reference muFFT header: every name, macro value and
 * signature below matches the reference (fft.h:44-65 flags and directions,
 * fft.h:86-121 1D, fft.h:131-207 convolution, fft.h:221-276 2D,
 * fft.h:286-297 memory) so a program written against muFFT recompiles and
 * links unchanged against libmufft_hip.so.
 *
 * Semantics kept from the reference:
 *  - transforms are unnormalised; forward = exp(-2 pi i nk/N) (fft.h:70-71);
 *  - r2c writes N/2+1 bins (N with MUFFT_FLAG_FULL_R2C); c2r reads N/2+1 bins;
 *  - 2D r2c/c2r rows hold Nx complex values, the first Nx/2+1 meaningful;
 *  - convolution output is normalised by 1/N;
 *  - creators return NULL on invalid sizes or allocation failure; execute
 *    functions are synchronous; free functions accept NULL;
 *  - the CPU SIMD selection bits (MUFFT_FLAG_CPU_*) are accepted and ignored.
 *
 * Buffers may be device memory (hipMalloc), pinned host memory (mufft_alloc)
 * or pageable host memory; host buffers are staged through device memory.
 * Batched, stream-ordered and device-resident entry points are declared in
 * mufft/mufft_hip.h.
 */
#ifndef MUFFT_HIP_FFT_H
#define MUFFT_HIP_FFT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__cplusplus)
#define MUFFT_RESTRICT
#elif __STDC_VERSION__ >= 199901L
#define MUFFT_RESTRICT restrict
#else
#define MUFFT_RESTRICT
#endif

/* Directions (reference fft.h:44-46). */
#define MUFFT_FORWARD (-1)
#define MUFFT_INVERSE (1)

/* Planning flags (reference fft.h:52-65). */
#define MUFFT_FLAG_CPU_ANY (0)
#define MUFFT_FLAG_CPU_NO_SIMD ((1 << 16) - 1)
#define MUFFT_FLAG_CPU_NO_AVX (1 << 0)
#define MUFFT_FLAG_CPU_NO_SSE3 (1 << 1)
#define MUFFT_FLAG_CPU_NO_SSE (1 << 2)
#define MUFFT_FLAG_FULL_R2C (1 << 16)
#define MUFFT_FLAG_ZERO_PAD_UPPER_HALF (1 << 17)

/* ---- 1D (reference fft.h:78-121) ---- */
typedef struct mufft_plan_1d mufft_plan_1d;

/* Replaces reference fft.h:86 (fft.c:626-662). N power of two, N >= 2. */
mufft_plan_1d *mufft_create_plan_1d_c2c(unsigned N, int direction, unsigned flags);
/* Replaces reference fft.h:99 (fft.c:476-513). N power of two, N >= 4. */
mufft_plan_1d *mufft_create_plan_1d_r2c(unsigned N, unsigned flags);
/* Replaces reference fft.h:111 (fft.c:515-547). N power of two, N >= 4. */
mufft_plan_1d *mufft_create_plan_1d_c2r(unsigned N, unsigned flags);
/* Replaces reference fft.h:117 (fft.c:811-850). Synchronous. */
void mufft_execute_plan_1d(mufft_plan_1d *plan, void *MUFFT_RESTRICT output,
                           const void *MUFFT_RESTRICT input);
/* Replaces reference fft.h:121 (fft.c:991-1002). NULL is a no-op. */
void mufft_free_plan_1d(mufft_plan_1d *plan);

/* ---- convolution (reference fft.h:131-207) ---- */
#define MUFFT_CONV_METHOD_FLAG_MONO_MONO 0
#define MUFFT_CONV_METHOD_FLAG_STEREO_MONO 1
#define MUFFT_CONV_METHOD_FLAG_ZERO_PAD_UPPER_HALF_FIRST (1 << 2)
#define MUFFT_CONV_METHOD_FLAG_ZERO_PAD_UPPER_HALF_SECOND (1 << 3)
#define MUFFT_CONV_BLOCK_FIRST 0
#define MUFFT_CONV_BLOCK_SECOND 1

typedef struct mufft_plan_conv mufft_plan_conv;

/* Replaces reference fft.h:168 (fft.c:565-624). */
mufft_plan_conv *mufft_create_plan_conv(unsigned N, unsigned flags, unsigned method);
/* Replaces reference fft.h:179 (fft.c:799-802). */
void mufft_execute_conv_input(mufft_plan_conv *plan, unsigned block, void *output, const void *input);
/* Replaces reference fft.h:185 (fft.c:794-797). */
size_t mufft_conv_get_transformed_block_size(mufft_plan_conv *plan);
/* Replaces reference fft.h:196 (fft.c:804-809); the multiply is fused into the inverse. */
void mufft_execute_conv_output(mufft_plan_conv *plan, void *output, const void *input_first,
                               const void *input_second);
/* Replaces reference fft.h:199 (fft.c:1019-1030). */
void mufft_free_plan_conv(mufft_plan_conv *plan);

typedef void (*mufft_convolve_func)(void *output, const void *a, const void *b,
                                    float normalization, unsigned samples);
/* Replaces reference fft.h:207 (fft.c:549-563): returns a GPU complex multiply. */
mufft_convolve_func mufft_get_convolve_func(unsigned flags);

/* ---- 2D (reference fft.h:221-276) ---- */
typedef struct mufft_plan_2d mufft_plan_2d;

/* Replaces reference fft.h:232 (fft.c:664-716). */
mufft_plan_2d *mufft_create_plan_2d_c2c(unsigned Nx, unsigned Ny, int direction, unsigned flags);
/* Replaces reference fft.h:248 (fft.c:718-759). */
mufft_plan_2d *mufft_create_plan_2d_r2c(unsigned Nx, unsigned Ny, unsigned flags);
/* Replaces reference fft.h:264 (fft.c:761-792). */
mufft_plan_2d *mufft_create_plan_2d_c2r(unsigned Nx, unsigned Ny, unsigned flags);
/* Replaces reference fft.h:270 (fft.c:852-989). Synchronous. */
void mufft_execute_plan_2d(mufft_plan_2d *plan, void *MUFFT_RESTRICT output,
                           const void *MUFFT_RESTRICT input);
/* Replaces reference fft.h:276 (fft.c:1004-1017). */
void mufft_free_plan_2d(mufft_plan_2d *plan);

/* ---- memory (reference fft.h:286-297) ---- */
/* Pinned host memory (hipHostMalloc): host-writable, DMA-capable, >= 64-byte aligned. */
void *mufft_alloc(size_t size);
void *mufft_calloc(size_t size);
void mufft_free(void *ptr);

#ifdef __cplusplus
}
#endif

#endif /* MUFFT_HIP_FFT_H */
