/* A C99 caller of the drop-in, built the way a reference user builds one:
 * #include "mufft/fft.h" and link libmufft_hip.so (the reference's callers
 * include fft.h and link the library, CMakeLists.txt:23-25, test.c:31-488).
 *
 *   dropin_check layout
 *       prints the layout of mufft_hip_conv_job and the macro values
 *       (no device needed);
 *   dropin_check run MANIFEST
 *       runs every line of MANIFEST through the 19 reference entry points:
 *         c2c N DIR IN OUT | r2c N FLAGS IN OUT | c2r N IN OUT
 *         c2c2d NX NY DIR IN OUT | r2c2d NX NY IN OUT | c2r2d NX NY IN OUT
 *         conv N METHOD A B OUT | convolve COUNT A B OUT
 *       IN/A/B are raw little-endian float32 files, OUT is written the same
 *       way.  Buffers come from mufft_alloc / mufft_calloc (pinned host
 *       memory), as the reference's test harness uses mufft_alloc.
 *       Also checks that invalid sizes give NULL plans (fft.c:478-481,
 *       628-631, 666-669) and that the free functions accept NULL.
 */
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mufft/fft.h"
#include "mufft/mufft_hip.h"

static int fail(const char *what) {
    fprintf(stderr, "dropin_check: %s (%s)\n", what, mufft_hip_get_last_error());
    return 1;
}

/* Reads a whole file into a new mufft_alloc buffer of at least min_bytes
 * (the remainder zeroed); *bytes gets the file size. */
static void *read_file(const char *path, size_t min_bytes, size_t *bytes) {
    FILE *f = fopen(path, "rb");
    long n;
    void *p;
    if (!f) return NULL;
    fseek(f, 0, SEEK_END);
    n = ftell(f);
    fseek(f, 0, SEEK_SET);
    p = mufft_calloc((size_t)n > min_bytes ? (size_t)n : min_bytes);
    if (p && fread(p, 1, (size_t)n, f) != (size_t)n) {
        mufft_free(p);
        p = NULL;
    }
    fclose(f);
    if (bytes) *bytes = (size_t)n;
    return p;
}

static int write_file(const char *path, const void *p, size_t bytes) {
    FILE *f = fopen(path, "wb");
    size_t w;
    if (!f) return 1;
    w = fwrite(p, 1, bytes, f);
    fclose(f);
    return w != bytes;
}

static int run_1d(const char *kind, unsigned N, int arg, const char *in, const char *out) {
    mufft_plan_1d *plan;
    size_t out_bytes;
    void *x, *y;
    int rc;
    if (!strcmp(kind, "c2c")) {
        plan = mufft_create_plan_1d_c2c(N, arg, 0);
        out_bytes = (size_t)N * 8;
    } else if (!strcmp(kind, "r2c")) {
        plan = mufft_create_plan_1d_r2c(N, (unsigned)arg);
        out_bytes = ((unsigned)arg & MUFFT_FLAG_FULL_R2C) ? (size_t)N * 8 : (size_t)(N / 2 + 1) * 8;
    } else {
        plan = mufft_create_plan_1d_c2r(N, 0);
        out_bytes = (size_t)N * 4;
    }
    if (!plan) return fail("1D plan creation");
    x = read_file(in, 0, NULL);
    y = mufft_calloc(out_bytes);
    if (!x || !y) return fail("1D buffers");
    mufft_execute_plan_1d(plan, y, x);
    rc = write_file(out, y, out_bytes);
    mufft_free(x);
    mufft_free(y);
    mufft_free_plan_1d(plan);
    return rc;
}

static int run_2d(const char *kind, unsigned Nx, unsigned Ny, int dir, const char *in, const char *out) {
    mufft_plan_2d *plan;
    size_t out_bytes;
    void *x, *y;
    int rc;
    if (!strcmp(kind, "c2c2d")) {
        plan = mufft_create_plan_2d_c2c(Nx, Ny, dir, 0);
        out_bytes = (size_t)Nx * Ny * 8;
    } else if (!strcmp(kind, "r2c2d")) {
        plan = mufft_create_plan_2d_r2c(Nx, Ny, 0);
        out_bytes = (size_t)Nx * Ny * 8;   /* Nx complex per row, Nx/2+1 meaningful */
    } else {
        plan = mufft_create_plan_2d_c2r(Nx, Ny, 0);
        out_bytes = (size_t)Nx * Ny * 4;
    }
    if (!plan) return fail("2D plan creation");
    x = read_file(in, 0, NULL);
    /* c2r: 2*Nx*Ny floats, as the reference documents for its scratch use (fft.h:258-260) */
    y = mufft_calloc(2 * out_bytes);
    if (!x || !y) return fail("2D buffers");
    mufft_execute_plan_2d(plan, y, x);
    rc = write_file(out, y, out_bytes);
    mufft_free(x);
    mufft_free(y);
    mufft_free_plan_2d(plan);
    return rc;
}

static int run_conv(unsigned N, unsigned method, const char *fa, const char *fb, const char *out) {
    mufft_plan_conv *plan = mufft_create_plan_conv(N, 0, method);
    size_t bs, out_bytes = (method & MUFFT_CONV_METHOD_FLAG_STEREO_MONO) ? (size_t)N * 8 : (size_t)N * 4;
    void *a, *b, *blk0, *blk1, *y;
    int rc;
    if (!plan) return fail("conv plan creation");
    bs = mufft_conv_get_transformed_block_size(plan);
    a = read_file(fa, 0, NULL);
    b = read_file(fb, 0, NULL);
    blk0 = mufft_alloc(bs);
    blk1 = mufft_alloc(bs);
    y = mufft_calloc(out_bytes);
    if (!a || !b || !blk0 || !blk1 || !y) return fail("conv buffers");
    /* the reference's calling pattern (bench.c:251-257) */
    mufft_execute_conv_input(plan, MUFFT_CONV_BLOCK_SECOND, blk1, b);
    mufft_execute_conv_input(plan, MUFFT_CONV_BLOCK_FIRST, blk0, a);
    mufft_execute_conv_output(plan, y, blk0, blk1);
    rc = write_file(out, y, out_bytes);
    mufft_free(a);
    mufft_free(b);
    mufft_free(blk0);
    mufft_free(blk1);
    mufft_free(y);
    mufft_free_plan_conv(plan);
    return rc;
}

static int run_convolve(unsigned count, const char *fa, const char *fb, const char *out) {
    mufft_convolve_func f = mufft_get_convolve_func(0);
    void *a, *b, *y;
    int rc;
    if (!f) return fail("mufft_get_convolve_func");
    a = read_file(fa, 0, NULL);
    b = read_file(fb, 0, NULL);
    y = mufft_alloc((size_t)count * 8);
    if (!a || !b || !y) return fail("convolve buffers");
    f(y, a, b, 0.25f, count);
    rc = write_file(out, y, (size_t)count * 8);
    mufft_free(a);
    mufft_free(b);
    mufft_free(y);
    return rc;
}

static int invalid_sizes(void) {
    /* NULL plans for bad sizes, as the reference's creators return (fft.c:478-489). */
    if (mufft_create_plan_1d_c2c(1000, MUFFT_FORWARD, 0)) return fail("c2c N=1000 accepted");
    if (mufft_create_plan_1d_c2c(1, MUFFT_FORWARD, 0)) return fail("c2c N=1 accepted");
    if (mufft_create_plan_1d_r2c(2, 0)) return fail("r2c N=2 accepted");
    if (mufft_create_plan_1d_c2r(6, 0)) return fail("c2r N=6 accepted");
    if (mufft_create_plan_2d_c2c(3, 8, MUFFT_FORWARD, 0)) return fail("2D 3x8 accepted");
    if (mufft_create_plan_conv(100, 0, 0)) return fail("conv N=100 accepted");
    mufft_free_plan_1d(NULL);
    mufft_free_plan_2d(NULL);
    mufft_free_plan_conv(NULL);
    mufft_free(NULL);
    return 0;
}

int main(int argc, char **argv) {
    if (argc >= 2 && !strcmp(argv[1], "layout")) {
        printf("conv_job %zu %zu %zu %zu %zu %zu\n", sizeof(mufft_hip_conv_job),
               offsetof(mufft_hip_conv_job, N), offsetof(mufft_hip_conv_job, method),
               offsetof(mufft_hip_conv_job, first), offsetof(mufft_hip_conv_job, second),
               offsetof(mufft_hip_conv_job, output));
        printf("macros %d %d %d %d %d %d %d %d %d %d\n", MUFFT_FORWARD, MUFFT_INVERSE,
               MUFFT_FLAG_CPU_NO_SIMD, MUFFT_FLAG_FULL_R2C, MUFFT_FLAG_ZERO_PAD_UPPER_HALF,
               MUFFT_CONV_METHOD_FLAG_STEREO_MONO, MUFFT_CONV_METHOD_FLAG_ZERO_PAD_UPPER_HALF_FIRST,
               MUFFT_CONV_METHOD_FLAG_ZERO_PAD_UPPER_HALF_SECOND, MUFFT_CONV_BLOCK_FIRST,
               MUFFT_CONV_BLOCK_SECOND);
        return 0;
    }
    if (argc >= 3 && !strcmp(argv[1], "run")) {
        FILE *m = fopen(argv[2], "r");
        char line[2048], op[32], f1[512], f2[512], f3[512];
        int n = 0;
        if (!m) return fail("manifest");
        if (mufft_hip_device_count() < 1) return fail("no GPU");
        if (invalid_sizes()) return 1;
        while (fgets(line, sizeof line, m)) {
            unsigned a = 0, b = 0;
            int c = 0, rc = 1;
            if (sscanf(line, "%31s", op) != 1) continue;
            if (!strcmp(op, "c2c") && sscanf(line, "%*s %u %d %511s %511s", &a, &c, f1, f2) == 4)
                rc = run_1d(op, a, c, f1, f2);
            else if (!strcmp(op, "r2c") && sscanf(line, "%*s %u %d %511s %511s", &a, &c, f1, f2) == 4)
                rc = run_1d(op, a, c, f1, f2);
            else if (!strcmp(op, "c2r") && sscanf(line, "%*s %u %511s %511s", &a, f1, f2) == 3)
                rc = run_1d(op, a, 0, f1, f2);
            else if (!strcmp(op, "c2c2d") && sscanf(line, "%*s %u %u %d %511s %511s", &a, &b, &c, f1, f2) == 5)
                rc = run_2d(op, a, b, c, f1, f2);
            else if ((!strcmp(op, "r2c2d") || !strcmp(op, "c2r2d")) &&
                     sscanf(line, "%*s %u %u %511s %511s", &a, &b, f1, f2) == 4)
                rc = run_2d(op, a, b, 0, f1, f2);
            else if (!strcmp(op, "conv") && sscanf(line, "%*s %u %u %511s %511s %511s", &a, &b, f1, f2, f3) == 5)
                rc = run_conv(a, b, f1, f2, f3);
            else if (!strcmp(op, "convolve") && sscanf(line, "%*s %u %511s %511s %511s", &a, f1, f2, f3) == 4)
                rc = run_convolve(a, f1, f2, f3);
            if (rc) {
                fprintf(stderr, "dropin_check: failed on: %s", line);
                return 1;
            }
            ++n;
        }
        fclose(m);
        printf("ran %d cases\n", n);
        return 0;
    }
    fprintf(stderr, "usage: dropin_check layout | run MANIFEST\n");
    return 2;
}
