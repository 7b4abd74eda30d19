/* Native-thread driver for the CPU baseline (test and measurement
 * infrastructure only): runs the benchmark workload on the compiled reference
 * library through its public API (fft.h), one set of plans per thread, with
 * the timed loop entirely in C so that no interpreter lock serialises the
 * threads.  The reference's own benchmark times one plan the same way
 * (bench.c:201-228: plan outside the loop, CLOCK_MONOTONIC); plans are not
 * re-entrant (fft.c:806, 815), hence one per thread.
 *
 * Workload per repetition and thread t: forward then inverse c2c of length
 * n_c2c_len on inputs t, t + threads, ... of c2c_in (n_c2c packed complex
 * transforms), and r2c then c2r of length n_real_len on inputs t, t + threads,
 * ... of real_in (n_real packed real signals).
 */
#include <pthread.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "fft.h"

/* Start gate: each thread prepares its plans and buffers, reports ready,
 * then waits for `go` (or `abort`, when not every thread could be started). */
typedef struct {
    pthread_mutex_t m;
    pthread_cond_t cv;
    int ready, go, abort;
} gate;

typedef struct {
    int tid, threads, nrep;
    unsigned n_c2c_len, n_real_len;
    const float *c2c_in, *real_in;
    int n_c2c, n_real;
    gate *g;
    int ok;
    struct timespec end;   /* when this thread's timed loop finished */
} job;

static void *run(void *arg) {
    job *j = (job *)arg;
    const unsigned nc = j->n_c2c_len, nr = j->n_real_len;
    mufft_plan_1d *fwd = mufft_create_plan_1d_c2c(nc, MUFFT_FORWARD, 0);
    mufft_plan_1d *inv = mufft_create_plan_1d_c2c(nc, MUFFT_INVERSE, 0);
    mufft_plan_1d *r2c = mufft_create_plan_1d_r2c(nr, 0);
    mufft_plan_1d *c2r = mufft_create_plan_1d_c2r(nr, 0);
    int mine_c = 0, mine_r = 0;
    for (int i = j->tid; i < j->n_c2c; i += j->threads) mine_c++;
    for (int i = j->tid; i < j->n_real; i += j->threads) mine_r++;
    /* aligned private copies of this thread's inputs (fft.h:178) */
    float *xc = mufft_alloc((size_t)(mine_c ? mine_c : 1) * nc * 2 * sizeof(float));
    float *xr = mufft_alloc((size_t)(mine_r ? mine_r : 1) * nr * sizeof(float));
    float *X = mufft_alloc((size_t)nc * 2 * sizeof(float));
    float *Y = mufft_alloc((size_t)nc * 2 * sizeof(float));
    float *S = mufft_alloc((size_t)(nr / 2 + 1) * 2 * sizeof(float));
    float *Z = mufft_alloc((size_t)nr * sizeof(float));
    j->ok = fwd && inv && r2c && c2r && xc && xr && X && Y && S && Z;
    if (j->ok) {
        int k = 0;
        for (int i = j->tid; i < j->n_c2c; i += j->threads, k++)
            memcpy(xc + (size_t)k * nc * 2, j->c2c_in + (size_t)i * nc * 2, (size_t)nc * 2 * sizeof(float));
        k = 0;
        for (int i = j->tid; i < j->n_real; i += j->threads, k++)
            memcpy(xr + (size_t)k * nr, j->real_in + (size_t)i * nr, (size_t)nr * sizeof(float));
    }
    pthread_mutex_lock(&j->g->m);
    j->g->ready++;
    pthread_cond_broadcast(&j->g->cv);
    while (!j->g->go && !j->g->abort) pthread_cond_wait(&j->g->cv, &j->g->m);
    const int run_it = j->g->go && j->ok;
    pthread_mutex_unlock(&j->g->m);
    if (run_it) {
        for (int rep = 0; rep < j->nrep; rep++) {
            for (int k = 0; k < mine_c; k++) {
                mufft_execute_plan_1d(fwd, X, xc + (size_t)k * nc * 2);
                mufft_execute_plan_1d(inv, Y, X);
            }
            for (int k = 0; k < mine_r; k++) {
                mufft_execute_plan_1d(r2c, S, xr + (size_t)k * nr);
                mufft_execute_plan_1d(c2r, Z, S);
            }
        }
    }
    clock_gettime(CLOCK_MONOTONIC, &j->end);
    mufft_free_plan_1d(fwd);
    mufft_free_plan_1d(inv);
    mufft_free_plan_1d(r2c);
    mufft_free_plan_1d(c2r);
    mufft_free(xc);
    mufft_free(xr);
    mufft_free(X);
    mufft_free(Y);
    mufft_free(S);
    mufft_free(Z);
    return NULL;
}

/* Wall-clock seconds from the common start of all threads to the end of the
 * last one's timed loop, or -1 if a plan or buffer could not be created. */
double mufft_ref_bench_run(int threads, int nrep, unsigned n_c2c_len, const float *c2c_in, int n_c2c,
                           unsigned n_real_len, const float *real_in, int n_real) {
    if (threads < 1 || threads > 256 || nrep < 1) return -1.0;
    pthread_t *th = calloc((size_t)threads, sizeof(pthread_t));
    job *jobs = calloc((size_t)threads, sizeof(job));
    if (!th || !jobs) {
        free(th);
        free(jobs);
        return -1.0;
    }
    gate g;
    pthread_mutex_init(&g.m, NULL);
    pthread_cond_init(&g.cv, NULL);
    g.ready = g.go = g.abort = 0;
    int started = 0;
    for (int t = 0; t < threads; t++) {
        jobs[t] = (job){t, threads, nrep, n_c2c_len, n_real_len, c2c_in, real_in, n_c2c, n_real, &g, 0, {0, 0}};
        if (pthread_create(&th[t], NULL, run, &jobs[t]) != 0) break;
        started++;
    }
    double secs = -1.0;
    struct timespec t0;
    pthread_mutex_lock(&g.m);
    while (g.ready < started) pthread_cond_wait(&g.cv, &g.m);
    if (started == threads) g.go = 1;
    else g.abort = 1;
    pthread_cond_broadcast(&g.cv);
    pthread_mutex_unlock(&g.m);
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int t = 0; t < started; t++) pthread_join(th[t], NULL);
    if (started == threads) {
        for (int t = 0; t < threads; t++) {
            const double s = (double)(jobs[t].end.tv_sec - t0.tv_sec) +
                             1e-9 * (double)(jobs[t].end.tv_nsec - t0.tv_nsec);
            if (s > secs) secs = s;
        }
        for (int t = 0; t < threads; t++)
            if (!jobs[t].ok) secs = -1.0;
    }
    pthread_cond_destroy(&g.cv);
    pthread_mutex_destroy(&g.m);
    free(th);
    free(jobs);
    return secs;
}

/* ---- Per-config timings beside the GPU figures (bench.c's method: plans
 * outside the timed loop, CLOCK_MONOTONIC around `nrep` repetitions). ---- */

static double now_s(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

/* One plan, one transform per call, one thread (config: c2c N=1024 x 1):
 * seconds for nrep calls, or -1.  kind 0 c2c (dir), 1 r2c, 2 c2r. */
double mufft_ref_time_1d(int kind, unsigned N, int dir, const void *in, int nrep) {
    mufft_plan_1d *p = kind == 0 ? mufft_create_plan_1d_c2c(N, dir, 0)
                     : kind == 1 ? mufft_create_plan_1d_r2c(N, 0) : mufft_create_plan_1d_c2r(N, 0);
    const size_t ib = kind == 1 ? (size_t)N * 4 : (size_t)N * 8;
    void *x = mufft_alloc(ib), *y = mufft_alloc((size_t)N * 8 + 64);
    double s = -1.0;
    if (p && x && y && nrep > 0) {
        memcpy(x, in, ib);
        mufft_execute_plan_1d(p, y, x);
        const double t0 = now_s();
        for (int r = 0; r < nrep; r++) mufft_execute_plan_1d(p, y, x);
        s = now_s() - t0;
    }
    mufft_free_plan_1d(p);
    mufft_free(x);
    mufft_free(y);
    return s;
}

/* 2D images on native threads, one plan per thread (plans are not
 * re-entrant, fft.c:860): thread t runs images t, t + threads, ...  kind 0
 * c2c (dir), 1 r2c, 2 c2r; in holds nimg images of the plan's input layout
 * (c2c: Nx*Ny complex; r2c: Nx*Ny reals; c2r: Nx*Ny complex, padded rows). */
typedef struct {
    int tid, threads, nrep, kind, dir, nimg;
    unsigned nx, ny;
    const char *in;
    size_t ib;
    gate *g;
    int ok;
    struct timespec end;
} job2d;

static void *run2d(void *arg) {
    job2d *j = (job2d *)arg;
    mufft_plan_2d *p = j->kind == 0 ? mufft_create_plan_2d_c2c(j->nx, j->ny, j->dir, 0)
                     : j->kind == 1 ? mufft_create_plan_2d_r2c(j->nx, j->ny, 0)
                                    : mufft_create_plan_2d_c2r(j->nx, j->ny, 0);
    int mine = 0;
    for (int i = j->tid; i < j->nimg; i += j->threads) mine++;
    char *x = mufft_alloc((size_t)(mine ? mine : 1) * j->ib);
    /* c2r writes 2*Nx*Ny floats (fft.h:258-260); c2c / r2c Nx*Ny complex */
    void *y = mufft_alloc((size_t)j->nx * j->ny * 8);
    j->ok = p && x && y;
    if (j->ok) {
        int k = 0;
        for (int i = j->tid; i < j->nimg; i += j->threads, k++)
            memcpy(x + (size_t)k * j->ib, j->in + (size_t)i * j->ib, j->ib);
    }
    pthread_mutex_lock(&j->g->m);
    j->g->ready++;
    pthread_cond_broadcast(&j->g->cv);
    while (!j->g->go && !j->g->abort) pthread_cond_wait(&j->g->cv, &j->g->m);
    const int run_it = j->g->go && j->ok;
    pthread_mutex_unlock(&j->g->m);
    if (run_it)
        for (int rep = 0; rep < j->nrep; rep++)
            for (int k = 0; k < mine; k++) mufft_execute_plan_2d(p, y, x + (size_t)k * j->ib);
    clock_gettime(CLOCK_MONOTONIC, &j->end);
    mufft_free_plan_2d(p);
    mufft_free(x);
    mufft_free(y);
    return NULL;
}

/* Conv rounds (fft.c:799-809) on native threads: thread t runs jobs t,
 * t + threads, ... with its own plan per (N, method) and its own blocks;
 * each job = execute_conv_input(block 0), (block 1), execute_conv_output. */
#define REF_CONV_MAX_PLANS 32
typedef struct {
    int tid, threads, nrep, njobs;
    const unsigned *N, *method;
    const void *const *first, *const *second;
    gate *g;
    int ok;
    struct timespec end;
} jobconv;

static void *runconv(void *arg) {
    jobconv *j = (jobconv *)arg;
    mufft_plan_conv *plans[REF_CONV_MAX_PLANS];
    unsigned keyN[REF_CONV_MAX_PLANS], keyM[REF_CONV_MAX_PLANS];
    void *b0[REF_CONV_MAX_PLANS], *b1[REF_CONV_MAX_PLANS], *outb[REF_CONV_MAX_PLANS];
    int np = 0;
    j->ok = 1;
    int *pi = calloc((size_t)(j->njobs ? j->njobs : 1), sizeof(int));
    /* aligned private copies of this thread's inputs (the SIMD kernels need
     * aligned buffers, fft.h:178) */
    void **xa = calloc((size_t)(j->njobs ? j->njobs : 1), sizeof(void *));
    void **xb = calloc((size_t)(j->njobs ? j->njobs : 1), sizeof(void *));
    if (!pi || !xa || !xb) j->ok = 0;
    for (int i = j->tid; j->ok && i < j->njobs; i += j->threads) {
        const unsigned n = j->N[i], m = j->method[i];
        const size_t elem1 = (m & MUFFT_CONV_METHOD_FLAG_STEREO_MONO) ? 8 : 4;
        const size_t b1 = ((m & MUFFT_CONV_METHOD_FLAG_ZERO_PAD_UPPER_HALF_FIRST) ? n / 2 : n) * elem1;
        const size_t b2 = ((m & MUFFT_CONV_METHOD_FLAG_ZERO_PAD_UPPER_HALF_SECOND) ? n / 2 : n) * 4;
        xa[i] = mufft_alloc(b1);
        xb[i] = mufft_alloc(b2);
        if (!xa[i] || !xb[i]) {
            j->ok = 0;
            break;
        }
        memcpy(xa[i], j->first[i], b1);
        memcpy(xb[i], j->second[i], b2);
    }
    for (int i = j->tid; j->ok && i < j->njobs; i += j->threads) {
        int f = -1;
        for (int q = 0; q < np; q++)
            if (keyN[q] == j->N[i] && keyM[q] == j->method[i]) f = q;
        if (f < 0) {
            if (np == REF_CONV_MAX_PLANS) {
                j->ok = 0;
                break;
            }
            plans[np] = mufft_create_plan_conv(j->N[i], 0, j->method[i]);
            if (!plans[np]) {
                j->ok = 0;
                break;
            }
            const size_t bs = mufft_conv_get_transformed_block_size(plans[np]);
            b0[np] = mufft_alloc(bs);
            b1[np] = mufft_alloc(bs);
            outb[np] = mufft_alloc((size_t)j->N[i] * 8);
            keyN[np] = j->N[i];
            keyM[np] = j->method[i];
            f = np++;
            if (!b0[f] || !b1[f] || !outb[f]) j->ok = 0;
        }
        pi[i] = f;
    }
    pthread_mutex_lock(&j->g->m);
    j->g->ready++;
    pthread_cond_broadcast(&j->g->cv);
    while (!j->g->go && !j->g->abort) pthread_cond_wait(&j->g->cv, &j->g->m);
    const int run_it = j->g->go && j->ok;
    pthread_mutex_unlock(&j->g->m);
    if (run_it)
        for (int rep = 0; rep < j->nrep; rep++)
            for (int i = j->tid; i < j->njobs; i += j->threads) {
                const int f = pi[i];
                mufft_execute_conv_input(plans[f], MUFFT_CONV_BLOCK_FIRST, b0[f], xa[i]);
                mufft_execute_conv_input(plans[f], MUFFT_CONV_BLOCK_SECOND, b1[f], xb[i]);
                mufft_execute_conv_output(plans[f], outb[f], b0[f], b1[f]);
            }
    clock_gettime(CLOCK_MONOTONIC, &j->end);
    for (int q = 0; q < np; q++) {
        mufft_free_plan_conv(plans[q]);
        mufft_free(b0[q]);
        mufft_free(b1[q]);
        mufft_free(outb[q]);
    }
    for (int i = j->tid; xa && xb && i < j->njobs; i += j->threads) {
        mufft_free(xa[i]);
        mufft_free(xb[i]);
    }
    free(xa);
    free(xb);
    free(pi);
    return NULL;
}

/* Starts `threads` workers over fn/args (element size `esz`), opens the gate
 * when all are ready, and returns the wall seconds to the last finish (-1 on
 * failure).  `ok_at` / `end_at` are the offsets of the ok flag and the end
 * time inside an argument record. */
static double run_gated(int threads, void *(*fn)(void *), char *args, size_t esz, gate *g,
                        size_t ok_at, size_t end_at) {
    pthread_t *th = calloc((size_t)threads, sizeof(pthread_t));
    if (!th) return -1.0;
    int started = 0;
    for (int t = 0; t < threads; t++) {
        if (pthread_create(&th[t], NULL, fn, args + (size_t)t * esz) != 0) break;
        started++;
    }
    pthread_mutex_lock(&g->m);
    while (g->ready < started) pthread_cond_wait(&g->cv, &g->m);
    if (started == threads) g->go = 1;
    else g->abort = 1;
    pthread_cond_broadcast(&g->cv);
    pthread_mutex_unlock(&g->m);
    struct timespec t0;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int t = 0; t < started; t++) pthread_join(th[t], NULL);
    double secs = -1.0;
    if (started == threads) {
        for (int t = 0; t < threads; t++) {
            const char *rec = args + (size_t)t * esz;
            const struct timespec *e = (const struct timespec *)(rec + end_at);
            const double s = (double)(e->tv_sec - t0.tv_sec) + 1e-9 * (double)(e->tv_nsec - t0.tv_nsec);
            if (s > secs) secs = s;
            if (!*(const int *)(rec + ok_at)) secs = -1.0;
        }
        for (int t = 0; t < threads; t++)
            if (!*(const int *)(args + (size_t)t * esz + ok_at)) secs = -1.0;
    }
    free(th);
    return secs;
}

double mufft_ref_time_2d(int kind, unsigned nx, unsigned ny, int dir, const void *in, int nimg,
                         int threads, int nrep) {
    if (threads < 1 || threads > 256 || nrep < 1 || nimg < 1) return -1.0;
    job2d *jobs = calloc((size_t)threads, sizeof(job2d));
    if (!jobs) return -1.0;
    gate g;
    pthread_mutex_init(&g.m, NULL);
    pthread_cond_init(&g.cv, NULL);
    g.ready = g.go = g.abort = 0;
    const size_t ib = kind == 1 ? (size_t)nx * ny * 4 : (size_t)nx * ny * 8;
    for (int t = 0; t < threads; t++)
        jobs[t] = (job2d){t, threads, nrep, kind, dir, nimg, nx, ny, (const char *)in, ib, &g, 0, {0, 0}};
    const double s = run_gated(threads, run2d, (char *)jobs, sizeof(job2d), &g, offsetof(job2d, ok),
                               offsetof(job2d, end));
    pthread_cond_destroy(&g.cv);
    pthread_mutex_destroy(&g.m);
    free(jobs);
    return s;
}

double mufft_ref_time_conv(int threads, int nrep, int njobs, const unsigned *N, const unsigned *method,
                           const void *const *first, const void *const *second) {
    if (threads < 1 || threads > 256 || nrep < 1 || njobs < 1) return -1.0;
    jobconv *jobs = calloc((size_t)threads, sizeof(jobconv));
    if (!jobs) return -1.0;
    gate g;
    pthread_mutex_init(&g.m, NULL);
    pthread_cond_init(&g.cv, NULL);
    g.ready = g.go = g.abort = 0;
    for (int t = 0; t < threads; t++)
        jobs[t] = (jobconv){t, threads, nrep, njobs, N, method, first, second, &g, 0, {0, 0}};
    const double s = run_gated(threads, runconv, (char *)jobs, sizeof(jobconv), &g, offsetof(jobconv, ok),
                               offsetof(jobconv, end));
    pthread_cond_destroy(&g.cv);
    pthread_mutex_destroy(&g.m);
    free(jobs);
    return s;
}
