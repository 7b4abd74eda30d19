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
