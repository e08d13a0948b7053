/* CPU-baseline timing harness around the reference library (oracle/_ref/libbridct.so).
 *
 * Test and benchmark infrastructure only: bench.py's cpu_baseline leg calls it to time the
 * reference CPU path on the host cores, and nothing in the product links it.
 *
 * Contract (restated from the reference bench/bench.c:41-145): plans and buffers are
 * allocated before timing; a warm-up pass; the repetition count doubles until one timed
 * pass takes at least target_ms; then `blocks` passes are timed and the median is reported
 * as nanoseconds per array.  No overhead is subtracted.
 *
 * Threading (the reference plan is not concurrently usable, docs/API.md:23): each thread
 * owns its own plans and a static, contiguous partition of the arrays.
 *
 * Mode 0 = FORWARD, 1 = INVERSE, 2 = fused C ROUNDTRIP, 3 = materialized FORWARD into a
 * spectrum buffer then INVERSE into the output (the NumPy Plan.roundtrip semantics,
 * python/bridct_numpy.c:114-133).
 */
#define _POSIX_C_SOURCE 200809L
#include "bridct.h"
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef struct {
    int count;                 /* arrays in this thread's partition */
    const int *h, *w;          /* per-array shape */
    const size_t *offset;      /* per-array float offset into in/out/spec */
    const float *in;
    float *out, *spec;
    int mode;
    uint64_t reps;
    bridct_plan *plans[8][8];  /* one plan per (log2 h - 3, log2 w - 3) */
    int status;
    double ns;
} job_t;

static uint64_t now_ns(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (uint64_t)t.tv_sec * 1000000000ull + (uint64_t)t.tv_nsec;
}

static int ilog2(int n)
{
    int k = 0;
    while ((1 << k) < n)
        k++;
    return k;
}

static int run_pass(job_t *j)
{
    for (int i = 0; i < j->count; i++) {
        bridct_plan *p = j->plans[ilog2(j->h[i]) - 3][ilog2(j->w[i]) - 3];
        const float *x = j->in + j->offset[i];
        float *y = j->out + j->offset[i];
        int rc;
        if (j->mode == 3) {
            float *s = j->spec + j->offset[i];
            rc = bridct_plan_apply(p, BRIDCT_FORWARD, x, s);
            if (!rc)
                rc = bridct_plan_apply(p, BRIDCT_INVERSE, s, y);
        } else {
            rc = bridct_plan_apply(p, (enum bridct_mode)j->mode, x, y);
        }
        if (rc)
            return rc;
    }
    return 0;
}

static void *worker(void *arg)
{
    job_t *j = arg;
    uint64_t t0 = now_ns();
    for (uint64_t r = 0; r < j->reps && !j->status; r++)
        j->status = run_pass(j);
    j->ns = (double)(now_ns() - t0);
    return NULL;
}

static int cmp_double(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return x < y ? -1 : x > y;
}

/* One timed pass over all threads; returns wall nanoseconds or a negative value. */
static double timed(job_t *jobs, int threads, uint64_t reps)
{
    pthread_t tid[256];
    for (int t = 0; t < threads; t++) {
        jobs[t].reps = reps;
        jobs[t].status = 0;
    }
    uint64_t t0 = now_ns();
    for (int t = 0; t < threads; t++)
        if (pthread_create(&tid[t], NULL, worker, &jobs[t]))
            return -1.0;
    int bad = 0;
    for (int t = 0; t < threads; t++) {
        pthread_join(tid[t], NULL);
        bad |= jobs[t].status != 0;
    }
    double ns = (double)(now_ns() - t0);
    return bad ? -2.0 : ns;
}

/* Times `count` arrays of (possibly mixed) shapes.  Returns 0 and writes the median
 * nanoseconds per full pass over all arrays (one repetition) to *ns_per_pass. */
int ref_bench_run(int count, const int *h, const int *w, const size_t *offset,
                  const float *in, float *out, float *spec, int mode, int threads,
                  int blocks, double target_ms, double *ns_per_pass)
{
    if (count < 1 || threads < 1 || threads > 256 || blocks < 1 || blocks > 101 ||
        mode < 0 || mode > 3 || (mode == 3 && !spec))
        return -1;
    if (threads > count)
        threads = count;
    job_t *jobs = calloc((size_t)threads, sizeof(job_t));
    if (!jobs)
        return -1;
    int rc = 0;
    for (int t = 0; t < threads && !rc; t++) {
        int lo = (int)((long long)count * t / threads);
        int hi = (int)((long long)count * (t + 1) / threads);
        job_t *j = &jobs[t];
        j->count = hi - lo;
        j->h = h + lo;
        j->w = w + lo;
        j->offset = offset + lo;
        j->in = in;
        j->out = out;
        j->spec = spec;
        j->mode = mode;
        for (int i = 0; i < j->count && !rc; i++) {
            int a = ilog2(j->h[i]) - 3, b = ilog2(j->w[i]) - 3;
            if (!bridct_supported(j->h[i], j->w[i]))
                rc = -1;
            else if (!j->plans[a][b] && !(j->plans[a][b] = bridct_plan_create(j->h[i], j->w[i])))
                rc = -1;
        }
    }
    double ns = 0.0;
    uint64_t reps = 1;
    if (!rc && timed(jobs, threads, 1) < 0)   /* warm-up */
        rc = -2;
    while (!rc) {
        ns = timed(jobs, threads, reps);
        if (ns < 0) {
            rc = -2;
            break;
        }
        if (ns >= target_ms * 1e6 || reps >= (1ull << 30))
            break;
        reps *= 2;
    }
    double samples[101];
    for (int b = 0; b < blocks && !rc; b++) {
        ns = timed(jobs, threads, reps);
        if (ns < 0)
            rc = -2;
        samples[b] = ns / (double)reps;
    }
    if (!rc) {
        qsort(samples, (size_t)blocks, sizeof(double), cmp_double);
        *ns_per_pass = samples[blocks / 2];
    }
    for (int t = 0; t < threads; t++)
        for (int a = 0; a < 8; a++)
            for (int b = 0; b < 8; b++)
                bridct_plan_destroy(jobs[t].plans[a][b]);
    free(jobs);
    return rc;
}

/* 1 when this host has AVX2 and FMA, i.e. when the reference library, built with its
 * runtime-selected AVX2 objects, takes its AVX2 routes (src/bridct_runtime.h:16-32). */
int ref_bench_avx2_fma(void)
{
#if defined(__x86_64__)
    return __builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma");
#else
    return 0;
#endif
}
