/* Host side of the MI355X BRiDCT plan API (C11).
 *
 * Mirrors the reference plan lifecycle (src/plan.c:12-163, src/bridct.c:13-75): the same
 * shape check, route strings, status codes, validation order and "validate before any
 * write" rule.  The arithmetic is issued to the GPU through the thin C ABI of
 * bridct_device.h; this file never calls HIP directly.
 *
 * A plan owns: one HIP stream, a device block of constant tables (twiddles, tangents and
 * the per-axis normalization for every axis length 8..1024, evaluated in FP64 and rounded
 * once to FP32), and device staging buffers for the host-pointer calls.  It has no
 * process-global mutable state (reference docs/API.md:23).
 */
#include "bridct.h"
#include "bridct_gpu.h"

#include "bridct_device.h"
#include "generated/sj_factors.h"
#include "copy_pool.h"
#include "host_internal.h"

#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static const double PI = 3.14159265358979323846;

/* Floats of staging per direction (host-pointer calls are copied in chunks of arrays). */
#define STAGE_FLOATS ((size_t)1 << 20)

/* Host calls of at least this many floats go through the pinned pipeline: two chunk slots,
 * each with a pinned input and output buffer, a device input and output buffer and two
 * events.  The host copies chunk k - 2 out of pinned memory and chunk k into it while the
 * GPU copies in, transforms and copies out chunk k - 1; the copies in run on a stream of
 * their own, so that they overlap the copies out on the other DMA engine.  Pageable
 * copies run at 31 GB/s (in) and 18 GB/s (out) here; pinned ones at 57 GB/s each way and
 * one thread's memcpy at 32 GB/s, so the host copies are shared by the threads of a
 * copy_pool (copy_pool.h). */
#define PIPELINE_MIN_FLOATS ((size_t)1 << 18)

struct bridct_plan {
    int height, width;
    size_t elements;      /* h * w */
    size_t chunk;         /* arrays per staging chunk */
    void *stream;         /* hipStream_t */
    float *d_tables;      /* BDX_TABLE_FLOATS */
    float *d_in, *d_out;  /* chunk * elements each */
    /* pinned pipeline for large host calls, allocated on first use */
    float *pin_in[2], *pin_out[2], *d_in2, *d_out2;
    void *ev[2];          /* chunk of the slot copied out (compute stream) */
    void *ev_in[2];       /* chunk of the slot copied in (copy stream) */
    void *copy_stream;    /* hipStream_t of the copies in */
    copy_pool *pool;
    int route;            /* BRIDCT_ROUTE_* */
    int pack;             /* arrays per workgroup of the whole-array kernel, 0 = automatic */
    float *d_dense;       /* dense route: Q_h', Q_h'^T, Q_w' */
    float dense_scale;    /* 1 / (s_h s_w) */
    float *d_work;        /* dense route intermediate */
    size_t work_floats;
    int device;           /* HIP device current at creation */
};

/* src/plan.c:33-37 */
int bridct_supported(int height, int width)
{
    return height >= 8 && height <= 1024 && width >= 8 && width <= 1024 &&
           !(height & (height - 1)) && !(width & (width - 1));
}

/* Johnson-Frigo scale s_{N,k}: 1 for N <= 4, else s_{N/4, k mod N/4} times the cosine
 * (k mod N/4 <= N/8) or sine of 2 pi (k mod N/4) / N (src/plan.c:39-47). */
static double jf_scale(int n, int k)
{
    if (n <= 4)
        return 1.0;
    k %= n / 4;
    const double angle = 2.0 * PI * k / n;
    return jf_scale(n / 4, k) * (k <= n / 8 ? cos(angle) : sin(angle));
}

static const double *sj_factor(int n)
{
    switch (n) {
    case 8: return bridct_sj_factor_8;
    case 16: return bridct_sj_factor_16;
    case 32: return bridct_sj_factor_32;
    default: return bridct_sj_factor_64;
    }
}

/* Per-axis normalization of the GPU core for axis length n, output index k, in FP64:
 * the orthonormal factor sqrt((k ? 2 : 1) / n) times the core's own output scale.
 *  - n <= 64: the core is diag(1/d_n) F_n with F_n = C_II / s_{4n,k} (src/plan.c:96),
 *    so the factor is s_{4n,k} d_n[k] sqrt(...).
 *  - n >= 128: the core returns C_II[k] / sigma_k with sigma_k = cos(pi k' / (2n)),
 *    k' = min(k, n - k), halved except at k in {0, n/4, n/2, 3n/4}. */
double bridct_axis_alpha(int n, int k)
{
    const double q = sqrt((k ? 2.0 : 1.0) / n);
    if (n <= 64)
        return q * jf_scale(4 * n, k) * sj_factor(n)[k];
    const int kk = k <= n / 2 ? k : n - k;
    const int special = k == 0 || k == n / 2 || k == n / 4 || k == 3 * n / 4;
    return q * cos(PI * kk / (2.0 * n)) * (special ? 1.0 : 0.5);
}

void bridct_build_tables(float *t)
{
    memset(t, 0, BDX_TABLE_FLOATS * sizeof(float));
    for (int n = 8; n <= 1024; n *= 2) {
        float *b = t + BDX_TABLE_BASE(n);
        for (int k = 0; k < n; k++) {
            const double a = 2.0 * PI * k / n;
            b[2 * k] = (float)cos(a);
            b[2 * k + 1] = (float)-sin(a);
        }
        for (int k = 0; k < n; k++)
            b[2 * n + k] = k <= n / 2 ? (float)tan(PI * k / (2.0 * n)) : 0.0f;
        for (int k = 0; k < n; k++) {
            const double a = bridct_axis_alpha(n, k);
            b[3 * n + k] = (float)a;
            b[4 * n + k] = (float)(a * a);
        }
        /* team twiddles exp(-2 pi i t k1 / N), N = n/2, stored at 2 (k1 T + t) */
        const int T = BDX_TEAM(n), N = n / 2, N1 = N / T;
        for (int k1 = 0; k1 < N1 && T > 1; k1++)
            for (int t = 0; t < T; t++) {
                const double a = 2.0 * PI * (double)((t * k1) % N) / N;
                b[5 * n + 2 * (k1 * T + t)] = (float)cos(a);
                b[5 * n + 2 * (k1 * T + t) + 1] = (float)-sin(a);
            }
    }
}

void bridct_plan_destroy(bridct_plan *plan)
{
    if (!plan)
        return;
    if (plan->copy_stream)
        bdx_stream_sync(plan->copy_stream);
    if (plan->stream)
        bdx_stream_sync(plan->stream);
    if (plan->d_dense)
        bdx_free(plan->d_dense);
    if (plan->d_work)
        bdx_free(plan->d_work);
    if (plan->d_tables)
        bdx_free(plan->d_tables);
    if (plan->d_in)
        bdx_free(plan->d_in);
    if (plan->d_out)
        bdx_free(plan->d_out);
    if (plan->d_in2)
        bdx_free(plan->d_in2);
    if (plan->d_out2)
        bdx_free(plan->d_out2);
    for (int i = 0; i < 2; i++) {
        if (plan->pin_in[i])
            bdx_host_free(plan->pin_in[i]);
        if (plan->pin_out[i])
            bdx_host_free(plan->pin_out[i]);
        if (plan->ev[i])
            bdx_event_destroy(plan->ev[i]);
        if (plan->ev_in[i])
            bdx_event_destroy(plan->ev_in[i]);
    }
    pool_destroy(plan->pool);
    if (plan->copy_stream)
        bdx_stream_destroy(plan->copy_stream);
    if (plan->stream)
        bdx_stream_destroy(plan->stream);
    free(plan);
}

bridct_plan *bridct_plan_create(int height, int width)
{
    if (!bridct_supported(height, width) || bdx_device_ready() != 0)
        return NULL;
    bridct_plan *plan = calloc(1, sizeof(*plan));
    if (!plan)
        return NULL;
    plan->height = height;
    plan->width = width;
    plan->elements = (size_t)height * width;
    plan->chunk = STAGE_FLOATS / plan->elements;
    if (plan->chunk < 1)
        plan->chunk = 1;
    if (bdx_current_device(&plan->device))
        goto fail;
    float *host = malloc(BDX_TABLE_FLOATS * sizeof(float));
    if (!host)
        goto fail;
    bridct_build_tables(host);
    const size_t stage = plan->chunk * plan->elements * sizeof(float);
    if (bdx_stream_create(&plan->stream) ||
        bdx_malloc((void **)&plan->d_tables, BDX_TABLE_FLOATS * sizeof(float)) ||
        bdx_malloc((void **)&plan->d_in, stage) || bdx_malloc((void **)&plan->d_out, stage) ||
        bdx_copy_h2d(plan->d_tables, host, BDX_TABLE_FLOATS * sizeof(float), plan->stream) ||
        bdx_stream_sync(plan->stream)) {
        free(host);
        goto fail;
    }
    free(host);
    /* per-shape route and packing chosen by the autotuner, when a table is present; a
     * route this plan cannot take (e.g. a table written for other shapes) is ignored */
    int route = BRIDCT_ROUTE_BUTTERFLY, pack = 0;
    if (bridct_schedule_lookup(height, width, &route, &pack) == 0) {
        if (bridct_plan_set_pack(plan, pack) != BRIDCT_OK)
            plan->pack = 0;
        if (route != BRIDCT_ROUTE_BUTTERFLY && bridct_plan_set_route(plan, route) != BRIDCT_OK)
            plan->route = BRIDCT_ROUTE_BUTTERFLY;
    }
    return plan;

fail:
    bridct_plan_destroy(plan);
    return NULL;
}

/* src/plan.c:117-124 without the x86-only AVX2 special case: the categories are
 * historical shape classes (docs/API.md:23). */
const char *bridct_plan_route(const bridct_plan *plan)
{
    if (!plan)
        return NULL;
    return plan->height == plan->width && plan->height <= 256 ? "paper8-square" : "blocked-extension";
}

static int distance_below(const void *a, const void *b, size_t bytes)
{
    const uintptr_t x = (uintptr_t)a, y = (uintptr_t)b;
    return (x <= y ? y - x : x - y) < bytes;
}

/* Power of two near sqrt(n) (2^floor(log2 sqrt n)): the dense route scales Q_n by it so
 * that the lo parts of the split matrix entries stay normal FP16 numbers. */
int bridct_dense_scale(int n)
{
    int s = 1;
    while ((long)(2 * s) * (2 * s) <= n)
        s *= 2;
    return s;
}

/* Q_n'[k][j] = s_n sqrt((k ? 2 : 1) / n) cos(pi k (2j + 1) / (2n)) in FP64, rounded once;
 * the cosine argument k (2j + 1) is reduced modulo 4n first.  Writes Q' and, when qt is
 * not NULL, its transpose. */
void bridct_dense_matrix(int n, float *q, float *qt)
{
    const double s = bridct_dense_scale(n);
    for (int k = 0; k < n; k++) {
        const double f = s * sqrt((k ? 2.0 : 1.0) / n);
        for (int j = 0; j < n; j++) {
            const long m = ((long)k * (2 * j + 1)) % (4L * n);
            const float v = (float)(f * cos(PI * (double)m / (2.0 * n)));
            q[(size_t)k * n + j] = v;
            if (qt)
                qt[(size_t)j * n + k] = v;
        }
    }
}

int bridct_plan_set_route(bridct_plan *plan, int route)
{
    if (!plan || route < BRIDCT_ROUTE_BUTTERFLY || route > BRIDCT_ROUTE_MFMA_BF16)
        return BRIDCT_INVALID_ARGUMENT;
    if (route != BRIDCT_ROUTE_BUTTERFLY && (plan->height < 64 || plan->width < 64))
        return BRIDCT_INVALID_ARGUMENT;
    if (route != BRIDCT_ROUTE_BUTTERFLY && !plan->d_dense) {
        const size_t hh = (size_t)plan->height * plan->height, ww = (size_t)plan->width * plan->width;
        float *host = malloc((2 * hh + ww) * sizeof(float));
        if (!host)
            return BRIDCT_DEVICE_ERROR;
        bridct_dense_matrix(plan->height, host, host + hh);
        bridct_dense_matrix(plan->width, host + 2 * hh, NULL);
        if (bdx_malloc((void **)&plan->d_dense, (2 * hh + ww) * sizeof(float)) ||
            bdx_copy_h2d(plan->d_dense, host, (2 * hh + ww) * sizeof(float), plan->stream) ||
            bdx_stream_sync(plan->stream)) {
            free(host);
            if (plan->d_dense)
                bdx_free(plan->d_dense);
            plan->d_dense = NULL;
            return BRIDCT_DEVICE_ERROR;
        }
        free(host);
        plan->dense_scale = 1.0f / (float)(bridct_dense_scale(plan->height) * bridct_dense_scale(plan->width));
    }
    plan->route = route;
    return BRIDCT_OK;
}

int bridct_plan_get_route(const bridct_plan *plan)
{
    return plan ? plan->route : BRIDCT_INVALID_ARGUMENT;
}

int bridct_plan_set_pack(bridct_plan *plan, int pack)
{
    if (!plan || pack < BRIDCT_PACK_TWO_PASS || pack > 4096 ||
        (pack == BRIDCT_PACK_TWO_PASS && !BDX_TWO_PASS_OK(plan->height, plan->width)))
        return BRIDCT_INVALID_ARGUMENT;
    plan->pack = pack;
    return BRIDCT_OK;
}

int bridct_plan_get_pack(const bridct_plan *plan)
{
    return plan ? plan->pack : BRIDCT_INVALID_ARGUMENT;
}

/* Dense route: FORWARD / INVERSE through `work` (count * elements floats), or through the
 * plan's own workspace (grown on demand) when work is NULL. */
static int dense(bridct_plan *plan, int mode, const float *d_in, float *d_out, size_t count, float *work)
{
    const size_t max_batch = (size_t)65535 * 64 / (size_t)plan->height < 65535
                                 ? (size_t)65535 * 64 / (size_t)plan->height : 65535;
    const size_t slice = count < max_batch ? count : max_batch;
    const size_t need = BDX_DENSE_WORK_FLOATS(slice, plan->height, plan->width);
    if (!work && need > plan->work_floats) {
        if (plan->d_work && (bdx_stream_sync(plan->stream) || bdx_free(plan->d_work)))
            return -1;
        plan->d_work = NULL;
        plan->work_floats = 0;
        if (bdx_malloc((void **)&plan->d_work, need * sizeof(float)))
            return -1;
        plan->work_floats = need;
    }
    if (!work)
        work = plan->d_work;
    const size_t hh = (size_t)plan->height * plan->height;
    const int prec = plan->route == BRIDCT_ROUTE_MFMA_F16 ? BDX_DENSE_F16 : BDX_DENSE_BF16;
    /* batches beyond one launch's grid limits go in slices */
    for (size_t done = 0; done < count;) {
        const size_t n = count - done < max_batch ? count - done : max_batch;
        const size_t off = done * plan->elements;
        if (bdx_dense_transform(d_in + off, d_out + off, work, plan->height, plan->width, n, mode, prec,
                                plan->d_dense, plan->d_dense + hh, plan->d_dense + 2 * hh, plan->dense_scale,
                                plan->stream))
            return -1;
        done += n;
    }
    return 0;
}

/* `work`: the dense route's intermediate (NULL: the plan's own). */
static int transform_mode(bridct_plan *plan, int mode, const float *d_in, float *d_out, size_t count, float *work)
{
    if (plan->route == BRIDCT_ROUTE_BUTTERFLY)
        return bdx_transform(d_in, d_out, plan->height, plan->width, count, plan->pack, mode, plan->d_tables,
                             plan->stream);
    if (mode == BDX_ROUNDTRIP)
        return dense(plan, BDX_FORWARD, d_in, d_out, count, work) ||
               dense(plan, BDX_INVERSE, d_out, d_out, count, work);
    return dense(plan, mode, d_in, d_out, count, work);
}

static int transform_with(bridct_plan *plan, int mode, const float *d_in, float *d_out, size_t count, float *work)
{
    if (plan->route != BRIDCT_ROUTE_BUTTERFLY) {
        if (mode == BRIDCT_GPU_FORWARD_THEN_INVERSE)
            return transform_mode(plan, BDX_FORWARD, d_in, d_out, count, work) ||
                   transform_mode(plan, BDX_INVERSE, d_out, d_out, count, work);
        return transform_mode(plan, mode, d_in, d_out, count, work);
    }
    if (mode == BRIDCT_GPU_FORWARD_THEN_INVERSE)
        return bdx_transform(d_in, d_out, plan->height, plan->width, count, plan->pack, BDX_FORWARD,
                             plan->d_tables, plan->stream) ||
               bdx_transform(d_out, d_out, plan->height, plan->width, count, plan->pack, BDX_INVERSE,
                             plan->d_tables, plan->stream);
    return bdx_transform(d_in, d_out, plan->height, plan->width, count, plan->pack, mode, plan->d_tables,
                         plan->stream);
}

static int transform(bridct_plan *plan, int mode, const float *d_in, float *d_out, size_t count)
{
    return transform_with(plan, mode, d_in, d_out, count, NULL);
}

static int pipeline_ready(bridct_plan *plan)
{
    if (plan->pool)
        return 0;
    const size_t bytes = plan->chunk * plan->elements * sizeof(float);
    for (int i = 0; i < 2; i++)
        if ((!plan->pin_in[i] && bdx_host_malloc((void **)&plan->pin_in[i], bytes)) ||
            (!plan->pin_out[i] && bdx_host_malloc((void **)&plan->pin_out[i], bytes)) ||
            (!plan->ev[i] && bdx_event_create(&plan->ev[i])) ||
            (!plan->ev_in[i] && bdx_event_create(&plan->ev_in[i])))
            return -1;
    if ((!plan->d_in2 && bdx_malloc((void **)&plan->d_in2, bytes)) ||
        (!plan->d_out2 && bdx_malloc((void **)&plan->d_out2, bytes)) ||
        (!plan->copy_stream && bdx_stream_create(&plan->copy_stream)))
        return -1;
    plan->pool = pool_create();   /* the last step: its presence marks the pipeline ready */
    return plan->pool ? 0 : -1;
}

/* Large host calls: chunks alternate between two slots (see PIPELINE_MIN_FLOATS). */
static int run_host_pipelined(bridct_plan *plan, int mode, const float *input, float *output, size_t batch)
{
    const size_t per = plan->chunk, nchunks = (batch + per - 1) / per;
    float *d_in[2] = {plan->d_in, plan->d_in2}, *d_out[2] = {plan->d_out, plan->d_out2};
    size_t floats_of[2] = {0, 0};
    int rc = BRIDCT_OK;
    for (size_t k = 0; k < nchunks + 2; k++) {
        const int s = (int)(k & 1);
        copy_job jobs[2];
        int njobs = 0;
        if (k >= 2) {   /* chunk k - 2 finished in slot s (copied in, transformed, copied out) */
            if (bdx_event_sync(plan->ev[s])) {
                rc = BRIDCT_DEVICE_ERROR;
                break;
            }
            jobs[njobs++] = (copy_job){output + (k - 2) * per * plan->elements, plan->pin_out[s],
                                       floats_of[s] * sizeof(float)};
        }
        size_t count = 0;
        if (k < nchunks) {
            count = batch - k * per < per ? batch - k * per : per;
            floats_of[s] = count * plan->elements;
            jobs[njobs++] = (copy_job){plan->pin_in[s], input + k * per * plan->elements,
                                       floats_of[s] * sizeof(float)};
        }
        pool_copy(plan->pool, jobs, njobs);
        if (k < nchunks) {
            const size_t bytes = floats_of[s] * sizeof(float);
            if (bdx_copy_h2d(d_in[s], plan->pin_in[s], bytes, plan->copy_stream) ||
                bdx_event_record(plan->ev_in[s], plan->copy_stream) ||
                bdx_stream_wait_event(plan->stream, plan->ev_in[s]) ||
                transform(plan, mode, d_in[s], d_out[s], count) ||
                bdx_copy_d2h(plan->pin_out[s], d_out[s], bytes, plan->stream) ||
                bdx_event_record(plan->ev[s], plan->stream)) {
                rc = BRIDCT_DEVICE_ERROR;
                break;
            }
        }
    }
    if (bdx_stream_sync(plan->copy_stream) || bdx_stream_sync(plan->stream))
        rc = BRIDCT_DEVICE_ERROR;
    return rc;
}

static int run_host(bridct_plan *plan, int mode, const float *input, float *output, size_t batch)
{
    if (batch * plan->elements >= PIPELINE_MIN_FLOATS && pipeline_ready(plan) == 0)
        return run_host_pipelined(plan, mode, input, output, batch);
    for (size_t done = 0; done < batch;) {
        size_t count = batch - done;
        if (count > plan->chunk)
            count = plan->chunk;
        const size_t floats = count * plan->elements;
        if (bdx_copy_h2d(plan->d_in, input + done * plan->elements, floats * sizeof(float),
                         plan->stream) ||
            transform(plan, mode, plan->d_in, plan->d_out, count) ||
            bdx_copy_d2h(output + done * plan->elements, plan->d_out, floats * sizeof(float),
                         plan->stream) ||
            bdx_stream_sync(plan->stream))
            return BRIDCT_DEVICE_ERROR;
        done += count;
    }
    return BRIDCT_OK;
}

/* src/plan.c:126-135: argument checks, then the overlap rule (distance below h*w*4). */
int bridct_plan_apply(bridct_plan *plan, enum bridct_mode mode, const float *input, float *output)
{
    if (!plan || !input || !output || (unsigned)mode > BRIDCT_ROUNDTRIP)
        return BRIDCT_INVALID_ARGUMENT;
    if (distance_below(input, output, plan->elements * sizeof(float)))
        return BRIDCT_OVERLAPPING_BUFFERS;
    return run_host(plan, (int)mode, input, output, 1);
}

int bridct_plan_apply_batch(bridct_plan *plan, int mode, const float *input, float *output,
                            size_t batch)
{
    if (!plan || !input || !output || (unsigned)mode > BRIDCT_GPU_FORWARD_THEN_INVERSE || batch < 1 ||
        batch > SIZE_MAX / sizeof(float) / plan->elements)
        return BRIDCT_INVALID_ARGUMENT;
    if (distance_below(input, output, batch * plan->elements * sizeof(float)))
        return BRIDCT_OVERLAPPING_BUFFERS;
    return run_host(plan, mode, input, output, batch);
}

int bridct_plan_apply_device(bridct_plan *plan, enum bridct_mode mode, const float *d_input,
                             float *d_output, size_t batch)
{
    if (!plan || !d_input || !d_output || (unsigned)mode > BRIDCT_ROUNDTRIP ||
        batch > SIZE_MAX / sizeof(float) / plan->elements)
        return BRIDCT_INVALID_ARGUMENT;
    if (((uintptr_t)d_input | (uintptr_t)d_output) % 16)
        return BRIDCT_BAD_ALIGNMENT;
    if (d_input != d_output &&
        distance_below(d_input, d_output, batch * plan->elements * sizeof(float)))
        return BRIDCT_OVERLAPPING_BUFFERS;
    if (transform(plan, (int)mode, d_input, d_output, batch))
        return BRIDCT_DEVICE_ERROR;
    return BRIDCT_OK;
}

/* The same call with nothing allocated: a dense-route plan takes its intermediate from the
 * library's static per-GPU scratch array instead of growing its own buffer. */
int bridct_plan_apply_device_static(bridct_plan *plan, enum bridct_mode mode, const float *d_input,
                                    float *d_output, size_t batch)
{
    if (!plan || !d_input || !d_output || (unsigned)mode > BRIDCT_ROUNDTRIP ||
        batch > SIZE_MAX / sizeof(float) / plan->elements)
        return BRIDCT_INVALID_ARGUMENT;
    if (((uintptr_t)d_input | (uintptr_t)d_output) % 16)
        return BRIDCT_BAD_ALIGNMENT;
    if (d_input != d_output &&
        distance_below(d_input, d_output, batch * plan->elements * sizeof(float)))
        return BRIDCT_OVERLAPPING_BUFFERS;
    float *work = NULL;
    if (plan->route != BRIDCT_ROUTE_BUTTERFLY) {
        size_t floats = 0;
        if (bdx_dense_static_work(plan->device, &work, &floats))
            return BRIDCT_DEVICE_ERROR;
        if (batch * plan->elements > BDX_STATIC_ARRAY_FLOATS ||
            BDX_DENSE_WORK_FLOATS(batch, plan->height, plan->width) > floats)
            return BRIDCT_WORKSPACE_TOO_SMALL;
    }
    if (transform_with(plan, (int)mode, d_input, d_output, batch, work))
        return BRIDCT_DEVICE_ERROR;
    return BRIDCT_OK;
}

int bridct_plan_apply_device_pass(bridct_plan *plan, enum bridct_mode mode, int pass, const float *d_input,
                                  float *d_output, size_t batch)
{
    if (!plan || !d_input || !d_output || (mode != BRIDCT_FORWARD && mode != BRIDCT_INVERSE) ||
        (pass != 0 && pass != 1) || (pass == 1 && d_input != d_output) ||
        batch > SIZE_MAX / sizeof(float) / plan->elements)
        return BRIDCT_INVALID_ARGUMENT;
    if (((uintptr_t)d_input | (uintptr_t)d_output) % 16)
        return BRIDCT_BAD_ALIGNMENT;
    if (d_input != d_output &&
        distance_below(d_input, d_output, batch * plan->elements * sizeof(float)))
        return BRIDCT_OVERLAPPING_BUFFERS;
    if (bdx_transform_pass(d_input, d_output, plan->height, plan->width, batch, plan->pack, (int)mode, pass,
                           plan->d_tables, plan->stream))
        return BRIDCT_DEVICE_ERROR;
    return BRIDCT_OK;
}

int bridct_plan_synchronize(bridct_plan *plan)
{
    if (!plan)
        return BRIDCT_INVALID_ARGUMENT;
    return bdx_stream_sync(plan->stream) ? BRIDCT_DEVICE_ERROR : BRIDCT_OK;
}

void *bridct_plan_stream(const bridct_plan *plan)
{
    return plan ? plan->stream : NULL;
}

/* src/bridct.c:13-21: the caller-workspace size of the legacy square API, kept for
 * callers that size buffers with it. */
size_t bridct_workspace_floats(int n)
{
    if (n != 8 && n != 16 && n != 32 && n != 64 && n != 128 && n != 256)
        return 0;
    const size_t e = (size_t)n * n;
    return n <= 16 ? 2 * e : 8 * e + 2 * 1100 * 4 + 4 * 4 * 256 + 4 * (size_t)n * (4 * n + 4);
}

static int overlaps(const void *a, size_t bytes_a, const void *b, size_t bytes_b)
{
    const uintptr_t x = (uintptr_t)a, y = (uintptr_t)b;
    return x <= y ? y - x < bytes_a : x - y < bytes_b;
}

/* GPU state of the legacy square API: one plan per size and calling thread, created by the
 * thread's first call of that size and destroyed when the thread exits (a thread that is
 * still running at process exit keeps its plans until then).  Per thread, so concurrent
 * callers never share a plan, as with the reference's caller-owned workspaces; the plan's
 * tables and staging buffers are built once instead of on every call. */
static pthread_key_t legacy_key;
static pthread_once_t legacy_once = PTHREAD_ONCE_INIT;
static int legacy_key_ok;

typedef struct {
    bridct_plan *plan[6];   /* n = 8 << i */
} legacy_plans;

static void legacy_release(void *arg)
{
    legacy_plans *l = arg;
    for (int i = 0; i < 6; i++)
        bridct_plan_destroy(l->plan[i]);
    free(l);
}

static void legacy_key_create(void)
{
    legacy_key_ok = pthread_key_create(&legacy_key, legacy_release) == 0;
}

static bridct_plan *legacy_plan(int n)
{
    pthread_once(&legacy_once, legacy_key_create);
    if (!legacy_key_ok)
        return NULL;
    legacy_plans *l = pthread_getspecific(legacy_key);
    if (!l) {
        if (!(l = calloc(1, sizeof(*l))))
            return NULL;
        if (pthread_setspecific(legacy_key, l)) {
            free(l);
            return NULL;
        }
    }
    int i = 0;
    while ((8 << i) < n)
        i++;
    if (!l->plan[i])
        l->plan[i] = bridct_plan_create(n, n);
    return l->plan[i];
}

/* src/bridct.c:30-45: legacy square transform with caller workspace.  The checks and
 * their precedence are the reference's; the workspace itself is not needed on the GPU. */
int bridct_apply(int n, enum bridct_mode mode, const float *input, float *output,
                 float *workspace, size_t workspace_count)
{
    const size_t needed = bridct_workspace_floats(n);
    if (!needed || !input || !output || !workspace || (unsigned)mode > BRIDCT_ROUNDTRIP)
        return BRIDCT_INVALID_ARGUMENT;
    if (workspace_count < needed)
        return BRIDCT_WORKSPACE_TOO_SMALL;
    if ((uintptr_t)workspace % 16)
        return BRIDCT_BAD_ALIGNMENT;
    const size_t bytes = (size_t)n * n * sizeof(float);
    if (overlaps(input, bytes, output, bytes) ||
        overlaps(input, bytes, workspace, needed * sizeof(float)) ||
        overlaps(output, bytes, workspace, needed * sizeof(float)))
        return BRIDCT_OVERLAPPING_BUFFERS;
    bridct_plan *plan = legacy_plan(n);
    if (!plan)
        return BRIDCT_DEVICE_ERROR;
    return run_host(plan, (int)mode, input, output, 1);
}

int bridct_gpu_available(void)
{
    return bdx_device_ready() == 0;
}

long bridct_debug_alloc_faults(long after)
{
    return bdx_debug_alloc_faults(after);
}

long bridct_debug_grid_limit(long blocks)
{
    return bdx_debug_grid_limit(blocks);
}

void *bridct_gpu_malloc(size_t bytes)
{
    void *p = NULL;
    if (bdx_device_ready() != 0 || bdx_malloc(&p, bytes ? bytes : 1))
        return NULL;
    return p;
}

int bridct_gpu_free(void *d_ptr)
{
    if (!d_ptr)
        return BRIDCT_OK;
    return bdx_free(d_ptr) ? BRIDCT_DEVICE_ERROR : BRIDCT_OK;
}

int bridct_gpu_upload(void *d_dst, const void *src, size_t bytes)
{
    if (!d_dst || !src)
        return BRIDCT_INVALID_ARGUMENT;
    if (bdx_copy_h2d(d_dst, src, bytes, NULL) || bdx_stream_sync(NULL))
        return BRIDCT_DEVICE_ERROR;
    return BRIDCT_OK;
}

int bridct_gpu_download(void *dst, const void *d_src, size_t bytes)
{
    if (!dst || !d_src)
        return BRIDCT_INVALID_ARGUMENT;
    if (bdx_copy_d2h(dst, d_src, bytes, NULL) || bdx_stream_sync(NULL))
        return BRIDCT_DEVICE_ERROR;
    return BRIDCT_OK;
}
