/* Ragged batches: arrays of mixed supported shapes transformed by one persistent launch
 * (C11 host side of bdx_batch.hip).
 *
 * The reference applies one plan per array shape (src/plan.c:126-163; callers loop over
 * arrays, e.g. python/bridct_numpy.c:96-112).  A batch object takes the whole list of
 * shapes once, lays the arrays out back to back, and turns every array into work items
 * of the persistent kernel:
 *   - arrays with both axes at most 128: SMALL items of up to BDX_BATCH_PER(h, w) whole
 *     arrays of one run (a run = consecutive arrays of one shape);
 *   - larger arrays: ROW items (bands of rows, input -> output), COL items (strips of
 *     columns, in place, each waiting for all ROW items of its array) and, for the fused
 *     round trip, ROW2 items (inverse rows, waiting for all COL items).
 * Queues: the work is cut into units, each two-pass array (by decreasing size) with all its
 * items, then each SMALL item.  The persistent kernel keeps one queue per XCD: the
 * workgroups of an XCD draw tickets from it, and slot j of the queue is the ROW items of the
 * unit the queue takes j-th (from a global unit counter), then the COL items of its
 * previous unit and, for the round trip, the ROW2 items of the one before.  All items of an
 * array thus run on one XCD, whose L2 holds the row pass's output for the column pass, and
 * a dependency only waits for items of earlier tickets of the same queue, taken by running
 * workgroups: the schedule cannot deadlock.
 *
 * A second strategy, GROUPED, runs the arrays of each distinct shape with the plan path's
 * own per-shape kernels in gather form (the arrays of a shape are not contiguous in the
 * batch buffer, so the kernels read an offset table): one or two launches per shape,
 * largest shapes first, capturable as one graph.  Each launch
 * is specialized for its shape (registers, LDS, team size) and needs no cross-workgroup
 * dependency waits, which makes it faster when every shape has enough pixels to fill the
 * GPU; the automatic choice takes it when the average shape has at least
 * GROUPED_MIN_PIXELS pixels.  Batches of tiny arrays always use the tiny kernel.  The
 * large two-pass shapes run one after the other on the batch stream; the whole-array
 * shapes and the small two-pass shapes, whose launches are short and leave most of the GPU
 * idle, go on two auxiliary streams (forked from and joined back to the batch stream with
 * events, so a graph capture records them all) and run beside them.  GROUPED also follows
 * the schedule table's route per shape: a shape the table sends to a split-MFMA route
 * (both axes >= 64) runs the dense GEMMs in gather form, so the automatic choice takes
 * GROUPED whenever such a shape is present (the persistent kernel runs butterflies only).
 *
 * The object owns its stream, device tables, item lists and counters, the grouped offset
 * table, and one captured graph (hipGraph) for replays of the same (mode, input, output);
 * it has no global state.
 */
#include "bridct_gpu.h"

#include "bridct_device.h"
#include "host_internal.h"
#include "host_stage.h"

#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Host-pointer calls of at least this many bytes are staged through pinned memory, in
 * chunks of STAGE_CHUNK_BYTES. */
#define STAGE_MIN_BYTES ((size_t)1 << 20)
#define STAGE_CHUNK_BYTES ((size_t)8 << 20)

/* Dependency waits give up (and flag an error) after this many sleeps of the waiting
 * workgroup, about a second: a legitimate wait is one array's row pass. */
#define SPIN_LIMIT (1u << 21)

/* Automatic strategy: GROUPED when the batch averages at least this many pixels per
 * distinct shape (each launch then has several microseconds of work, so launch gaps stay
 * small against it). */
#define GROUPED_MIN_PIXELS ((size_t)1 << 21)

/* GROUPED: auxiliary streams beside the batch stream, and the pixel count below which a
 * two-pass shape goes on one of them: its launches (a few hundred workgroups, 10-40 us)
 * do not fill the GPU.  Measured on the 10,000-array mix (per direction): all two-pass
 * shapes on the batch stream and one auxiliary stream for the whole-array shapes
 * 2.16-2.18 ms; two auxiliary streams with two-pass shapes below 24 or 32 Mpx 1.93-1.96;
 * below 16 Mpx 2.05; below 48 or 64 Mpx 1.98-2.01; three streams 1.94-2.01; one stream
 * with the small two-pass shapes 2.18. */
#define GROUPED_AUX_STREAMS 2
#define GROUPED_SMALL_TWO_PASS ((size_t)32 << 20)

/* One distinct shape of a batch: its arrays' offsets are d_goffs[first, first + count).
 * Dense shapes: Q_h', Q_h'^T, Q_w' at d_dense + qoff, and the exact scale 1 / (s_h s_w). */
typedef struct {
    int h, w, pack;
    int route;               /* BRIDCT_ROUTE_* from the schedule table */
    size_t count, first;
    size_t qoff;
    float dscale;
} group;

/* Arrays per launch of the dense GEMMs of an h-row shape (grid limits). */
static size_t dense_slice(int h)
{
    const size_t by_rows = (size_t)65535 * 64 / (size_t)h;
    return by_rows < 65535 ? by_rows : 65535;
}

typedef struct {
    bdx_item *items;
    unsigned n, cap;
} item_list;

struct bridct_batch {
    size_t count, elements;
    unsigned spin_limit;     /* dependency-wait bound of the persistent kernel */
    size_t *offsets;         /* element offset of each array */
    unsigned nruns, ntwo;    /* runs, two-pass arrays */
    unsigned blocks;
    int tiny;                /* all arrays <= BDX_TINY_MAX on both axes: tiny kernel */
    int strategy;            /* BRIDCT_BATCH_PERSISTENT or BRIDCT_BATCH_GROUPED */
    int auto_strategy;       /* what BRIDCT_BATCH_AUTO resolves to */
    group *groups;           /* distinct shapes, largest total first */
    unsigned ngroups;
    size_t *d_goffs;         /* element offsets of the arrays, grouped by shape */
    float *d_dense;          /* matrices of the dense shapes (GROUPED) */
    float *d_dwork;          /* their intermediate, one slice of the largest dense shape */
    int ndense;
    void *aux[GROUPED_AUX_STREAMS];       /* whole-array and small shapes of the grouped strategy */
    void *ev_fork, *ev_join[GROUPED_AUX_STREAMS];
    host_stage *stage;       /* pinned staging of large host-pointer calls, on first use */
    void *stream;
    float *d_tables;
    bdx_run *d_runs;
    bdx_item *d_items[2];    /* [0] forward / inverse order, [1] fused round trip */
    unsigned nitems[2];
    bdx_unit *d_units;       /* the persistent kernel's units */
    unsigned nunits, qbase;
    unsigned *d_ctr;
    unsigned nctr;
    float *d_in, *d_out;     /* staging for host-pointer calls, allocated on first use */
    void *graph;             /* captured exec and the call it replays */
    int graph_mode;
    const float *graph_in;
    float *graph_out;
    /* arguments of the call being captured */
    int cap_mode;
    const float *cap_in;
    float *cap_out;
};

static int push(item_list *l, unsigned run, unsigned idx, unsigned unit, unsigned kind, unsigned ctr)
{
    if (l->n == l->cap) {
        const unsigned cap = l->cap ? 2 * l->cap : 1024;
        bdx_item *p = realloc(l->items, (size_t)cap * sizeof(*p));
        if (!p)
            return -1;
        l->items = p;
        l->cap = cap;
    }
    bdx_item *it = &l->items[l->n++];
    it->run = run;
    it->idx = idx;
    it->unit = unit;
    it->kind_ctr = kind << 28 | ctr;
    return 0;
}

typedef struct {
    unsigned run, idx;
    int h, w;
} two_pass;

static int by_size_desc(const void *a, const void *b)
{
    const two_pass *x = a, *y = b;
    const long sx = (long)x->h * x->w, sy = (long)y->h * y->w;
    if (sx != sy)
        return sx < sy ? 1 : -1;
    /* stable: keep batch order among arrays of one size */
    if (x->run != y->run)
        return x->run < y->run ? -1 : 1;
    return x->idx < y->idx ? -1 : x->idx > y->idx;
}

/* Builds the units and the item lists of both mode families ([0] forward / inverse, [1] round
 * trip): each two-pass array (tp, by decreasing size) is a unit of its ROW items, its COL
 * items and, in the round-trip list, its ROW2 items; then each SMALL item (up to
 * BDX_BATCH_PER whole arrays of one run, BDX_TINY_PER for tiny batches) is a unit. */
static int build_units(bdx_schedule *s, const two_pass *tp, item_list lists[2])
{
    unsigned nsmall = 0;
    for (unsigned r = 0; r < s->nruns; r++) {
        const int h = (int)s->runs[r].h, w = (int)s->runs[r].w;
        if (BDX_BATCH_SMALL(h, w)) {
            const unsigned per = (unsigned)(s->tiny ? BDX_TINY_PER(h, w) : BDX_BATCH_PER(h, w));
            nsmall += (s->runs[r].count + per - 1) / per;
        }
    }
    s->units = malloc((size_t)(s->ntwo + nsmall) * sizeof(bdx_unit));
    if (!s->units)
        return -1;
    int rc = 0;
    for (unsigned k = 0; k < s->ntwo && !rc; k++) {
        bdx_unit *u = &s->units[s->nunits++];
        const unsigned nrow = (unsigned)BDX_CEIL(tp[k].h, BDX_ROW_UNIT(tp[k].w));
        const unsigned ncol = (unsigned)BDX_CEIL(tp[k].w, BDX_COL_UNIT(tp[k].h));
        u->nrow = (unsigned short)nrow;
        u->ncol = (unsigned short)ncol;
        u->nrow2 = (unsigned short)nrow;
        u->pad = 0;
        for (int f = 0; f < 2 && !rc; f++) {
            u->item0[f] = lists[f].n;
            for (unsigned i = 0; i < nrow && !rc; i++)
                rc = push(&lists[f], tp[k].run, tp[k].idx, i, BDX_ITEM_ROW, k);
            for (unsigned i = 0; i < ncol && !rc; i++)
                rc = push(&lists[f], tp[k].run, tp[k].idx, i, BDX_ITEM_COL, k);
            for (unsigned i = 0; i < nrow && f == 1 && !rc; i++)
                rc = push(&lists[f], tp[k].run, tp[k].idx, i, BDX_ITEM_ROW2, k);
        }
    }
    for (unsigned r = 0; r < s->nruns && !rc; r++) {
        const int h = (int)s->runs[r].h, w = (int)s->runs[r].w;
        if (!BDX_BATCH_SMALL(h, w))
            continue;
        const unsigned per = (unsigned)(s->tiny ? BDX_TINY_PER(h, w) : BDX_BATCH_PER(h, w)), n = s->runs[r].count;
        const unsigned parts = (n + per - 1) / per;
        for (unsigned p = 0, first = 0; p < parts && !rc; p++) {
            const unsigned take = (n - first) / (parts - p) + ((n - first) % (parts - p) != 0);
            bdx_unit *u = &s->units[s->nunits++];
            u->nrow = 1;
            u->ncol = 0;
            u->nrow2 = 0;
            u->pad = 0;
            for (int f = 0; f < 2 && !rc; f++) {
                u->item0[f] = lists[f].n;
                rc = push(&lists[f], r, first, take, BDX_ITEM_SMALL, 0);
            }
            first += take;
        }
    }
    return rc;
}

void bridct_batch_destroy(bridct_batch *b)
{
    if (!b)
        return;
    /* both streams drained before any buffer they may still read is freed */
    if (b->stream)
        bdx_stream_sync(b->stream);
    for (int k = 0; k < GROUPED_AUX_STREAMS; k++)
        if (b->aux[k])
            bdx_stream_sync(b->aux[k]);
    if (b->graph)
        bdx_graph_destroy(b->graph);
    stage_destroy(b->stage);
    void *bufs[] = {b->d_tables, b->d_runs, b->d_items[0], b->d_items[1], b->d_units, b->d_ctr,
                    b->d_in,     b->d_out,  b->d_goffs,    b->d_dense,    b->d_dwork};
    for (size_t i = 0; i < sizeof(bufs) / sizeof(bufs[0]); i++)
        if (bufs[i])
            bdx_free(bufs[i]);
    for (int k = 0; k < GROUPED_AUX_STREAMS; k++) {
        if (b->aux[k])
            bdx_stream_destroy(b->aux[k]);
        if (b->ev_join[k])
            bdx_event_destroy(b->ev_join[k]);
    }
    if (b->ev_fork)
        bdx_event_destroy(b->ev_fork);
    if (b->stream)
        bdx_stream_destroy(b->stream);
    free(b->offsets);
    free(b->groups);
    free(b);
}

typedef struct {
    long key;     /* h * 2048 + w */
    size_t idx;   /* array index */
} shape_key;

static int by_shape(const void *pa, const void *pb)
{
    const shape_key *a = pa, *b = pb;
    return a->key < b->key ? -1 : a->key > b->key ? 1 : a->idx < b->idx ? -1 : a->idx > b->idx;
}

static int by_pixels_desc(const void *pa, const void *pb)
{
    const group *a = pa, *b = pb;
    const size_t na = (size_t)a->h * a->w * a->count, nb = (size_t)b->h * b->w * b->count;
    return na > nb ? -1 : na < nb ? 1 : 0;
}

/* Groups the arrays by shape; host offsets into `goffs` (count entries). */
static int build_groups(bridct_batch *b, const int *heights, const int *widths, size_t *goffs)
{
    shape_key *idx = malloc(b->count * sizeof(shape_key));
    if (!idx)
        return -1;
    for (size_t i = 0; i < b->count; i++) {
        idx[i].key = (long)heights[i] * 2048 + widths[i];
        idx[i].idx = i;
    }
    qsort(idx, b->count, sizeof(shape_key), by_shape);
    unsigned n = 0;
    for (size_t i = 0; i < b->count; i++)
        if (i == 0 || idx[i].key != idx[i - 1].key)
            n++;
    b->groups = calloc(n, sizeof(group));
    if (!b->groups) {
        free(idx);
        return -1;
    }
    for (size_t i = 0; i < b->count; i++) {
        const size_t a = idx[i].idx;
        goffs[i] = b->offsets[a];
        if (i == 0 || idx[i].key != idx[i - 1].key) {
            group *g = &b->groups[b->ngroups++];
            g->h = heights[a];
            g->w = widths[a];
            g->first = i;
            int route = BRIDCT_ROUTE_BUTTERFLY, pack = 0;
            if (bridct_schedule_lookup(g->h, g->w, &route, &pack) != 0)
                route = BRIDCT_ROUTE_BUTTERFLY, pack = 0;
            if (route != BRIDCT_ROUTE_BUTTERFLY && (g->h < 64 || g->w < 64))
                route = BRIDCT_ROUTE_BUTTERFLY;   /* the dense route needs both axes >= 64 */
            if (pack == BDX_PACK_TWO_PASS && !BDX_TWO_PASS_OK(g->h, g->w))
                pack = 0;                         /* a shape the two-pass kernels do not take */
            g->route = route;
            g->pack = pack;
        }
        b->groups[b->ngroups - 1].count++;
    }
    free(idx);
    qsort(b->groups, b->ngroups, sizeof(group), by_pixels_desc);
    return 0;
}

/* The grouped strategy's auxiliary streams and their events. */
static int aux_setup(bridct_batch *b)
{
    if (bdx_event_create(&b->ev_fork))
        return -1;
    for (int k = 0; k < GROUPED_AUX_STREAMS; k++)
        if (bdx_stream_create(&b->aux[k]) || bdx_event_create(&b->ev_join[k]))
            return -1;
    return 0;
}

static int upload(void **dst, const void *src, size_t bytes, void *stream)
{
    if (bdx_malloc(dst, bytes ? bytes : 16))
        return -1;
    return bytes ? bdx_copy_h2d(*dst, src, bytes, stream) : 0;
}

/* Matrices and workspace of the dense shapes (route from the schedule table). */
static int dense_setup(bridct_batch *b)
{
    size_t floats = 0, work = 0;
    for (unsigned i = 0; i < b->ngroups; i++) {
        group *g = &b->groups[i];
        if (g->route == BRIDCT_ROUTE_BUTTERFLY)
            continue;
        g->qoff = floats;
        floats += 2 * (size_t)g->h * g->h + (size_t)g->w * g->w;
        g->dscale = 1.0f / (float)(bridct_dense_scale(g->h) * bridct_dense_scale(g->w));
        const size_t n = g->count < dense_slice(g->h) ? g->count : dense_slice(g->h);
        if (BDX_DENSE_WORK_FLOATS(n, g->h, g->w) > work)
            work = BDX_DENSE_WORK_FLOATS(n, g->h, g->w);
    }
    float *host = malloc(floats * sizeof(float));
    if (!host)
        return -1;
    for (unsigned i = 0; i < b->ngroups; i++) {
        const group *g = &b->groups[i];
        if (g->route == BRIDCT_ROUTE_BUTTERFLY)
            continue;
        const size_t hh = (size_t)g->h * g->h;
        bridct_dense_matrix(g->h, host + g->qoff, host + g->qoff + hh);
        bridct_dense_matrix(g->w, host + g->qoff + 2 * hh, NULL);
    }
    /* the copy from pageable memory must be done before `host` is freed */
    const int rc = upload((void **)&b->d_dense, host, floats * sizeof(float), b->stream) ||
                   bdx_stream_sync(b->stream) || bdx_malloc((void **)&b->d_dwork, work * sizeof(float));
    free(host);
    return rc ? -1 : 0;
}

void bdx_schedule_free(bdx_schedule *s)
{
    if (!s)
        return;
    free(s->runs);
    free(s->items[0]);
    free(s->items[1]);
    free(s->units);
    free(s->offsets);
    memset(s, 0, sizeof(*s));
}

int bdx_batch_is_tiny(size_t count, const int *heights, const int *widths)
{
    for (size_t i = 0; i < count; i++)
        if (heights[i] > BDX_TINY_MAX || widths[i] > BDX_TINY_MAX)
            return 0;
    return count > 0;
}

int bdx_batch_schedule(size_t count, const int *heights, const int *widths, bdx_schedule *s)
{
    memset(s, 0, sizeof(*s));
    if (count == 0 || count > UINT32_MAX / 2 || !heights || !widths)
        return -1;
    for (size_t i = 0; i < count; i++)
        if (!bridct_supported(heights[i], widths[i]))
            return -1;
    two_pass *tp = malloc(count * sizeof(two_pass));
    item_list lists[2] = {{0}, {0}};
    s->offsets = malloc(count * sizeof(size_t));
    s->runs = malloc(count * sizeof(bdx_run));
    if (!tp || !s->offsets || !s->runs) {
        free(tp);
        bdx_schedule_free(s);
        return -1;
    }
    /* runs of consecutive equal shapes; two-pass arrays */
    size_t off = 0;
    for (size_t i = 0; i < count; i++) {
        const int h = heights[i], w = widths[i];
        s->offsets[i] = off;
        if (s->nruns && (int)s->runs[s->nruns - 1].h == h && (int)s->runs[s->nruns - 1].w == w) {
            s->runs[s->nruns - 1].count++;
        } else {
            bdx_run *r = &s->runs[s->nruns++];
            r->h = (unsigned)h;
            r->w = (unsigned)w;
            r->count = 1;
            r->pad = 0;
            r->off = off;
        }
        if (!BDX_BATCH_SMALL(h, w)) {
            two_pass *a = &tp[s->ntwo++];
            a->run = s->nruns - 1;
            a->idx = s->runs[s->nruns - 1].count - 1;
            a->h = h;
            a->w = w;
        }
        off += (size_t)h * w;
    }
    s->count = count;
    s->elements = off;
    qsort(tp, s->ntwo, sizeof(two_pass), by_size_desc);
    s->tiny = bdx_batch_is_tiny(count, heights, widths);
    const int rc = build_units(s, tp, lists);
    free(tp);
    s->items[0] = lists[0].items;
    s->items[1] = lists[1].items;
    s->nitems[0] = lists[0].n;
    s->nitems[1] = lists[1].n;
    s->qbase = BDX_QBASE(s->ntwo);
    s->nctr = s->qbase + BDX_QWORDS(s->nunits);
    if (rc) {
        bdx_schedule_free(s);
        return -1;
    }
    return 0;
}

bridct_batch *bridct_batch_create(size_t count, const int *heights, const int *widths)
{
    if (bdx_device_ready() != 0)
        return NULL;
    const int cus = bdx_device_cus();
    if (cus <= 0)
        return NULL;
    bridct_batch *b = calloc(1, sizeof(*b));
    if (!b)
        return NULL;
    /* two 256-lane workgroups per CU (LDS and registers); the tiny kernel launches one
     * workgroup per item instead */
    b->blocks = 2u * (unsigned)cus;
    b->spin_limit = SPIN_LIMIT;
    bdx_schedule sc;
    if (bdx_batch_schedule(count, heights, widths, &sc)) {
        free(b);
        return NULL;
    }
    b->count = sc.count;
    b->elements = sc.elements;
    b->offsets = sc.offsets;
    sc.offsets = NULL;
    b->nruns = sc.nruns;
    b->ntwo = sc.ntwo;
    b->nitems[0] = sc.nitems[0];
    b->nitems[1] = sc.nitems[1];
    b->nctr = sc.nctr;
    b->nunits = sc.nunits;
    b->qbase = sc.qbase;
    b->tiny = sc.tiny;
    float *host_tables = malloc(BDX_TABLE_FLOATS * sizeof(float));
    unsigned *zeros = calloc(b->nctr, sizeof(unsigned));   /* counters start at zero */
    size_t *goffs = malloc(b->count * sizeof(size_t));
    if (!host_tables || !zeros || !goffs || build_groups(b, heights, widths, goffs))
        goto fail;
    for (unsigned i = 0; i < b->ngroups; i++)
        b->ndense += b->groups[i].route != BRIDCT_ROUTE_BUTTERFLY;
    b->auto_strategy = !b->tiny && (b->ndense || b->elements / b->ngroups >= GROUPED_MIN_PIXELS)
                           ? BRIDCT_BATCH_GROUPED
                           : BRIDCT_BATCH_PERSISTENT;
    b->strategy = b->auto_strategy;
    bridct_build_tables(host_tables);
    if (bdx_stream_create(&b->stream) ||
        upload((void **)&b->d_tables, host_tables, BDX_TABLE_FLOATS * sizeof(float), b->stream) ||
        upload((void **)&b->d_runs, sc.runs, sc.nruns * sizeof(bdx_run), b->stream) ||
        upload((void **)&b->d_items[0], sc.items[0], sc.nitems[0] * sizeof(bdx_item), b->stream) ||
        upload((void **)&b->d_items[1], sc.items[1], sc.nitems[1] * sizeof(bdx_item), b->stream) ||
        upload((void **)&b->d_units, sc.units, sc.nunits * sizeof(bdx_unit), b->stream) ||
        upload((void **)&b->d_ctr, zeros, b->nctr * sizeof(unsigned), b->stream) ||
        upload((void **)&b->d_goffs, goffs, b->count * sizeof(size_t), b->stream) ||
        aux_setup(b) ||
        (b->ndense && dense_setup(b)) || bdx_stream_sync(b->stream))
        goto fail;
    free(host_tables);
    free(zeros);
    free(goffs);
    bdx_schedule_free(&sc);
    return b;

fail:
    free(host_tables);
    free(zeros);
    free(goffs);
    bdx_schedule_free(&sc);
    bridct_batch_destroy(b);
    return NULL;
}

size_t bridct_batch_elements(const bridct_batch *b)
{
    return b ? b->elements : 0;
}

size_t bridct_batch_offset(const bridct_batch *b, size_t index)
{
    return b && index < b->count ? b->offsets[index] : 0;
}

size_t bridct_batch_items(const bridct_batch *b, int mode)
{
    if (!b || mode < 0 || mode > BRIDCT_GPU_FORWARD_THEN_INVERSE)
        return 0;
    return mode == BRIDCT_ROUNDTRIP ? b->nitems[1] : mode == BRIDCT_GPU_FORWARD_THEN_INVERSE ? 2 * b->nitems[0] : b->nitems[0];
}

static int launch(bridct_batch *b, int mode, const float *d_in, float *d_out)
{
    const int fused = mode == BDX_ROUNDTRIP;
    int flags = 0;
    if (b->tiny)
        flags = BDX_LAUNCH_TINY | (2 * b->elements * sizeof(float) > BDX_STREAM_MIN_BYTES ? BDX_LAUNCH_STREAM : 0);
    return bdx_batch_launch(d_in, d_out, b->d_runs, b->d_items[fused], b->nitems[fused], b->d_units, b->nunits,
                            b->qbase, b->d_ctr, b->nctr, mode, b->d_tables, b->blocks, b->spin_limit, flags,
                            b->stream);
}

/* A dense shape on the batch stream, in slices of the GEMMs' grid limits; the round trip
 * is the forward transform then the inverse in place, as on the plan path. */
static int launch_dense(bridct_batch *b, const group *g, int mode, const float *d_in, float *d_out)
{
    const int prec = g->route == BRIDCT_ROUTE_MFMA_F16 ? BDX_DENSE_F16 : BDX_DENSE_BF16;
    const size_t hh = (size_t)g->h * g->h, slice = dense_slice(g->h);
    const float *qh = b->d_dense + g->qoff, *qht = qh + hh, *qw = qht + hh;
    for (size_t done = 0; done < g->count; done += slice) {
        const size_t n = g->count - done < slice ? g->count - done : slice;
        const size_t *offs = b->d_goffs + g->first + done;
        if (mode == BDX_ROUNDTRIP) {
            if (bdx_dense_transform_gather(d_in, d_out, b->d_dwork, g->h, g->w, n, offs, BDX_FORWARD, prec, qh, qht,
                                           qw, g->dscale, b->stream) ||
                bdx_dense_transform_gather(d_out, d_out, b->d_dwork, g->h, g->w, n, offs, BDX_INVERSE, prec, qh,
                                           qht, qw, g->dscale, b->stream))
                return -1;
        } else if (bdx_dense_transform_gather(d_in, d_out, b->d_dwork, g->h, g->w, n, offs, mode, prec, qh, qht, qw,
                                              g->dscale, b->stream)) {
            return -1;
        }
    }
    return 0;
}

/* GROUPED: the plan path's kernels, per distinct shape, in gather form.  Two-pass shapes of
 * at least GROUPED_SMALL_TWO_PASS pixels go on the batch stream, one after the other;
 * whole-array shapes and smaller two-pass shapes, whose launches leave most of the GPU
 * idle, go on the auxiliary streams beside them (each to the one with the least work so
 * far; the batch stream takes its share too when no shape is that large), forked from and
 * joined back to the batch stream with events. */
static int aux_group(const group *g)
{
    return g->route == BRIDCT_ROUTE_BUTTERFLY &&
           (BDX_WHOLE_ARRAY(g->h, g->w, g->pack) || (size_t)g->h * g->w * g->count < GROUPED_SMALL_TWO_PASS);
}

static int launch_grouped(bridct_batch *b, int mode, const float *d_in, float *d_out)
{
    unsigned naux = 0;
    for (unsigned i = 0; i < b->ngroups; i++)
        naux += aux_group(&b->groups[i]);
    /* Streams the small groups are dealt to: the auxiliary ones, and the batch stream too
     * when no group needs it for itself. */
    void *cand[GROUPED_AUX_STREAMS + 1];
    const int first_aux = naux == b->ngroups;   /* index of aux[0] in cand */
    int nc = 0;
    if (first_aux)
        cand[nc++] = b->stream;
    for (int k = 0; k < GROUPED_AUX_STREAMS; k++)
        cand[nc++] = b->aux[k];
    size_t load[GROUPED_AUX_STREAMS + 1] = {0};
    int used[GROUPED_AUX_STREAMS + 1] = {0};
    const int fork = naux > 0;
    int rc = fork ? bdx_event_record(b->ev_fork, b->stream) : 0;
    for (unsigned i = 0; i < b->ngroups && !rc; i++) {   /* largest first */
        const group *g = &b->groups[i];
        if (g->route != BRIDCT_ROUTE_BUTTERFLY) {
            rc = launch_dense(b, g, mode, d_in, d_out);
            continue;
        }
        void *s = b->stream;
        if (aux_group(g)) {
            const int whole = BDX_WHOLE_ARRAY(g->h, g->w, g->pack);
            const size_t px = (size_t)g->h * g->w * g->count;
            int k = 0;
            for (int j = 1; j < nc; j++)
                if (load[j] < load[k])
                    k = j;
            s = cand[k];
            if (!used[k] && s != b->stream)
                rc = bdx_stream_wait_event(s, b->ev_fork);
            used[k] = 1;
            load[k] += (whole ? 1 : 2) * (px + ((size_t)1 << 21));   /* a launch costs about 2 Mpx */
        }
        if (!rc)
            rc = bdx_transform_gather(d_in, d_out, g->h, g->w, g->count, b->d_goffs + g->first, g->pack, mode,
                                      b->d_tables, s);
    }
    /* the auxiliary streams this call used are joined, even after a failed launch, so that
     * the batch stream covers everything issued */
    for (int k = 0; k < GROUPED_AUX_STREAMS; k++)
        if (used[first_aux + k] &&
            (bdx_event_record(b->ev_join[k], b->aux[k]) || bdx_stream_wait_event(b->stream, b->ev_join[k])))
            rc = -1;
    return rc;
}

static int launch_any(bridct_batch *b, int mode, const float *d_in, float *d_out)
{
    return b->strategy == BRIDCT_BATCH_GROUPED ? launch_grouped(b, mode, d_in, d_out) : launch(b, mode, d_in, d_out);
}

static int issue(bridct_batch *b, int mode, const float *d_in, float *d_out)
{
    if (mode == BRIDCT_GPU_FORWARD_THEN_INVERSE)
        return launch_any(b, BDX_FORWARD, d_in, d_out) || launch_any(b, BDX_INVERSE, d_out, d_out);
    return launch_any(b, mode, d_in, d_out);
}

static int check_device_args(const bridct_batch *b, int mode, const float *d_in, const float *d_out)
{
    if (!b || !d_in || !d_out || mode < 0 || mode > BRIDCT_GPU_FORWARD_THEN_INVERSE)
        return BRIDCT_INVALID_ARGUMENT;
    if (((uintptr_t)d_in | (uintptr_t)d_out) % 16)
        return BRIDCT_BAD_ALIGNMENT;
    const uintptr_t x = (uintptr_t)d_in, y = (uintptr_t)d_out;
    if (d_in != d_out && (x <= y ? y - x : x - y) < b->elements * sizeof(float))
        return BRIDCT_OVERLAPPING_BUFFERS;
    return BRIDCT_OK;
}

int bridct_batch_apply_device(bridct_batch *b, int mode, const float *d_input, float *d_output)
{
    const int st = check_device_args(b, mode, d_input, d_output);
    if (st)
        return st;
    return issue(b, mode, d_input, d_output) ? BRIDCT_DEVICE_ERROR : BRIDCT_OK;
}

static int capture_fn(void *arg)
{
    bridct_batch *b = arg;
    return issue(b, b->cap_mode, b->cap_in, b->cap_out);
}

int bridct_batch_graph_apply(bridct_batch *b, int mode, const float *d_input, float *d_output)
{
    const int st = check_device_args(b, mode, d_input, d_output);
    if (st)
        return st;
    if (!b->graph || b->graph_mode != mode || b->graph_in != d_input || b->graph_out != d_output) {
        if (b->graph) {
            if (bdx_stream_sync(b->stream))
                return BRIDCT_DEVICE_ERROR;
            bdx_graph_destroy(b->graph);
            b->graph = NULL;
        }
        b->cap_mode = mode;
        b->cap_in = d_input;
        b->cap_out = d_output;
        if (bdx_graph_capture(b->stream, capture_fn, b, &b->graph)) {
            b->graph = NULL;
            return BRIDCT_DEVICE_ERROR;
        }
        b->graph_mode = mode;
        b->graph_in = d_input;
        b->graph_out = d_output;
    }
    return bdx_graph_launch(b->graph, b->stream) ? BRIDCT_DEVICE_ERROR : BRIDCT_OK;
}

int bridct_batch_synchronize(bridct_batch *b)
{
    if (!b)
        return BRIDCT_INVALID_ARGUMENT;
    unsigned err = 0;
    for (int k = 0; k < GROUPED_AUX_STREAMS; k++)
        if (bdx_stream_sync(b->aux[k]))
            return BRIDCT_DEVICE_ERROR;
    if (bdx_stream_sync(b->stream) ||
        bdx_copy_d2h(&err, b->d_ctr + BDX_CTR_ERROR, sizeof(err), b->stream) || bdx_stream_sync(b->stream))
        return BRIDCT_DEVICE_ERROR;
    if (err) {   /* reported once: grouped calls do not reset the flag themselves */
        const unsigned zero = 0;
        if (bdx_copy_h2d(b->d_ctr + BDX_CTR_ERROR, &zero, sizeof(zero), b->stream))
            return BRIDCT_DEVICE_ERROR;
        bdx_stream_sync(b->stream);
    }
    return err ? BRIDCT_DEVICE_ERROR : BRIDCT_OK;
}

int bridct_batch_apply(bridct_batch *b, int mode, const float *input, float *output)
{
    if (!b || !input || !output || mode < 0 || mode > BRIDCT_GPU_FORWARD_THEN_INVERSE)
        return BRIDCT_INVALID_ARGUMENT;
    const size_t bytes = b->elements * sizeof(float);
    const uintptr_t x = (uintptr_t)input, y = (uintptr_t)output;
    if ((x <= y ? y - x : x - y) < bytes)
        return BRIDCT_OVERLAPPING_BUFFERS;
    if (!b->d_in && (bdx_malloc((void **)&b->d_in, bytes) || bdx_malloc((void **)&b->d_out, bytes))) {
        if (b->d_in)
            bdx_free(b->d_in);
        b->d_in = NULL;
        return BRIDCT_DEVICE_ERROR;
    }
    /* large calls go through pinned staging (about 57 GB/s each way against 31 / 18 GB/s
     * for pageable copies); without it, or for small calls, the copies are direct */
    if (bytes >= STAGE_MIN_BYTES && !b->stage)
        b->stage = stage_create(STAGE_CHUNK_BYTES);
    const int staged = bytes >= STAGE_MIN_BYTES && b->stage;
    if ((staged ? stage_h2d(b->stage, b->d_in, input, bytes, b->stream)
                : bdx_copy_h2d(b->d_in, input, bytes, b->stream)) ||
        issue(b, mode, b->d_in, b->d_out) ||
        (staged ? stage_d2h(b->stage, output, b->d_out, bytes, b->stream)
                : bdx_copy_d2h(output, b->d_out, bytes, b->stream)))
        return BRIDCT_DEVICE_ERROR;
    return bridct_batch_synchronize(b);
}

int bridct_batch_set_spin_limit(bridct_batch *b, unsigned limit)
{
    if (!b || limit == 0)
        return BRIDCT_INVALID_ARGUMENT;
    if (b->graph) {   /* the captured launches carry the old bound */
        if (bdx_stream_sync(b->stream))
            return BRIDCT_DEVICE_ERROR;
        bdx_graph_destroy(b->graph);
        b->graph = NULL;
    }
    b->spin_limit = limit;
    return BRIDCT_OK;
}

void *bridct_batch_stream(const bridct_batch *b)
{
    return b ? b->stream : NULL;
}

int bridct_batch_set_strategy(bridct_batch *b, int strategy)
{
    if (!b || strategy < BRIDCT_BATCH_AUTO || strategy > BRIDCT_BATCH_GROUPED)
        return BRIDCT_INVALID_ARGUMENT;
    const int resolved = strategy == BRIDCT_BATCH_AUTO ? b->auto_strategy : strategy;
    if (resolved != b->strategy && b->graph) {   /* the captured sequence changes */
        if (bdx_stream_sync(b->stream))
            return BRIDCT_DEVICE_ERROR;
        bdx_graph_destroy(b->graph);
        b->graph = NULL;
    }
    b->strategy = resolved;
    return BRIDCT_OK;
}

int bridct_batch_get_strategy(const bridct_batch *b)
{
    return b ? b->strategy : BRIDCT_INVALID_ARGUMENT;
}
