/* Thin C ABI between the C host library (plan.c) and the HIP side (bridct_kernels.hip).
 *
 * Every function returns 0 on success and a negative value on a HIP failure.  Pointers
 * named d_* are device memory; streams are opaque hipStream_t values.  Nothing here
 * allocates except bdx_malloc and bdx_stream_create.
 */
#ifndef BRIDCT_DEVICE_H
#define BRIDCT_DEVICE_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Number of floats of the constant-table block a plan uploads (all axis lengths 8..1024). */
#define BDX_TABLE_FLOATS (6 * (2048 - 8))
/* Offset of the block for axis length n: omega (2n), tan (n), alpha (n), alpha^2 (n),
 * team twiddles (n: complex exp(-2 pi i t k1 / (n/2)) at 2 (k1 T + t), T = team size). */
#define BDX_TABLE_BASE(n) (6 * ((n) - 8))
/* Team size of the multi-lane 1D transform of length n (1 below 128). */
#define BDX_TEAM(n) ((n) <= 64 ? 1 : (n) >= 1024 ? 16 : (n) >= 256 ? 8 : 4)

/* ---- persistent batch kernel (shared geometry, host C and device) ----
 * A batch is a sequence of runs; a run is `count` contiguous h x w arrays at element
 * offset `off` of the batch buffers.  Arrays with both axes at most 128 are transformed
 * whole (several per unit); the others by row items of up to BDX_ROW_UNIT(w) rows and
 * column items of up to BDX_COL_UNIT(h) columns (256-lane workgroups; the last item of
 * an axis may be partial). */
#define BDX_BATCH_THREADS 256
#define BDX_BATCH_LDS 20400   /* floats: two workgroups fit the 160 KiB LDS of a CU */
#define BDX_BATCH_SMALL(h, w) ((h) <= 128 && (w) <= 128 && (long)(h) * ((w) + 1) <= BDX_BATCH_LDS)
#define BDX_BATCH_PER(h, w) (BDX_BATCH_LDS / ((h) * ((w) + 1)))
/* one row_unit call: a band of rows; one col_unit call: a strip of columns */
#define BDX_ROW_BAND(w) ((w) >= 128 ? BDX_BATCH_THREADS / BDX_TEAM(w) : BDX_BATCH_THREADS)
#define BDX_COL_STRIP(h) ((h) >= 128 ? BDX_BATCH_THREADS / BDX_TEAM(h) : BDX_BATCH_THREADS)
/* one work item: BDX_ITEM_BANDS bands / strips for lines of 512 and more, else one.
 * Larger items mean fewer dependency waits per byte but fewer items per array; on the
 * 10,000-array ragged mix two bands per item measured 1-2% slower than one: one. */
#ifndef BDX_ITEM_BANDS
#define BDX_ITEM_BANDS 1
#endif
#define BDX_ROW_UNIT(w) (BDX_ROW_BAND(w) * ((w) >= 512 ? BDX_ITEM_BANDS : 1))
#define BDX_COL_UNIT(h) (BDX_COL_STRIP(h) * ((h) >= 512 ? BDX_ITEM_BANDS : 1))
#define BDX_CEIL(a, b) (((a) + (b) - 1) / (b))
/* Calls whose input plus output exceed this many bytes (the 256 MB Infinity Cache of an
 * MI355X) stream their array data with non-temporal loads and stores: the two-pass
 * kernels and the tiny batch kernel. */
#define BDX_STREAM_MIN_BYTES ((size_t)256 << 20)

/* A two-pass call whose intermediate (batch x h x w floats) reaches BDX_CHUNK_MIN_BYTES runs
 * in chunks of whole arrays of at most BDX_CHUNK_BYTES each, every chunk's row pass followed
 * by its column pass.  The column pass then finds most of its input in the Infinity Cache;
 * in one call over a batch that large, the first arrays' rows are evicted before their
 * columns run (1024^2 batches of 128 and 166 arrays: 2-8% faster; 512 MiB batches of
 * 512 x 1024, 1024 x 512 and 512^2: 2-6%; at 312 MiB, chunks gain nothing). */
#define BDX_CHUNK_MIN_BYTES ((size_t)512 << 20)
#define BDX_CHUNK_BYTES ((size_t)256 << 20)

/* Whole-array route of the plan path: an h x w array with h (w + 1) <= BDX_SMALL_LDS_FLOATS
 * is transformed by one workgroup in LDS. */
#define BDX_SMALL_LDS_FLOATS 36864
#define BDX_SMALL_ROUTE(h, w) ((long)(h) * ((w) + 1) <= BDX_SMALL_LDS_FLOATS)
/* Shapes of that route that the row and column pass kernels also take: whole bands and
 * strips.  A row band is 256 / team rows (256 rows for lines up to 64, one lane per line);
 * a column strip is 64 columns (32 at height 1024; 256 for lines up to 64).  The pass
 * kernels check these sizes against their own at compile time. */
#define BDX_PASS_ROWS(w) ((w) >= 128 ? 256 / BDX_TEAM(w) : 256)
#define BDX_PASS_COLS(h) ((h) >= 1024 ? 32 : (h) >= 128 ? 64 : 256)
#define BDX_TWO_PASS_OK(h, w) ((h) % BDX_PASS_ROWS(w) == 0 && (w) % BDX_PASS_COLS(h) == 0)
/* Pack value that sends such a shape to the two-pass kernels (BRIDCT_PACK_TWO_PASS). */
#define BDX_PACK_TWO_PASS (-1)
/* Whether an h x w shape with packing `pack` runs on the whole-array kernel. */
#define BDX_WHOLE_ARRAY(h, w, pack) \
    (BDX_SMALL_ROUTE(h, w) && !((pack) == BDX_PACK_TWO_PASS && BDX_TWO_PASS_OK(h, w)))

/* Batches whose arrays all have both axes at most BDX_TINY_MAX have no dependent items: they
 * run the tiny kernel instead of the persistent queue, one workgroup per item, about
 * BDX_TINY_PIXELS pixels per item (the whole-array packing that measured fastest for these
 * shapes) and BDX_TINY_LDS floats of LDS, so that many workgroups stay resident per CU. */
#define BDX_TINY_MAX 32
#define BDX_TINY_PIXELS 2048
#define BDX_TINY_LDS 2304      /* BDX_TINY_PIXELS * (w + 1) / w floats at w = 8 */
#define BDX_TINY_PER(h, w) (BDX_TINY_PIXELS / ((h) * (w)))
/* work item kinds */
#define BDX_ITEM_SMALL 0u
#define BDX_ITEM_ROW 1u
#define BDX_ITEM_COL 2u
#define BDX_ITEM_ROW2 3u
/* counters block: [0] unused, [1] error flag, then per two-pass array c
 * [2 + 2 c] row items done, [3 + 2 c] column items done, then (from BDX_QBASE(ntwo), an even
 * word offset) the per-XCD queues of the persistent kernel:
 *   BDX_XQ ticket counters, one global unit counter, one pad word, then per queue q
 *   BDX_SLOTS(nunits) 64-bit slot words (slot j of queue q at word
 *   BDX_QSLOT(base, nunits, q, j)): 0 while unpublished, else (unit + 1) << 32 | end ticket,
 *   unit + 1 = BDX_UNIT_END when the units are exhausted. */
#define BDX_CTR_HEAD 0
#define BDX_CTR_ERROR 1
#define BDX_CTR_BASE 2
#define BDX_XQ 16                               /* queues, indexed by the XCD id (4 bits) */
#define BDX_QBASE(ntwo) ((BDX_CTR_BASE + 2u * (ntwo) + 1u) & ~1u)
/* Lag of the queues, in slots: slot j holds the ROW items of the unit u_j a queue takes
 * j-th, the COL items of u_{j-BDX_LAG} and the ROW2 items of u_{j-2 BDX_LAG}. */
#ifndef BDX_LAG
#define BDX_LAG 3
#endif
#define BDX_SLOTS(nunits) ((nunits) + 2u * BDX_LAG + 1u)   /* slots a queue can reach */
#define BDX_QSLOT(base, nunits, q, j) ((base) + BDX_XQ + 2u + 2u * ((q) * BDX_SLOTS(nunits) + (j)))
#define BDX_QWORDS(nunits) (BDX_XQ + 2u + 2u * BDX_XQ * BDX_SLOTS(nunits))
#define BDX_UNIT_END 0xFFFFFFFFu

typedef struct {
    unsigned h, w, count, pad;
    unsigned long long off;
} bdx_run;

/* kind_ctr = kind << 28 | two-pass array counter id (row/column kinds).
 * small: idx = first array of the run, unit = number of arrays.
 * row/column: idx = array within the run, unit = unit index. */
typedef struct {
    unsigned run, idx, unit, kind_ctr;
} bdx_item;

/* A unit of the persistent kernel: one two-pass array (its ROW items, then its COL items,
 * then for the fused round trip its ROW2 items) or one SMALL item.  Its items are
 * consecutive in the item list of each mode family from item0[family] on (family 0:
 * forward / inverse, 1: fused round trip); nrow2 = nrow for two-pass arrays, 0 for SMALL. */
typedef struct {
    unsigned item0[2];
    unsigned short nrow, ncol, nrow2, pad;
} bdx_unit;

/* Launches the persistent batch kernel (after a counter reset) on `stream`: `blocks`
 * resident workgroups take the `nunits` units through the per-XCD queues (every item of a
 * unit runs on the XCD whose queue took it, so dependencies stay inside one L2).
 * spin_limit bounds each wait (a wait that runs out sets the error flag instead of
 * hanging). */
/* flags of bdx_batch_launch: the tiny kernel instead of the persistent one, and streaming
 * (non-temporal) array accesses in it */
#define BDX_LAUNCH_TINY 1
#define BDX_LAUNCH_STREAM 2
int bdx_batch_launch(const float *d_in, float *d_out, const bdx_run *d_runs, const bdx_item *d_items,
                     unsigned nitems, const bdx_unit *d_units, unsigned nunits, unsigned qbase,
                     unsigned *d_counters, unsigned ncounters, int mode, const float *d_tables, unsigned blocks,
                     unsigned spin_limit, int flags, void *stream);
int bdx_device_cus(void);

/* Host-side schedule of a batch (no device needed): runs of consecutive equal shapes,
 * the units (two-pass arrays by decreasing size, then the SMALL items) and their item lists
 * ([0] forward / inverse, [1] fused round trip), and the counter words (per-array
 * counters and the queues).  Returns 0, or -1 for an invalid shape list or allocation
 * failure. */
typedef struct {
    size_t count, elements;
    size_t *offsets;
    bdx_run *runs;
    unsigned nruns, ntwo;
    bdx_item *items[2];
    unsigned nitems[2];
    bdx_unit *units;
    unsigned nunits;
    unsigned qbase;           /* word offset of the queues in the counters block */
    unsigned nctr;            /* counter words in all */
    int tiny;                 /* every array fits the tiny kernel; SMALL items hold BDX_TINY_PER */
} bdx_schedule;
int bdx_batch_schedule(size_t count, const int *heights, const int *widths, bdx_schedule *s);
/* 1 when every array of the shape list fits the tiny kernel (both axes <= BDX_TINY_MAX);
 * its schedule then packs BDX_TINY_PER(h, w) arrays per item at most. */
int bdx_batch_is_tiny(size_t count, const int *heights, const int *widths);
void bdx_schedule_free(bdx_schedule *s);

/* Graph capture of a sequence of calls issued by `fn(arg)` on `stream`. */
int bdx_graph_capture(void *stream, int (*fn)(void *), void *arg, void **graph_exec);
int bdx_graph_launch(void *graph_exec, void *stream);
int bdx_graph_destroy(void *graph_exec);

/* Transform modes seen by the kernels (same values as enum bridct_mode). */
#define BDX_FORWARD 0
#define BDX_INVERSE 1
#define BDX_ROUNDTRIP 2

int bdx_device_ready(void);
int bdx_current_device(int *device);
/* The dense route's static intermediate on `device`: a file-scope device array of
 * *floats floats (BDX_STATIC_WORK_FLOATS: BDX_STATIC_ARRAY_FLOATS of array data plus room
 * for the per-array prescale exponents), shared by every call on that GPU that uses it;
 * its address is resolved once per device.  0, or -1 when it cannot be resolved. */
#define BDX_STATIC_ARRAY_FLOATS ((size_t)8 << 20)
#define BDX_STATIC_WORK_FLOATS (BDX_STATIC_ARRAY_FLOATS + 65536)
int bdx_dense_static_work(int device, float **ptr, size_t *floats);
/* Failure injection and resource count (bridct_debug_alloc_faults in bridct_gpu.h). */
long bdx_debug_alloc_faults(long after);
/* Test hook (bridct_debug_grid_limit in bridct_gpu.h): workgroups per launch of the
 * launchers that split large grids; blocks <= 0 restores the hardware limit.  Returns the
 * previous setting (0 for the hardware limit). */
long bdx_debug_grid_limit(long blocks);
int bdx_malloc(void **ptr, size_t bytes);
int bdx_free(void *ptr);
int bdx_host_malloc(void **ptr, size_t bytes);   /* pinned host memory */
int bdx_host_free(void *ptr);
int bdx_event_create(void **event);
int bdx_event_destroy(void *event);
int bdx_event_record(void *event, void *stream);
int bdx_event_sync(void *event);
int bdx_stream_wait_event(void *stream, void *event);   /* later work on stream waits for event */
int bdx_stream_create(void **stream);
int bdx_stream_destroy(void *stream);
int bdx_stream_sync(void *stream);
int bdx_copy_h2d(void *dst, const void *src, size_t bytes, void *stream);
int bdx_copy_d2h(void *dst, const void *src, size_t bytes, void *stream);
int bdx_copy_d2d(void *dst, const void *src, size_t bytes, void *stream);
/* Streaming float4 device copy (floats a multiple of 4, 16-byte aligned): the measured
 * HBM ceiling for one read and one write per element. */
int bdx_copy_stream(const float *d_src, float *d_dst, size_t floats, void *stream);

/* Orthonormal 2D transform of `batch` contiguous row-major h x w float32 arrays.
 * d_in and d_out must be 16-byte aligned and either identical or disjoint.  pack: arrays
 * per workgroup of the whole-array kernel (0 = automatic; ignored by two-pass shapes). */
int bdx_transform(const float *d_in, float *d_out, int h, int w, size_t batch, int pack, int mode,
                  const float *d_tables, void *stream);
/* The same transform of `count` h x w arrays at element offsets d_offsets[0, count) of
 * d_in and d_out (device array; each array 16-byte aligned, arrays disjoint): the arrays
 * of one shape inside a ragged batch, with the plan path's kernels in gather form. */
int bdx_transform_gather(const float *d_in, float *d_out, int h, int w, size_t count, const size_t *d_offsets,
                         int pack, int mode, const float *d_tables, void *stream);
/* Dense split-MFMA route: operand precision of the hi/lo split. */
#define BDX_DENSE_F16 1
#define BDX_DENSE_BF16 2
/* Y = scale Q_h' X Q_w'^T (FORWARD) or X = scale Q_h'^T Y Q_w' (INVERSE) for `batch`
 * contiguous h x w arrays, h, w >= 64, batch <= 65535 and batch h / 64 <= 65535.
 * d_qh, d_qht, d_qw: the scaled DCT-II matrices Q_h' (h x h), its transpose, Q_w' (w x w),
 * row-major float32; d_work: BDX_DENSE_WORK_FLOATS(batch, h, w) floats for the
 * intermediate and the arrays' prescale exponents; in place allowed. */
#define BDX_DENSE_WORK_FLOATS(count, h, w) ((size_t)(count) * (size_t)(h) * (size_t)(w) + (size_t)(count))
int bdx_dense_transform(const float *d_in, float *d_out, float *d_work, int h, int w, size_t batch, int mode,
                        int prec, const float *d_qh, const float *d_qht, const float *d_qw, float scale,
                        void *stream);

/* The same for `count` h x w arrays at the element offsets d_offsets[0, count) of d_in and
 * d_out (the arrays of one shape inside a ragged batch); count <= 65535 and
 * count h / 64 <= 65535; d_work: BDX_DENSE_WORK_FLOATS(count, h, w) floats, contiguous. */
int bdx_dense_transform_gather(const float *d_in, float *d_out, float *d_work, int h, int w, size_t count,
                               const size_t *d_offsets, int mode, int prec, const float *d_qh, const float *d_qht,
                               const float *d_qw, float scale, void *stream);

/* 8x8 block codec step over a height x width u8 raster (multiples of 8; row stride
 * `stride` bytes, a multiple of 8; 8-byte aligned) and its block-contiguous int16
 * coefficients (64 per block, 16-byte aligned).  decode = 0: d_src is the raster, d_dst the
 * coefficients; decode = 1: the reverse.  d_factors: 64 encode factors
 * alpha_u alpha_v / Q_uv, then 64 decode factors Q_uv alpha_u alpha_v (alpha: the
 * length-8 core normalization). */
int bdx_codec8(int decode, const void *d_src, void *d_dst, size_t height, size_t width, size_t stride,
               const float *d_factors, void *stream);

/* One kernel of a FORWARD or INVERSE transform: pass 0 = row pass (d_in -> d_out; the
 * whole transform for shapes with a whole-array kernel), pass 1 = column pass in place
 * (d_in == d_out expected; nothing for whole-array shapes). */
int bdx_transform_pass(const float *d_in, float *d_out, int h, int w, size_t batch, int pack, int mode,
                       int pass, const float *d_tables, void *stream);

#ifdef __cplusplus
}
#endif
#endif
