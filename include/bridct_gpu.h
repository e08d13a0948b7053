/* GPU extensions of the BRiDCT plan API (MI355X).  No counterpart in the reference
 * interface; they expose what a device-resident caller needs: device pointers, batches,
 * streams, and a few device-memory helpers so that C and Python callers need no other
 * GPU runtime binding.
 */
#ifndef BRIDCT_GPU_H
#define BRIDCT_GPU_H

#include "bridct.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Status of a failed GPU runtime call (allocation, copy, launch). */
#define BRIDCT_DEVICE_ERROR (-5)

/* 1 when a gfx950 device is present and usable, else 0. */
BRIDCT_API int bridct_gpu_available(void);

/* Applies the plan to `batch` contiguous h x w arrays in device memory on the plan's
 * stream, asynchronously.  d_input and d_output must be 16-byte aligned and either the
 * same pointer (in place) or disjoint.  Statuses as bridct_plan_apply, plus
 * BRIDCT_BAD_ALIGNMENT and BRIDCT_DEVICE_ERROR. */
BRIDCT_API int bridct_plan_apply_device(bridct_plan *plan, enum bridct_mode mode,
                                        const float *d_input, float *d_output, size_t batch);

/* bridct_plan_apply_device with nothing allocated.  No counterpart in the reference: for
 * callers that cannot allocate device memory during the call (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).  The butterfly route
 * needs no memory beyond the plan's; a dense-route plan (bridct_plan_set_route) takes its
 * intermediate (and its per-array prescale exponents) from one static array per GPU of
 * 8,454,144 floats (32.25 MiB, held by each GPU the library runs on until the process
 * ends) instead of allocating its own buffer, so batch * height * width may not exceed
 * 8,388,608 there (else BRIDCT_WORKSPACE_TOO_SMALL, nothing run).  Every such call on one GPU shares that array: dense-route calls that can
 * overlap on one GPU (other streams, other threads) race and must be serialized by the
 * caller.  Results equal bridct_plan_apply_device's bitwise. */
BRIDCT_API int bridct_plan_apply_device_static(bridct_plan *plan, enum bridct_mode mode,
                                               const float *d_input, float *d_output, size_t batch);

/* Routes of a plan.  BUTTERFLY (default): the FP32 Shao-Johnson butterflies.
 * MFMA_F16 / MFMA_BF16: the dense route Y = C_h X C_w^T on the matrix cores, both
 * operands split into hi + lo half-precision parts (FP16 or BF16), three products, FP32
 * accumulation; both axes must be at least 64.  Its error against the FP64 definition is
 * larger than the butterflies' (measured per shape).  Each array is prescaled by the power
 * of two of its largest magnitude before the split and scaled back after, so the error
 * does not depend on the array's scale (2^k X gives exactly 2^k times the result of X).
 * The dense route keeps an intermediate buffer (one slice of batch h w floats plus one
 * exponent per array) in the plan, allocated on the first call that needs it. */
#define BRIDCT_ROUTE_BUTTERFLY 0
#define BRIDCT_ROUTE_MFMA_F16 1
#define BRIDCT_ROUTE_MFMA_BF16 2
BRIDCT_API int bridct_plan_set_route(bridct_plan *plan, int route);
BRIDCT_API int bridct_plan_get_route(const bridct_plan *plan);

/* Arrays per workgroup of the whole-array kernel (shapes whose array fits one
 * workgroup's LDS): 0 = automatic, else 1..4096 (clamped to what fits, and lowered for
 * batches too small to give about 512 workgroups at that packing).  BRIDCT_PACK_TWO_PASS
 * runs such a shape through the row and column pass kernels instead, as larger shapes
 * are; it is accepted for the shapes those kernels take in whole row bands (256 rows for
 * widths up to 64, else 256 / team size: 64, 32, 32, 16 rows at widths 128..1024) and
 * whole column strips (256 columns for heights up to 64, 64 above, 32 at height 1024).
 * The autotuner picks it where it measured faster. */
#define BRIDCT_PACK_TWO_PASS (-1)
BRIDCT_API int bridct_plan_set_pack(bridct_plan *plan, int pack);
BRIDCT_API int bridct_plan_get_pack(const bridct_plan *plan);

/* A new plan takes its route and packing from the schedule table written by the
 * autotuner (bridct_amd/autotune.py): the file named by $BRIDCT_SCHEDULE (empty or
 * "none" for no table), else schedule_gfx950.txt beside the library.  Lines read
 * "<height> <width> <butterfly|mfma_f16|mfma_bf16> <pack>". */

/* One kernel of bridct_plan_apply_device, for profiling a transform kernel by kernel:
 * pass 0 runs the row pass (d_input -> d_output; for shapes transformed by one whole-array
 * kernel, the whole transform), pass 1 the column pass in place on d_output (d_input must
 * equal d_output; nothing to do for whole-array shapes).  FORWARD or INVERSE only.
 * Pass 0 then pass 1 equals bridct_plan_apply_device. */
BRIDCT_API int bridct_plan_apply_device_pass(bridct_plan *plan, enum bridct_mode mode, int pass,
                                             const float *d_input, float *d_output, size_t batch);

/* Mode value accepted by bridct_plan_apply_batch only: FORWARD then INVERSE with the
 * normalized spectrum materialized in FP32 between them (the NumPy Plan.roundtrip
 * semantics of the reference, python/bridct_numpy.c:114-133).  Bitwise equal to the two
 * separate calls. */
#define BRIDCT_GPU_FORWARD_THEN_INVERSE 3

/* Host-pointer batch: `batch` contiguous arrays, copied through the plan's device
 * buffers in chunks.  Same validation as bridct_plan_apply over the whole batch;
 * mode may also be BRIDCT_GPU_FORWARD_THEN_INVERSE. */
BRIDCT_API int bridct_plan_apply_batch(bridct_plan *plan, int mode, const float *input,
                                       float *output, size_t batch);

/* Waits for all work issued on the plan's stream. */
BRIDCT_API int bridct_plan_synchronize(bridct_plan *plan);

/* The plan's stream (a hipStream_t), for callers that order their own work after it. */
BRIDCT_API void *bridct_plan_stream(const bridct_plan *plan);

/* ---- ragged batches ----
 * No counterpart in the reference, whose callers loop over arrays with one plan per
 * shape (src/plan.c:126-163, python/bridct_numpy.c:96-112).  A batch describes `count`
 * arrays of shapes heights[i] x widths[i] stored back to back in that order (array i
 * starts at element bridct_batch_offset(batch, i)); one call transforms all of them with
 * one persistent kernel launch (two for BRIDCT_GPU_FORWARD_THEN_INVERSE).  Each array's
 * result is that of a plan of its shape.  A batch object, like a plan, must not be used
 * from two threads at once. */
typedef struct bridct_batch bridct_batch;

/* NULL when count is 0, a shape is unsupported, no GPU is usable, or allocation fails. */
BRIDCT_API bridct_batch *bridct_batch_create(size_t count, const int *heights, const int *widths);
BRIDCT_API void bridct_batch_destroy(bridct_batch *batch);   /* NULL is a no-op */
/* Total floats of one batch buffer; element offset of array `index` (0 if out of range). */
BRIDCT_API size_t bridct_batch_elements(const bridct_batch *batch);
BRIDCT_API size_t bridct_batch_offset(const bridct_batch *batch, size_t index);
/* Work items one call of `mode` queues (diagnostics). */
BRIDCT_API size_t bridct_batch_items(const bridct_batch *batch, int mode);
/* Device buffers, asynchronous on the batch stream.  mode: FORWARD, INVERSE, ROUNDTRIP or
 * BRIDCT_GPU_FORWARD_THEN_INVERSE.  Buffers 16-byte aligned, identical or disjoint. */
BRIDCT_API int bridct_batch_apply_device(bridct_batch *batch, int mode, const float *d_input,
                                         float *d_output);
/* Same work replayed from a captured HIP graph; the graph is captured on the first call and
 * again whenever (mode, d_input, d_output) changes. */
BRIDCT_API int bridct_batch_graph_apply(bridct_batch *batch, int mode, const float *d_input,
                                        float *d_output);
/* Waits for the batch stream; BRIDCT_DEVICE_ERROR if a call since the last synchronize
 * failed on the device (a dependency wait of the persistent kernel that ran out). */
BRIDCT_API int bridct_batch_synchronize(bridct_batch *batch);
/* Host buffers of bridct_batch_elements floats each (disjoint); synchronous. */
BRIDCT_API int bridct_batch_apply(bridct_batch *batch, int mode, const float *input, float *output);
BRIDCT_API void *bridct_batch_stream(const bridct_batch *batch);
/* Bound on each dependency wait of the persistent kernel, in sleeps of the waiting
 * workgroup (default 2^21, about a second).  A wait that runs out sets the batch's error
 * flag (reported by bridct_batch_synchronize) and proceeds, so a call never hangs; a
 * small bound makes that path testable.  Takes effect from the next call. */
BRIDCT_API int bridct_batch_set_spin_limit(bridct_batch *batch, unsigned limit);

/* ---- 8x8 block codec step ----
 * No counterpart in the reference: the stage a block codec runs right after the forward
 * transform (encode) and right before the inverse (decode), fused with the 8x8 transform.
 *   encode: a u8 raster image, height x width pixels (both multiples of 8) with rows
 *           `stride` bytes apart -> for every 8x8 block, in raster order of the block
 *           grid, 64 int16 coefficients q[u*8+v] = round-half-even(Y[u][v] / Q[u*8+v]),
 *           saturated to int16, where Y is the orthonormal 2D DCT-II of the block minus
 *           128 (the plan FORWARD transform of an 8x8 array);
 *   decode: the coefficients -> X = orthonormal 2D DCT-III of (q * Q) (the plan INVERSE
 *           transform) -> pixels round-half-even(X + 128) saturated to 0..255.
 * Q (64 quantizer steps, row-major u*8+v) is fixed at creation; every step must be finite
 * and positive.  Arithmetic is FP32, with the transform normalization and 1/Q (or Q)
 * folded into one factor per coefficient, so a value within FP32 rounding of a tie may
 * round to either neighbour.  A codec object, like a plan, must not be used from two
 * threads at once. */
typedef struct bridct_codec8 bridct_codec8;
/* NULL when a step is not finite and positive, no GPU is usable, or allocation fails. */
BRIDCT_API bridct_codec8 *bridct_codec8_create(const float quant[64]);
BRIDCT_API void bridct_codec8_destroy(bridct_codec8 *codec);   /* NULL is a no-op */
/* Device buffers, asynchronous on the codec's stream: d_image 8-byte aligned with `stride`
 * a multiple of 8 and at least width; d_coeffs 16-byte aligned, (height/8)(width/8)*64
 * int16.  BRIDCT_INVALID_ARGUMENT for sizes that are zero or not multiples of 8,
 * BRIDCT_BAD_ALIGNMENT for misaligned pointers or stride. */
BRIDCT_API int bridct_codec8_encode_device(bridct_codec8 *codec, const unsigned char *d_image, size_t height,
                                           size_t width, size_t stride, short *d_coeffs);
BRIDCT_API int bridct_codec8_decode_device(bridct_codec8 *codec, const short *d_coeffs, size_t height,
                                           size_t width, size_t stride, unsigned char *d_image);
/* Host buffers, synchronous (copied through device buffers kept by the codec). */
BRIDCT_API int bridct_codec8_encode(bridct_codec8 *codec, const unsigned char *image, size_t height,
                                    size_t width, size_t stride, short *coeffs);
BRIDCT_API int bridct_codec8_decode(bridct_codec8 *codec, const short *coeffs, size_t height, size_t width,
                                    size_t stride, unsigned char *image);
BRIDCT_API int bridct_codec8_synchronize(bridct_codec8 *codec);
BRIDCT_API void *bridct_codec8_stream(const bridct_codec8 *codec);

/* Execution strategy of a batch.  PERSISTENT: one persistent kernel launch per direction
 * (the tiny kernel when every array is at most 32 x 32), butterfly route only.  GROUPED:
 * the plan path's kernels per distinct shape, reading the arrays of that shape through an
 * offset table, one or two launches per shape on the batch stream (one graph with
 * bridct_batch_graph_apply), with the route the schedule table gives each shape (the
 * split-MFMA GEMMs for a dense route, both axes >= 64); its results are bitwise those of
 * per-shape plans.  AUTO (the default at creation): GROUPED when the batch is not tiny and
 * either averages at least 2^21 pixels per distinct shape or has a shape the table sends
 * to a dense route, else PERSISTENT.  get returns the strategy in effect (never AUTO). */
#define BRIDCT_BATCH_AUTO 0
#define BRIDCT_BATCH_PERSISTENT 1
#define BRIDCT_BATCH_GROUPED 2
BRIDCT_API int bridct_batch_set_strategy(bridct_batch *batch, int strategy);
BRIDCT_API int bridct_batch_get_strategy(const bridct_batch *batch);

/* Failure injection, for tests of the allocation-failure paths (plan, batch and codec
 * creation, and the buffers some calls create on first use).  after >= 0: the library's
 * next `after` GPU-runtime acquisitions (device and pinned host buffers, streams, events)
 * succeed and the one after fails as if the runtime had refused it; then injection is off
 * again.  after < 0: off.  Returns the number of such resources the library holds at the
 * call, so a test can check that a failed creation released what it had acquired.
 * Process-wide; not for concurrent use with other library calls. */
BRIDCT_API long bridct_debug_alloc_faults(long after);

/* Test hook for the split of large launches: the whole-array kernels and the tiny batch
 * kernel launch at most 2^32 - 1 work-items per grid (the hardware limit; one 256-lane
 * workgroup per 8 KiB or so of arrays, so calls from about 128 GiB on split).  blocks > 0
 * lowers that to `blocks` workgroups per launch, so that a test can split small calls;
 * blocks <= 0 restores the hardware limit.  Returns the previous setting (0: hardware
 * limit).  Process-wide; not for concurrent use with other library calls.  No counterpart
 * in the reference (its CPU loops have no grid). */
BRIDCT_API long bridct_debug_grid_limit(long blocks);

/* Device memory helpers (synchronous). */
BRIDCT_API void *bridct_gpu_malloc(size_t bytes);
BRIDCT_API int bridct_gpu_free(void *d_ptr);
BRIDCT_API int bridct_gpu_upload(void *d_dst, const void *src, size_t bytes);
BRIDCT_API int bridct_gpu_download(void *dst, const void *d_src, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif
