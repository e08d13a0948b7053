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

/* Transform modes seen by the kernels (same values as enum bridct_mode). */
#define BDX_FORWARD 0
#define BDX_INVERSE 1
#define BDX_ROUNDTRIP 2

int bdx_device_ready(void);
int bdx_malloc(void **ptr, size_t bytes);
int bdx_free(void *ptr);
int bdx_stream_create(void **stream);
int bdx_stream_destroy(void *stream);
int bdx_stream_sync(void *stream);
int bdx_copy_h2d(void *dst, const void *src, size_t bytes, void *stream);
int bdx_copy_d2h(void *dst, const void *src, size_t bytes, void *stream);
int bdx_copy_d2d(void *dst, const void *src, size_t bytes, void *stream);

/* Orthonormal 2D transform of `batch` contiguous row-major h x w float32 arrays.
 * d_in and d_out must be 16-byte aligned and either identical or disjoint. */
int bdx_transform(const float *d_in, float *d_out, int h, int w, size_t batch, int mode,
                  const float *d_tables, void *stream);

#ifdef __cplusplus
}
#endif
#endif
