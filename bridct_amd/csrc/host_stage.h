/* Staged host <-> device copies for large host-pointer calls: the data goes through two
 * pinned chunk buffers, the host side of each chunk copied by a copy_pool (copy_pool.h),
 * while the DMA engine moves the previous chunk.  Pageable copies run at 31 GB/s (in) and
 * 18 GB/s (out) on MI355X hosts; pinned ones at 57 GB/s each way. */
#pragma once
#include <stddef.h>

typedef struct host_stage host_stage;

/* Two pinned buffers of chunk_bytes (a multiple of 256), two events and a copy pool;
 * NULL on failure. */
__attribute__((visibility("hidden"))) host_stage *stage_create(size_t chunk_bytes);

/* Waits for the stage's copies in flight, then frees it; NULL is a no-op. */
__attribute__((visibility("hidden"))) void stage_destroy(host_stage *s);

/* Copies bytes from host memory to the device on `stream`.  Returns once the host buffer
 * has been read; the last copies may still be in flight (work issued later on `stream`
 * is ordered after them).  0 or -1. */
__attribute__((visibility("hidden"))) int stage_h2d(host_stage *s, void *d_dst, const void *h_src, size_t bytes,
                                                     void *stream);

/* Copies bytes from the device to host memory after the work issued so far on `stream`;
 * returns once the host buffer holds them.  0 or -1. */
__attribute__((visibility("hidden"))) int stage_d2h(host_stage *s, void *h_dst, const void *d_src, size_t bytes,
                                                     void *stream);
