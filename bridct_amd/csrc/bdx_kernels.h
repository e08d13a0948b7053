// Launchers of the per-shape kernels (host side of the HIP translation units).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "bridct_device.h"

namespace bdx {

// Line operations: FA forward * alpha, FA2 forward * alpha^2, IA alpha * inverse, I0 inverse.
enum Op { FA = 1, FA2 = 2, IA = 3, I0 = 4 };

constexpr int kThreads = 256;
// XCD count of the current device, read once per device (the two-pass kernels order their
// workgroups by XCD; each XCD has its own L2).
unsigned xcd_count();
// Most workgroups of `threads` lanes one launch may have: 2^32 - 1 work-items per grid
// dimension, or the lower limit bdx_debug_grid_limit set.  Launchers whose grid grows with
// the batch (whole-array kernels, the tiny batch kernel) split larger calls.
size_t max_grid_blocks(int threads);
// Floats of LDS for one workgroup of the whole-array kernel (144 KiB of the 160 KiB).
constexpr int kSmallLds = BDX_SMALL_LDS_FLOATS;
// Floats of LDS per workgroup of the row / column pass kernels (two workgroups per CU).
constexpr int kPassLds = 16384;

// Whole-array ("small") route: h * (w + 1) <= kSmallLds.
constexpr bool small_route(int h, int w) { return (long)h * (w + 1) <= kSmallLds; }

// Rows per block of the row pass and columns per strip of the column pass.
constexpr int rows_per_block(int w) { return kPassLds / w; }
constexpr int cols_per_block(int h) { return kPassLds / h > 256 ? 256 : kPassLds / h; }

// mode: 0 forward, 1 inverse, 2 fused round trip.  Return 0 or a negative value.  With
// `offs`, array i of the batch is at element offs[i] of in and out (gather form) instead of
// at i * h * w.
int launch_small_h8(int w, const float *in, float *out, size_t batch, int pack, int mode, const float *tab,
                      hipStream_t s, const size_t *offs = nullptr);
int launch_small_h16(int w, const float *in, float *out, size_t batch, int pack, int mode, const float *tab,
                      hipStream_t s, const size_t *offs = nullptr);
int launch_small_h32(int w, const float *in, float *out, size_t batch, int pack, int mode, const float *tab,
                      hipStream_t s, const size_t *offs = nullptr);
int launch_small_h64(int w, const float *in, float *out, size_t batch, int pack, int mode, const float *tab,
                      hipStream_t s, const size_t *offs = nullptr);
int launch_small_h128(int w, const float *in, float *out, size_t batch, int pack, int mode, const float *tab,
                      hipStream_t s, const size_t *offs = nullptr);
int launch_small_h256(int w, const float *in, float *out, size_t batch, int pack, int mode, const float *tab,
                      hipStream_t s, const size_t *offs = nullptr);
int launch_small_h512(int w, const float *in, float *out, size_t batch, int pack, int mode, const float *tab,
                      hipStream_t s, const size_t *offs = nullptr);
int launch_small_h1024(int w, const float *in, float *out, size_t batch, int pack, int mode, const float *tab,
                      hipStream_t s, const size_t *offs = nullptr);

// Row pass over nrows rows of length w: op 1 = forward*alpha, 2 = forward*alpha^2,
// 3 = alpha*inverse, 4 = inverse.
// nt: non-temporal (streaming) loads and stores of the array data.
int launch_rows(int w, const float *in, float *out, size_t nrows, int op, const float *tab, hipStream_t s,
                const size_t *offs = nullptr, int h = 0, bool nt = false);
// Column pass over `batch` h x w arrays, in place allowed: mode 0 forward*alpha,
// 1 alpha*inverse, 2 forward*alpha^2 then inverse.
int launch_cols(int h, int w, const float *in, float *out, size_t batch, int mode, const float *tab,
                hipStream_t s, const size_t *offs = nullptr, bool nt = false);
// The same for lines of 8, 16 and 32 (bdx_passes_lines.hip); launch_rows and launch_cols
// hand those lengths on to them.
int launch_rows_lines(int w, const float *in, float *out, size_t nrows, int op, const float *tab, hipStream_t s,
                      const size_t *offs, int h, bool nt);
int launch_cols_lines(int h, int w, const float *in, float *out, size_t batch, int mode, const float *tab,
                      hipStream_t s, const size_t *offs, bool nt);

// Persistent batch kernel, one per mode (bdx_batch.hip); flags: BDX_LAUNCH_*.
int launch_batch_m0(int flags, const float *in, float *out, const bdx_run *runs, const bdx_item *items, unsigned nitems,
                    const bdx_unit *units, unsigned nunits, unsigned qbase, unsigned *ctr, const float *tab,
                    unsigned blocks, unsigned spin_limit, hipStream_t s);
int launch_batch_m1(int flags, const float *in, float *out, const bdx_run *runs, const bdx_item *items, unsigned nitems,
                    const bdx_unit *units, unsigned nunits, unsigned qbase, unsigned *ctr, const float *tab,
                    unsigned blocks, unsigned spin_limit, hipStream_t s);
int launch_batch_m2(int flags, const float *in, float *out, const bdx_run *runs, const bdx_item *items, unsigned nitems,
                    const bdx_unit *units, unsigned nunits, unsigned qbase, unsigned *ctr, const float *tab,
                    unsigned blocks, unsigned spin_limit, hipStream_t s);

}  // namespace bdx
