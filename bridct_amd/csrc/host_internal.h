/* Internal to the host library (not exported). */
#ifndef BRIDCT_HOST_INTERNAL_H
#define BRIDCT_HOST_INTERNAL_H

/* Fills t[0, BDX_TABLE_FLOATS) with the device constant tables of every axis length
 * 8..1024 (twiddles, tangents, per-axis normalization and its square, team twiddles),
 * evaluated in FP64 and rounded once. */
__attribute__((visibility("hidden"))) void bridct_build_tables(float *t);

/* Per-axis normalization of the length-n core, output index k, in FP64 (the values the
 * device tables hold rounded to FP32). */
__attribute__((visibility("hidden"))) double bridct_axis_alpha(int n, int k);

/* Dense route: s_n = 2^floor(log2 sqrt n), and Q_n' = s_n Q_n (n x n, row-major, FP64
 * rounded once) with, when qt is not NULL, its transpose. */
__attribute__((visibility("hidden"))) int bridct_dense_scale(int n);
__attribute__((visibility("hidden"))) void bridct_dense_matrix(int n, float *q, float *qt);

/* Route and whole-array packing of shape (h, w) from the schedule table; 0 if found. */
__attribute__((visibility("hidden"))) int bridct_schedule_lookup(int h, int w, int *route, int *pack);

/* The same lookup in the table file `path` (exported for tests of the table format). */
int bdx_schedule_lookup_file(const char *path, int h, int w, int *route, int *pack);

#endif
