/* Internal to the host library (not exported). */
#ifndef BRIDCT_HOST_INTERNAL_H
#define BRIDCT_HOST_INTERNAL_H

/* Fills t[0, BDX_TABLE_FLOATS) with the device constant tables of every axis length
 * 8..1024 (twiddles, tangents, per-axis normalization and its square, team twiddles),
 * evaluated in FP64 and rounded once. */
__attribute__((visibility("hidden"))) void bridct_build_tables(float *t);

#endif
