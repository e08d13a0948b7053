/* FP64 CPU oracle for the 2D DCT plan path.  Test infrastructure only: tests, the smoke
 * check and bench.py's CPU-baseline leg may call it; the product never links it. */
#ifndef BRIDCT_ORACLE_H
#define BRIDCT_ORACLE_H
#include <stddef.h>

int oracle_supported(int h, int w);
double oracle_sj_scale(int n, int k);
double oracle_axis_scale(int n, int k);
int oracle_norm_tables(int h, int w, float *norm, float *corr);
int oracle_qmatrix(int n, double *q);
int oracle_core(int n, int inverse, const double *x, double *y);
int oracle_dct2d(int h, int w, int mode, const double *x, double *y);

/* 8x8 block codec step in FP64 (the contract of bridct_codec8_* in bridct_gpu.h).
 * Raster: height x width u8 pixels, rows `stride` bytes apart; coefficients: 64 int16 per
 * block, blocks in raster order of the block grid.  `quotient` (Y / Q per coefficient) and
 * `value` (X + 128 per pixel, raster order with row pitch width) receive the unrounded
 * values when not NULL. */
int oracle_codec8_encode(const unsigned char *img, size_t height, size_t width, size_t stride,
                         const double *quant, double *quotient, short *coeffs);
int oracle_codec8_decode(const short *coeffs, size_t height, size_t width, size_t stride,
                         const double *quant, double *value, unsigned char *img);

#endif
