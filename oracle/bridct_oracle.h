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

#endif
