/* FP64 CPU oracle for the 2D DCT plan path (forward DCT-II, inverse DCT-III, round trip).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, the smoke check and bench.py's CPU-baseline leg
 * may load this library, and only as the checker.  The product path never calls it.
 *
 * Every function restates the reference's definition, cited as file:line of the
 * reference tree:
 *   - transform and normalization: docs/API.md:7-13 (Q_n), dataset/corpus.py:10-12;
 *   - shape support: src/plan.c:33-37;
 *   - Shao-Johnson scale s_{N,k}: src/plan.c:39-47;
 *   - per-plan normalization/correction tables: src/plan.c:86-109;
 *   - scaled 1D core F_n = diag(1/s_{4n,k}) C_II (what the generated butterflies compute
 *     before normalization): src/plan.c:38 comment and :96.
 * Cosines are evaluated as cos(pi*m/(2n)) with m = k*(2j+1) reduced modulo 4n, so the
 * argument stays in [0, 2*pi) and libm's double cosine is accurate to an ulp.
 */
#include "bridct_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

static const double PI = 3.14159265358979323846;

/* src/plan.c:33-37 */
int oracle_supported(int h, int w)
{
    return h >= 8 && h <= 1024 && w >= 8 && w <= 1024 && !(h & (h - 1)) && !(w & (w - 1));
}

/* src/plan.c:39-47: s_{N,k} = 1 for N <= 4, else
 * s_{N/4, k mod N/4} * (k mod N/4 <= N/8 ? cos(2 pi k4/N) : sin(2 pi k4/N)). */
double oracle_sj_scale(int n, int k)
{
    if (n <= 4)
        return 1.0;
    k %= n / 4;
    double angle = 2.0 * PI * k / n;
    return oracle_sj_scale(n / 4, k) * (k <= n / 8 ? cos(angle) : sin(angle));
}

/* src/plan.c:96: nu_n[k] = s_{4n,k} * sqrt((k ? 2 : 1) / n). */
double oracle_axis_scale(int n, int k)
{
    return oracle_sj_scale(4 * n, k) * sqrt((k ? 2.0 : 1.0) / n);
}

/* src/plan.c:86-109: normalization[r][c] = (float)(nu_h[r] nu_w[c]) and
 * correction = (float)(that product squared), each rounded once from double. */
int oracle_norm_tables(int h, int w, float *norm, float *corr)
{
    if (!oracle_supported(h, w))
        return -1;
    for (int r = 0; r < h; r++) {
        double a = 1.0 / oracle_axis_scale(h, r);
        for (int c = 0; c < w; c++) {
            double b = 1.0 / oracle_axis_scale(w, c);
            double s = 1.0 / (a * b);
            if (norm)
                norm[(size_t)r * w + c] = (float)s;
            if (corr)
                corr[(size_t)r * w + c] = (float)(s * s);
        }
    }
    return 0;
}

static double cos_term(int n, long k, long j)
{
    long m = (k * (2 * j + 1)) % (4L * n);
    return cos(PI * (double)m / (2.0 * n));
}

/* docs/API.md:11: Q_n[k,j] = sqrt((k ? 2 : 1)/n) cos(pi k (j + 1/2) / n), row-major n x n. */
int oracle_qmatrix(int n, double *q)
{
    if (n < 1 || !q)
        return -1;
    for (int k = 0; k < n; k++) {
        double a = sqrt((k ? 2.0 : 1.0) / n);
        for (int j = 0; j < n; j++)
            q[(size_t)k * n + j] = a * cos_term(n, k, j);
    }
    return 0;
}

/* Scaled 1D core: forward y[k] = sum_j cos(pi k (2j+1)/(2n)) x[j] / s_{4n,k};
 * inverse is the transpose, y[j] = sum_k cos(...) x[k] / s_{4n,k}. */
int oracle_core(int n, int inverse, const double *x, double *y)
{
    if (n < 1 || !x || !y)
        return -1;
    for (int o = 0; o < n; o++) {
        double acc = 0.0;
        for (int i = 0; i < n; i++) {
            int k = inverse ? i : o, j = inverse ? o : i;
            acc += cos_term(n, k, j) * x[i] / oracle_sj_scale(4 * n, k);
        }
        y[o] = acc;
    }
    return 0;
}

/* 2D closed form on an h x w row-major array (docs/API.md:9-12):
 * mode 0: Q_h X Q_w^T; mode 1: Q_h^T X Q_w; mode 2: the identity composition, whose exact
 * value is X (tests/verify_portability.py:76-77 uses the input itself as the target). */
int oracle_dct2d(int h, int w, int mode, const double *x, double *y)
{
    if (!oracle_supported(h, w) || mode < 0 || mode > 2 || !x || !y)
        return -1;
    size_t hw = (size_t)h * w;
    if (mode == 2) {
        memcpy(y, x, hw * sizeof(double));
        return 0;
    }
    double *qh = malloc((size_t)h * h * sizeof(double));
    double *qw = malloc((size_t)w * w * sizeof(double));
    double *t = malloc(hw * sizeof(double));
    if (!qh || !qw || !t) {
        free(qh);
        free(qw);
        free(t);
        return -2;
    }
    oracle_qmatrix(h, qh);
    oracle_qmatrix(w, qw);
    /* rows: t[r][c] = sum_j x[r][j] * A_w[c][j], with A_w = Q_w (fwd) or Q_w^T (inv) */
    for (int r = 0; r < h; r++)
        for (int c = 0; c < w; c++) {
            double acc = 0.0;
            for (int j = 0; j < w; j++)
                acc += x[(size_t)r * w + j] * (mode ? qw[(size_t)j * w + c] : qw[(size_t)c * w + j]);
            t[(size_t)r * w + c] = acc;
        }
    /* columns: y[k][c] = sum_r A_h[k][r] * t[r][c] */
    for (int k = 0; k < h; k++)
        for (int c = 0; c < w; c++) {
            double acc = 0.0;
            for (int r = 0; r < h; r++)
                acc += (mode ? qh[(size_t)r * h + k] : qh[(size_t)k * h + r]) * t[(size_t)r * w + c];
            y[(size_t)k * w + c] = acc;
        }
    free(qh);
    free(qw);
    free(t);
    return 0;
}

/* ---- 8x8 block codec step ------------------------------------------------------------
 * Not a reference algorithm: the caller-side stage of a block codec around the 8x8 plan
 * transform (docs/API.md:9-12 for the transform).  Quantization as in baseline JPEG
 * (ITU-T T.81, A.3.4): coefficient Y / Q rounded to nearest, ties to even here, saturated
 * to int16; decode multiplies back, inverts, adds 128, rounds and saturates to 0..255. */
static double round_half_even(double v)
{
    double f = floor(v), d = v - f;
    if (d > 0.5)
        return f + 1.0;
    if (d < 0.5)
        return f;
    return fmod(f, 2.0) == 0.0 ? f : f + 1.0;
}

static double clampd(double v, double lo, double hi)
{
    return v < lo ? lo : v > hi ? hi : v;
}

int oracle_codec8_encode(const unsigned char *img, size_t height, size_t width, size_t stride,
                         const double *quant, double *quotient, short *coeffs)
{
    if (!img || !quant || !coeffs || height % 8 || width % 8 || stride < width)
        return -1;
    double q8[64], x[64], y[64];
    oracle_qmatrix(8, q8);
    const size_t bw = width / 8, nb = (height / 8) * bw;
    for (size_t b = 0; b < nb; b++) {
        const size_t by = b / bw, bx = b % bw;
        for (int r = 0; r < 8; r++)
            for (int c = 0; c < 8; c++)
                x[r * 8 + c] = (double)img[(by * 8 + r) * stride + bx * 8 + c] - 128.0;
        /* Y = Q X Q^T */
        for (int u = 0; u < 8; u++)
            for (int v = 0; v < 8; v++) {
                double acc = 0.0;
                for (int r = 0; r < 8; r++)
                    for (int c = 0; c < 8; c++)
                        acc += q8[u * 8 + r] * x[r * 8 + c] * q8[v * 8 + c];
                y[u * 8 + v] = acc;
            }
        for (int i = 0; i < 64; i++) {
            const double t = y[i] / quant[i];
            if (quotient)
                quotient[b * 64 + i] = t;
            coeffs[b * 64 + i] = (short)clampd(round_half_even(t), -32768.0, 32767.0);
        }
    }
    return 0;
}

int oracle_codec8_decode(const short *coeffs, size_t height, size_t width, size_t stride,
                         const double *quant, double *value, unsigned char *img)
{
    if (!img || !quant || !coeffs || height % 8 || width % 8 || stride < width)
        return -1;
    double q8[64], y[64];
    oracle_qmatrix(8, q8);
    const size_t bw = width / 8, nb = (height / 8) * bw;
    for (size_t b = 0; b < nb; b++) {
        const size_t by = b / bw, bx = b % bw;
        for (int i = 0; i < 64; i++)
            y[i] = (double)coeffs[b * 64 + i] * quant[i];
        /* X = Q^T Y Q */
        for (int r = 0; r < 8; r++)
            for (int c = 0; c < 8; c++) {
                double acc = 0.0;
                for (int u = 0; u < 8; u++)
                    for (int v = 0; v < 8; v++)
                        acc += q8[u * 8 + r] * y[u * 8 + v] * q8[v * 8 + c];
                const double t = acc + 128.0;
                const size_t row = by * 8 + r, col = bx * 8 + c;
                if (value)
                    value[row * width + col] = t;
                img[row * stride + col] = (unsigned char)clampd(round_half_even(t), 0.0, 255.0);
            }
    }
    return 0;
}
