// HIP kernels of the MI355X BRiDCT 2D plan path (gfx950, wave64).
//
// Reference path: src/plan.c:126-163 (bridct_plan_apply) -> src/rect.c:5-23 (column pass,
// banded row pass, normalization) and src/generated/*.c (Shao-Johnson 1D cores).  The
// transform is the orthonormal 2D DCT-II (forward), DCT-III (inverse) and the fused
// normalized round trip (docs/API.md:7-13).
//
// Layout: arrays are row-major h x w float32, a batch is contiguous.  Every kernel stages
// an LDS tile (pitch = line length + 1 floats, which keeps both row- and column-wise lane
// accesses on distinct banks) and runs 1D "lines" over it:
//   * n <= 64: one lane per line, the generated Shao-Johnson core sjf_n / sjt_n;
//   * n >= 128: a team of T lanes of one wave per line (T = 4, 8, 8, 16 for
//     n = 128, 256, 512, 1024): Makhoul packing into N = n/2 complex points, a
//     Cooley-Tukey split N = N1 * T with Johnson-Frigo split-radix codelets (dft_N1 in
//     each lane, twiddle, in-place LDS exchange, dft_T), then the real-to-DCT
//     post-process with the tangent twiddle.  Each lane owns the k1 pairs {k1, N1-k1}, so
//     the post-process pairs (k, N-k) never leave the lane.
// The per-axis normalization alpha (computed in FP64 on the host and rounded once) also
// absorbs each core's own output scaling; forward multiplies it on the way out of a line,
// inverse on the way in, and the fused round trip uses alpha^2 between the two halves
// (reference plan.c:86-109, "U = E H, so H^T E^2 H").
//
// Kernels: k_fused (whole arrays, up to 128 KiB each, several small arrays per
// workgroup), k_rows (bands of full rows) and k_cols (strips of full columns, in place)
// for larger arrays.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "bridct_device.h"
#include "generated/sj_codelets.h"

namespace {

constexpr int kThreads = 256;
constexpr int kLdsFloats = 36864;   // 144 KiB of the 160 KiB LDS per CU

enum Op { F0 = 0, FA = 1, FA2 = 2, IA = 3, I0 = 4 };

template <int N, int S> struct Dft;
#define BDX_DFT(N)                                                                         \
    template <> struct Dft<N, -1> {                                                       \
        static __device__ __forceinline__ void run(float (&r)[N], float (&i)[N]) { dft_##N##_m(r, i); } \
    };                                                                                     \
    template <> struct Dft<N, 1> {                                                        \
        static __device__ __forceinline__ void run(float (&r)[N], float (&i)[N]) { dft_##N##_p(r, i); } \
    };
BDX_DFT(2)
BDX_DFT(4)
BDX_DFT(8)
BDX_DFT(16)
BDX_DFT(32)

template <int n> struct Sj;
#define BDX_SJ(n)                                                                          \
    template <> struct Sj<n> {                                                            \
        static __device__ __forceinline__ void fwd(float (&x)[n]) { sjf_##n(x); }          \
        static __device__ __forceinline__ void inv(float (&x)[n]) { sjt_##n(x); }          \
    };
BDX_SJ(8)
BDX_SJ(16)
BDX_SJ(32)
BDX_SJ(64)

template <int n> constexpr int team_size() { return n <= 64 ? 1 : n >= 1024 ? 16 : n >= 256 ? 8 : 4; }

// Lanes of one team live in one wave, so LDS traffic between them only needs ordering
// within the wave.
__device__ __forceinline__ void team_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ const float *alpha_of(const float *tab, int n, int which)
{
    return tab + BDX_TABLE_BASE(n) + 3 * n + (which == 2 ? n : 0);
}

// ---- per-lane Shao-Johnson line (n <= 64) -----------------------------------------------

template <int n, int OP>
__device__ __forceinline__ void sj_line(float *p, int s, const float *__restrict__ tab)
{
    float x[n];
#pragma unroll
    for (int i = 0; i < n; i++) {
        float v = p[i * s];
        if (OP == IA)
            v *= alpha_of(tab, n, 1)[i];
        x[i] = v;
    }
    if (OP == IA || OP == I0)
        Sj<n>::inv(x);
    else
        Sj<n>::fwd(x);
#pragma unroll
    for (int i = 0; i < n; i++) {
        float v = x[i];
        if (OP == FA || OP == FA2)
            v *= alpha_of(tab, n, OP == FA ? 1 : 2)[i];
        p[i * s] = v;
    }
}

// ---- team line (n >= 128) --------------------------------------------------------------

// Output index of slot (k2, which) of group g in lane-owned order.
template <int n, int T>
__device__ __forceinline__ int team_out_index(bool g0, int g, int k2, int which)
{
    constexpr int N = n / 2, N1 = N / T;
    if (g0 && k2 == T / 2) {
        const int sp[4] = {0, N, n / 4, 3 * n / 4};
        return sp[which];
    }
    const int k = g0 ? (k2 < T / 2 ? N1 / 2 + N1 * k2 : N1 * (k2 - T / 2)) : g + N1 * k2;
    return which == 0 ? k : which == 1 ? n - k : which == 2 ? N - k : N + k;
}

template <int n, int OP>
__device__ __forceinline__ void team_fwd(float *p, int s, int t, const float *__restrict__ tab)
{
    constexpr int T = team_size<n>(), N = n / 2, N1 = N / T, Q = N1 / (2 * T);
    const float *om = tab + BDX_TABLE_BASE(n);
    const float *tn = om + 2 * n;
    float re[N1], im[N1];
#pragma unroll
    for (int j1 = 0; j1 < N1; j1++) {
        const int m = t + T * j1;
        const int i0 = j1 < N1 / 2 ? 4 * m : 2 * n - 4 * m - 1;
        const int i1 = j1 < N1 / 2 ? i0 + 2 : i0 - 2;
        re[j1] = p[i0 * s];
        im[j1] = p[i1 * s];
    }
    team_sync();
    Dft<N1, -1>::run(re, im);
#pragma unroll
    for (int k1 = 1; k1 < N1; k1++) {
        const int e = (2 * t * k1) & (n - 1);
        const float wr = om[2 * e], wi = om[2 * e + 1];
        const float a = re[k1], b = im[k1];
        re[k1] = __builtin_fmaf(a, wr, -b * wi);
        im[k1] = __builtin_fmaf(a, wi, b * wr);
    }
#pragma unroll
    for (int k1 = 0; k1 < N1; k1++) {
        const int c = 2 * (k1 * T + t);
        p[c * s] = re[k1];
        p[(c + 1) * s] = im[k1];
    }
    team_sync();
    float o[Q][T][4];
#pragma unroll
    for (int r = 0; r < Q; r++) {
        const int g = t + T * r;
        const bool g0 = g == 0;
        const int ka = g0 ? 0 : g, kb = g0 ? N1 / 2 : N1 - g;
        float d0r[T], d0i[T], d1r[T], d1i[T];
#pragma unroll
        for (int j2 = 0; j2 < T; j2++) {
            d0r[j2] = p[2 * (ka * T + j2) * s];
            d0i[j2] = p[(2 * (ka * T + j2) + 1) * s];
            d1r[j2] = p[2 * (kb * T + j2) * s];
            d1i[j2] = p[(2 * (kb * T + j2) + 1) * s];
        }
        Dft<T, -1>::run(d0r, d0i);
        Dft<T, -1>::run(d1r, d1i);
#pragma unroll
        for (int k2 = 0; k2 < T; k2++) {
            float ar, ai, br, bi;
            int k;
            if (k2 < T / 2) {
                ar = g0 ? d1r[k2] : d0r[k2];
                ai = g0 ? d1i[k2] : d0i[k2];
                br = d1r[T - 1 - k2];
                bi = d1i[T - 1 - k2];
                k = g0 ? N1 / 2 + N1 * k2 : g + N1 * k2;
            } else if (k2 > T / 2) {
                const int j = k2 - T / 2;
                ar = g0 ? d0r[j] : d0r[k2];
                ai = g0 ? d0i[j] : d0i[k2];
                br = g0 ? d0r[T - j] : d1r[T - 1 - k2];
                bi = g0 ? d0i[T - j] : d1i[T - 1 - k2];
                k = g0 ? N1 * j : g + N1 * k2;
            } else {
                ar = d0r[k2];
                ai = d0i[k2];
                br = d1r[T - 1 - k2];
                bi = d1i[T - 1 - k2];
                k = g + N1 * k2;   // garbage for g0, replaced below
            }
            const float er = ar + br, ei = ai - bi, orr = ai + bi, oi = br - ar;
            const float wr = om[2 * k], wi = om[2 * k + 1];
            const float P = __builtin_fmaf(wr, orr, -wi * oi);
            const float Qv = __builtin_fmaf(wr, oi, wi * orr);
            const float vr = er + P, vi = ei + Qv, ur = er - P, ui = Qv - ei;
            const float t1 = tn[k], t2 = tn[N - k];
            float o0 = __builtin_fmaf(t1, vi, vr);
            float o1 = __builtin_fmaf(t1, vr, -vi);
            float o2 = __builtin_fmaf(t2, ui, ur);
            float o3 = __builtin_fmaf(t2, ur, -ui);
            if (k2 == T / 2) {
                // lane g = 0: Z_0 gives outputs 0 and N, Z_{N/2} gives n/4 and 3n/4
                const float th = tn[n / 4];
                const float z0r = d0r[0], z0i = d0i[0], zhr = d0r[T / 2], zhi = d0i[T / 2];
                o0 = g0 ? z0r + z0i : o0;
                o1 = g0 ? z0r - z0i : o1;
                o2 = g0 ? __builtin_fmaf(-th, zhi, zhr) : o2;
                o3 = g0 ? __builtin_fmaf(th, zhr, zhi) : o3;
            }
            o[r][k2][0] = o0;
            o[r][k2][1] = o1;
            o[r][k2][2] = o2;
            o[r][k2][3] = o3;
        }
    }
    team_sync();
#pragma unroll
    for (int r = 0; r < Q; r++) {
        const int g = t + T * r;
#pragma unroll
        for (int k2 = 0; k2 < T; k2++)
#pragma unroll
            for (int w = 0; w < 4; w++) {
                const int idx = team_out_index<n, T>(g == 0, g, k2, w);
                float v = o[r][k2][w];
                if (OP == FA || OP == FA2)
                    v *= alpha_of(tab, n, OP == FA ? 1 : 2)[idx];
                p[idx * s] = v;
            }
    }
}

template <int n, int OP>
__device__ __forceinline__ void team_inv(float *p, int s, int t, const float *__restrict__ tab)
{
    constexpr int T = team_size<n>(), N = n / 2, N1 = N / T, Q = N1 / (2 * T);
    const float *om = tab + BDX_TABLE_BASE(n);
    const float *tn = om + 2 * n;
    float o[Q][T][4];
#pragma unroll
    for (int r = 0; r < Q; r++) {
        const int g = t + T * r;
#pragma unroll
        for (int k2 = 0; k2 < T; k2++)
#pragma unroll
            for (int w = 0; w < 4; w++) {
                const int idx = team_out_index<n, T>(g == 0, g, k2, w);
                float v = p[idx * s];
                if (OP == IA)
                    v *= alpha_of(tab, n, 1)[idx];
                o[r][k2][w] = v;
            }
    }
    team_sync();
#pragma unroll
    for (int r = 0; r < Q; r++) {
        const int g = t + T * r;
        const bool g0 = g == 0;
        const int ka = g0 ? 0 : g, kb = g0 ? N1 / 2 : N1 - g;
        float zar[T], zai[T], zbr[T], zbi[T];
        float z0r = 0.f, z0i = 0.f, zhr = 0.f, zhi = 0.f;
#pragma unroll
        for (int k2 = 0; k2 < T; k2++) {
            int k;
            if (k2 < T / 2)
                k = g0 ? N1 / 2 + N1 * k2 : g + N1 * k2;
            else if (k2 > T / 2)
                k = g0 ? N1 * (k2 - T / 2) : g + N1 * k2;
            else
                k = g + N1 * k2;
            const float o0 = o[r][k2][0], o1 = o[r][k2][1], o2 = o[r][k2][2], o3 = o[r][k2][3];
            const float t1 = tn[k], t2 = tn[N - k];
            const float vr = __builtin_fmaf(t1, o1, o0), vi = __builtin_fmaf(t1, o0, -o1);
            const float ur = __builtin_fmaf(t2, o3, o2), ui = __builtin_fmaf(t2, o2, -o3);
            const float er = vr + ur, P = vr - ur, ei = vi - ui, Qv = vi + ui;
            const float wr = om[2 * k], wi = om[2 * k + 1];
            const float orr = __builtin_fmaf(wr, P, wi * Qv);
            const float oi = __builtin_fmaf(wr, Qv, -wi * P);
            zar[k2] = er - oi;
            zai[k2] = ei + orr;
            zbr[k2] = er + oi;
            zbi[k2] = orr - ei;
            if (k2 == T / 2) {
                const float th = tn[n / 4];
                z0r = o0 + o1;
                z0i = o0 - o1;
                zhr = __builtin_fmaf(th, o3, o2);
                zhi = __builtin_fmaf(-th, o2, o3);
            }
        }
        float d0r[T], d0i[T], d1r[T], d1i[T];
#pragma unroll
        for (int x = 0; x < T; x++) {
            float gr, gi;   // lane g = 0 mapping of D0[x]
            if (x == 0) {
                gr = z0r;
                gi = z0i;
            } else if (x == T / 2) {
                gr = zhr;
                gi = zhi;
            } else if (x < T / 2) {
                gr = zar[T / 2 + x];
                gi = zai[T / 2 + x];
            } else {
                gr = zbr[T / 2 + T - x];
                gi = zbi[T / 2 + T - x];
            }
            d0r[x] = g0 ? gr : zar[x];
            d0i[x] = g0 ? gi : zai[x];
            if (x < T / 2) {
                d1r[x] = g0 ? zar[x] : zbr[T - 1 - x];
                d1i[x] = g0 ? zai[x] : zbi[T - 1 - x];
            } else {
                d1r[x] = zbr[T - 1 - x];
                d1i[x] = zbi[T - 1 - x];
            }
        }
        Dft<T, 1>::run(d0r, d0i);
        Dft<T, 1>::run(d1r, d1i);
#pragma unroll
        for (int j2 = 0; j2 < T; j2++) {
            p[2 * (ka * T + j2) * s] = d0r[j2];
            p[(2 * (ka * T + j2) + 1) * s] = d0i[j2];
            p[2 * (kb * T + j2) * s] = d1r[j2];
            p[(2 * (kb * T + j2) + 1) * s] = d1i[j2];
        }
    }
    team_sync();
    float re[N1], im[N1];
#pragma unroll
    for (int k1 = 0; k1 < N1; k1++) {
        const int c = 2 * (k1 * T + t);
        re[k1] = p[c * s];
        im[k1] = p[(c + 1) * s];
    }
    team_sync();
#pragma unroll
    for (int k1 = 1; k1 < N1; k1++) {
        const int e = (2 * t * k1) & (n - 1);
        const float wr = om[2 * e], wi = om[2 * e + 1];
        const float a = re[k1], b = im[k1];
        re[k1] = __builtin_fmaf(a, wr, b * wi);
        im[k1] = __builtin_fmaf(b, wr, -a * wi);
    }
    Dft<N1, 1>::run(re, im);
#pragma unroll
    for (int j1 = 0; j1 < N1; j1++) {
        const int m = t + T * j1;
        const int i0 = j1 < N1 / 2 ? 4 * m : 2 * n - 4 * m - 1;
        const int i1 = j1 < N1 / 2 ? i0 + 2 : i0 - 2;
        p[i0 * s] = re[j1];
        p[i1 * s] = im[j1];
    }
}

template <int n, int OP>
__device__ __forceinline__ void line_op(float *p, int s, int t, const float *__restrict__ tab)
{
    if constexpr (n <= 64)
        sj_line<n, OP>(p, s, tab);
    else if constexpr (OP == IA || OP == I0)
        team_inv<n, OP>(p, s, t, tab);
    else
        team_fwd<n, OP>(p, s, t, tab);
}

// Lines of one pass over an LDS tile.  Row lines: line L at tile + L * pitch, stride 1.
// Column lines: line L = (array a, column c) at tile + a * n * pitch + c, stride pitch.
template <int n, int OP, int OP2>
__device__ __forceinline__ void run_lines(float *tile, int nlines, bool cols, int pitch, int w,
                                          const float *__restrict__ tab)
{
    constexpr int T = team_size<n>();
    constexpr int teams = kThreads / T;
    const int team = threadIdx.x / T, t = threadIdx.x % T;
    for (int L = team; L < nlines; L += teams) {
        float *p;
        int s;
        if (cols) {
            const int a = L / w, c = L - a * w;
            p = tile + a * n * pitch + c;
            s = pitch;
        } else {
            p = tile + L * pitch;
            s = 1;
        }
        line_op<n, OP>(p, s, t, tab);
        if constexpr (OP2 >= 0) {
            team_sync();
            line_op<n, OP2>(p, s, t, tab);
        }
    }
}

template <int OP, int OP2 = -1>
__device__ __forceinline__ void pass(float *tile, int n, int nlines, bool cols, int pitch, int w,
                                     const float *__restrict__ tab)
{
    switch (n) {
    case 8: run_lines<8, OP, OP2>(tile, nlines, cols, pitch, w, tab); break;
    case 16: run_lines<16, OP, OP2>(tile, nlines, cols, pitch, w, tab); break;
    case 32: run_lines<32, OP, OP2>(tile, nlines, cols, pitch, w, tab); break;
    case 64: run_lines<64, OP, OP2>(tile, nlines, cols, pitch, w, tab); break;
    case 128: run_lines<128, OP, OP2>(tile, nlines, cols, pitch, w, tab); break;
    case 256: run_lines<256, OP, OP2>(tile, nlines, cols, pitch, w, tab); break;
    case 512: run_lines<512, OP, OP2>(tile, nlines, cols, pitch, w, tab); break;
    default: run_lines<1024, OP, OP2>(tile, nlines, cols, pitch, w, tab); break;
    }
}

// Copies `rows` rows of `cols` floats (global row stride gstride) into LDS (pitch) and back.
__device__ __forceinline__ void tile_load(float *tile, int pitch, const float *__restrict__ src,
                                          size_t gstride, int rows, int cols)
{
    const int per_row = cols / 4, nv = rows * per_row;
    for (int i = threadIdx.x; i < nv; i += kThreads) {
        const int r = i / per_row, c = (i - r * per_row) * 4;
        const float4 v = *reinterpret_cast<const float4 *>(src + r * gstride + c);
        float *d = tile + r * pitch + c;
        d[0] = v.x;
        d[1] = v.y;
        d[2] = v.z;
        d[3] = v.w;
    }
}

__device__ __forceinline__ void tile_store(const float *tile, int pitch, float *__restrict__ dst,
                                           size_t gstride, int rows, int cols)
{
    const int per_row = cols / 4, nv = rows * per_row;
    for (int i = threadIdx.x; i < nv; i += kThreads) {
        const int r = i / per_row, c = (i - r * per_row) * 4;
        const float *d = tile + r * pitch + c;
        *reinterpret_cast<float4 *>(dst + r * gstride + c) = make_float4(d[0], d[1], d[2], d[3]);
    }
}

// Whole arrays in one workgroup: `per` arrays of h x w per block.
__global__ __launch_bounds__(kThreads) void k_fused(const float *in, float *out, int h, int w,
                                                     int per, size_t batch, int mode,
                                                     const float *__restrict__ tab)
{
    extern __shared__ float tile[];
    const int pitch = w + 1;
    const size_t a0 = (size_t)blockIdx.x * per;
    const int na = (int)min((size_t)per, batch - a0);
    const int rows = na * h;
    const size_t off = a0 * (size_t)h * w;
    tile_load(tile, pitch, in + off, w, rows, w);
    __syncthreads();
    if (mode == BDX_FORWARD) {
        pass<FA>(tile, w, rows, false, pitch, w, tab);
        __syncthreads();
        pass<FA>(tile, h, na * w, true, pitch, w, tab);
    } else if (mode == BDX_INVERSE) {
        pass<IA>(tile, w, rows, false, pitch, w, tab);
        __syncthreads();
        pass<IA>(tile, h, na * w, true, pitch, w, tab);
    } else {
        pass<FA2>(tile, w, rows, false, pitch, w, tab);
        __syncthreads();
        pass<FA2, I0>(tile, h, na * w, true, pitch, w, tab);
        __syncthreads();
        pass<I0>(tile, w, rows, false, pitch, w, tab);
    }
    __syncthreads();
    tile_store(tile, pitch, out + off, w, rows, w);
}

// Bands of `per` full rows of length w; rows of all arrays of the batch are one sequence.
__global__ __launch_bounds__(kThreads) void k_rows(const float *in, float *out, int w, int per,
                                                    size_t nrows, int op,
                                                    const float *__restrict__ tab)
{
    extern __shared__ float tile[];
    const int pitch = w + 1;
    const size_t r0 = (size_t)blockIdx.x * per;
    const int nr = (int)min((size_t)per, nrows - r0);
    tile_load(tile, pitch, in + r0 * w, w, nr, w);
    __syncthreads();
    if (op == FA)
        pass<FA>(tile, w, nr, false, pitch, w, tab);
    else if (op == FA2)
        pass<FA2>(tile, w, nr, false, pitch, w, tab);
    else if (op == IA)
        pass<IA>(tile, w, nr, false, pitch, w, tab);
    else
        pass<I0>(tile, w, nr, false, pitch, w, tab);
    __syncthreads();
    tile_store(tile, pitch, out + r0 * w, w, nr, w);
}

// Strips of `cw` full columns of one h x w array.  op = FA, IA, or FA2 (forward, alpha^2,
// then inverse: the middle of the fused round trip).
__global__ __launch_bounds__(kThreads) void k_cols(const float *in, float *out, int h, int w,
                                                    int cw, int op, const float *__restrict__ tab)
{
    extern __shared__ float tile[];
    const int pitch = cw + 1;
    const int strips = w / cw;
    const size_t a = blockIdx.x / strips;
    const int c0 = (blockIdx.x - (int)a * strips) * cw;
    const size_t off = a * (size_t)h * w + c0;
    tile_load(tile, pitch, in + off, w, h, cw);
    __syncthreads();
    if (op == FA)
        pass<FA>(tile, h, cw, true, pitch, cw, tab);
    else if (op == IA)
        pass<IA>(tile, h, cw, true, pitch, cw, tab);
    else
        pass<FA2, I0>(tile, h, cw, true, pitch, cw, tab);
    __syncthreads();
    tile_store(tile, pitch, out + off, w, h, cw);
}

bool g_attr_done = false;

int set_lds_limits()
{
    if (g_attr_done)
        return 0;
    const int bytes = kLdsFloats * 4 + 4096;
    if (hipFuncSetAttribute((const void *)k_fused, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess ||
        hipFuncSetAttribute((const void *)k_rows, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess ||
        hipFuncSetAttribute((const void *)k_cols, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess)
        return -1;
    g_attr_done = true;
    return 0;
}

}  // namespace

extern "C" {

int bdx_device_ready(void)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1)
        return -1;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess)
        return -1;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return -1;
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return -2;
    return set_lds_limits();
}

int bdx_malloc(void **ptr, size_t bytes)
{
    return hipMalloc(ptr, bytes) == hipSuccess ? 0 : -1;
}

int bdx_free(void *ptr)
{
    return hipFree(ptr) == hipSuccess ? 0 : -1;
}

int bdx_stream_create(void **stream)
{
    hipStream_t s;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
        return -1;
    *stream = (void *)s;
    return 0;
}

int bdx_stream_destroy(void *stream)
{
    return hipStreamDestroy((hipStream_t)stream) == hipSuccess ? 0 : -1;
}

int bdx_stream_sync(void *stream)
{
    return hipStreamSynchronize((hipStream_t)stream) == hipSuccess ? 0 : -1;
}

int bdx_copy_h2d(void *dst, const void *src, size_t bytes, void *stream)
{
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream) == hipSuccess ? 0 : -1;
}

int bdx_copy_d2h(void *dst, const void *src, size_t bytes, void *stream)
{
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream) == hipSuccess ? 0 : -1;
}

int bdx_copy_d2d(void *dst, const void *src, size_t bytes, void *stream)
{
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream) == hipSuccess ? 0 : -1;
}

int bdx_transform(const float *d_in, float *d_out, int h, int w, size_t batch, int mode,
                  const float *d_tables, void *stream)
{
    if (batch == 0)
        return 0;
    if (set_lds_limits())
        return -1;
    hipStream_t st = (hipStream_t)stream;
    const size_t hw = (size_t)h * w;
    if ((size_t)h * (w + 1) <= (size_t)kLdsFloats) {
        int per = kLdsFloats / (h * (w + 1));
        if ((size_t)per > batch)
            per = (int)batch;
        const size_t blocks = (batch + per - 1) / per;
        const size_t lds = (size_t)per * h * (w + 1) * 4;
        hipLaunchKernelGGL(k_fused, dim3((unsigned)blocks), dim3(kThreads), lds, st, d_in, d_out, h, w,
                           per, batch, mode, d_tables);
        return hipGetLastError() == hipSuccess ? 0 : -1;
    }
    // Two-pass route: rows in -> out, then columns in place on out.
    const size_t nrows = batch * h;
    int per = kLdsFloats / (w + 1);
    if (per > 64)
        per = 64;
    const size_t rblocks = (nrows + per - 1) / per;
    const size_t rlds = (size_t)per * (w + 1) * 4;
    int cw = 64;   // power of two, so strips tile w exactly and stay 16-byte aligned
    while (cw > 8 && (size_t)h * (cw + 1) > (size_t)kLdsFloats)
        cw /= 2;
    if (cw > w)
        cw = w;
    const size_t cblocks = batch * (size_t)(w / cw);
    const size_t clds = (size_t)h * (cw + 1) * 4;
    const int rop = mode == BDX_FORWARD ? FA : mode == BDX_INVERSE ? IA : FA2;
    const int cop = mode == BDX_FORWARD ? FA : mode == BDX_INVERSE ? IA : FA2;
    hipLaunchKernelGGL(k_rows, dim3((unsigned)rblocks), dim3(kThreads), rlds, st, d_in, d_out, w, per,
                       nrows, rop, d_tables);
    hipLaunchKernelGGL(k_cols, dim3((unsigned)cblocks), dim3(kThreads), clds, st, d_out, d_out, h, w, cw,
                       cop, d_tables);
    if (mode == BDX_ROUNDTRIP)
        hipLaunchKernelGGL(k_rows, dim3((unsigned)rblocks), dim3(kThreads), rlds, st, d_out, d_out, w,
                           per, nrows, (int)I0, d_tables);
    (void)hw;
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // extern "C"
