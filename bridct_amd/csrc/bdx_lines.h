// Device-side 1D lines of the MI355X BRiDCT kernels (gfx950, wave64).
//
// A "line" is one 1D transform of length n stored in LDS at p[i * S], i < n, with the
// stride S known at compile time (so every LDS access is base + immediate offset).
//   * n <= 64: one lane per line runs the generated Shao-Johnson core sjf_n / sjt_n
//     (reference 1D cores: src/generated/rect_generated.inc col_f*/col_t*).
//   * n >= 128: a team of T lanes of one wave per line (T = 4, 8, 8, 16 for
//     n = 128, 256, 512, 1024).  Makhoul packing into N = n/2 complex points, a
//     Cooley-Tukey split N = N1 * T with Johnson-Frigo split-radix codelets (dft_N1 in
//     each lane, twiddle, in-place exchange through the line's own LDS storage, dft_T),
//     then the real-to-DCT post-process with the tangent twiddle (the tangent form of
//     the reference post8_body, rect_generated.inc:89-96).  Lane t owns the k1 pairs
//     {k1, N1 - k1}, so the post-process pairs (k, N - k) never leave the lane.
// Scaling: each core's own output scale and the orthonormal factor are folded into the
// per-axis vector alpha (host FP64, rounded once).  OP selects where it is applied:
//   FA  forward, multiply alpha on output;   FA2 forward, multiply alpha^2 on output;
//   IA  inverse, multiply alpha on input;    I0  inverse, no scaling.
#pragma once
#include <hip/hip_runtime.h>

#include "bdx_kernels.h"
#include "bridct_device.h"
#include "generated/sj_codelets.h"

namespace bdx {

template <int N, int SG> struct Dft;
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
#undef BDX_DFT

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
#undef BDX_SJ

template <int n> constexpr int team_size() { return BDX_TEAM(n); }

// Global loads and stores of array data, with the non-temporal hint when NT (streaming
// through the caches).  The kernels take NT for calls whose input and output exceed the
// 256 MB Infinity Cache (BDX_STREAM_MIN_BYTES): a float4 copy kernel reaches 6.7 TB/s with
// the hint against 6.2 TB/s without it on MI355X, but batches that fit the cache ran up to
// 13% slower with it.  The row pass of a two-pass transform streams its loads only (its
// output is read again by the column pass).
typedef float f4v __attribute__((ext_vector_type(4)));
template <bool NT> __device__ __forceinline__ float4 ld4(const float *p)
{
    if constexpr (NT) {
        const f4v v = __builtin_nontemporal_load(reinterpret_cast<const f4v *>(p));
        return make_float4(v.x, v.y, v.z, v.w);
    } else {
        return *reinterpret_cast<const float4 *>(p);
    }
}
template <bool NT> __device__ __forceinline__ void st4(float *p, float4 v)
{
    if constexpr (NT) {
        f4v u = {v.x, v.y, v.z, v.w};
        __builtin_nontemporal_store(u, reinterpret_cast<f4v *>(p));
    } else {
        *reinterpret_cast<float4 *>(p) = v;
    }
}
template <bool NT> __device__ __forceinline__ float ld1(const float *p)
{
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}
template <bool NT> __device__ __forceinline__ void st1(float *p, float v)
{
    if constexpr (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}


// Lanes of one team live in one wave; LDS traffic between them needs ordering only
// within the wave.
__device__ __forceinline__ void team_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int n> __device__ __forceinline__ const float *alpha_of(const float *tab, int which)
{
    return tab + BDX_TABLE_BASE(n) + 3 * n + (which == 2 ? n : 0);
}

// ---- one lane per line: Shao-Johnson core --------------------------------------------

template <int n, int S, int OP>
__device__ __forceinline__ void sj_line(float *p, const float *__restrict__ tab)
{
    float x[n];
#pragma unroll
    for (int i = 0; i < n; i++) {
        float v = p[i * S];
        if constexpr (OP == IA)
            v *= alpha_of<n>(tab, 1)[i];
        x[i] = v;
    }
    if constexpr (OP == IA || OP == I0)
        Sj<n>::inv(x);
    else
        Sj<n>::fwd(x);
#pragma unroll
    for (int i = 0; i < n; i++) {
        float v = x[i];
        if constexpr (OP == FA || OP == FA2)
            v *= alpha_of<n>(tab, OP == FA ? 1 : 2)[i];
        p[i * S] = v;
    }
}

// ---- team of T lanes per line ----------------------------------------------------------

// Output index of slot (k2, which) of pair group g, in lane-owned order.
template <int n, int T>
__device__ __forceinline__ int team_out_index(bool g0, int g, int k2, int which)
{
    constexpr int N = n / 2, N1 = N / T;
    if (g0 && k2 == T / 2)
        return which == 0 ? 0 : which == 1 ? N : which == 2 ? n / 4 : 3 * n / 4;
    const int k = g0 ? (k2 < T / 2 ? N1 / 2 + N1 * k2 : N1 * (k2 - T / 2)) : g + N1 * k2;
    return which == 0 ? k : which == 1 ? n - k : which == 2 ? N - k : N + k;
}

// Makhoul-packed input pair of complex point m = t + T*j1: (x[4m], x[4m+2]) in the first
// half, (x[2n-4m-1], x[2n-4m-3]) in the second.
template <int n, int T>
__device__ __forceinline__ void team_pack_index(int t, int j1, int &i0, int &i1)
{
    constexpr int N1 = n / 2 / T;
    const int m = t + T * j1;
    i0 = j1 < N1 / 2 ? 4 * m : 2 * n - 4 * m - 1;
    i1 = j1 < N1 / 2 ? i0 + 2 : i0 - 2;
}

template <int n, int S, int OP>
__device__ __forceinline__ void team_fwd(float *p, int t, const float *__restrict__ tab)
{
    constexpr int T = team_size<n>(), N = n / 2, N1 = N / T, Q = N1 / (2 * T);
    const float *om = tab + BDX_TABLE_BASE(n);
    const float *tn = om + 2 * n;
    float re[N1], im[N1];
#pragma unroll
    for (int j1 = 0; j1 < N1; j1++) {
        int i0, i1;
        team_pack_index<n, T>(t, j1, i0, i1);
        re[j1] = p[i0 * S];
        im[j1] = p[i1 * S];
    }
    team_sync();
    Dft<N1, -1>::run(re, im);
    __builtin_amdgcn_sched_barrier(0);
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
        p[c * S] = re[k1];
        p[(c + 1) * S] = im[k1];
    }
    team_sync();
    // All exchange reads of the lane's pair groups first, then one sync, so results can
    // go straight back into the line without holding them.
    float d0r[Q][T], d0i[Q][T], d1r[Q][T], d1i[Q][T];
#pragma unroll
    for (int r = 0; r < Q; r++) {
        const int g = t + T * r;
        const int ka = g == 0 ? 0 : g, kb = g == 0 ? N1 / 2 : N1 - g;
#pragma unroll
        for (int j2 = 0; j2 < T; j2++) {
            d0r[r][j2] = p[2 * (ka * T + j2) * S];
            d0i[r][j2] = p[(2 * (ka * T + j2) + 1) * S];
            d1r[r][j2] = p[2 * (kb * T + j2) * S];
            d1i[r][j2] = p[(2 * (kb * T + j2) + 1) * S];
        }
    }
    team_sync();
#pragma unroll
    for (int r = 0; r < Q; r++) {
        const int g = t + T * r;
        const bool g0 = g == 0;
        Dft<T, -1>::run(d0r[r], d0i[r]);
        Dft<T, -1>::run(d1r[r], d1i[r]);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int k2 = 0; k2 < T; k2++) {
            float ar, ai, br, bi;
            int k;
            if (k2 < T / 2) {
                ar = g0 ? d1r[r][k2] : d0r[r][k2];
                ai = g0 ? d1i[r][k2] : d0i[r][k2];
                br = d1r[r][T - 1 - k2];
                bi = d1i[r][T - 1 - k2];
                k = g0 ? N1 / 2 + N1 * k2 : g + N1 * k2;
            } else if (k2 > T / 2) {
                const int j = k2 - T / 2;
                ar = g0 ? d0r[r][j] : d0r[r][k2];
                ai = g0 ? d0i[r][j] : d0i[r][k2];
                br = g0 ? d0r[r][T - j] : d1r[r][T - 1 - k2];
                bi = g0 ? d0i[r][T - j] : d1i[r][T - 1 - k2];
                k = g0 ? N1 * j : g + N1 * k2;
            } else {
                ar = d0r[r][k2];
                ai = d0i[r][k2];
                br = d1r[r][T - 1 - k2];
                bi = d1i[r][T - 1 - k2];
                k = g0 ? 1 : g + N1 * k2;   // lane g = 0 uses the special forms below
            }
            const float er = ar + br, ei = ai - bi, orr = ai + bi, oi = br - ar;
            const float wr = om[2 * k], wi = om[2 * k + 1];
            const float P = __builtin_fmaf(wr, orr, -wi * oi);
            const float Qv = __builtin_fmaf(wr, oi, wi * orr);
            const float vr = er + P, vi = ei + Qv, ur = er - P, ui = Qv - ei;
            const float t1 = tn[k], t2 = tn[N - k];
            float o[4];
            o[0] = __builtin_fmaf(t1, vi, vr);
            o[1] = __builtin_fmaf(t1, vr, -vi);
            o[2] = __builtin_fmaf(t2, ui, ur);
            o[3] = __builtin_fmaf(t2, ur, -ui);
            if (k2 == T / 2) {
                // lane g = 0: Z_0 gives outputs 0 and N, Z_{N/2} gives n/4 and 3n/4
                const float th = tn[n / 4];
                const float z0r = d0r[r][0], z0i = d0i[r][0];
                const float zhr = d0r[r][T / 2], zhi = d0i[r][T / 2];
                o[0] = g0 ? z0r + z0i : o[0];
                o[1] = g0 ? z0r - z0i : o[1];
                o[2] = g0 ? __builtin_fmaf(-th, zhi, zhr) : o[2];
                o[3] = g0 ? __builtin_fmaf(th, zhr, zhi) : o[3];
            }
#pragma unroll
            for (int w = 0; w < 4; w++) {
                const int idx = team_out_index<n, T>(g0, g, k2, w);
                float v = o[w];
                if constexpr (OP == FA || OP == FA2)
                    v *= alpha_of<n>(tab, OP == FA ? 1 : 2)[idx];
                p[idx * S] = v;
            }
        }
    }
}

template <int n, int S, int OP>
__device__ __forceinline__ void team_inv(float *p, int t, const float *__restrict__ tab)
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
                float v = p[idx * S];
                if constexpr (OP == IA)
                    v *= alpha_of<n>(tab, 1)[idx];
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
                k = g0 ? 1 : g + N1 * k2;
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
            float gr, gi;   // lane g = 0 placement of D0[x]
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
            p[2 * (ka * T + j2) * S] = d0r[j2];
            p[(2 * (ka * T + j2) + 1) * S] = d0i[j2];
            p[2 * (kb * T + j2) * S] = d1r[j2];
            p[(2 * (kb * T + j2) + 1) * S] = d1i[j2];
        }
    }
    team_sync();
    float re[N1], im[N1];
#pragma unroll
    for (int k1 = 0; k1 < N1; k1++) {
        const int c = 2 * (k1 * T + t);
        re[k1] = p[c * S];
        im[k1] = p[(c + 1) * S];
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
        int i0, i1;
        team_pack_index<n, T>(t, j1, i0, i1);
        p[i0 * S] = re[j1];
        p[i1 * S] = im[j1];
    }
}

template <int n, int S, int OP>
__device__ __forceinline__ void line_op(float *p, int t, const float *__restrict__ tab)
{
    if constexpr (n <= 64)
        sj_line<n, S, OP>(p, tab);
    else if constexpr (OP == IA || OP == I0)
        team_inv<n, S, OP>(p, t, tab);
    else
        team_fwd<n, S, OP>(p, t, tab);
}

// Runs line_op over `nlines` lines of an LDS tile with `threads` lanes.  Line L starts
// at tile + base(L); teams of team_size<n>() consecutive lanes share a line.
template <int n, int S, int OP, int THREADS, typename Base>
__device__ __forceinline__ void run_lines(float *tile, int nlines, Base base,
                                          const float *__restrict__ tab)
{
    constexpr int T = team_size<n>();
    constexpr int teams = THREADS / T;
    const int team = threadIdx.x / T, t = threadIdx.x % T;
    for (int L = team; L < nlines; L += teams)
        line_op<n, S, OP>(tile + base(L), t, tab);
}

// Copies `rows` rows of COLS floats into an LDS tile of pitch PITCH, and back; row r of
// the global side starts at row(r) (16-byte aligned).  Global side is float4-coalesced.
template <int COLS, int PITCH, int THREADS, bool NT = false, typename RowPtr>
__device__ __forceinline__ void tile_load_rows(float *tile, RowPtr row, int rows)
{
    constexpr int per_row = COLS / 4;
    const int nv = rows * per_row;
    for (int i = threadIdx.x; i < nv; i += THREADS) {
        const int r = i / per_row, c = (i - r * per_row) * 4;
        const float4 v = ld4<NT>(row(r) + c);
        float *d = tile + r * PITCH + c;
        d[0] = v.x;
        d[1] = v.y;
        d[2] = v.z;
        d[3] = v.w;
    }
}

template <int COLS, int PITCH, int THREADS, bool NT = false, typename RowPtr>
__device__ __forceinline__ void tile_store_rows(const float *tile, RowPtr row, int rows)
{
    constexpr int per_row = COLS / 4;
    const int nv = rows * per_row;
    for (int i = threadIdx.x; i < nv; i += THREADS) {
        const int r = i / per_row, c = (i - r * per_row) * 4;
        const float *d = tile + r * PITCH + c;
        st4<NT>(row(r) + c, make_float4(d[0], d[1], d[2], d[3]));
    }
}

// Contiguous rows, global row stride gstride floats.
template <int COLS, int PITCH, int THREADS, bool NT = false>
__device__ __forceinline__ void tile_load(float *tile, const float *__restrict__ src, size_t gstride,
                                          int rows)
{
    tile_load_rows<COLS, PITCH, THREADS, NT>(tile, [=](int r) { return src + r * gstride; }, rows);
}

template <int COLS, int PITCH, int THREADS, bool NT = false>
__device__ __forceinline__ void tile_store(const float *tile, float *__restrict__ dst, size_t gstride,
                                           int rows)
{
    tile_store_rows<COLS, PITCH, THREADS, NT>(tile, [=](int r) { return dst + r * gstride; }, rows);
}

}  // namespace bdx
