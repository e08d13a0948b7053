// Row-style team lines read and written straight from global memory.
//
// A team of T consecutive lanes of one wave transforms one contiguous row of length n
// (T = team_size<n>(), aligned to T lanes).  Lane t loads float4 number c of the row at
// element 4 t + 4 T c: a team reads 16 T contiguous bytes per instruction.  The Makhoul
// packing needs, for the second half of the packed points, elements 1 and 3 of the
// float4s of the mirrored lane T - 1 - t; one DPP mirror (row_mirror for T = 16,
// row_half_mirror for 8, quad_perm [3,2,1,0] for 4) per value brings them over, with no
// LDS traffic.  The exchange between the two FFT phases goes through the team's own
// n-float LDS region; outputs go straight to global memory.  Algorithm and index maps as
// in bdx_lines.h and bdx_post.h.
#pragma once
#include "bdx_lines.h"
#include "bdx_post.h"

namespace bdx {

template <int T> __device__ __forceinline__ float team_mirror(float x)
{
    constexpr int ctrl = T == 16 ? 0x140 : T == 8 ? 0x141 : T == 4 ? 0x1B : 0xB1;
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), ctrl, 0xF, 0xF, false));
}

// Forward: row `src` (n floats, 16-byte aligned) -> `dst` (may alias src only if the
// whole team has loaded before any lane stores, which holds: stores follow two team syncs).
template <int n, int OP>
__device__ __forceinline__ void rteam_fwd(const float *src, float *dst, int t, float *xl,
                                          const float *__restrict__ tab)
{
    constexpr int T = team_size<n>(), N = n / 2, N1 = N / T, Q = N1 / (2 * T), C = N1 / 2;
    const float *om = tab + BDX_TABLE_BASE(n);
    const float *tn = om + 2 * n;
    const float *al = alpha_of<n>(tab, OP == FA2 ? 2 : 1);
    float re[N1], im[N1];
    float4 v[C];
#pragma unroll
    for (int c = 0; c < C; c++)
        v[c] = *reinterpret_cast<const float4 *>(src + 4 * t + 4 * T * c);
#pragma unroll
    for (int c = 0; c < C; c++) {
        re[c] = v[c].x;
        im[c] = v[c].z;
        re[N1 - 1 - c] = team_mirror<T>(v[c].w);
        im[N1 - 1 - c] = team_mirror<T>(v[c].y);
    }
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
        xl[c] = re[k1];
        xl[c + 1] = im[k1];
    }
    team_sync();
    float dr[Q][2][T], di[Q][2][T];
#pragma unroll
    for (int r = 0; r < Q; r++) {
        const int g = t + T * r;
        const int ka = g == 0 ? 0 : g, kb = g == 0 ? N1 / 2 : N1 - g;
#pragma unroll
        for (int j2 = 0; j2 < T; j2++) {
            dr[r][0][j2] = xl[2 * (ka * T + j2)];
            di[r][0][j2] = xl[2 * (ka * T + j2) + 1];
            dr[r][1][j2] = xl[2 * (kb * T + j2)];
            di[r][1][j2] = xl[2 * (kb * T + j2) + 1];
        }
    }
    team_sync();   // the team's region is free again once every lane has read its groups
#pragma unroll
    for (int r = 0; r < Q; r++) {
        const int g = t + T * r;
        const bool g0 = g == 0;
        Dft<T, -1>::run(dr[r][0], di[r][0]);
        Dft<T, -1>::run(dr[r][1], di[r][1]);
#pragma unroll
        for (int k2 = 0; k2 < T; k2++) {
            float ar, ai, br, bi;
            int k;
            if (k2 < T / 2) {
                ar = g0 ? dr[r][1][k2] : dr[r][0][k2];
                ai = g0 ? di[r][1][k2] : di[r][0][k2];
                br = dr[r][1][T - 1 - k2];
                bi = di[r][1][T - 1 - k2];
                k = g0 ? N1 / 2 + N1 * k2 : g + N1 * k2;
            } else if (k2 > T / 2) {
                const int j = k2 - T / 2;
                ar = g0 ? dr[r][0][j] : dr[r][0][k2];
                ai = g0 ? di[r][0][j] : di[r][0][k2];
                br = g0 ? dr[r][0][T - j] : dr[r][1][T - 1 - k2];
                bi = g0 ? di[r][0][T - j] : di[r][1][T - 1 - k2];
                k = g0 ? N1 * j : g + N1 * k2;
            } else {
                ar = dr[r][0][k2];
                ai = di[r][0][k2];
                br = dr[r][1][T - 1 - k2];
                bi = di[r][1][T - 1 - k2];
                k = g0 ? 1 : g + N1 * k2;
            }
            float o[4];
            post_fwd(ar, ai, br, bi, om[2 * k], om[2 * k + 1], tn[k], tn[N - k], o);
            if (k2 == T / 2) {
                float s[4];
                post_fwd_special(dr[r][0][0], di[r][0][0], dr[r][0][T / 2], di[r][0][T / 2], tn[n / 4], s);
#pragma unroll
                for (int w = 0; w < 4; w++)
                    o[w] = g0 ? s[w] : o[w];
            }
#pragma unroll
            for (int w = 0; w < 4; w++) {
                const int idx = team_out_index<n, T>(g0, g, k2, w);
                float y = o[w];
                if constexpr (OP == FA || OP == FA2)
                    y *= al[idx];
                dst[idx] = y;
            }
        }
    }
}

// Inverse: inputs at the lane-owned output indices of the forward, outputs as float4
// (the mirror of the forward load).
template <int n, int OP>
__device__ __forceinline__ void rteam_inv(const float *src, float *dst, int t, float *xl,
                                          const float *__restrict__ tab)
{
    constexpr int T = team_size<n>(), N = n / 2, N1 = N / T, Q = N1 / (2 * T), C = N1 / 2;
    const float *om = tab + BDX_TABLE_BASE(n);
    const float *tn = om + 2 * n;
    const float *al = alpha_of<n>(tab, 1);
    float o[Q][T][4];
#pragma unroll
    for (int r = 0; r < Q; r++) {
        const int g = t + T * r;
#pragma unroll
        for (int k2 = 0; k2 < T; k2++)
#pragma unroll
            for (int w = 0; w < 4; w++) {
                const int idx = team_out_index<n, T>(g == 0, g, k2, w);
                float y = src[idx];
                if constexpr (OP == IA)
                    y *= al[idx];
                o[r][k2][w] = y;
            }
    }
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
            post_inv(o[r][k2], om[2 * k], om[2 * k + 1], tn[k], tn[N - k], zar[k2], zai[k2], zbr[k2], zbi[k2]);
            if (k2 == T / 2)
                post_inv_special(o[r][k2], tn[n / 4], z0r, z0i, zhr, zhi);
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
            xl[2 * (ka * T + j2)] = d0r[j2];
            xl[2 * (ka * T + j2) + 1] = d0i[j2];
            xl[2 * (kb * T + j2)] = d1r[j2];
            xl[2 * (kb * T + j2) + 1] = d1i[j2];
        }
    }
    team_sync();
    float re[N1], im[N1];
#pragma unroll
    for (int k1 = 0; k1 < N1; k1++) {
        const int c = 2 * (k1 * T + t);
        re[k1] = xl[c];
        im[k1] = xl[c + 1];
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
    for (int c = 0; c < C; c++) {
        const float4 v = make_float4(re[c], team_mirror<T>(im[N1 - 1 - c]), im[c],
                                     team_mirror<T>(re[N1 - 1 - c]));
        *reinterpret_cast<float4 *>(dst + 4 * t + 4 * T * c) = v;
    }
}

}  // namespace bdx
