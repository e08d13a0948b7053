// Column-style team lines: lane = line, team member = wave.
//
// A block of T waves (T = team_size<n>()) transforms 64 lines of length n at once, one per
// lane; wave t is team member t for every lane.  All twiddles, tangents, normalization
// factors and element indices depend only on (t, compile-time slot), so they are
// wave-uniform (scalar loads, no per-lane selects), and the loads and stores of element
// i of the 64 lines are one coalesced 256-byte access when the lines are adjacent
// columns.  The exchange between members goes through LDS with block barriers, in
// ROUNDS slices of 64 / ROUNDS lanes to bound the LDS footprint to n * 64 / ROUNDS floats.
// Algorithm and index maps as in bdx_lines.h (team of T lanes per line) and bdx_post.h.
#pragma once
#include "bdx_lines.h"
#include "bdx_post.h"

namespace bdx {

template <int n> constexpr int wteam_rounds() { return n * 64 > 32768 ? 2 : 1; }
template <int n> constexpr int wteam_lds() { return n * 64 / wteam_rounds<n>(); }

// Element index of output slot (k2, which) of pair group g (wave-uniform g).
template <int n, int T>
__device__ __forceinline__ int wteam_index(int g, int k2, int which)
{
    constexpr int N = n / 2, N1 = N / T;
    int k;
    if (g == 0) {
        if (k2 == T / 2)
            return which == 0 ? 0 : which == 1 ? N : which == 2 ? n / 4 : 3 * n / 4;
        k = k2 < T / 2 ? N1 / 2 + N1 * k2 : N1 * (k2 - T / 2);
    } else {
        k = g + N1 * k2;
    }
    return which == 0 ? k : which == 1 ? n - k : which == 2 ? N - k : N + k;
}

// Forward core.  ld(i) returns element i of the lane's line; emit(r, k2, w, v) receives
// the unscaled output of index wteam_index(t + T r, k2, w) as soon as it is computed.
template <int n, typename Ld, typename Emit>
__device__ __forceinline__ void wteam_fwd(int t, Ld ld, float *xl, const float *__restrict__ tab, Emit emit)
{
    constexpr int T = team_size<n>(), N = n / 2, N1 = N / T, Q = N1 / (2 * T);
    constexpr int ROUNDS = wteam_rounds<n>(), SL = 64 / ROUNDS;
    const int lane = threadIdx.x & 63, sl = lane % SL, rnd = lane / SL;
    const float *om = tab + BDX_TABLE_BASE(n);
    const float *tn = om + 2 * n;
    float re[N1], im[N1];
#pragma unroll
    for (int j1 = 0; j1 < N1; j1++) {
        int i0, i1;
        team_pack_index<n, T>(t, j1, i0, i1);
        re[j1] = ld(i0);
        im[j1] = ld(i1);
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
    float dr[Q][2][T], di[Q][2][T];
#pragma unroll
    for (int rr = 0; rr < ROUNDS; rr++) {
        if (rnd == rr) {
#pragma unroll
            for (int k1 = 0; k1 < N1; k1++) {
                const int c = k1 * T + t;
                xl[(2 * c) * SL + sl] = re[k1];
                xl[(2 * c + 1) * SL + sl] = im[k1];
            }
        }
        __syncthreads();
        if (rnd == rr) {
#pragma unroll
            for (int r = 0; r < Q; r++) {
                const int g = t + T * r;
                const int ka = g == 0 ? 0 : g, kb = g == 0 ? N1 / 2 : N1 - g;
#pragma unroll
                for (int j2 = 0; j2 < T; j2++) {
                    dr[r][0][j2] = xl[(2 * (ka * T + j2)) * SL + sl];
                    di[r][0][j2] = xl[(2 * (ka * T + j2) + 1) * SL + sl];
                    dr[r][1][j2] = xl[(2 * (kb * T + j2)) * SL + sl];
                    di[r][1][j2] = xl[(2 * (kb * T + j2) + 1) * SL + sl];
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < Q; r++) {
        const int g = t + T * r;
        Dft<T, -1>::run(dr[r][0], di[r][0]);
        Dft<T, -1>::run(dr[r][1], di[r][1]);
        if (g == 0) {
#pragma unroll
            for (int k2 = 0; k2 < T; k2++) {
                float o[4];
                if (k2 == T / 2) {
                    post_fwd_special(dr[r][0][0], di[r][0][0], dr[r][0][T / 2], di[r][0][T / 2],
                                     tn[n / 4], o);
                } else {
                    const bool lo = k2 < T / 2;
                    const int j = k2 - T / 2;
                    const int k = lo ? N1 / 2 + N1 * k2 : N1 * j;
                    const float ar = lo ? dr[r][1][k2] : dr[r][0][j];
                    const float ai = lo ? di[r][1][k2] : di[r][0][j];
                    const float br = lo ? dr[r][1][T - 1 - k2] : dr[r][0][T - j];
                    const float bi = lo ? di[r][1][T - 1 - k2] : di[r][0][T - j];
                    post_fwd(ar, ai, br, bi, om[2 * k], om[2 * k + 1], tn[k], tn[N - k], o);
                }
#pragma unroll
                for (int w = 0; w < 4; w++)
                    emit(r, k2, w, o[w]);
            }
        } else {
#pragma unroll
            for (int k2 = 0; k2 < T; k2++) {
                const int k = g + N1 * k2;
                float o[4];
                post_fwd(dr[r][0][k2], di[r][0][k2], dr[r][1][T - 1 - k2], di[r][1][T - 1 - k2],
                         om[2 * k], om[2 * k + 1], tn[k], tn[N - k], o);
#pragma unroll
                for (int w = 0; w < 4; w++)
                    emit(r, k2, w, o[w]);
            }
        }
    }
}

// Inverse core: o[r][k2][w] holds the (already scaled) input of index wteam_index(...);
// st(i, v) stores element i of the lane's line.
template <int n, typename St>
__device__ __forceinline__ void wteam_inv(int t, float (&o)[n / 2 / team_size<n>() / (2 * team_size<n>())][team_size<n>()][4],
                                          St st, float *xl, const float *__restrict__ tab)
{
    constexpr int T = team_size<n>(), N = n / 2, N1 = N / T, Q = N1 / (2 * T);
    constexpr int ROUNDS = wteam_rounds<n>(), SL = 64 / ROUNDS;
    const int lane = threadIdx.x & 63, sl = lane % SL, rnd = lane / SL;
    const float *om = tab + BDX_TABLE_BASE(n);
    const float *tn = om + 2 * n;
    float dr[Q][2][T], di[Q][2][T];
#pragma unroll
    for (int r = 0; r < Q; r++) {
        const int g = t + T * r;
        if (g == 0) {
#pragma unroll
            for (int k2 = 0; k2 < T; k2++) {
                if (k2 == T / 2) {
                    post_inv_special(o[r][k2], tn[n / 4], dr[r][0][0], di[r][0][0], dr[r][0][T / 2],
                                     di[r][0][T / 2]);
                    continue;
                }
                const bool lo = k2 < T / 2;
                const int j = k2 - T / 2;
                const int k = lo ? N1 / 2 + N1 * k2 : N1 * j;
                float ar, ai, br, bi;
                post_inv(o[r][k2], om[2 * k], om[2 * k + 1], tn[k], tn[N - k], ar, ai, br, bi);
                if (lo) {
                    dr[r][1][k2] = ar;
                    di[r][1][k2] = ai;
                    dr[r][1][T - 1 - k2] = br;
                    di[r][1][T - 1 - k2] = bi;
                } else {
                    dr[r][0][j] = ar;
                    di[r][0][j] = ai;
                    dr[r][0][T - j] = br;
                    di[r][0][T - j] = bi;
                }
            }
        } else {
#pragma unroll
            for (int k2 = 0; k2 < T; k2++) {
                const int k = g + N1 * k2;
                post_inv(o[r][k2], om[2 * k], om[2 * k + 1], tn[k], tn[N - k], dr[r][0][k2], di[r][0][k2],
                         dr[r][1][T - 1 - k2], di[r][1][T - 1 - k2]);
            }
        }
        Dft<T, 1>::run(dr[r][0], di[r][0]);
        Dft<T, 1>::run(dr[r][1], di[r][1]);
    }
    float re[N1], im[N1];
#pragma unroll
    for (int rr = 0; rr < ROUNDS; rr++) {
        if (rnd == rr) {
#pragma unroll
            for (int r = 0; r < Q; r++) {
                const int g = t + T * r;
                const int ka = g == 0 ? 0 : g, kb = g == 0 ? N1 / 2 : N1 - g;
#pragma unroll
                for (int j2 = 0; j2 < T; j2++) {
                    xl[(2 * (ka * T + j2)) * SL + sl] = dr[r][0][j2];
                    xl[(2 * (ka * T + j2) + 1) * SL + sl] = di[r][0][j2];
                    xl[(2 * (kb * T + j2)) * SL + sl] = dr[r][1][j2];
                    xl[(2 * (kb * T + j2) + 1) * SL + sl] = di[r][1][j2];
                }
            }
        }
        __syncthreads();
        if (rnd == rr) {
#pragma unroll
            for (int k1 = 0; k1 < N1; k1++) {
                const int c = k1 * T + t;
                re[k1] = xl[(2 * c) * SL + sl];
                im[k1] = xl[(2 * c + 1) * SL + sl];
            }
        }
        __syncthreads();
    }
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
        st(i0, re[j1]);
        st(i1, im[j1]);
    }
}

}  // namespace bdx
