// Column-style team lines: lane = line, team member = (part of) a wave.
//
// A block transforms L = 64 / M lines of length n, one per lane, with the T members of
// each line's team spread over T / M waves (M members per wave: lanes [0, L) of a wave
// are member M w, lanes [L, 2L) member M w + 1).  With M = 1 every twiddle, tangent,
// normalization factor and element index is wave-uniform (scalar loads, no selects);
// with M = 2 (used for n = 1024, where one member per wave would need 1024-lane blocks
// and leave one block per CU) the constants come from an LDS copy of the table block
// and two half-wave values per wave.  Loads and stores of element i of adjacent columns
// are 256-byte (M = 1) or 128-byte (M = 2) row segments.  The exchange between members
// goes through LDS with block barriers, in ROUNDS slices of the lines to bound the LDS
// footprint.  Algorithm and index maps as in bdx_lines.h and bdx_post.h.
#pragma once
#include "bdx_lines.h"
#include "bdx_post.h"
#include "bdx_rteam.h"

namespace bdx {

// Geometry of a column team of length n run by blocks of THREADS lanes.
template <int n, int THREADS> struct CTeam {
    static constexpr int T = team_size<n>(), N = n / 2, N1 = N / T, Q = N1 / (2 * T);
    static constexpr int M = 64 * T / THREADS;            // members per wave
    static constexpr int L = 64 / M;                      // lines per block
    // Exchange in two rounds from 16384 floats up: the exchange then takes at most 32 KiB,
    // and three 512-lane workgroups of the 256 team fit a CU (64 KiB allowed two; equal
    // results bitwise, within 2% in speed).
    static constexpr int ROUNDS = n * L >= 16384 ? 2 : 1;
    static constexpr int SL = L / ROUNDS;                 // lines per exchange round
    static constexpr int XLDS = n * SL;                   // exchange floats
    static constexpr int TLDS = M > 1 ? RowTab<n>::SIZE : 1;   // LDS constant table
    static_assert(M >= 1 && 64 % M == 0, "team members must tile the wave");
};

// Element index of output slot (k2, which) of pair group g.
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

// Constants of member t: from the global table block (M = 1, wave-uniform) or from the
// LDS copy (M > 1).
template <int n, int M> struct CConst {
    static constexpr int T = team_size<n>();
    const float *g, *l;
    __device__ __forceinline__ float tw_re(int k1, int t) const
    {
        return M == 1 ? g[2 * ((2 * t * k1) & (n - 1))] : l[RowTab<n>::TW + 2 * (k1 * T + t)];
    }
    __device__ __forceinline__ float tw_im(int k1, int t) const
    {
        return M == 1 ? g[2 * ((2 * t * k1) & (n - 1)) + 1] : l[RowTab<n>::TW + 2 * (k1 * T + t) + 1];
    }
    __device__ __forceinline__ float om_re(int k) const { return M == 1 ? g[2 * k] : l[RowTab<n>::OM + 2 * k]; }
    __device__ __forceinline__ float om_im(int k) const { return M == 1 ? g[2 * k + 1] : l[RowTab<n>::OM + 2 * k + 1]; }
    __device__ __forceinline__ float tn(int k) const { return M == 1 ? g[2 * n + k] : l[RowTab<n>::TN + k]; }
};

// Forward core.  ld(i) returns element i of the lane's line; emit(r, k2, w, v) receives
// the unscaled output of index wteam_index(t + T r, k2, w) as soon as it is computed.
template <int n, int THREADS, typename Ld, typename Emit>
__device__ __forceinline__ void wteam_fwd(int t, int line, Ld ld, float *xl, CConst<n, CTeam<n, THREADS>::M> cc,
                                          Emit emit)
{
    using C = CTeam<n, THREADS>;
    constexpr int T = C::T, N = C::N, N1 = C::N1, Q = C::Q, SL = C::SL, ROUNDS = C::ROUNDS;
    const int sl = line % SL, rnd = line / SL;
    float re[N1], im[N1];
#pragma unroll
    for (int j1 = 0; j1 < N1; j1++) {
        int i0, i1;
        team_pack_index<n, T>(t, j1, i0, i1);
        re[j1] = ld(i0);
        im[j1] = ld(i1);
    }
    Dft<N1, -1>::run(re, im);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k1 = 1; k1 < N1; k1++) {
        const float wr = cc.tw_re(k1, t), wi = cc.tw_im(k1, t);
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
    __builtin_amdgcn_sched_barrier(0);
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
                                     cc.tn(n / 4), o);
                } else {
                    const bool lo = k2 < T / 2;
                    const int j = k2 - T / 2;
                    const int k = lo ? N1 / 2 + N1 * k2 : N1 * j;
                    const float ar = lo ? dr[r][1][k2] : dr[r][0][j];
                    const float ai = lo ? di[r][1][k2] : di[r][0][j];
                    const float br = lo ? dr[r][1][T - 1 - k2] : dr[r][0][T - j];
                    const float bi = lo ? di[r][1][T - 1 - k2] : di[r][0][T - j];
                    post_fwd(ar, ai, br, bi, cc.om_re(k), cc.om_im(k), cc.tn(k), cc.tn(N - k), o);
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
                         cc.om_re(k), cc.om_im(k), cc.tn(k), cc.tn(N - k), o);
#pragma unroll
                for (int w = 0; w < 4; w++)
                    emit(r, k2, w, o[w]);
            }
        }
    }
}

// Inverse core: o[r][k2][w] holds the (already scaled) input of index wteam_index(...);
// st(i, v) stores element i of the lane's line.
template <int n, int THREADS, typename St>
__device__ __forceinline__ void wteam_inv(int t, int line, float (&o)[CTeam<n, THREADS>::Q][CTeam<n, THREADS>::T][4],
                                          St st, float *xl, CConst<n, CTeam<n, THREADS>::M> cc)
{
    using C = CTeam<n, THREADS>;
    constexpr int T = C::T, N = C::N, N1 = C::N1, Q = C::Q, SL = C::SL, ROUNDS = C::ROUNDS;
    const int sl = line % SL, rnd = line / SL;
    float dr[Q][2][T], di[Q][2][T];
#pragma unroll
    for (int r = 0; r < Q; r++) {
        const int g = t + T * r;
        if (g == 0) {
#pragma unroll
            for (int k2 = 0; k2 < T; k2++) {
                if (k2 == T / 2) {
                    post_inv_special(o[r][k2], cc.tn(n / 4), dr[r][0][0], di[r][0][0], dr[r][0][T / 2],
                                     di[r][0][T / 2]);
                    continue;
                }
                const bool lo = k2 < T / 2;
                const int j = k2 - T / 2;
                const int k = lo ? N1 / 2 + N1 * k2 : N1 * j;
                float ar, ai, br, bi;
                post_inv(o[r][k2], cc.om_re(k), cc.om_im(k), cc.tn(k), cc.tn(N - k), ar, ai, br, bi);
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
                post_inv(o[r][k2], cc.om_re(k), cc.om_im(k), cc.tn(k), cc.tn(N - k), dr[r][0][k2],
                         di[r][0][k2], dr[r][1][T - 1 - k2], di[r][1][T - 1 - k2]);
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
        const float wr = cc.tw_re(k1, t), wi = cc.tw_im(k1, t);
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
