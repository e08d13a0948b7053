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
//
// Addressing.  Every row a member reads or writes is a lane part plus a wave-uniform part:
// the Makhoul-packed inputs m = t + T j1 sit at rows 4t + 4T j1 (+2) in the first half and
// 4(T-1-t) + (2n - 1 - 4T j1 - 4(T-1)) (-2) in the second (families P and R), and the
// outputs of pair group g = t + T r at rows t + (T r + N1 k2), (T-1-t) + (n - T + 1 - T r -
// N1 k2), (T-1-t) + (N - T + 1 - T r - N1 k2) and t + (N + T r + N1 k2) (families C, D, D,
// C).  So a lane keeps four byte offsets (one per family) and each access adds a uniform
// row offset: the address is a scalar base plus one of those four registers, with no
// per-access vector arithmetic.  The member t = 0 (the lead) also owns the self-paired
// groups 0 and N1/2.  With one member per wave it takes a branch of its own for them
// (absolute rows); with several it follows the general slot formula through two more
// families (CL, DL: lane part N1/2 instead of t below slot T/2), so that no wave runs a
// second, divergent post-process.
//
// Registers.  The exchange writes a lane's N1 complex values to LDS and reads the values
// of its pair groups back into the same registers (Q 2 T = N1), so one set of 2 N1 floats
// is live across it, not two.
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

// Constants of member t: from the global table block (M = 1, wave-uniform) or from the
// LDS copy (M > 1).  With M > 1 the member index varies across the wave, so the lane keeps
// one pointer per family of constants, each at the lowest address the family reaches, and
// every access is that pointer plus an immediate offset (as SlotPtrs in bdx_rteam.h; the
// compiler had turned g + N1 k2 into g | N1 k2, which does not fold into an offset).
//   twiddle (k1, t)                          tw_lane + 2 T k1
//   slot k = t + T r + N1 k2: om[2k], tan(k)  om + 2 u, up + u      (u = T r + N1 k2)
//   tan(N - k)                               dn + SPAN - u
//   inverse pair twiddles (k1, j2):          twa + 2 T T r + 2 j2                 (k1 = g)
//                                            twb0 + 2 j2 (r = 0), twb + 2 T T (Q-1-r) + 2 j2
//                                                                  (k1 = N1 - g, N1/2 at g = 0)
template <int n, int M> struct CConst {
    using R = RowTab<n>;
    static constexpr int T = team_size<n>(), N = n / 2, N1 = N / T, Q = N1 / (2 * T);
    static constexpr int SPAN = T * (Q - 1) + N1 * (T - 1);   // largest u
    const float *g, *l;
    int t;
    const float *tw_lane, *om, *up, *dn, *twa, *twb0, *twb;
    __device__ __forceinline__ CConst(const float *g_, const float *l_, int t_)
        : g(g_), l(l_), t(t_), tw_lane(l_ + R::TW + 2 * t_), om(l_ + R::OM + 2 * t_), up(l_ + R::TN + t_),
          dn(l_ + R::TN + N - t_ - SPAN), twa(l_ + R::TW + 2 * T * t_),
          twb0(l_ + R::TW + 2 * T * (t_ == 0 ? N1 / 2 : N1 - t_)), twb(l_ + R::TW + 2 * T * (N1 - t_ - T * (Q - 1)))
    {
    }
    static constexpr int u(int r, int k2) { return T * r + N1 * k2; }
    // twiddle (k1, t) of the lane's own member t
    __device__ __forceinline__ float twl_re(int k1) const
    {
        return M == 1 ? g[2 * ((2 * t * k1) & (n - 1))] : tw_lane[2 * T * k1];
    }
    __device__ __forceinline__ float twl_im(int k1) const
    {
        return M == 1 ? g[2 * ((2 * t * k1) & (n - 1)) + 1] : tw_lane[2 * T * k1 + 1];
    }
    // constants of slot k = t + T r + N1 k2
    __device__ __forceinline__ float oms_re(int r, int k2) const
    {
        return M == 1 ? g[2 * (t + u(r, k2))] : om[2 * u(r, k2)];
    }
    __device__ __forceinline__ float oms_im(int r, int k2) const
    {
        return M == 1 ? g[2 * (t + u(r, k2)) + 1] : om[2 * u(r, k2) + 1];
    }
    __device__ __forceinline__ float tns(int r, int k2) const { return M == 1 ? g[2 * n + t + u(r, k2)] : up[u(r, k2)]; }
    __device__ __forceinline__ float tnns(int r, int k2) const
    {
        return M == 1 ? g[2 * n + N - t - u(r, k2)] : dn[SPAN - u(r, k2)];
    }
    // twiddle (k1, j2) of the inverse's pair group s of group r: k1 = g (s = 0) or N1 - g,
    // N1/2 for g = 0 (s = 1), g = t + T r
    __device__ __forceinline__ float twp(int r, int s, int j2, int im) const
    {
        if constexpr (M == 1) {
            const int gg = t + T * r, k1 = s == 0 ? gg : (gg == 0 ? N1 / 2 : N1 - gg);
            return g[2 * ((2 * j2 * k1) & (n - 1)) + im];
        } else {
            return s == 0 ? twa[2 * T * T * r + 2 * j2 + im]
                 : r == 0 ? twb0[2 * j2 + im]
                          : twb[2 * T * T * (Q - 1 - r) + 2 * j2 + im];
        }
    }
    // absolute index k (compile-time: the group g = 0 of member 0)
    __device__ __forceinline__ float om_re(int k) const { return M == 1 ? g[2 * k] : l[R::OM + 2 * k]; }
    __device__ __forceinline__ float om_im(int k) const { return M == 1 ? g[2 * k + 1] : l[R::OM + 2 * k + 1]; }
    __device__ __forceinline__ float tn(int k) const { return M == 1 ? g[2 * n + k] : l[R::TN + k]; }
};

// Row families (lane parts): 0 = P (4t), 1 = R (4(T-1-t)), 2 = C (t), 3 = D (T-1-t), and
// the low slots of the lead member with several members per wave: 4 = CL (tl), 5 = DL
// (T-1-tl), tl = N1/2 for t = 0, else t.
enum { FAM_P = 0, FAM_R = 1, FAM_C = 2, FAM_D = 3, FAM_CL = 4, FAM_DL = 5 };

// Family and uniform part of output row (k2, which) of pair group g = t + T r, g != 0.
template <int n, int T> __device__ __forceinline__ constexpr int out_family(int which)
{
    return which == 1 || which == 2 ? FAM_D : FAM_C;
}
template <int n, int T> __device__ __forceinline__ constexpr int out_family_lo(int which)
{
    return which == 1 || which == 2 ? FAM_DL : FAM_CL;
}
template <int n, int T> __device__ __forceinline__ constexpr int out_urow(int r, int k2, int which)
{
    constexpr int N = n / 2, N1 = N / T;
    return which == 0 ? T * r + N1 * k2
         : which == 1 ? n - T + 1 - T * r - N1 * k2
         : which == 2 ? N - T + 1 - T * r - N1 * k2
                      : N + T * r + N1 * k2;
}
// Smallest uniform row of the DL family (outputs 1 and 2 of the low slots r = 0, k2 < T/2):
// the column unit keeps that family biased by N1/2 rows, which this must not undercut.
template <int n, int T> __device__ __forceinline__ constexpr int min_dl_urow()
{
    int m = n;
    for (int k2 = 0; k2 < T / 2; k2++)
        for (int w = 1; w <= 2; w++)
            m = out_urow<n, T>(0, k2, w) < m ? out_urow<n, T>(0, k2, w) : m;
    return m;
}
// Absolute output row (k2, which) of the self-paired group g = 0 (family C, lane part 0).
template <int n, int T> __device__ __forceinline__ constexpr int out_row_g0(int k2, int which)
{
    constexpr int N = n / 2, N1 = N / T;
    if (k2 == T / 2)
        return which == 0 ? 0 : which == 1 ? N : which == 2 ? n / 4 : 3 * n / 4;
    const int k = k2 < T / 2 ? N1 / 2 + N1 * k2 : N1 * (k2 - T / 2);
    return which == 0 ? k : which == 1 ? n - k : which == 2 ? N - k : N + k;
}
// Family and uniform part of the packed input pair j1 (u: its first row; the second is u + 2
// in family P, u - 2 in family R).
template <int n, int T> __device__ __forceinline__ constexpr int in_family(int j1)
{
    return j1 < n / 2 / T / 2 ? FAM_P : FAM_R;
}
template <int n, int T> __device__ __forceinline__ constexpr int in_urow(int j1)
{
    return j1 < n / 2 / T / 2 ? 4 * T * j1 : 2 * n - 1 - 4 * T * j1 - 4 * (T - 1);
}

// Exchange: the lane's N1 complex values (k1 = 0..N1-1 at slot k1 T + t) go to LDS; the
// values of its pair groups (ka = g or 0, kb = N1 - g or N1/2, j2 = 0..T-1) come back into
// the same registers, dr[r][s][j2] = v[(2 r + s) T + j2].  TO_PAIRS: forward direction
// (lane order in, pair order out); else the reverse.
template <int n, int THREADS, bool TO_PAIRS>
__device__ __forceinline__ void cteam_exchange(int t, int line, float (&re)[CTeam<n, THREADS>::N1],
                                               float (&im)[CTeam<n, THREADS>::N1], float *xl)
{
    using C = CTeam<n, THREADS>;
    constexpr int T = C::T, N1 = C::N1, Q = C::Q, SL = C::SL, ROUNDS = C::ROUNDS;
    // Real parts in the first half of the area, imaginary parts in the second (at IM), slot
    // (k1, j) at (k1 T + j) SL + sl: a lane's writes of k1 and k1 + 1 are 64 floats apart times
    // a whole number, and its reads of j2 and j2 + 1 SL floats apart, so every access is one
    // base register plus an immediate offset (interleaved re/im pairs had needed an address
    // add per slot).
    constexpr int IM = N1 * T * SL;
    const int sl = line % SL, rnd = line / SL;
    const auto lane_slot = [&](int k1) { return (k1 * T + t) * SL + sl; };
    const auto pair_slot = [&](int r, int s, int j2) {
        const int g = t + T * r;
        const int k = s == 0 ? (g == 0 ? 0 : g) : (g == 0 ? N1 / 2 : N1 - g);
        return (k * T + j2) * SL + sl;
    };
#pragma unroll
    for (int rr = 0; rr < ROUNDS; rr++) {
        if (rnd == rr) {
            if constexpr (TO_PAIRS) {
#pragma unroll
                for (int k1 = 0; k1 < N1; k1++) {
                    xl[lane_slot(k1)] = re[k1];
                    xl[IM + lane_slot(k1)] = im[k1];
                }
            } else {
#pragma unroll
                for (int r = 0; r < Q; r++)
#pragma unroll
                    for (int s = 0; s < 2; s++)
#pragma unroll
                        for (int j2 = 0; j2 < T; j2++) {
                            xl[pair_slot(r, s, j2)] = re[(2 * r + s) * T + j2];
                            xl[IM + pair_slot(r, s, j2)] = im[(2 * r + s) * T + j2];
                        }
            }
        }
        __syncthreads();
        if (rnd == rr) {
            if constexpr (TO_PAIRS) {
#pragma unroll
                for (int r = 0; r < Q; r++)
#pragma unroll
                    for (int s = 0; s < 2; s++)
#pragma unroll
                        for (int j2 = 0; j2 < T; j2++) {
                            re[(2 * r + s) * T + j2] = xl[pair_slot(r, s, j2)];
                            im[(2 * r + s) * T + j2] = xl[IM + pair_slot(r, s, j2)];
                        }
            } else {
#pragma unroll
                for (int k1 = 0; k1 < N1; k1++) {
                    re[k1] = xl[lane_slot(k1)];
                    im[k1] = xl[IM + lane_slot(k1)];
                }
            }
        }
        __syncthreads();
    }
}

// i with its log2(N) bits reversed (N a power of two)
template <int N> __device__ __forceinline__ constexpr int bit_reverse(int i)
{
    int r = 0;
    for (int b = 1; b < N; b <<= 1)
        r = (r << 1) | ((i & b) ? 1 : 0);
    return r;
}

template <int T> __device__ __forceinline__ float (&sub(float *v, int i))[T]
{
    return *reinterpret_cast<float(*)[T]>(v + i * T);
}

// Forward core.  ld(family, urow) returns the lane's element of that row; emit(r, k2, which,
// family, urow, v) receives the unscaled output of that row as soon as it is computed
// (family FAM_C with the absolute row for the group g = 0).  With several members per wave,
// which = 4 and 5 carry out[0] and out[N] of the lead (t = 0); they reach emit on every
// lane, and only the lead's are outputs.
template <int n, int THREADS, typename Ld, typename Emit>
__device__ __forceinline__ void cteam_fwd(int t, int line, Ld ld, float *xl, CConst<n, CTeam<n, THREADS>::M> cc,
                                          Emit emit)
{
    using C = CTeam<n, THREADS>;
    constexpr int T = C::T, N = C::N, N1 = C::N1, Q = C::Q;
    float re[N1], im[N1];
    // Loads in bit-reversed order, the order in which the DFT_N1 codelet takes its inputs
    // (its first butterflies combine x[0], x[N1/2], x[N1/4], ...): they can start while
    // the later loads are in flight.
#pragma unroll
    for (int i = 0; i < N1; i++) {
        const int j1 = bit_reverse<N1>(i);
        const int fam = in_family<n, T>(j1), u = in_urow<n, T>(j1);
        re[j1] = ld(fam, u);
        im[j1] = ld(fam, fam == FAM_P ? u + 2 : u - 2);
    }
    Dft<N1, -1>::run(re, im);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k1 = 1; k1 < N1; k1++) {
        const float wr = cc.twl_re(k1), wi = cc.twl_im(k1);
        const float a = re[k1], b = im[k1];
        re[k1] = __builtin_fmaf(a, wr, -b * wi);
        im[k1] = __builtin_fmaf(a, wi, b * wr);
    }
    cteam_exchange<n, THREADS, true>(t, line, re, im, xl);
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (C::M > 1) {
        // Several members per wave: the lead (t = 0) takes the general slot formula too, so
        // that its wave does not run a second, divergent post-process.  Its low slots (r = 0,
        // k2 < T/2) take k = N1/2 + N1 k2 from (D1[k2], D1[T-1-k2]) through the lo offsets
        // (lane part N1/2 instead of t); its high slots k = N1 k2 from (D0[k2], D0[T-k2]),
        // the pair (N - k, k) seen from its other side, which post_fwd maps to the same four
        // outputs bitwise; its slot T/2 (k = N/2, rows n/4 and 3n/4, twice each) carries
        // out[n/4] and out[3n/4] of Z_0 and Z_{N/2}, and out[0], out[N] go out apart
        // (emit outputs 4 and 5).
        const bool lead = t == 0;
        const CConst<n, C::M> cl(cc.g, cc.l, lead ? N1 / 2 : t);
        float lead_first = 0.f, lead_second = 0.f;
#pragma unroll
        for (int r = 0; r < Q; r++) {
            float (&d0r)[T] = sub<T>(re, 2 * r), (&d0i)[T] = sub<T>(im, 2 * r);
            float (&d1r)[T] = sub<T>(re, 2 * r + 1), (&d1i)[T] = sub<T>(im, 2 * r + 1);
            Dft<T, -1>::run(d0r, d0i);
            Dft<T, -1>::run(d1r, d1i);
#pragma unroll
            for (int k2 = 0; k2 < T; k2++) {
                const bool lo = r == 0 && k2 < T / 2, hi_swap = r == 0 && k2 > T / 2;
                const CConst<n, C::M> &P = lo ? cl : cc;
                const float ar = lo && lead ? d1r[k2] : d0r[k2];
                const float ai = lo && lead ? d1i[k2] : d0i[k2];
                const float br = hi_swap && lead ? d0r[hi_swap ? T - k2 : 0] : d1r[T - 1 - k2];
                const float bi = hi_swap && lead ? d0i[hi_swap ? T - k2 : 0] : d1i[T - 1 - k2];
                float o[4];
                post_fwd(ar, ai, br, bi, P.oms_re(r, k2), P.oms_im(r, k2), P.tns(r, k2), P.tnns(r, k2), o);
                if (r == 0 && k2 == T / 2) {
                    float sp[4];
                    post_fwd_special(d0r[0], d0i[0], d0r[T / 2], d0i[T / 2], cc.tn(n / 4), sp);
#pragma unroll
                    for (int w = 0; w < 4; w++)
                        o[w] = lead ? sp[2 + (w & 1)] : o[w];
                    lead_first = sp[0];
                    lead_second = sp[1];
                }
#pragma unroll
                for (int w = 0; w < 4; w++)
                    emit(r, k2, w, lo ? out_family_lo<n, T>(w) : out_family<n, T>(w), out_urow<n, T>(r, k2, w),
                         o[w]);
            }
        }
        // on every lane (no divergent branch): emit keeps outputs 4 and 5 of lanes t != 0
        // out of memory
        emit(0, 0, 4, FAM_C, 0, lead_first);
        emit(0, 0, 5, FAM_C, N, lead_second);
        return;
    }
#pragma unroll
    for (int r = 0; r < Q; r++) {
        float (&d0r)[T] = sub<T>(re, 2 * r), (&d0i)[T] = sub<T>(im, 2 * r);
        float (&d1r)[T] = sub<T>(re, 2 * r + 1), (&d1i)[T] = sub<T>(im, 2 * r + 1);
        Dft<T, -1>::run(d0r, d0i);
        Dft<T, -1>::run(d1r, d1i);
        if (r == 0 && t == 0) {
#pragma unroll
            for (int k2 = 0; k2 < T; k2++) {
                float o[4];
                if (k2 == T / 2) {
                    post_fwd_special(d0r[0], d0i[0], d0r[T / 2], d0i[T / 2], cc.tn(n / 4), o);
                } else {
                    const bool lo = k2 < T / 2;
                    const int j = k2 - T / 2;
                    const int k = lo ? N1 / 2 + N1 * k2 : N1 * j;
                    const float ar = lo ? d1r[k2] : d0r[j];
                    const float ai = lo ? d1i[k2] : d0i[j];
                    const float br = lo ? d1r[T - 1 - k2] : d0r[T - j];
                    const float bi = lo ? d1i[T - 1 - k2] : d0i[T - j];
                    post_fwd(ar, ai, br, bi, cc.om_re(k), cc.om_im(k), cc.tn(k), cc.tn(N - k), o);
                }
#pragma unroll
                for (int w = 0; w < 4; w++)
                    emit(r, k2, w, FAM_C, out_row_g0<n, T>(k2, w), o[w]);
            }
            // Keeps the compiler from sinking the last stores of both branches into one
            // block, which would turn their row offsets into a scratch-memory table.
            __asm__ volatile("" ::: "memory");
        } else {
#pragma unroll
            for (int k2 = 0; k2 < T; k2++) {
                float o[4];
                post_fwd(d0r[k2], d0i[k2], d1r[T - 1 - k2], d1i[T - 1 - k2], cc.oms_re(r, k2), cc.oms_im(r, k2),
                         cc.tns(r, k2), cc.tnns(r, k2), o);
#pragma unroll
                for (int w = 0; w < 4; w++)
                    emit(r, k2, w, out_family<n, T>(w), out_urow<n, T>(r, k2, w), o[w]);
            }
        }
    }
}

// Inverse core.  in(r, k2, which, family, urow) returns the (already scaled) input of that
// output row; st(family, urow, v) stores the lane's element of that row.
template <int n, int THREADS, typename In, typename St>
__device__ __forceinline__ void cteam_inv(int t, int line, In in, St st, float *xl, CConst<n, CTeam<n, THREADS>::M> cc)
{
    using C = CTeam<n, THREADS>;
    constexpr int T = C::T, N = C::N, N1 = C::N1, Q = C::Q;
    float re[N1], im[N1];
    // Several members per wave: the slots of cteam_fwd, lead included (in(.., 4 / 5, ..) are
    // in[0] and in[N], read by the lead only).
    const bool lead = t == 0;
    const CConst<n, C::M> cl(cc.g, cc.l, C::M > 1 && lead ? N1 / 2 : t);
    float lead_in0 = 0.f, lead_inN = 0.f;
    // The whole first wave reads them (a uniform branch): under a lane-divergent one the
    // compiler built every later row address with 64-bit vector arithmetic instead of a
    // scalar base plus a 32-bit offset (1634 against 1328 VALU instructions at n = 1024).
    if (C::M > 1 && __builtin_amdgcn_readfirstlane(t / C::M) == 0) {
        lead_in0 = in(0, 0, 4, FAM_C, 0);
        lead_inN = in(0, 0, 5, FAM_C, N);
    }
#pragma unroll
    for (int r = 0; r < Q; r++) {
        float (&d0r)[T] = sub<T>(re, 2 * r), (&d0i)[T] = sub<T>(im, 2 * r);
        float (&d1r)[T] = sub<T>(re, 2 * r + 1), (&d1i)[T] = sub<T>(im, 2 * r + 1);
        if constexpr (C::M > 1) {
            float ar[T], ai[T], br[T], bi[T];
            float z0r = 0.f, z0i = 0.f, zhr = 0.f, zhi = 0.f;
#pragma unroll
            for (int k2 = 0; k2 < T; k2++) {
                const bool lo = r == 0 && k2 < T / 2;
                const CConst<n, C::M> &P = lo ? cl : cc;
                float o[4];
#pragma unroll
                for (int w = 0; w < 4; w++)
                    o[w] = in(r, k2, w, lo ? out_family_lo<n, T>(w) : out_family<n, T>(w), out_urow<n, T>(r, k2, w));
                post_inv(o, P.oms_re(r, k2), P.oms_im(r, k2), P.tns(r, k2), P.tnns(r, k2), ar[k2], ai[k2], br[k2],
                         bi[k2]);
                if (r == 0 && k2 == T / 2) {
                    const float os[4] = {lead_in0, lead_inN, o[0], o[1]};
                    post_inv_special(os, cc.tn(n / 4), z0r, z0i, zhr, zhi);
                }
            }
            // D0[x] = Z_{g + N1 x}, D1[x] = Z_{N1 - g + N1 x}: from slot x and the b side of
            // slot T-1-x; the lead's group 0 takes D0[0] = Z_0, D0[T/2] = Z_{N/2}, D0[x] from
            // the b side of its high slot T - x (x < T/2), and D1[x] from the a side of its
            // low slot x (x < T/2).
            const bool l0 = r == 0 && lead;
#pragma unroll
            for (int x = 0; x < T; x++) {
                float pr = ar[x], pi = ai[x];
                if (r == 0 && x == 0) {
                    pr = l0 ? z0r : pr;
                    pi = l0 ? z0i : pi;
                } else if (r == 0 && x == T / 2) {
                    pr = l0 ? zhr : pr;
                    pi = l0 ? zhi : pi;
                } else if (r == 0 && x < T / 2) {
                    pr = l0 ? br[T - x] : pr;
                    pi = l0 ? bi[T - x] : pi;
                }
                d0r[x] = pr;
                d0i[x] = pi;
                d1r[x] = x < T / 2 && l0 ? ar[x] : br[T - 1 - x];
                d1i[x] = x < T / 2 && l0 ? ai[x] : bi[T - 1 - x];
            }
        } else if (r == 0 && t == 0) {
#pragma unroll
            for (int k2 = 0; k2 < T; k2++) {
                float o[4];
#pragma unroll
                for (int w = 0; w < 4; w++)
                    o[w] = in(r, k2, w, FAM_C, out_row_g0<n, T>(k2, w));
                if (k2 == T / 2) {
                    post_inv_special(o, cc.tn(n / 4), d0r[0], d0i[0], d0r[T / 2], d0i[T / 2]);
                    continue;
                }
                const bool lo = k2 < T / 2;
                const int j = k2 - T / 2;
                const int k = lo ? N1 / 2 + N1 * k2 : N1 * j;
                float ar, ai, br, bi;
                post_inv(o, cc.om_re(k), cc.om_im(k), cc.tn(k), cc.tn(N - k), ar, ai, br, bi);
                if (lo) {
                    d1r[k2] = ar;
                    d1i[k2] = ai;
                    d1r[T - 1 - k2] = br;
                    d1i[T - 1 - k2] = bi;
                } else {
                    d0r[j] = ar;
                    d0i[j] = ai;
                    d0r[T - j] = br;
                    d0i[T - j] = bi;
                }
            }
            __asm__ volatile("" ::: "memory");   // no sinking across the branches (as above)
        } else {
#pragma unroll
            for (int k2 = 0; k2 < T; k2++) {
                float o[4];
#pragma unroll
                for (int w = 0; w < 4; w++)
                    o[w] = in(r, k2, w, out_family<n, T>(w), out_urow<n, T>(r, k2, w));
                post_inv(o, cc.oms_re(r, k2), cc.oms_im(r, k2), cc.tns(r, k2), cc.tnns(r, k2), d0r[k2], d0i[k2],
                         d1r[T - 1 - k2], d1i[T - 1 - k2]);
            }
        }
        Dft<T, 1>::run(d0r, d0i);
        Dft<T, 1>::run(d1r, d1i);
    }
    // The twiddles of the DFT_N1 stage are applied in pair order, before the exchange: value
    // (k1, member j2) of the lane's groups (k1 = ka or kb) times conj(w(k1, j2)), the same
    // products the member j2 would form after the exchange, k1 = 0 untouched.  After the
    // exchange the twiddle loads would be scheduled next to the full DFT_N1 working set
    // (150 VGPRs at two members per wave against 120 here).
#pragma unroll
    for (int r = 0; r < Q; r++)
#pragma unroll
        for (int s = 0; s < 2; s++)
#pragma unroll
            for (int j2 = 0; j2 < T; j2++) {
                const int g = t + T * r;
                const int k1 = s == 0 ? g : (g == 0 ? N1 / 2 : N1 - g);
                const int v = (2 * r + s) * T + j2;
                const float wr = cc.twp(r, s, j2, 0), wi = cc.twp(r, s, j2, 1);
                const float a = re[v], b = im[v];
                const float x = __builtin_fmaf(a, wr, b * wi), y = __builtin_fmaf(b, wr, -a * wi);
                re[v] = k1 == 0 ? a : x;
                im[v] = k1 == 0 ? b : y;
            }
    cteam_exchange<n, THREADS, false>(t, line, re, im, xl);
    Dft<N1, 1>::run(re, im);
#pragma unroll
    for (int j1 = 0; j1 < N1; j1++) {
        const int fam = in_family<n, T>(j1), u = in_urow<n, T>(j1);
        st(fam, u, re[j1]);
        st(fam, fam == FAM_P ? u + 2 : u - 2, im[j1]);
    }
}

}  // namespace bdx
