// Row-style team lines read and written straight from global memory.
//
// A team of T consecutive lanes of one wave transforms one contiguous row of length n
// (T = team_size<n>(), aligned to T lanes).  Lane t loads float4 number c of the row at
// element 4 t + 4 T c: a team reads 16 T contiguous bytes per instruction.  The Makhoul
// packing needs, for the second half of the packed points, elements 1 and 3 of the
// float4s of the mirrored lane T - 1 - t; one DPP mirror (row_mirror for T = 16,
// row_half_mirror for 8, quad_perm [3,2,1,0] for 4) per value brings them over, with no
// LDS traffic.  The exchange between the two FFT phases goes through a per-team LDS
// region, real parts then imaginary parts, in columns of T slots padded by one so that
// both the writes and the reads of a wave hit distinct banks with immediate-offset
// addresses (an XOR swizzle had cost three to four address instructions per slot).  The forward
// outputs are scattered over the row, so they are collected in the same region and
// written back as whole float4 lines (scattered 64-byte stores write partial cache lines
// and cost extra HBM write traffic).  Per-lane constants (team twiddles, post-process
// twiddles and tangents, normalization) come from a per-block LDS copy of the table
// block (RowTab).  Algorithm and index maps as in bdx_lines.h and bdx_post.h.
//
// With NT (streaming calls) only the loads carry the non-temporal hint.  The stores stay
// cacheable, because a row pass's output is the input of the column pass that follows at
// once.  On 64 x 1024^2 transforms this measured 197 / 194 us per direction against
// 222 / 214 us with streaming stores.
#pragma once
#include "bdx_lines.h"
#include "bdx_post.h"

namespace bdx {

template <int T> __device__ __forceinline__ float team_mirror(float x)
{
    constexpr int ctrl = T == 16 ? 0x140 : T == 8 ? 0x141 : T == 4 ? 0x1B : 0xB1;
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), ctrl, 0xF, 0xF, false));
}

// LDS copy of the constants of axis length n used by a team; XCH is the per-team LDS
// region (forward: FFT exchange, then the output row; inverse: the input row, then the
// FFT exchange).  It holds half a row: the FFT exchange moves n/2 values at a time (in
// N1 columns of T slots padded to T + 1, see xpad), and the output (input) row is staged
// in two rounds of half a row -- outputs 0 and 2 of every slot lie in the first half, 1
// and 3 in the second.  At n = 1024 that brings a 16-row workgroup from 80.9 to about 50
// KB of LDS, three workgroups per CU instead of two.  The region size is rounded up to
// T modulo 32 floats, so that neighbouring teams start T banks apart and their exchange
// accesses cover distinct banks (rounding to T modulo 64 measured the same).
template <int n> struct RowTab {
    static constexpr int N = n / 2;
    static constexpr int TW = 0;                // team twiddles, 2 (k1 T + t)
    static constexpr int OM = n;                // exp(-2 pi i k / n), k < N, re/im
    static constexpr int TN = 2 * n;            // tan(pi k / (2n)), k <= N
    static constexpr int AL = 2 * n + N + 4;    // normalization (alpha or alpha^2)
    static constexpr int SIZE = AL + n;
    static constexpr int XMIN = n / 2 + n / 2 / BDX_TEAM(n);   // N1 (T + 1) exchange slots >= half a row
    static constexpr int XCH = XMIN + ((BDX_TEAM(n) - XMIN) % 32 + 32) % 32;
    static_assert(SIZE % 4 == 0 && XCH % 4 == 0, "team regions must stay 16-byte aligned");
};

template <int n, int OP, int THREADS>
__device__ __forceinline__ void load_row_tables(float *lt, const float *__restrict__ tab)
{
    using R = RowTab<n>;
    const float *g = tab + BDX_TABLE_BASE(n);
    for (int i = threadIdx.x; i < n; i += THREADS) {
        lt[R::TW + i] = g[5 * n + i];
        lt[R::OM + i] = g[i];
        lt[R::AL + i] = g[(OP == FA2 ? 4 : 3) * n + i];
    }
    for (int i = threadIdx.x; i <= R::N; i += THREADS)
        lt[R::TN + i] = g[2 * n + i];
}

// Exchange slot of (column k1, member j): k1 (T + 1) + j.  A lane writes one column per
// member slot (k1 varies at compile time, j = its own member index) and reads the T slots
// of its pair columns (j varies at compile time, k1 = its pair group), so every address is
// a per-lane base plus an immediate offset.  Banks: the T lanes of a team write T
// consecutive slots, and read slots (T + 1) k1 + j with k1 differing by 1 between
// neighbouring lanes, which spreads them over distinct banks (T + 1 is odd).
template <int T> __device__ __forceinline__ constexpr int xpad(int k1, int j)
{
    return k1 * (T + 1) + j;
}

// Per-lane pointers of the post-process slots of a team.  Slot k2 of pair group
// g = base + T r (r < Q) handles the pair (k, N - k), k = base + T r + N1 k2.  Each pointer
// sits at the lowest address its family reaches, so every access is a pointer plus a
// non-negative immediate offset:
//   om[2k], om[2k+1]           om + 2 u           (u = T r + N1 k2)
//   tan(k), al[k], al[N + k]    up + {0, AL - TN, AL - TN + N} + u
//   tan(N - k), al[N - k],      dn + {0, AL - TN, AL - TN + N} + d   (d = SPAN - u)
//   al[n - k]
//   half-row slots [k], [N - k] xu + u, xd + d
template <int n> struct SlotPtrs {
    using R = RowTab<n>;
    static constexpr int T = BDX_TEAM(n), N = n / 2, N1 = N / T, Q = N1 / (2 * T);
    static constexpr int SPAN = T * (Q - 1) + N1 * (T - 1);   // largest u
    static constexpr int A = R::AL - R::TN;
    static_assert(N - T - SPAN >= 0, "the downward pointers stay inside their regions");
    const float *om, *up, *dn;
    float *xu, *xd;
    __device__ __forceinline__ SlotPtrs(const float *lt, float *xl, int base)
        : om(lt + R::OM + 2 * base), up(lt + R::TN + base), dn(lt + R::TN + N - base - SPAN), xu(xl + base),
          xd(xl + N - base - SPAN)
    {
    }
    static constexpr int u(int r, int k2) { return T * r + N1 * k2; }
    static constexpr int d(int r, int k2) { return SPAN - u(r, k2); }
    __device__ __forceinline__ float om_re(int r, int k2) const { return om[2 * u(r, k2)]; }
    __device__ __forceinline__ float om_im(int r, int k2) const { return om[2 * u(r, k2) + 1]; }
    __device__ __forceinline__ float tn_k(int r, int k2) const { return up[u(r, k2)]; }
    __device__ __forceinline__ float tn_nk(int r, int k2) const { return dn[d(r, k2)]; }
    // normalization of output w of the slot: out[k], out[n - k], out[N - k], out[N + k]
    __device__ __forceinline__ float al(int r, int k2, int w) const
    {
        return w == 0 ? up[A + u(r, k2)] : w == 1 ? dn[A + N + d(r, k2)] : w == 2 ? dn[A + d(r, k2)] : up[A + N + u(r, k2)];
    }
    // half-row slot of output w (outputs 1 and 3 in the second half, local index - N)
    __device__ __forceinline__ float &x(int r, int k2, int w) const
    {
        return w == 0 || w == 3 ? xu[u(r, k2)] : xd[d(r, k2)];
    }
};

template <int n, int OP, bool NT = false>
__device__ __forceinline__ void rteam_fwd(const float *src, float *dst, int t, float *xl, const float *lt)
{
    using R = RowTab<n>;
    constexpr int T = team_size<n>(), N = n / 2, N1 = N / T, Q = N1 / (2 * T), C = N1 / 2;
    float re[N1], im[N1];
    {
        float4 v[C];
#pragma unroll
        for (int c = 0; c < C; c++)
            v[c] = ld4<NT>(src + 4 * t + 4 * T * c);
#pragma unroll
        for (int c = 0; c < C; c++) {
            re[c] = v[c].x;
            im[c] = v[c].z;
            re[N1 - 1 - c] = team_mirror<T>(v[c].w);
            im[N1 - 1 - c] = team_mirror<T>(v[c].y);
        }
    }
    Dft<N1, -1>::run(re, im);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k1 = 1; k1 < N1; k1++) {
        const float wr = lt[R::TW + 2 * (k1 * T + t)], wi = lt[R::TW + 2 * (k1 * T + t) + 1];
        const float a = re[k1], b = im[k1];
        re[k1] = __builtin_fmaf(a, wr, -b * wi);
        im[k1] = __builtin_fmaf(a, wi, b * wr);
    }
    float dr[Q][2][T], di[Q][2][T];
#pragma unroll
    for (int k1 = 0; k1 < N1; k1++)
        xl[xpad<T>(k1, t)] = re[k1];
    team_sync();
#pragma unroll
    for (int r = 0; r < Q; r++) {
        const int g = t + T * r;
        const int ka = g == 0 ? 0 : g, kb = g == 0 ? N1 / 2 : N1 - g;
#pragma unroll
        for (int j2 = 0; j2 < T; j2++) {
            dr[r][0][j2] = xl[xpad<T>(ka, j2)];
            dr[r][1][j2] = xl[xpad<T>(kb, j2)];
        }
    }
    team_sync();
#pragma unroll
    for (int k1 = 0; k1 < N1; k1++)
        xl[xpad<T>(k1, t)] = im[k1];
    team_sync();
#pragma unroll
    for (int r = 0; r < Q; r++) {
        const int g = t + T * r;
        const int ka = g == 0 ? 0 : g, kb = g == 0 ? N1 / 2 : N1 - g;
#pragma unroll
        for (int j2 = 0; j2 < T; j2++) {
            di[r][0][j2] = xl[xpad<T>(ka, j2)];
            di[r][1][j2] = xl[xpad<T>(kb, j2)];
        }
    }
    team_sync();   // the region is free for the next line once every lane has read
    __builtin_amdgcn_sched_barrier(0);
    // Slot k2 of pair group g = t + T r yields the outputs k, n - k, N - k and N + k of the
    // pair (k, N - k), k = g + N1 k2, from (D0[k2], D1[T-1-k2]).  Lane t = 0 of group r = 0
    // owns the self-paired groups 0 and N1/2 instead (RowSlots): its low slots take k = N1/2 +
    // N1 k2 from (D1[k2], D1[T-1-k2]), its high slots k = N1 k2 from (D0[k2], D0[T-k2]) --
    // the pair (N - k, k) seen from its other side, which post_fwd maps to the same four
    // outputs bitwise -- and its slot T/2 holds Z_0 and Z_{N/2}.  Every slot's k is thus a
    // per-lane base plus N1 k2, and every table and output address a per-lane base plus an
    // immediate offset.
    const bool lead = t == 0;
    const SlotPtrs<n> lo(lt, xl, lead ? N1 / 2 : t), hi(lt, xl, t);   // slots k2 < T/2 of group 0; the others
    float second[Q][T][2];
    float lead_first = 0.f, lead_second = 0.f;   // out[0] and out[N] (lane t = 0 only)
#pragma unroll
    for (int r = 0; r < Q; r++) {
        Dft<T, -1>::run(dr[r][0], di[r][0]);
        Dft<T, -1>::run(dr[r][1], di[r][1]);
#pragma unroll
        for (int k2 = 0; k2 < T; k2++) {
            const SlotPtrs<n> &P = r == 0 && k2 < T / 2 ? lo : hi;
            const bool swap_a = r == 0 && k2 < T / 2, swap_b = r == 0 && k2 > T / 2;
            const float ar = swap_a && lead ? dr[r][1][k2] : dr[r][0][k2];
            const float ai = swap_a && lead ? di[r][1][k2] : di[r][0][k2];
            const float br = swap_b && lead ? dr[r][0][T - k2] : dr[r][1][T - 1 - k2];
            const float bi = swap_b && lead ? di[r][0][T - k2] : di[r][1][T - 1 - k2];
            float o[4];
            post_fwd(ar, ai, br, bi, P.om_re(r, k2), P.om_im(r, k2), P.tn_k(r, k2), P.tn_nk(r, k2), o);
            if (r == 0 && k2 == T / 2) {
                // lane 0: k = N/2, so its four outputs land on n/4, 3n/4, n/4, 3n/4; it
                // writes out[n/4] and out[3n/4] there (twice each) and out[0], out[N] apart
                float sp[4];
                post_fwd_special(dr[0][0][0], di[0][0][0], dr[0][0][T / 2], di[0][0][T / 2],
                                 lt[R::TN + n / 4], sp);
#pragma unroll
                for (int w = 0; w < 4; w++)
                    o[w] = lead ? sp[2 + (w & 1)] : o[w];
                lead_first = sp[0];
                lead_second = sp[1];
                if constexpr (OP == FA || OP == FA2) {
                    lead_first *= lt[R::AL];
                    lead_second *= lt[R::AL + N];
                }
            }
            // outputs 0 and 2 of a slot lie in the first half of the row, 1 and 3 in the
            // second: the first go to the half-row region now, the second wait in registers
#pragma unroll
            for (int w = 0; w < 4; w++) {
                float y = o[w];
                if constexpr (OP == FA || OP == FA2)
                    y *= P.al(r, k2, w);
                if (w % 2 == 0)
                    P.x(r, k2, w) = y;
                else
                    second[r][k2][w / 2] = y;
            }
        }
    }
    if (lead)
        xl[0] = lead_first;
    team_sync();
#pragma unroll
    for (int c = 0; c < N / (4 * T); c++) {
        const int e = 4 * t + 4 * T * c;
        st4<false>(dst + e, make_float4(xl[e], xl[e + 1], xl[e + 2], xl[e + 3]));
    }
    team_sync();
#pragma unroll
    for (int r = 0; r < Q; r++)
#pragma unroll
        for (int k2 = 0; k2 < T; k2++) {
            const SlotPtrs<n> &P = r == 0 && k2 < T / 2 ? lo : hi;
            P.x(r, k2, 1) = second[r][k2][0];
            P.x(r, k2, 3) = second[r][k2][1];
        }
    if (lead)
        xl[0] = lead_second;
    team_sync();
#pragma unroll
    for (int c = 0; c < N / (4 * T); c++) {
        const int e = 4 * t + 4 * T * c;
        st4<false>(dst + N + e, make_float4(xl[e], xl[e + 1], xl[e + 2], xl[e + 3]));
    }
    team_sync();
}

// Inverse: inputs at the lane-owned output indices of the forward, outputs as float4
// (the mirror of the forward load).
template <int n, int OP, bool NT = false>
__device__ __forceinline__ void rteam_inv(const float *src, float *dst, int t, float *xl, const float *lt)
{
    using R = RowTab<n>;
    constexpr int T = team_size<n>(), N = n / 2, N1 = N / T, Q = N1 / (2 * T), C = N1 / 2;
    // The spectrum is read in pairs (k, n - k) spread over the row: stage the row in the
    // team's LDS region with whole-line float4 loads first (scattered global loads of half
    // lines would be fetched twice once the rows in flight outgrow the L2).
    // The staging runs in two rounds of half a row: inputs 0 and 2 of a slot lie in the
    // first half, 1 and 3 in the second.
    float4 v[n / (4 * T)];
#pragma unroll
    for (int c = 0; c < n / (4 * T); c++)
        v[c] = ld4<NT>(src + 4 * t + 4 * T * c);
#pragma unroll
    for (int c = 0; c < N / (4 * T); c++)
        *reinterpret_cast<float4 *>(xl + 4 * t + 4 * T * c) = v[c];
    team_sync();
    // Slots as in rteam_fwd: slot k2 of pair group g = t + T r reads the inputs at k, n - k,
    // N - k and N + k, k = per-lane base + N1 k2; lane t = 0 of group 0 takes the base N1/2
    // below T/2, and at slot T/2 (k = N/2) finds in[n/4] and in[3n/4], with in[0] and in[N]
    // read apart.
    const bool lead = t == 0;
    const SlotPtrs<n> lo(lt, xl, lead ? N1 / 2 : t), hi(lt, xl, t);
    const auto scaled = [&](float y, float a) {
        if constexpr (OP == IA)
            y *= a;
        return y;
    };
    float first[Q][T][2];
#pragma unroll
    for (int r = 0; r < Q; r++)
#pragma unroll
        for (int k2 = 0; k2 < T; k2++) {
            const SlotPtrs<n> &P = r == 0 && k2 < T / 2 ? lo : hi;
            first[r][k2][0] = scaled(P.x(r, k2, 0), P.al(r, k2, 0));
            first[r][k2][1] = scaled(P.x(r, k2, 2), P.al(r, k2, 2));
        }
    const float lead_in0 = scaled(xl[0], lt[R::AL]);   // in[0] (used by lane 0 only)
    team_sync();
#pragma unroll
    for (int c = 0; c < N / (4 * T); c++)
        *reinterpret_cast<float4 *>(xl + 4 * t + 4 * T * c) = v[N / (4 * T) + c];
    team_sync();
    const float lead_inN = scaled(xl[0], lt[R::AL + N]);   // in[N]
    float dr[Q][2][T], di[Q][2][T];
#pragma unroll
    for (int r = 0; r < Q; r++) {
        // A, B: Z_k and Z_{N-k} of slot k2
        float ar[T], ai[T], br[T], bi[T];
        float z0r = 0.f, z0i = 0.f, zhr = 0.f, zhi = 0.f;
#pragma unroll
        for (int k2 = 0; k2 < T; k2++) {
            const SlotPtrs<n> &P = r == 0 && k2 < T / 2 ? lo : hi;
            const float o[4] = {first[r][k2][0], scaled(P.x(r, k2, 1), P.al(r, k2, 1)), first[r][k2][1],
                                scaled(P.x(r, k2, 3), P.al(r, k2, 3))};
            post_inv(o, P.om_re(r, k2), P.om_im(r, k2), P.tn_k(r, k2), P.tn_nk(r, k2), ar[k2], ai[k2], br[k2],
                     bi[k2]);
            if (r == 0 && k2 == T / 2) {
                const float os[4] = {lead_in0, lead_inN, o[0], o[1]};
                post_inv_special(os, lt[R::TN + n / 4], z0r, z0i, zhr, zhi);
            }
        }
        // D0[x] = Z_{g + N1 x}, D1[x] = Z_{N1 - g + N1 x}: from slot x and from the b side of
        // slot T-1-x; lane 0 of group 0 (groups 0 and N1/2) takes D0[0] = Z_0, D0[T/2] =
        // Z_{N/2}, D0[x] from the b side of its high slot T - x (x < T/2), D1[x] from the a
        // side of its low slot x (x < T/2).
        const bool l0 = r == 0 && lead;
#pragma unroll
        for (int x = 0; x < T; x++) {
            float pr = ar[x], pi = ai[x];
            if (x == 0) {
                pr = l0 ? z0r : pr;
                pi = l0 ? z0i : pi;
            } else if (x == T / 2) {
                pr = l0 ? zhr : pr;
                pi = l0 ? zhi : pi;
            } else if (x < T / 2) {
                pr = l0 ? br[T - x] : pr;
                pi = l0 ? bi[T - x] : pi;
            }
            dr[r][0][x] = pr;
            di[r][0][x] = pi;
            dr[r][1][x] = x < T / 2 && l0 ? ar[x] : br[T - 1 - x];
            di[r][1][x] = x < T / 2 && l0 ? ai[x] : bi[T - 1 - x];
        }
        Dft<T, 1>::run(dr[r][0], di[r][0]);
        Dft<T, 1>::run(dr[r][1], di[r][1]);
    }
    __builtin_amdgcn_sched_barrier(0);
    team_sync();   // every lane has read the staged row before the exchange reuses the region
    float re[N1], im[N1];
#pragma unroll
    for (int r = 0; r < Q; r++) {
        const int g = t + T * r;
        const int ka = g == 0 ? 0 : g, kb = g == 0 ? N1 / 2 : N1 - g;
#pragma unroll
        for (int j2 = 0; j2 < T; j2++) {
            xl[xpad<T>(ka, j2)] = dr[r][0][j2];
            xl[xpad<T>(kb, j2)] = dr[r][1][j2];
        }
    }
    team_sync();
#pragma unroll
    for (int k1 = 0; k1 < N1; k1++)
        re[k1] = xl[xpad<T>(k1, t)];
    team_sync();
#pragma unroll
    for (int r = 0; r < Q; r++) {
        const int g = t + T * r;
        const int ka = g == 0 ? 0 : g, kb = g == 0 ? N1 / 2 : N1 - g;
#pragma unroll
        for (int j2 = 0; j2 < T; j2++) {
            xl[xpad<T>(ka, j2)] = di[r][0][j2];
            xl[xpad<T>(kb, j2)] = di[r][1][j2];
        }
    }
    team_sync();
#pragma unroll
    for (int k1 = 0; k1 < N1; k1++)
        im[k1] = xl[xpad<T>(k1, t)];
    team_sync();
#pragma unroll
    for (int k1 = 1; k1 < N1; k1++) {
        const float wr = lt[R::TW + 2 * (k1 * T + t)], wi = lt[R::TW + 2 * (k1 * T + t) + 1];
        const float a = re[k1], b = im[k1];
        re[k1] = __builtin_fmaf(a, wr, b * wi);
        im[k1] = __builtin_fmaf(b, wr, -a * wi);
    }
    Dft<N1, 1>::run(re, im);
#pragma unroll
    for (int c = 0; c < C; c++) {
        const float4 v = make_float4(re[c], team_mirror<T>(im[N1 - 1 - c]), im[c],
                                     team_mirror<T>(re[N1 - 1 - c]));
        st4<false>(dst + 4 * t + 4 * T * c, v);
    }
}

}  // namespace bdx
