// Block-level work units of the 2D transform, shared by the per-shape kernels
// (bdx_small.hip, bdx_passes.hip) and the persistent batch kernel (bdx_batch.hip).
//
// Every unit runs on one workgroup of THREADS lanes, uses the LDS area `smem` it is given
// (at most unit_lds<...>() floats) and touches only its own part of the output, so units
// of the same pass can run in any order and concurrently.
//   small_unit   count whole h x w arrays: load, rows, columns (and the inverse halves of
//                the fused round trip), store; one HBM read and write per pixel.
//   row_unit     a band of full rows of a two-pass array (row pass input -> output).
//   col_unit     a strip of full columns of a two-pass array, in place on the output.
// Reference equivalents: src/rect.c:5-23 (column pass, banded row pass, normalization),
// src/generated/small_optimized.c (whole small arrays).
#pragma once
#include "bdx_kernels.h"
#include "bdx_lines.h"
#include "bdx_rteam.h"
#include "bdx_wteam.h"

namespace bdx {

// ---- whole arrays --------------------------------------------------------------------

template <int H, int W, int THREADS> constexpr int small_per_unit(int lds_floats) { return lds_floats / (H * (W + 1)); }

// `count` whole arrays; row r of the sequence of their rows (array r / H, row r % H) is read
// at src(r) and written at dst(r).
template <int H, int W, int MODE, int THREADS, bool NT = false, typename Src, typename Dst>
__device__ __forceinline__ void small_unit_rows(Src src, Dst dst, int count, float *tile,
                                                const float *__restrict__ tab)
{
    constexpr int P = W + 1;
    const auto row_base = [](int L) { return L * P; };
    const auto col_base = [](int L) {
        const int a = L / W, c = L - a * W;
        return a * H * P + c;
    };
    tile_load_rows<W, P, THREADS, NT>(tile, src, count * H);
    __syncthreads();
    if constexpr (MODE == 0) {
        run_lines<W, 1, FA, THREADS>(tile, count * H, row_base, tab);
        __syncthreads();
        run_lines<H, P, FA, THREADS>(tile, count * W, col_base, tab);
    } else if constexpr (MODE == 1) {
        run_lines<W, 1, IA, THREADS>(tile, count * H, row_base, tab);
        __syncthreads();
        run_lines<H, P, IA, THREADS>(tile, count * W, col_base, tab);
    } else {
        run_lines<W, 1, FA2, THREADS>(tile, count * H, row_base, tab);
        __syncthreads();
        run_lines<H, P, FA2, THREADS>(tile, count * W, col_base, tab);
        __syncthreads();
        run_lines<H, P, I0, THREADS>(tile, count * W, col_base, tab);
        __syncthreads();
        run_lines<W, 1, I0, THREADS>(tile, count * H, row_base, tab);
    }
    __syncthreads();
    tile_store_rows<W, P, THREADS, NT>(tile, dst, count * H);
}

// `count` contiguous whole arrays starting at element `off`.
template <int H, int W, int MODE, int THREADS, bool NT = false>
__device__ __forceinline__ void small_unit(const float *in, float *out, size_t off, int count, float *tile,
                                           const float *__restrict__ tab)
{
    small_unit_rows<H, W, MODE, THREADS, NT>([=](int r) { return in + off + (size_t)r * W; },
                                             [=](int r) { return out + off + (size_t)r * W; }, count, tile, tab);
}

// ---- row pass ------------------------------------------------------------------------

// Rows per unit: one team per row for the team path, one lane per row otherwise.
template <int W, int THREADS> constexpr int row_unit_rows() { return W >= 128 ? THREADS / team_size<W>() : THREADS; }

template <int W, int OP, int THREADS> constexpr int row_unit_lds()
{
    if constexpr (W >= 128)
        return RowTab<W>::SIZE + row_unit_rows<W, THREADS>() * RowTab<W>::XCH;
    else
        return THREADS * (W + 1);
}

// Rows [row0, row0 + nrows) of the row sequence (all rows of all arrays of the same width
// are one sequence), nrows <= row_unit_rows.  With `tables_ready` the LDS constant table
// of width W is already in smem.  NT streams the loads only: the output stays cacheable
// for the column pass that reads it next.
template <int W, int OP, int THREADS, bool NT = false>
__device__ __forceinline__ void row_unit(const float *in, float *out, size_t row0, int nrows, float *smem,
                                         const float *__restrict__ tab, bool tables_ready)
{
    if constexpr (W >= 128) {
        using R = RowTab<W>;
        constexpr int T = team_size<W>();
        constexpr int X = R::XCH;
        float *lt = smem, *xl = smem + R::SIZE;
        if (!tables_ready) {
            load_row_tables<W, OP, THREADS>(lt, tab);
            __syncthreads();
        }
        const int team = threadIdx.x / T, t = threadIdx.x % T;
        const size_t row = row0 + team;
        if (team < nrows) {   // a team works alone on its row: no block barrier follows
            if constexpr (OP == FA || OP == FA2)
                rteam_fwd<W, OP, NT>(in + row * W, out + row * W, t, xl + team * X, lt);
            else
                rteam_inv<W, OP, NT>(in + row * W, out + row * W, t, xl + team * X, lt);
        }
    } else {
        constexpr int P = W + 1;
        tile_load<W, P, THREADS, NT>(smem, in + row0 * W, W, nrows);
        __syncthreads();
        run_lines<W, 1, OP, THREADS>(smem, nrows, [](int L) { return L * P; }, tab);
        __syncthreads();
        tile_store<W, P, THREADS, false>(smem, out + row0 * W, W, nrows);   // read again by the column pass
    }
}

// ---- column pass ---------------------------------------------------------------------

template <int H, int THREADS> constexpr int col_unit_cols()
{
    if constexpr (H >= 128)
        return CTeam<H, THREADS>::L;
    else
        return THREADS;
}

template <int H, int THREADS> constexpr int col_unit_lds()
{
    if constexpr (H >= 128)
        return CTeam<H, THREADS>::XLDS + CTeam<H, THREADS>::TLDS;
    else
        return 1;
}

// Columns [c0, c0 + ncols) of the h x w array at `base` (in and out may alias),
// 1 <= ncols <= col_unit_cols.  The lines of lanes past a partial strip repeat its last
// column: they read it, compute the same bits and store them to the same addresses (every
// load of a block precedes its exchange barriers, every store follows them, so the
// in-place repeat reads only input).  No access is predicated, so a strip width known only
// at run time costs nothing (per-lane predicates on every access had taken the length-1024
// unit from 196 to 248 VGPRs as a function of the persistent batch kernel).
// MODE 0 forward * alpha, 1 alpha * inverse, 2 forward * alpha^2 then inverse.
template <int H, int MODE, int THREADS, bool NT = false>
__device__ __forceinline__ void col_unit(const float *in, float *out, int w, size_t base, int c0, int ncols,
                                         float *smem, const float *__restrict__ tab, bool tables_ready)
{
    if constexpr (H >= 128) {
        using C = CTeam<H, THREADS>;
        constexpr int T = C::T, Q = C::Q, L = C::L, M = C::M;
        // constant table first: its offsets then fit the 16-bit immediate of DS
        // instructions (after the 64 KiB exchange area they did not, and every table read
        // took an address instruction)
        float *lt = smem, *xl = smem + (M > 1 ? C::TLDS : 0);
        if (M > 1 && !tables_ready) {
            load_row_tables<H, MODE == 2 ? FA2 : FA, THREADS>(lt, tab);
            __syncthreads();
        }
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const int t = M == 1 ? __builtin_amdgcn_readfirstlane(wave) : wave * M + lane / L;
        const int line = lane % L;
        const int col = line < ncols ? line : ncols - 1;
        // Row r of the strip is at the uniform strip base + r w floats.  The lane keeps one
        // byte offset per row family (lane part of the row, plus its column); each access
        // adds a uniform row offset in scalar registers (an array is at most 4 MiB, so the
        // offsets fit 32 bits).
        const unsigned w4 = 4u * (unsigned)w, l4 = 4u * (unsigned)col;
        const int row_p = 4 * t, row_r = 4 * (T - 1 - t), row_c = t, row_d = T - 1 - t;
        // Low slots of the lead member (several members per wave, t = 0): lane part N1/2.  The
        // DL part can be negative, so it is kept biased by BL rows (and its uniform rows
        // lowered by as many): every offset stays a non-negative 32-bit value.
        constexpr int BL = C::N1 / 2;
        static_assert(min_dl_urow<H, T>() >= BL, "lowered DL rows must stay non-negative");
        const int row_cl = M > 1 && t == 0 ? C::N1 / 2 : t, row_dl = T - 1 - row_cl + BL;
        const unsigned off_p = (unsigned)row_p * w4 + l4, off_r = (unsigned)row_r * w4 + l4;
        const unsigned off_c = (unsigned)row_c * w4 + l4, off_d = (unsigned)row_d * w4 + l4;
        const unsigned off_cl = (unsigned)row_cl * w4 + l4, off_dl = (unsigned)row_dl * w4 + l4;
        const auto off = [=](int fam) {
            return fam == FAM_P ? off_p
                 : fam == FAM_R ? off_r
                 : fam == FAM_C ? off_c
                 : fam == FAM_D ? off_d
                 : fam == FAM_CL ? off_cl
                                 : off_dl;
        };
        // (the DL rows a lead reads or writes, N - k and n - k with k < N/2, exceed BL: the
        // lowered row is still non-negative, and unsigned keeps the scalar-base form)
        const auto urow_of = [](int fam, int urow) { return (size_t)(unsigned)(fam == FAM_DL ? urow - BL : urow); };
        const char *src = reinterpret_cast<const char *>(in + base + c0);
        char *dst = reinterpret_cast<char *>(out + base + c0);
        const auto ld = [&](int fam, int urow) {
            return ld1<NT>(reinterpret_cast<const float *>(src + urow_of(fam, urow) * w4 + off(fam)));
        };
        const auto st = [&](int fam, int urow, float v) {
            st1<NT>(reinterpret_cast<float *>(dst + urow_of(fam, urow) * w4 + off(fam)), v);
        };
        const float *al = M == 1 ? alpha_of<H>(tab, MODE == 2 ? 2 : 1) : lt + RowTab<H>::AL;
        // one pointer per row family, so that each factor is that pointer plus an immediate
        const float *al_p = al + row_p, *al_r = al + row_r, *al_c = al + row_c, *al_d = al + row_d;
        const float *al_cl = al + row_cl, *al_dl = al + row_dl;
        const auto alpha = [&](int fam, int urow) {
            return (fam == FAM_P    ? al_p
                    : fam == FAM_R  ? al_r
                    : fam == FAM_C  ? al_c
                    : fam == FAM_D  ? al_d
                    : fam == FAM_CL ? al_cl
                                    : al_dl)[urow_of(fam, urow)];
        };
        const CConst<H, M> cc(tab + BDX_TABLE_BASE(H), lt, t);
        if constexpr (MODE == 1) {
            cteam_inv<H, THREADS>(
                t, line, [&](int, int, int, int fam, int urow) { return ld(fam, urow) * alpha(fam, urow); }, st, xl,
                cc);
        } else if constexpr (MODE == 0) {
            cteam_fwd<H, THREADS>(t, line, ld, xl, cc, [&](int, int, int c, int fam, int urow, float v) {
                if (c < 4 || t == 0)   // outputs 4 and 5 belong to the lead only
                    st(fam, urow, v * alpha(fam, urow));
            });
        } else {
            // outputs 4 and 5 (the lead's rows 0 and N with several members per wave) in ex
            float o[Q][T][4], ex[2] = {0.f, 0.f};
            cteam_fwd<H, THREADS>(t, line, ld, xl, cc, [&](int r, int k2, int c, int fam, int urow, float v) {
                if (c >= 4)
                    ex[c - 4] = v * alpha(fam, urow);
                else
                    o[r][k2][c] = v * alpha(fam, urow);
            });
            cteam_inv<H, THREADS>(
                t, line, [&](int r, int k2, int c, int, int) { return c >= 4 ? ex[c - 4] : o[r][k2][c]; }, st, xl,
                cc);
        }
    } else {
        // one lane per column: the generated Shao-Johnson core on H values in registers
        if ((int)threadIdx.x >= ncols)
            return;
        const size_t off = base + c0 + threadIdx.x;
        const float *al = alpha_of<H>(tab, 1);
        const float *al2 = alpha_of<H>(tab, 2);
        float x[H];
#pragma unroll
        for (int i = 0; i < H; i++)
            x[i] = ld1<NT>(in + off + (size_t)i * w) * (MODE == 1 ? al[i] : 1.0f);
        if constexpr (MODE == 1) {
            Sj<H>::inv(x);
        } else {
            Sj<H>::fwd(x);
#pragma unroll
            for (int i = 0; i < H; i++)
                x[i] *= MODE == 0 ? al[i] : al2[i];
            if constexpr (MODE == 2)
                Sj<H>::inv(x);
        }
#pragma unroll
        for (int i = 0; i < H; i++)
            st1<NT>(out + off + (size_t)i * w, x[i]);
    }
}

}  // namespace bdx
