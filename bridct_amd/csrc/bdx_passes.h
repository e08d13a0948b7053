// Templates of the row-pass and column-pass kernels and their launchers, instantiated by
// bdx_passes.hip (lines of 64 and more) and bdx_passes_lines.hip (lines of 8 to 32, used
// only where the schedule table sends a whole-array shape to the two-pass kernels).  The
// short lines sit in their own code object so that adding them leaves the layout of the
// hot kernels' code object as it was: with them in the same one, every kernel timed the
// same alone but the batch-64 four-shape step on four streams ran about 1% slower.
// Per-shape row-pass and column-pass kernels for arrays larger than one workgroup's LDS
// (reference equivalent: the banded row pass and column pass of src/rect.c:5-23).  The
// block bodies are the work units of bdx_units.h:
//   k_rows<W, OP>:  row_unit bands; rows of all arrays of a batch are one sequence.
//   k_cols<H, MODE>: col_unit strips of one array, in place on the output; one wave per
//                   team member (column teams), one lane per column below 128.
//   k_cols_narrow:  the same column teams on strips of 32 or 16 columns, for grids that
//                   would leave CUs idle.
// Streaming calls (input plus output beyond the 256 MB Infinity Cache) use non-temporal
// loads in the row pass, whose output stays cacheable for the column pass, and
// non-temporal loads and stores in the column pass.
// The gather forms (G) take array i of the batch at element offs[i] of in and out instead
// of at i * h * w (the arrays of one shape inside a ragged batch).  A row band never
// crosses arrays: every two-pass height is a multiple of the band.
#pragma once
#include "bdx_units.h"

namespace bdx {
namespace {

// Lanes per column-pass workgroup: one wave per team member up to 512 lanes (uniform member
// index: scalar row addresses and constants), else two members per wave.  Measured on
// 64 x 1024^2 (length-1024 team, 16 members):
//   forward: 512 lanes (two members per wave, 103 VGPRs, two workgroups per CU) 99.7 us,
//            1024 lanes (one member per wave, 97 VGPRs, one workgroup per CU) 115.8 us;
//   inverse: 512 lanes (120 VGPRs, two workgroups per CU) 104.5 us; 1024 lanes (one
//            workgroup per CU) 116.0 us.
// The fused round trip (141 VGPRs) runs one 512-lane workgroup per CU; at 1024 lanes its
// length-1024 team spills.
template <int H, int MODE> constexpr int col_threads()
{
    return H >= 128 ? (64 * team_size<H>() < 512 ? 64 * team_size<H>() : 512) : kThreads;
}

// Rows of 512 and 1024 fit three workgroups per CU by LDS; the bound keeps their
// registers within three waves per SIMD (168 VGPRs).
template <int W> constexpr int rows_min_blocks() { return W >= 512 ? 3 : 1; }

// Block order that gives each XCD one contiguous range of the grid.  Workgroups are
// dispatched to the `xcds` XCDs of the device round-robin (workgroup b on XCD b % xcds),
// so block b takes position (b % xcds) (nb / xcds) + b / xcds: a permutation of [0, nb)
// when nb is a multiple of xcds, else the plain order.  `xcds` is the device's XCD count
// (8 on an MI355X in SPX mode; 1 in CPX mode, where the order is the plain one).
__device__ __forceinline__ unsigned xcd_order(unsigned b, unsigned nb, unsigned xcds)
{
    return xcds > 1 && nb % xcds == 0 ? (b % xcds) * (nb / xcds) + b / xcds : b;
}

template <int W, int OP, bool G, bool NT>
__global__ __launch_bounds__(kThreads, rows_min_blocks<W>()) void k_rows(const float *in, float *out,
                                                                          const size_t *__restrict__ offs, int h,
                                                                          const float *__restrict__ tab,
                                                                          unsigned xcds)
{
    __shared__ float smem[row_unit_lds<W, OP, kThreads>()];
    constexpr int R = row_unit_rows<W, kThreads>();
    // The XCD-aware order of the column pass: each XCD takes a contiguous range of bands,
    // i.e. the same arrays its column strips take next.  2-3% faster on 64 x 128x1024 and
    // 64 x 512x128 (cache-resident calls), and the batch-64 four-shape step runs 4% faster,
    // although the 1024^2 column pass timed alone after it is 5% slower.
    const unsigned bid = xcd_order(blockIdx.x, gridDim.x, xcds);
    const size_t row0 = (size_t)bid * R;
    if constexpr (G) {
        const size_t a = row0 / (unsigned)h, base = offs[a] + (row0 - a * h) * W;
        row_unit<W, OP, kThreads, NT>(in + base, out + base, 0, R, smem, tab, false);
    } else {
        row_unit<W, OP, kThreads, NT>(in, out, row0, R, smem, tab, false);
    }
}

// Strips of TH / 64 * (64 / M) columns per workgroup of TH lanes.
template <int H, int MODE, bool G, bool NT, int TH>
__device__ __forceinline__ void cols_body(const float *in, float *out, int w, const size_t *__restrict__ offs,
                                          const float *__restrict__ tab, unsigned xcds)
{
    constexpr int L = col_unit_cols<H, TH>();
    __shared__ float smem[col_unit_lds<H, TH>()];
    const int strips = w / L;
    // Each XCD takes one contiguous range of the strips (measured: 1024^2 transforms 5%
    // faster than with consecutive strips on alternating XCDs).
    const unsigned bid = xcd_order(blockIdx.x, gridDim.x, xcds);
    const size_t a = bid / strips;
    const int c0 = (int)(bid - a * strips) * L;
    col_unit<H, MODE, TH, NT>(in, out, w, G ? offs[a] : a * (size_t)H * w, c0, L, smem, tab, false);
}

template <int H, int MODE, bool G, bool NT>
__global__ __launch_bounds__((col_threads<H, MODE>())) void k_cols(const float *in, float *out, int w,
                                                                  const size_t *__restrict__ offs,
                                                                  const float *__restrict__ tab, unsigned xcds)
{
    cols_body<H, MODE, G, NT, col_threads<H, MODE>()>(in, out, w, offs, tab, xcds);
}

// The same column pass with narrower strips (TH = col_threads / 2 or / 4: 32 or 16
// columns, several team members per wave), for grids that would otherwise leave CUs idle:
// e.g. 64 arrays of 512 x 128 are 128 workgroups of 64 columns.  Team members and
// arithmetic are those of k_cols; only the strip width differs, so the results are equal
// bitwise (the LDS copy of the team twiddles holds the same floats as the global table).
template <int H, int MODE, bool G, int TH>
__global__ __launch_bounds__(TH) void k_cols_narrow(const float *in, float *out, int w,
                                                     const size_t *__restrict__ offs,
                                                     const float *__restrict__ tab, unsigned xcds)
{
    cols_body<H, MODE, G, false, TH>(in, out, w, offs, tab, xcds);
}

// Work-items per grid dimension of a launch.  The transform splits calls with more than
// 512 MiB of intermediate into chunks of 256 MiB, far below this; single-pass calls
// (bdx_transform_pass) are checked against it.
constexpr size_t kMaxGrid = 0xffffffffu;

// Workgroups below which a column pass takes narrower strips: two per CU.
constexpr size_t kColsMinGrid = 512;

// The band and strip sizes the host side assumes (BDX_TWO_PASS_OK) are these kernels'.
template <int N> constexpr bool pass_sizes_match()
{
    return BDX_PASS_ROWS(N) == row_unit_rows<N, kThreads>() &&
           BDX_PASS_COLS(N) == col_unit_cols<N, col_threads<N, 0>()>() &&
           BDX_PASS_COLS(N) == col_unit_cols<N, col_threads<N, 1>()>() &&
           BDX_PASS_COLS(N) == col_unit_cols<N, col_threads<N, 2>()>();
}
static_assert(pass_sizes_match<8>() && pass_sizes_match<16>() && pass_sizes_match<32>() && pass_sizes_match<64>() &&
                  pass_sizes_match<128>() && pass_sizes_match<256>() && pass_sizes_match<512>() &&
                  pass_sizes_match<1024>(),
              "band and strip sizes of BDX_TWO_PASS_OK");


template <int W, bool G, bool NT>
int rows_wg(const float *in, float *out, size_t nrows, int op, const float *tab, hipStream_t s, const size_t *offs,
            int h)
{
    constexpr int R = row_unit_rows<W, kThreads>();
    if (nrows % R || (G && (h <= 0 || h % R)) || nrows / R > kMaxGrid / kThreads)
        return -1;
    const dim3 grid((unsigned)(nrows / R)), block(kThreads);
    switch (op) {
    case FA: hipLaunchKernelGGL((k_rows<W, FA, G, NT>), grid, block, 0, s, in, out, offs, h, tab, xcd_count()); break;
    case FA2: hipLaunchKernelGGL((k_rows<W, FA2, G, NT>), grid, block, 0, s, in, out, offs, h, tab, xcd_count()); break;
    case IA: hipLaunchKernelGGL((k_rows<W, IA, G, NT>), grid, block, 0, s, in, out, offs, h, tab, xcd_count()); break;
    case I0: hipLaunchKernelGGL((k_rows<W, I0, G, NT>), grid, block, 0, s, in, out, offs, h, tab, xcd_count()); break;
    default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int W>
int rows_w(const float *in, float *out, size_t nrows, int op, const float *tab, hipStream_t s, const size_t *offs,
           int h, bool nt)
{
    if (offs)
        return nt ? rows_wg<W, true, true>(in, out, nrows, op, tab, s, offs, h)
                  : rows_wg<W, true, false>(in, out, nrows, op, tab, s, offs, h);
    return nt ? rows_wg<W, false, true>(in, out, nrows, op, tab, s, nullptr, 0)
              : rows_wg<W, false, false>(in, out, nrows, op, tab, s, nullptr, 0);
}

template <int H, bool G, int TH>
int cols_narrow(int w, const float *in, float *out, size_t batch, int mode, const float *tab, hipStream_t s,
                const size_t *offs)
{
    constexpr int L = col_unit_cols<H, TH>();
    if (batch * (size_t)(w / L) > kMaxGrid / TH)
        return -1;
    const dim3 grid((unsigned)(batch * (size_t)(w / L))), block(TH);
    switch (mode) {
    case 0: hipLaunchKernelGGL((k_cols_narrow<H, 0, G, TH>), grid, block, 0, s, in, out, w, offs, tab, xcd_count()); break;
    case 1: hipLaunchKernelGGL((k_cols_narrow<H, 1, G, TH>), grid, block, 0, s, in, out, w, offs, tab, xcd_count()); break;
    case 2: hipLaunchKernelGGL((k_cols_narrow<H, 2, G, TH>), grid, block, 0, s, in, out, w, offs, tab, xcd_count()); break;
    default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int H, int MODE, bool G, bool NT>
int cols_mode(int w, const float *in, float *out, size_t batch, const float *tab, hipStream_t s, const size_t *offs)
{
    constexpr int TH = col_threads<H, MODE>(), L = col_unit_cols<H, TH>();
    if (w % L)
        return -1;
    // Small grids (never streaming ones: those exceed the 256 MB cache) take strips of half
    // or a quarter of the width, down to 16 columns, until there are two workgroups per CU.
    if constexpr (!NT && H >= 128) {
        constexpr int L2 = col_unit_cols<H, TH / 2>(), L4 = col_unit_cols<H, TH / 4>();
        static_assert(L2 == L / 2 && L4 == L / 4, "narrow strips split the team's lines");
        if (batch * (size_t)(w / L) < kColsMinGrid) {
            if constexpr (L4 >= 16) {
                if (batch * (size_t)(w / L2) >= kColsMinGrid)
                    return cols_narrow<H, G, TH / 2>(w, in, out, batch, MODE, tab, s, offs);
                return cols_narrow<H, G, TH / 4>(w, in, out, batch, MODE, tab, s, offs);
            } else if constexpr (L2 >= 16) {
                return cols_narrow<H, G, TH / 2>(w, in, out, batch, MODE, tab, s, offs);
            }
        }
    }
    if (batch * (size_t)(w / L) > kMaxGrid / TH)
        return -1;
    const dim3 grid((unsigned)(batch * (size_t)(w / L))), block(TH);
    hipLaunchKernelGGL((k_cols<H, MODE, G, NT>), grid, block, 0, s, in, out, w, offs, tab, xcd_count());
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int H, bool G, bool NT>
int cols_hg(int w, const float *in, float *out, size_t batch, int mode, const float *tab, hipStream_t s,
            const size_t *offs)
{
    switch (mode) {
    case 0: return cols_mode<H, 0, G, NT>(w, in, out, batch, tab, s, offs);
    case 1: return cols_mode<H, 1, G, NT>(w, in, out, batch, tab, s, offs);
    case 2: return cols_mode<H, 2, G, NT>(w, in, out, batch, tab, s, offs);
    default: return -1;
    }
}

template <int H>
int cols_h(int w, const float *in, float *out, size_t batch, int mode, const float *tab, hipStream_t s,
           const size_t *offs, bool nt)
{
    if (offs)
        return nt ? cols_hg<H, true, true>(w, in, out, batch, mode, tab, s, offs)
                  : cols_hg<H, true, false>(w, in, out, batch, mode, tab, s, offs);
    return nt ? cols_hg<H, false, true>(w, in, out, batch, mode, tab, s, nullptr)
              : cols_hg<H, false, false>(w, in, out, batch, mode, tab, s, nullptr);
}

}  // namespace
}  // namespace bdx
