// Row-pass and column-pass kernels for arrays larger than one workgroup's LDS
// (reference equivalent: the banded row pass and column pass of src/rect.c:5-23).
//
// k_rows_t<W, OP> (W >= 128): one team of team_size<W>() lanes per row, float4 loads
//   straight from global memory, DPP mirror for the Makhoul packing, outputs stored
//   directly (bdx_rteam.h).  Rows of all arrays of a batch are one sequence.
// k_rows64<OP> (W = 64): one lane per row over an LDS-staged band of 256 rows.
// k_cols_w<H, MODE> (H >= 128): 64 adjacent columns per block, one lane per column and
//   one wave per team member (bdx_wteam.h): every element access is a 256-byte row
//   segment and every constant is wave-uniform.  MODE 2 (the middle of the fused round
//   trip) keeps the forward results in registers and runs the inverse on them.
// k_cols64<MODE> (H = 64): one lane per column, 256 columns per block.
// The column kernels may run in place: every lane loads all its inputs before the first
// block barrier and stores only after the last one.
#include "bdx_kernels.h"
#include "bdx_lines.h"
#include "bdx_rteam.h"
#include "bdx_wteam.h"

namespace bdx {
namespace {

template <int W, int OP>
__global__ __launch_bounds__(kThreads) void k_rows_t(const float *in, float *out,
                                                      const float *__restrict__ tab)
{
    using R = RowTab<W>;
    constexpr int T = team_size<W>(), TEAMS = kThreads / T;
    __shared__ float lt[R::SIZE];
    __shared__ float xl[TEAMS * R::XCH];
    load_row_tables<W, OP, kThreads>(lt, tab);
    __syncthreads();
    const int team = threadIdx.x / T, t = threadIdx.x % T;
    const size_t row = (size_t)blockIdx.x * TEAMS + team;
    const float *src = in + row * W;
    float *dst = out + row * W;
    if constexpr (OP == FA || OP == FA2)
        rteam_fwd<W, OP>(src, dst, t, xl + team * R::XCH, lt);
    else
        rteam_inv<W, OP>(src, dst, t, xl + team * R::XCH, lt);
}

template <int OP>
__global__ __launch_bounds__(kThreads) void k_rows64(const float *in, float *out, const float *__restrict__ tab)
{
    constexpr int R = kThreads, P = 65;
    __shared__ float tile[R * P];
    const size_t r0 = (size_t)blockIdx.x * R;
    tile_load<64, P, kThreads>(tile, in + r0 * 64, 64, R);
    __syncthreads();
    run_lines<64, 1, OP, kThreads>(tile, R, [](int L) { return L * P; }, tab);
    __syncthreads();
    tile_store<64, P, kThreads>(tile, out + r0 * 64, 64, R);
}

template <int H, int MODE>
__global__ __launch_bounds__(64 * team_size<H>()) void k_cols_w(const float *in, float *out, int w,
                                                                 const float *__restrict__ tab)
{
    constexpr int T = team_size<H>(), Q = H / 2 / T / (2 * T);
    __shared__ float xl[wteam_lds<H>()];
    const int t = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int strips = w >> 6;
    const size_t a = blockIdx.x / strips;
    const int c0 = (int)(blockIdx.x - a * strips) * 64;
    // uniform row base (scalar registers) plus the lane's column offset
    const size_t off = a * (size_t)H * w + c0;
    const float *src = in + off;
    float *dst = out + off;
    const auto ld = [&](int i) { return (src + (size_t)i * w)[lane]; };
    const auto st = [&](int i, float v) { (dst + (size_t)i * w)[lane] = v; };
    float o[Q][T][4];
    if constexpr (MODE == 1) {
        const float *al = alpha_of<H>(tab, 1);
#pragma unroll
        for (int r = 0; r < Q; r++)
#pragma unroll
            for (int k2 = 0; k2 < T; k2++)
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    const int idx = wteam_index<H, T>(t + T * r, k2, c);
                    o[r][k2][c] = ld(idx) * al[idx];
                }
        wteam_inv<H>(t, o, st, xl, tab);
    } else {
        const float *al = alpha_of<H>(tab, MODE == 0 ? 1 : 2);
        wteam_fwd<H>(t, ld, xl, tab, [&](int r, int k2, int c, float v) {
            const int idx = wteam_index<H, T>(t + T * r, k2, c);
            if constexpr (MODE == 0)
                st(idx, v * al[idx]);
            else
                o[r][k2][c] = v * al[idx];
        });
        if constexpr (MODE == 2)
            wteam_inv<H>(t, o, st, xl, tab);
    }
}

template <int MODE>
__global__ __launch_bounds__(kThreads) void k_cols64(const float *in, float *out, int w,
                                                      const float *__restrict__ tab)
{
    const int strips = w / kThreads;
    const size_t a = blockIdx.x / strips;
    const int c0 = (int)(blockIdx.x - a * strips) * kThreads;
    const size_t off = a * (size_t)64 * w + c0 + threadIdx.x;
    const float *al = alpha_of<64>(tab, 1);
    const float *al2 = alpha_of<64>(tab, 2);
    float x[64];
#pragma unroll
    for (int i = 0; i < 64; i++)
        x[i] = in[off + (size_t)i * w] * (MODE == 1 ? al[i] : 1.0f);
    if constexpr (MODE == 1) {
        Sj<64>::inv(x);
    } else {
        Sj<64>::fwd(x);
#pragma unroll
        for (int i = 0; i < 64; i++)
            x[i] *= MODE == 0 ? al[i] : al2[i];
        if constexpr (MODE == 2)
            Sj<64>::inv(x);
    }
#pragma unroll
    for (int i = 0; i < 64; i++)
        out[off + (size_t)i * w] = x[i];
}

template <int W>
int rows_w(const float *in, float *out, size_t nrows, int op, const float *tab, hipStream_t s)
{
    constexpr int R = kThreads / team_size<W>();
    if (nrows % R)
        return -1;
    const dim3 grid((unsigned)(nrows / R)), block(kThreads);
    switch (op) {
    case FA: hipLaunchKernelGGL((k_rows_t<W, FA>), grid, block, 0, s, in, out, tab); break;
    case FA2: hipLaunchKernelGGL((k_rows_t<W, FA2>), grid, block, 0, s, in, out, tab); break;
    case IA: hipLaunchKernelGGL((k_rows_t<W, IA>), grid, block, 0, s, in, out, tab); break;
    case I0: hipLaunchKernelGGL((k_rows_t<W, I0>), grid, block, 0, s, in, out, tab); break;
    default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int rows_64(const float *in, float *out, size_t nrows, int op, const float *tab, hipStream_t s)
{
    if (nrows % kThreads)
        return -1;
    const dim3 grid((unsigned)(nrows / kThreads)), block(kThreads);
    switch (op) {
    case FA: hipLaunchKernelGGL((k_rows64<FA>), grid, block, 0, s, in, out, tab); break;
    case FA2: hipLaunchKernelGGL((k_rows64<FA2>), grid, block, 0, s, in, out, tab); break;
    case IA: hipLaunchKernelGGL((k_rows64<IA>), grid, block, 0, s, in, out, tab); break;
    case I0: hipLaunchKernelGGL((k_rows64<I0>), grid, block, 0, s, in, out, tab); break;
    default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int H>
int cols_h(int w, const float *in, float *out, size_t batch, int mode, const float *tab, hipStream_t s)
{
    if (w % 64)
        return -1;
    const dim3 grid((unsigned)(batch * (size_t)(w / 64))), block(64 * team_size<H>());
    switch (mode) {
    case 0: hipLaunchKernelGGL((k_cols_w<H, 0>), grid, block, 0, s, in, out, w, tab); break;
    case 1: hipLaunchKernelGGL((k_cols_w<H, 1>), grid, block, 0, s, in, out, w, tab); break;
    case 2: hipLaunchKernelGGL((k_cols_w<H, 2>), grid, block, 0, s, in, out, w, tab); break;
    default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int cols_64(int w, const float *in, float *out, size_t batch, int mode, const float *tab, hipStream_t s)
{
    if (w % kThreads)
        return -1;
    const dim3 grid((unsigned)(batch * (size_t)(w / kThreads))), block(kThreads);
    switch (mode) {
    case 0: hipLaunchKernelGGL((k_cols64<0>), grid, block, 0, s, in, out, w, tab); break;
    case 1: hipLaunchKernelGGL((k_cols64<1>), grid, block, 0, s, in, out, w, tab); break;
    case 2: hipLaunchKernelGGL((k_cols64<2>), grid, block, 0, s, in, out, w, tab); break;
    default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace

int launch_rows(int w, const float *in, float *out, size_t nrows, int op, const float *tab, hipStream_t s)
{
    switch (w) {
    case 64: return rows_64(in, out, nrows, op, tab, s);
    case 128: return rows_w<128>(in, out, nrows, op, tab, s);
    case 256: return rows_w<256>(in, out, nrows, op, tab, s);
    case 512: return rows_w<512>(in, out, nrows, op, tab, s);
    case 1024: return rows_w<1024>(in, out, nrows, op, tab, s);
    default: return -1;
    }
}

int launch_cols(int h, int w, const float *in, float *out, size_t batch, int mode, const float *tab,
                hipStream_t s)
{
    switch (h) {
    case 64: return cols_64(w, in, out, batch, mode, tab, s);
    case 128: return cols_h<128>(w, in, out, batch, mode, tab, s);
    case 256: return cols_h<256>(w, in, out, batch, mode, tab, s);
    case 512: return cols_h<512>(w, in, out, batch, mode, tab, s);
    case 1024: return cols_h<1024>(w, in, out, batch, mode, tab, s);
    default: return -1;
    }
}

}  // namespace bdx
