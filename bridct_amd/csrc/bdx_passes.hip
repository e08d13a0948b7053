// Row-pass and column-pass kernels for arrays larger than one workgroup's LDS
// (reference equivalent: the banded row pass and column pass of src/rect.c:5-23).
//
// k_rows<W, OP>: a block transforms rows_per_block(W) full rows (16 KiB floats of LDS);
//   rows of all arrays of a batch form one sequence.
// k_cols<H, MODE>: a block transforms a strip of cols_per_block(H) full columns of one
//   array, in place allowed (the intermediate of the two-pass route lives in the output
//   buffer, so a forward or inverse reads the input once and writes the output once plus
//   one intermediate round trip).
#include "bdx_kernels.h"
#include "bdx_lines.h"

namespace bdx {
namespace {

template <int W, int OP>
__global__ __launch_bounds__(kThreads) void k_rows(const float *in, float *out, size_t nrows,
                                                    const float *__restrict__ tab)
{
    constexpr int R = rows_per_block(W), P = W + 1;
    __shared__ float tile[R * P];
    const size_t r0 = (size_t)blockIdx.x * R;
    const int nr = (int)min((size_t)R, nrows - r0);
    tile_load<W, P, kThreads>(tile, in + r0 * W, W, nr);
    __syncthreads();
    run_lines<W, 1, OP, kThreads>(tile, nr, [](int L) { return L * P; }, tab);
    __syncthreads();
    tile_store<W, P, kThreads>(tile, out + r0 * W, W, nr);
}

template <int H, int MODE>
__global__ __launch_bounds__(kThreads) void k_cols(const float *in, float *out, int w,
                                                    const float *__restrict__ tab)
{
    constexpr int C = cols_per_block(H), P = C + 1;
    __shared__ float tile[H * P];
    const int strips = w / C;
    const size_t a = blockIdx.x / strips;
    const int c0 = (int)(blockIdx.x - a * strips) * C;
    const size_t off = a * (size_t)H * w + c0;
    tile_load<C, P, kThreads>(tile, in + off, w, H);
    __syncthreads();
    const auto base = [](int L) { return L; };
    if constexpr (MODE == 0) {
        run_lines<H, P, FA, kThreads>(tile, C, base, tab);
    } else if constexpr (MODE == 1) {
        run_lines<H, P, IA, kThreads>(tile, C, base, tab);
    } else {
        run_lines<H, P, FA2, kThreads>(tile, C, base, tab);
        __syncthreads();
        run_lines<H, P, I0, kThreads>(tile, C, base, tab);
    }
    __syncthreads();
    tile_store<C, P, kThreads>(tile, out + off, w, H);
}

template <int W>
int rows_w(const float *in, float *out, size_t nrows, int op, const float *tab, hipStream_t s)
{
    constexpr int R = rows_per_block(W);
    const dim3 grid((unsigned)((nrows + R - 1) / R)), block(kThreads);
    switch (op) {
    case FA: hipLaunchKernelGGL((k_rows<W, FA>), grid, block, 0, s, in, out, nrows, tab); break;
    case FA2: hipLaunchKernelGGL((k_rows<W, FA2>), grid, block, 0, s, in, out, nrows, tab); break;
    case IA: hipLaunchKernelGGL((k_rows<W, IA>), grid, block, 0, s, in, out, nrows, tab); break;
    case I0: hipLaunchKernelGGL((k_rows<W, I0>), grid, block, 0, s, in, out, nrows, tab); break;
    default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int H>
int cols_h(int w, const float *in, float *out, size_t batch, int mode, const float *tab, hipStream_t s)
{
    constexpr int C = cols_per_block(H);
    if (w % C)
        return -1;
    const dim3 grid((unsigned)(batch * (size_t)(w / C))), block(kThreads);
    switch (mode) {
    case 0: hipLaunchKernelGGL((k_cols<H, 0>), grid, block, 0, s, in, out, w, tab); break;
    case 1: hipLaunchKernelGGL((k_cols<H, 1>), grid, block, 0, s, in, out, w, tab); break;
    case 2: hipLaunchKernelGGL((k_cols<H, 2>), grid, block, 0, s, in, out, w, tab); break;
    default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace

int launch_rows(int w, const float *in, float *out, size_t nrows, int op, const float *tab, hipStream_t s)
{
    switch (w) {
    case 64: return rows_w<64>(in, out, nrows, op, tab, s);
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
    case 64: return cols_h<64>(w, in, out, batch, mode, tab, s);
    case 128: return cols_h<128>(w, in, out, batch, mode, tab, s);
    case 256: return cols_h<256>(w, in, out, batch, mode, tab, s);
    case 512: return cols_h<512>(w, in, out, batch, mode, tab, s);
    case 1024: return cols_h<1024>(w, in, out, batch, mode, tab, s);
    default: return -1;
    }
}

}  // namespace bdx
