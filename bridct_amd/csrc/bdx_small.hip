// Whole-array kernels for arrays that fit one workgroup's LDS (h * (w + 1) floats at most
// 144 KiB): load, row pass, column pass (and the inverse halves for the fused round
// trip), store -- one HBM read and one HBM write per pixel.  Several small arrays share a
// workgroup.  This file is compiled once per array height BDX_H; each (BDX_H, w) pair is
// its own kernel, so strides and array geometry are compile-time constants.
#include "bdx_units.h"

#ifndef BDX_H
#error "compile with -DBDX_H=<height>"
#endif

namespace bdx {
namespace {

template <int H, int W, int MODE>
__global__ __launch_bounds__(kThreads) void k_small(const float *in, float *out, int per, size_t batch,
                                                     const float *__restrict__ tab)
{
    extern __shared__ float tile[];
    const size_t a0 = (size_t)blockIdx.x * per;
    const int na = (int)min((size_t)per, batch - a0);
    small_unit<H, W, MODE, kThreads>(in, out, a0 * H * W, na, tile, tab);
}

template <int H, int W>
int launch_hw(const float *in, float *out, size_t batch, int pack, int mode, const float *tab, hipStream_t s)
{
    if constexpr (!small_route(H, W)) {
        return -1;
    } else {
        constexpr int P = W + 1;
        constexpr int per_max = kSmallLds / (H * P);
        // automatic (pack 0): enough workgroups to cover the GPU several times before
        // packing arrays together; otherwise the requested arrays per workgroup (tuned on
        // large batches), reduced for small batches so that about 512 workgroups remain
        size_t per = (batch + 1023) / 1024;
        if (pack > 0) {
            per = (size_t)pack;
            const size_t keep = (batch + 511) / 512;
            if (per > keep)
                per = keep;
        }
        if (per < 1)
            per = 1;
        if (per > (size_t)per_max)
            per = per_max;
        const size_t blocks = (batch + per - 1) / per;
        const size_t lds = per * H * P * sizeof(float);
        const void *k = mode == 0 ? (const void *)k_small<H, W, 0>
                      : mode == 1 ? (const void *)k_small<H, W, 1> : (const void *)k_small<H, W, 2>;
        if (lds > 65536 &&
            hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
            return -1;
        if (mode == 0)
            hipLaunchKernelGGL((k_small<H, W, 0>), dim3((unsigned)blocks), dim3(kThreads), lds, s, in,
                               out, (int)per, batch, tab);
        else if (mode == 1)
            hipLaunchKernelGGL((k_small<H, W, 1>), dim3((unsigned)blocks), dim3(kThreads), lds, s, in,
                               out, (int)per, batch, tab);
        else
            hipLaunchKernelGGL((k_small<H, W, 2>), dim3((unsigned)blocks), dim3(kThreads), lds, s, in,
                               out, (int)per, batch, tab);
        return hipGetLastError() == hipSuccess ? 0 : -1;
    }
}

}  // namespace

#define BDX_CAT2(a, b) a##b
#define BDX_CAT(a, b) BDX_CAT2(a, b)

int BDX_CAT(launch_small_h, BDX_H)(int w, const float *in, float *out, size_t batch, int pack, int mode,
                                    const float *tab, hipStream_t s)
{
    switch (w) {
    case 8: return launch_hw<BDX_H, 8>(in, out, batch, pack, mode, tab, s);
    case 16: return launch_hw<BDX_H, 16>(in, out, batch, pack, mode, tab, s);
    case 32: return launch_hw<BDX_H, 32>(in, out, batch, pack, mode, tab, s);
    case 64: return launch_hw<BDX_H, 64>(in, out, batch, pack, mode, tab, s);
    case 128: return launch_hw<BDX_H, 128>(in, out, batch, pack, mode, tab, s);
    case 256: return launch_hw<BDX_H, 256>(in, out, batch, pack, mode, tab, s);
    case 512: return launch_hw<BDX_H, 512>(in, out, batch, pack, mode, tab, s);
    case 1024: return launch_hw<BDX_H, 1024>(in, out, batch, pack, mode, tab, s);
    default: return -1;
    }
}

}  // namespace bdx
