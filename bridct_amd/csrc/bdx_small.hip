// Whole-array kernels for arrays that fit one workgroup's LDS (h * (w + 1) floats at most
// 144 KiB): load, row pass, column pass (and the inverse halves for the fused round
// trip), store -- one HBM read and one HBM write per pixel.  Several small arrays share a
// workgroup.  This file is compiled once per array height BDX_H; each (BDX_H, w) pair is
// its own kernel, so strides and array geometry are compile-time constants.  The gather
// form (G) takes the arrays at the element offsets offs[0, batch) of one buffer instead of
// back to back (the arrays of one shape inside a ragged batch).
#include "bdx_units.h"

#ifndef BDX_H
#error "compile with -DBDX_H=<height>"
#endif

namespace bdx {
namespace {

template <int H, int W, int MODE, bool G>
__global__ __launch_bounds__(kThreads) void k_small(const float *in, float *out, int per, size_t batch,
                                                     const size_t *__restrict__ offs,
                                                     const float *__restrict__ tab)
{
    extern __shared__ float tile[];
    const size_t a0 = (size_t)blockIdx.x * per;
    const int na = (int)min((size_t)per, batch - a0);
    if constexpr (G) {
        const size_t *o = offs + a0;
        small_unit_rows<H, W, MODE, kThreads>([=](int r) { return in + o[r / H] + (size_t)(r % H) * W; },
                                              [=](int r) { return out + o[r / H] + (size_t)(r % H) * W; }, na,
                                              tile, tab);
    } else {
        small_unit<H, W, MODE, kThreads>(in, out, a0 * H * W, na, tile, tab);
    }
}

template <int H, int W, bool G>
int launch_hwg(const float *in, float *out, size_t batch, size_t per, const size_t *offs, int mode,
               const float *tab, hipStream_t s)
{
    const size_t lds = per * H * (W + 1) * sizeof(float);
    const void *k = mode == 0 ? (const void *)k_small<H, W, 0, G>
                  : mode == 1 ? (const void *)k_small<H, W, 1, G> : (const void *)k_small<H, W, 2, G>;
    if (lds > 65536 && hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
        return -1;
    // a batch whose grid would pass the launch limit runs in several launches
    const size_t step = max_grid_blocks(kThreads) * per;
    for (size_t a = 0; a < batch; a += step) {
        const size_t n = batch - a < step ? batch - a : step, blocks = (n + per - 1) / per;
        const float *src = G ? in : in + a * H * W;
        float *dst = G ? out : out + a * H * W;
        const size_t *og = G ? offs + a : nullptr;
        if (mode == 0)
            hipLaunchKernelGGL((k_small<H, W, 0, G>), dim3((unsigned)blocks), dim3(kThreads), lds, s, src, dst,
                               (int)per, n, og, tab);
        else if (mode == 1)
            hipLaunchKernelGGL((k_small<H, W, 1, G>), dim3((unsigned)blocks), dim3(kThreads), lds, s, src, dst,
                               (int)per, n, og, tab);
        else
            hipLaunchKernelGGL((k_small<H, W, 2, G>), dim3((unsigned)blocks), dim3(kThreads), lds, s, src, dst,
                               (int)per, n, og, tab);
        if (hipGetLastError() != hipSuccess)
            return -1;
    }
    return 0;
}

template <int H, int W>
int launch_hw(const float *in, float *out, size_t batch, int pack, int mode, const float *tab, hipStream_t s,
              const size_t *offs)
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
        return offs ? launch_hwg<H, W, true>(in, out, batch, per, offs, mode, tab, s)
                    : launch_hwg<H, W, false>(in, out, batch, per, nullptr, mode, tab, s);
    }
}

}  // namespace

#define BDX_CAT2(a, b) a##b
#define BDX_CAT(a, b) BDX_CAT2(a, b)

int BDX_CAT(launch_small_h, BDX_H)(int w, const float *in, float *out, size_t batch, int pack, int mode,
                                    const float *tab, hipStream_t s, const size_t *offs)
{
    switch (w) {
    case 8: return launch_hw<BDX_H, 8>(in, out, batch, pack, mode, tab, s, offs);
    case 16: return launch_hw<BDX_H, 16>(in, out, batch, pack, mode, tab, s, offs);
    case 32: return launch_hw<BDX_H, 32>(in, out, batch, pack, mode, tab, s, offs);
    case 64: return launch_hw<BDX_H, 64>(in, out, batch, pack, mode, tab, s, offs);
    case 128: return launch_hw<BDX_H, 128>(in, out, batch, pack, mode, tab, s, offs);
    case 256: return launch_hw<BDX_H, 256>(in, out, batch, pack, mode, tab, s, offs);
    case 512: return launch_hw<BDX_H, 512>(in, out, batch, pack, mode, tab, s, offs);
    case 1024: return launch_hw<BDX_H, 1024>(in, out, batch, pack, mode, tab, s, offs);
    default: return -1;
    }
}

}  // namespace bdx
