// 8x8 block codec step on gfx950: the work a block codec does right after (encode) and
// right before (decode) the transform, fused with it so that the image crosses HBM once
// as bytes and the coefficients once as int16.
//
//   encode: u8 raster -> x - 128 -> orthonormal 2D DCT-II of every 8x8 block ->
//           round-half-even(Y / Q) saturated to int16, block-contiguous (64 per block,
//           row-major u * 8 + v, blocks in raster order of the block grid).
//   decode: int16 coefficients -> q * Q -> orthonormal 2D DCT-III -> round-half-even(x + 128)
//           saturated to 0..255, u8 raster.
//
// One lane owns one 8x8 block in registers: the eight row lines and the eight column
// lines run the generated length-8 Shao-Johnson cores (the reference's 8-point cores,
// src/generated/sj_body.inc one_f8 / one_t8), rows first.  The cores' per-axis
// normalization alpha is constant along each line, so it commutes with the other axis'
// core and is folded with the quantizer step into one factor per coefficient
// (alpha_u alpha_v / Q_uv after the passes, Q_uv alpha_u alpha_v before them): the kernels
// are bound by HBM, not by a correctly rounded division per coefficient.  Consecutive
// lanes own horizontally adjacent blocks, so each row of a wave's loads (encode) or stores
// (decode) is 512 contiguous bytes of the raster.  The reference has no codec step; it is
// the caller-side stage of a block codec (quantization as in baseline JPEG, ITU-T T.81
// A.3.4, without the entropy coder).
#include "bdx_lines.h"

namespace bdx {
namespace {

constexpr int kCodecThreads = 256;

__device__ __forceinline__ void unpack8(uint2 v, float *x)
{
#pragma unroll
    for (int k = 0; k < 4; k++) {
        x[k] = (float)((v.x >> (8 * k)) & 0xFFu) - 128.f;
        x[4 + k] = (float)((v.y >> (8 * k)) & 0xFFu) - 128.f;
    }
}

__device__ __forceinline__ unsigned pixel(float v)
{
    return (unsigned)fminf(fmaxf(__builtin_rintf(v + 128.f), 0.f), 255.f);
}

__global__ __launch_bounds__(kCodecThreads) void k_encode8(const unsigned char *__restrict__ img, size_t stride,
                                                           unsigned bw, size_t nblocks,
                                                           const float *__restrict__ factor,
                                                           short *__restrict__ coeffs)
{
    const size_t b = (size_t)blockIdx.x * kCodecThreads + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const size_t b0 = b - lane;   // first block of the wave
    if (b0 >= nblocks)            // whole wave past the end
        return;
    const bool live = b < nblocks;
    const size_t by = live ? b / bw : 0, bx = live ? b - by * bw : 0;
    const unsigned char *src = img + by * 8 * stride + bx * 8;
    float x[64];
#pragma unroll
    for (int r = 0; r < 8; r++)
        unpack8(*reinterpret_cast<const uint2 *>(src + r * stride), x + 8 * r);
#pragma unroll
    for (int r = 0; r < 8; r++) {
        float l[8];
#pragma unroll
        for (int c = 0; c < 8; c++)
            l[c] = x[8 * r + c];
        Sj<8>::fwd(l);
#pragma unroll
        for (int c = 0; c < 8; c++)
            x[8 * r + c] = l[c];
    }
#pragma unroll
    for (int c = 0; c < 8; c++) {
        float l[8];
#pragma unroll
        for (int u = 0; u < 8; u++)
            l[u] = x[8 * u + c];
        Sj<8>::fwd(l);
#pragma unroll
        for (int u = 0; u < 8; u++)
            x[8 * u + c] = l[u];
    }
    // Each lane's 128 coefficient bytes are contiguous in the output, so direct 16-byte
    // stores of a wave would touch 64 lines per instruction; the wave's 8 KiB are staged in
    // LDS instead (lane pitch 9 int4 against bank conflicts) and stored as whole lines.
    __shared__ int4 stage[kCodecThreads / 64][64 * 9];
    int4 *ws = stage[threadIdx.x >> 6];
#pragma unroll
    for (int r = 0; r < 8; r++) {
        short q[8];
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const float v = __builtin_rintf(__fmul_rn(x[8 * r + c], factor[8 * r + c]));
            q[c] = (short)(int)fminf(fmaxf(v, -32768.f), 32767.f);
        }
        int4 o;
        o.x = (int)((unsigned short)q[0] | (unsigned)(unsigned short)q[1] << 16);
        o.y = (int)((unsigned short)q[2] | (unsigned)(unsigned short)q[3] << 16);
        o.z = (int)((unsigned short)q[4] | (unsigned)(unsigned short)q[5] << 16);
        o.w = (int)((unsigned short)q[6] | (unsigned)(unsigned short)q[7] << 16);
        ws[lane * 9 + r] = o;
    }
    team_sync();
    // the wave's live blocks b0 .. b0 + nw - 1 are consecutive; chunk k (16 bytes) of the
    // wave's output belongs to lane k / 8, row k % 8
    const int nw = (int)min((size_t)64, nblocks - b0);
    int4 *dst = reinterpret_cast<int4 *>(coeffs + b0 * 64);
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int k = lane + 64 * i;
        if (k < 8 * nw)
            dst[k] = ws[(k >> 3) * 9 + (k & 7)];
    }
}

__global__ __launch_bounds__(kCodecThreads) void k_decode8(const short *__restrict__ coeffs, size_t stride,
                                                           unsigned bw, size_t nblocks,
                                                           const float *__restrict__ factor,
                                                           unsigned char *__restrict__ img)
{
    const size_t b = (size_t)blockIdx.x * kCodecThreads + threadIdx.x;
    if (b >= nblocks)
        return;
    const size_t by = b / bw, bx = b - by * bw;
    const float *g = factor + 64;
    const short *src = coeffs + b * 64;
    float x[64];
#pragma unroll
    for (int r = 0; r < 8; r++) {
        const int4 v = *reinterpret_cast<const int4 *>(src + 8 * r);
        const int w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            x[8 * r + 2 * k] = __fmul_rn((float)(short)(w[k] & 0xFFFF), g[8 * r + 2 * k]);
            x[8 * r + 2 * k + 1] = __fmul_rn((float)(short)((unsigned)w[k] >> 16), g[8 * r + 2 * k + 1]);
        }
    }
#pragma unroll
    for (int r = 0; r < 8; r++) {
        float l[8];
#pragma unroll
        for (int c = 0; c < 8; c++)
            l[c] = x[8 * r + c];
        Sj<8>::inv(l);
#pragma unroll
        for (int c = 0; c < 8; c++)
            x[8 * r + c] = l[c];
    }
#pragma unroll
    for (int c = 0; c < 8; c++) {
        float l[8];
#pragma unroll
        for (int u = 0; u < 8; u++)
            l[u] = x[8 * u + c];
        Sj<8>::inv(l);
#pragma unroll
        for (int u = 0; u < 8; u++)
            x[8 * u + c] = l[u];
    }
    unsigned char *dst = img + by * 8 * stride + bx * 8;
#pragma unroll
    for (int r = 0; r < 8; r++) {
        uint2 o;
        o.x = pixel(x[8 * r]) | pixel(x[8 * r + 1]) << 8 | pixel(x[8 * r + 2]) << 16 | pixel(x[8 * r + 3]) << 24;
        o.y = pixel(x[8 * r + 4]) | pixel(x[8 * r + 5]) << 8 | pixel(x[8 * r + 6]) << 16 |
              pixel(x[8 * r + 7]) << 24;
        *reinterpret_cast<uint2 *>(dst + r * stride) = o;
    }
}

}  // namespace
}  // namespace bdx

extern "C" int bdx_codec8(int decode, const void *d_src, void *d_dst, size_t height, size_t width,
                          size_t stride, const float *d_factors, void *stream)
{
    const size_t bw = width / 8, nblocks = (height / 8) * bw;
    if (nblocks == 0)
        return 0;
    const size_t grid = (nblocks + bdx::kCodecThreads - 1) / bdx::kCodecThreads;
    if (grid > 0x7FFFFFFFu || bw > 0xFFFFFFFFu)
        return -1;
    hipStream_t s = (hipStream_t)stream;
    if (decode)
        hipLaunchKernelGGL(bdx::k_decode8, dim3((unsigned)grid), dim3(bdx::kCodecThreads), 0, s,
                           (const short *)d_src, stride, (unsigned)bw, nblocks, d_factors,
                           (unsigned char *)d_dst);
    else
        hipLaunchKernelGGL(bdx::k_encode8, dim3((unsigned)grid), dim3(bdx::kCodecThreads), 0, s,
                           (const unsigned char *)d_src, stride, (unsigned)bw, nblocks, d_factors,
                           (short *)d_dst);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
