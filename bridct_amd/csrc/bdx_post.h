// Real-to-DCT post-process of the packed half-length FFT and its transpose.
//
// With Z = DFT_N(z), z[m] = v[2m] + i v[2m+1] (v the Makhoul reordering of the line,
// N = n/2), the pair (k, N - k), 0 < k < N/2, gives four DCT-II outputs, each divided by
// cos(theta) of its own angle (tangent form; theta_k = pi k / (2n)) and by 2:
//   E = Z_k + conj Z_{N-k},  O = (Z_k - conj Z_{N-k}) / i,  V = E + w O,  U = E - w O,
//   out[k] = Re V + t1 Im V,  out[n-k] = t1 Re V - Im V,
//   out[N-k] = Re U + t2 Im U, out[N+k] = t2 Re U - Im U,
// with w = exp(-2 pi i k / n), t1 = tan theta_k, t2 = tan theta_{N-k}.  The self-paired
// points are Z_0 (outputs 0 and N = n/2) and Z_{N/2} (outputs n/4 and 3n/4).
// The inverse direction applies the exact transposes of these maps.
#pragma once
#include <hip/hip_runtime.h>

namespace bdx {

__device__ __forceinline__ void post_fwd(float ar, float ai, float br, float bi, float wr, float wi,
                                         float t1, float t2, float (&o)[4])
{
    const float er = ar + br, ei = ai - bi, orr = ai + bi, oi = br - ar;
    const float P = __builtin_fmaf(wr, orr, -wi * oi);
    const float Q = __builtin_fmaf(wr, oi, wi * orr);
    const float vr = er + P, vi = ei + Q, ur = er - P, ui = Q - ei;
    o[0] = __builtin_fmaf(t1, vi, vr);
    o[1] = __builtin_fmaf(t1, vr, -vi);
    o[2] = __builtin_fmaf(t2, ui, ur);
    o[3] = __builtin_fmaf(t2, ur, -ui);
}

// Z_0 -> (out[0], out[N]); Z_{N/2} -> (out[n/4], out[3n/4]) with th = tan(pi / 8).
__device__ __forceinline__ void post_fwd_special(float z0r, float z0i, float zhr, float zhi, float th,
                                                 float (&o)[4])
{
    o[0] = z0r + z0i;
    o[1] = z0r - z0i;
    o[2] = __builtin_fmaf(-th, zhi, zhr);
    o[3] = __builtin_fmaf(th, zhr, zhi);
}

// Transpose of post_fwd: from (out[k], out[n-k], out[N-k], out[N+k]) to Z_k (a), Z_{N-k} (b).
__device__ __forceinline__ void post_inv(const float (&o)[4], float wr, float wi, float t1, float t2,
                                         float &ar, float &ai, float &br, float &bi)
{
    const float vr = __builtin_fmaf(t1, o[1], o[0]), vi = __builtin_fmaf(t1, o[0], -o[1]);
    const float ur = __builtin_fmaf(t2, o[3], o[2]), ui = __builtin_fmaf(t2, o[2], -o[3]);
    const float er = vr + ur, P = vr - ur, ei = vi - ui, Q = vi + ui;
    const float orr = __builtin_fmaf(wr, P, wi * Q);
    const float oi = __builtin_fmaf(wr, Q, -wi * P);
    ar = er - oi;
    ai = ei + orr;
    br = er + oi;
    bi = orr - ei;
}

// The sums are never contracted with the products that scale o (inputs times normalization
// factors): whether the compiler fused them would depend on the block the inputs were read
// in, and kernels that read in[0], in[N] apart must give the same bits as those that do not.
__device__ __forceinline__ void post_inv_special(const float (&o)[4], float th, float &z0r, float &z0i,
                                                 float &zhr, float &zhi)
{
#pragma clang fp contract(off)
    z0r = o[0] + o[1];
    z0i = o[0] - o[1];
    zhr = __builtin_fmaf(th, o[3], o[2]);
    zhi = __builtin_fmaf(-th, o[2], o[3]);
}

}  // namespace bdx
