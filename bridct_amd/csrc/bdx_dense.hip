// Dense (split-MFMA) route of the 2D transform: Y = C_h X C_w^T as two batched GEMMs on
// the matrix cores, each operand split into a hi and a lo half-precision part
// (FP16 or BF16), three products per multiply (A_lo B_hi + A_hi B_lo + A_hi B_hi) with FP32
// accumulation; the two small cross products accumulate apart from the hi x hi products
// and are added once at the end, so that their contributions are not rounded against the
// large running sum at every step.  The reference has no dense route (it only has the butterflies of
// src/generated/*.c); the matrices are the orthonormal DCT-II matrices of docs/API.md:7-13.
//
//   forward:  T = X Q_w'^T          (rows, M = batch*h, N = K = w)
//             Y = Q_h' T / (s_h s_w) (columns, per array, M = K = h, N = w)
//   inverse:  T = Y Q_w'            X = Q_h'^T T / (s_h s_w)
// with Q_n' = s_n Q_n, s_n a power of two near sqrt(n) (so that the lo parts of the
// matrix entries stay normal half-precision numbers); the final scale is exact.
//
// Kernel: 256 lanes, a 64 x 64 output tile per workgroup, four waves of 32 x 32, k steps
// of 32 staged in LDS as hi/lo halves ([row][k] for A, [column][k] for B, so that every
// lane reads its eight k values of a v_mfma_f32_32x32x16 operand with one 16-byte load).
// MFMA lane layout (checked on gfx950): lane l holds A[l % 32][8 (l / 32) + i],
// B[8 (l / 32) + i][l % 32], and result element a at row 8 (a / 4) + 4 (l / 32) + a % 4,
// column l % 32.
#include <hip/hip_runtime.h>

#include "bridct_device.h"

namespace bdx {
namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

constexpr int BM = 64, BN = 64, BK = 32, PITCH = BK + 8;   // pitch in half-precision elements

template <int PREC> struct Half;
template <> struct Half<BDX_DENSE_F16> {
    typedef _Float16 T;
    typedef h8 V;
    static __device__ __forceinline__ f16v mfma(V a, V b, f16v c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
};
template <> struct Half<BDX_DENSE_BF16> {
    typedef __bf16 T;
    typedef b8 V;
    static __device__ __forceinline__ f16v mfma(V a, V b, f16v c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
};

// x = hi + lo (+ a remainder below the lo part's precision)
template <typename T> __device__ __forceinline__ void split(float x, T &hi, T &lo)
{
    hi = (T)x;
    lo = (T)(x - (float)hi);
}

// C[b] = scale * A[b] B[b]; A is M x K row-major; B is K x N row-major, or N x K row-major
// when BT (i.e. B^T given).  M, N multiples of 64, K a multiple of 32; batch = gridDim.z.
template <int PREC, bool BT>
__global__ __launch_bounds__(256) void k_gemm(const float *__restrict__ A, const float *__restrict__ B,
                                              float *__restrict__ C, int N, int K, long long sa, long long sb,
                                              long long sc, float scale)
{
    using H = Half<PREC>;
    using T = typename H::T;
    __shared__ __attribute__((aligned(16))) T a_hi[BM * PITCH], a_lo[BM * PITCH];
    __shared__ __attribute__((aligned(16))) T b_hi[BN * PITCH], b_lo[BN * PITCH];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const long long m0 = (long long)blockIdx.y * BM, n0 = (long long)blockIdx.x * BN;
    A += blockIdx.z * sa + m0 * K;
    B += blockIdx.z * sb;
    C += blockIdx.z * sc;
    f16v acc = {}, acc_lo = {};   // hi x hi products, and the two small cross products
    for (int k0 = 0; k0 < K; k0 += BK) {
#pragma unroll
        for (int j = 0; j < 2; j++) {
            // A tile: 64 rows x 32 floats, float4 per lane
            const int idx = tid + 256 * j, r = idx >> 3, c = (idx & 7) * 4;
            const float4 v = *reinterpret_cast<const float4 *>(A + (long long)r * K + k0 + c);
            T *ph = a_hi + r * PITCH + c, *pl = a_lo + r * PITCH + c;
            split(v.x, ph[0], pl[0]);
            split(v.y, ph[1], pl[1]);
            split(v.z, ph[2], pl[2]);
            split(v.w, ph[3], pl[3]);
        }
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int idx = tid + 256 * j;
            if constexpr (BT) {
                // B^T tile: 64 rows (n) x 32 floats (k)
                const int r = idx >> 3, c = (idx & 7) * 4;
                const float4 v = *reinterpret_cast<const float4 *>(B + (n0 + r) * K + k0 + c);
                T *ph = b_hi + r * PITCH + c, *pl = b_lo + r * PITCH + c;
                split(v.x, ph[0], pl[0]);
                split(v.y, ph[1], pl[1]);
                split(v.z, ph[2], pl[2]);
                split(v.w, ph[3], pl[3]);
            } else {
                // B tile: 32 rows (k) x 64 floats (n), stored transposed
                const int r = idx >> 4, c = (idx & 15) * 4;
                const float4 v = *reinterpret_cast<const float4 *>(B + (long long)(k0 + r) * N + n0 + c);
                split(v.x, b_hi[(c + 0) * PITCH + r], b_lo[(c + 0) * PITCH + r]);
                split(v.y, b_hi[(c + 1) * PITCH + r], b_lo[(c + 1) * PITCH + r]);
                split(v.z, b_hi[(c + 2) * PITCH + r], b_lo[(c + 2) * PITCH + r]);
                split(v.w, b_hi[(c + 3) * PITCH + r], b_lo[(c + 3) * PITCH + r]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            const int ao = (wm * 32 + (lane & 31)) * PITCH + kk + 8 * (lane >> 5);
            const int bo = (wn * 32 + (lane & 31)) * PITCH + kk + 8 * (lane >> 5);
            const auto ah = *reinterpret_cast<const typename H::V *>(a_hi + ao);
            const auto al = *reinterpret_cast<const typename H::V *>(a_lo + ao);
            const auto bh = *reinterpret_cast<const typename H::V *>(b_hi + bo);
            const auto bl = *reinterpret_cast<const typename H::V *>(b_lo + bo);
            acc_lo = H::mfma(al, bh, acc_lo);
            acc_lo = H::mfma(ah, bl, acc_lo);
            acc = H::mfma(ah, bh, acc);
        }
        __syncthreads();
    }
    const long long col = n0 + wn * 32 + (lane & 31);
#pragma unroll
    for (int a = 0; a < 16; a++) {
        const long long row = m0 + wm * 32 + 8 * (a >> 2) + 4 * (lane >> 5) + (a & 3);
        C[row * N + col] = (acc[a] + acc_lo[a]) * scale;
    }
}

template <int PREC>
int gemm(bool bt, const float *A, const float *B, float *C, int M, int N, int K, int batch, long long sa,
         long long sb, long long sc, float scale, hipStream_t s)
{
    if (M % BM || N % BN || K % BK || batch < 1 || batch > 65535 || M / BM > 65535 || N / BN > 65535)
        return -1;
    const dim3 grid((unsigned)(N / BN), (unsigned)(M / BM), (unsigned)batch), block(256);
    if (bt)
        hipLaunchKernelGGL((k_gemm<PREC, true>), grid, block, 0, s, A, B, C, N, K, sa, sb, sc, scale);
    else
        hipLaunchKernelGGL((k_gemm<PREC, false>), grid, block, 0, s, A, B, C, N, K, sa, sb, sc, scale);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int gemm_any(int prec, bool bt, const float *A, const float *B, float *C, int M, int N, int K, int batch,
             long long sa, long long sb, long long sc, float scale, hipStream_t s)
{
    if (prec == BDX_DENSE_F16)
        return gemm<BDX_DENSE_F16>(bt, A, B, C, M, N, K, batch, sa, sb, sc, scale, s);
    if (prec == BDX_DENSE_BF16)
        return gemm<BDX_DENSE_BF16>(bt, A, B, C, M, N, K, batch, sa, sb, sc, scale, s);
    return -1;
}

}  // namespace
}  // namespace bdx

extern "C" int bdx_dense_transform(const float *d_in, float *d_out, float *d_work, int h, int w, size_t batch,
                                   int mode, int prec, const float *d_qh, const float *d_qht, const float *d_qw,
                                   float scale, void *stream)
{
    using namespace bdx;
    if (batch == 0)
        return 0;
    if (h < 64 || w < 64 || (mode != BDX_FORWARD && mode != BDX_INVERSE) || batch > 65535)
        return -1;
    hipStream_t s = (hipStream_t)stream;
    const long long hw = (long long)h * w;
    // rows: T = X Q_w'^T (forward; B^T = Q_w') or T = Y Q_w' (inverse; B = Q_w')
    if (gemm_any(prec, mode == BDX_FORWARD, d_in, d_qw, d_work, (int)(batch * h), w, w, 1, 0, 0, 0, 1.0f, s))
        return -1;
    // columns: Y = Q_h' T (forward) or X = Q_h'^T T (inverse), per array
    return gemm_any(prec, false, mode == BDX_FORWARD ? d_qh : d_qht, d_work, d_out, h, w, h, (int)batch, 0, hw,
                    hw, scale, s);
}
