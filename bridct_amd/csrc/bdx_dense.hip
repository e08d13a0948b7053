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
// Each array is also prescaled by a power of two, 2^-e with e the exponent of its largest
// magnitude, before its values are split (and its result scaled back by 2^e at the end):
// the split then sees values in [-1, 1] whatever the array's scale, so the lo parts keep
// their precision for tiny arrays (no half-precision subnormals) and the hi parts cannot
// overflow for large ones.  Both scalings are exact, so the route is scale-equivariant:
// 2^k X gives exactly 2^k times the result of X.
//
// Kernel: 256 lanes, a 64 x 64 or 128 x 128 output tile per workgroup, four waves, k steps
// of 32 staged in LDS as hi/lo halves ([row][k] for A, [column][k] for B, so that every
// lane reads its eight k values of a v_mfma_f32_32x32x16 operand with one 16-byte load);
// the next step's FP32 operands are loaded into registers during the current products.
// MFMA lane layout (checked on gfx950): lane l holds A[l % 32][8 (l / 32) + i],
// B[8 (l / 32) + i][l % 32], and result element a at row 8 (a / 4) + 4 (l / 32) + a % 4,
// column l % 32.
#include <hip/hip_runtime.h>

#include <mutex>

#include <type_traits>

#include "bridct_device.h"

namespace bdx {
namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

constexpr int BK = 32, PITCH = BK + 8;   // k step; LDS row pitch in half-precision elements

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

// Prescale exponent of an array from the bit pattern of its largest magnitude: e with
// max|x| 2^-e in [0.5, 1) (frexp); 0 -- no prescale -- for an array of zeros or one holding
// a NaN or an infinity (pattern at or above that of infinity).
__device__ __forceinline__ int prescale_exponent(unsigned bits)
{
    if (bits == 0u || bits >= 0x7f800000u)
        return 0;
    int e;
    (void)__builtin_frexpf(__uint_as_float(bits), &e);
    return e;
}

// x = hi + lo (+ a remainder below the lo part's precision)
template <typename T> __device__ __forceinline__ void split(float x, T &hi, T &lo)
{
    hi = (T)x;
    lo = (T)(x - (float)hi);
}

// Eight consecutive k values of one row (A) or column (B) -> hi and lo halves, each
// written with one 16-byte LDS store (row pitch 80 B: sixteen lanes of consecutive rows
// hit sixteen different 4-bank groups).
template <typename V> __device__ __forceinline__ void stage8(const float (&x)[8], V *hi, V *lo)
{
    V h, l;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        typename std::remove_reference<decltype(h[0])>::type a, b;
        split(x[i], a, b);
        h[i] = a;
        l[i] = b;
    }
    *hi = h;
    *lo = l;
}

// C[b] = scale * A[b] B[b]; A is M x K row-major; B is K x N row-major, or N x K row-major
// when BT (i.e. B^T given).  M, N multiples of TILE (64 or 128), K a multiple of 32;
// batch = gridDim.z.  Four waves of (TILE/2) x (TILE/2), i.e. S x S MFMA tiles of 32 x 32
// each (S = TILE / 64), so that every operand fragment read from LDS feeds S tiles.
// Lane l stages k values 8 (l % 4) .. +7 of tile rows l / 4 + 64 j (A, B^T: two float4
// loads each) or of tile columns l % 64 + 64 j (B: eight loads, each wave reading 64
// consecutive floats of a row).
// GM: 0 contiguous operands; 1 (single GEMM) row m of A at A + offs[m / hrows] +
// (m % hrows) K, i.e. the rows of the arrays of one shape inside a ragged batch; 2
// (batched) C of batch entry z at C + offs[z].
// Prescale (expo: one word per array, the bit pattern of its largest magnitude; may be
// null): with ea the rows of A belong to array m / hrows and are multiplied by 2^-e of that
// array before the split; otherwise C of batch entry z is multiplied by 2^e of array z at
// the store (e = prescale_exponent(word)).
template <int PREC, bool BT, int TILE, int GM>
__global__ __launch_bounds__(256, TILE == 128 ? 2 : 1) void k_gemm(const float *__restrict__ A, const float *__restrict__ B,
                                              float *__restrict__ C, int N, int K, long long sa, long long sb,
                                              long long sc, float scale, const size_t *__restrict__ offs,
                                              int hrows, const unsigned *__restrict__ expo, bool ea)
{
    using H = Half<PREC>;
    using T = typename H::T;
    using V = typename H::V;
    constexpr int S = TILE / 64;
    __shared__ __attribute__((aligned(16))) T a_hi[TILE * PITCH], a_lo[TILE * PITCH];
    __shared__ __attribute__((aligned(16))) T b_hi[TILE * PITCH], b_lo[TILE * PITCH];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const long long m0 = (long long)blockIdx.y * TILE, n0 = (long long)blockIdx.x * TILE;
    const float *arow[S];   // A rows of this lane's stage slots
    int ashift[S];          // their prescale exponents (-e of their array, or 0)
#pragma unroll
    for (int j = 0; j < S; j++) {
        const long long m = m0 + (tid >> 2) + 64 * j;
        arow[j] = GM == 1 ? A + offs[m / hrows] + (m % hrows) * K : A + blockIdx.z * sa + m * K;
        ashift[j] = expo && ea ? -prescale_exponent(expo[m / hrows]) : 0;
    }
    const int cshift = expo && !ea ? prescale_exponent(expo[blockIdx.z]) : 0;
    B += blockIdx.z * sb;
    C += GM == 2 ? (long long)offs[blockIdx.z] : blockIdx.z * sc;
    f16v acc[S][S] = {}, acc_lo[S][S] = {};   // hi x hi products, and the two small cross products
    const int ar = tid >> 2, ak = (tid & 3) * 8;            // A / B^T: row, first k
    const int bn = tid & 63, bk = (tid >> 6) * 8;           // B: column, first k
    // The next k step's operands are loaded into registers before the current step's
    // products, so that their latency hides behind the MFMAs.
    float xa[S][8], xb[S][8];
    const auto load = [&](int k0) {
#pragma unroll
        for (int j = 0; j < S; j++) {
            const float *src = arow[j] + k0 + ak;
            const float4 u = *reinterpret_cast<const float4 *>(src);
            const float4 v = *reinterpret_cast<const float4 *>(src + 4);
            xa[j][0] = u.x, xa[j][1] = u.y, xa[j][2] = u.z, xa[j][3] = u.w;
            xa[j][4] = v.x, xa[j][5] = v.y, xa[j][6] = v.z, xa[j][7] = v.w;
            if (ashift[j])
#pragma unroll
                for (int i = 0; i < 8; i++)
                    xa[j][i] = __builtin_ldexpf(xa[j][i], ashift[j]);
        }
#pragma unroll
        for (int j = 0; j < S; j++) {
            if constexpr (BT) {
                const float *src = B + (n0 + ar + 64 * j) * K + k0 + ak;
                const float4 u = *reinterpret_cast<const float4 *>(src);
                const float4 v = *reinterpret_cast<const float4 *>(src + 4);
                xb[j][0] = u.x, xb[j][1] = u.y, xb[j][2] = u.z, xb[j][3] = u.w;
                xb[j][4] = v.x, xb[j][5] = v.y, xb[j][6] = v.z, xb[j][7] = v.w;
            } else {
                const float *src = B + (long long)(k0 + bk) * N + n0 + bn + 64 * j;
#pragma unroll
                for (int i = 0; i < 8; i++)
                    xb[j][i] = src[(long long)i * N];
            }
        }
    };
    load(0);
    for (int k0 = 0; k0 < K; k0 += BK) {
#pragma unroll
        for (int j = 0; j < S; j++) {
            const int o = (ar + 64 * j) * PITCH + ak;
            stage8(xa[j], reinterpret_cast<V *>(a_hi + o), reinterpret_cast<V *>(a_lo + o));
        }
#pragma unroll
        for (int j = 0; j < S; j++) {
            const int o = BT ? (ar + 64 * j) * PITCH + ak : (bn + 64 * j) * PITCH + bk;
            stage8(xb[j], reinterpret_cast<V *>(b_hi + o), reinterpret_cast<V *>(b_lo + o));
        }
        __syncthreads();
        if (k0 + BK < K)
            load(k0 + BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            V ah[S], al[S], bh[S], bl[S];
#pragma unroll
            for (int i = 0; i < S; i++) {
                const int ao = (wm * (TILE / 2) + 32 * i + (lane & 31)) * PITCH + kk + 8 * (lane >> 5);
                const int bo = (wn * (TILE / 2) + 32 * i + (lane & 31)) * PITCH + kk + 8 * (lane >> 5);
                ah[i] = *reinterpret_cast<const V *>(a_hi + ao);
                al[i] = *reinterpret_cast<const V *>(a_lo + ao);
                bh[i] = *reinterpret_cast<const V *>(b_hi + bo);
                bl[i] = *reinterpret_cast<const V *>(b_lo + bo);
            }
#pragma unroll
            for (int i = 0; i < S; i++)
#pragma unroll
                for (int j = 0; j < S; j++) {
                    acc_lo[i][j] = H::mfma(al[i], bh[j], acc_lo[i][j]);
                    acc_lo[i][j] = H::mfma(ah[i], bl[j], acc_lo[i][j]);
                    acc[i][j] = H::mfma(ah[i], bh[j], acc[i][j]);
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < S; i++)
#pragma unroll
        for (int j = 0; j < S; j++) {
            const long long col = n0 + wn * (TILE / 2) + 32 * j + (lane & 31);
#pragma unroll
            for (int a = 0; a < 16; a++) {
                const long long row = m0 + wm * (TILE / 2) + 32 * i + 8 * (a >> 2) + 4 * (lane >> 5) + (a & 3);
                C[row * N + col] = __builtin_ldexpf((acc[i][j][a] + acc_lo[i][j][a]) * scale, cshift);
            }
        }
}

struct Expo {
    const unsigned *e;   // one max-magnitude word per array, or null
    bool on_a;      // prescale A's rows (row stage) rather than scale C (column stage)
};

template <int PREC, int TILE, int GM>
void launch(bool bt, dim3 grid, hipStream_t s, const float *A, const float *B, float *C, int N, int K, long long sa,
            long long sb, long long sc, float scale, const size_t *offs, int hrows, Expo x)
{
    if (bt)
        hipLaunchKernelGGL((k_gemm<PREC, true, TILE, GM>), grid, dim3(256), 0, s, A, B, C, N, K, sa, sb, sc, scale,
                           offs, hrows, x.e, x.on_a);
    else
        hipLaunchKernelGGL((k_gemm<PREC, false, TILE, GM>), grid, dim3(256), 0, s, A, B, C, N, K, sa, sb, sc, scale,
                           offs, hrows, x.e, x.on_a);
}

template <int PREC, int TILE>
void launch_gm(int gm, bool bt, dim3 grid, hipStream_t s, const float *A, const float *B, float *C, int N, int K,
               long long sa, long long sb, long long sc, float scale, const size_t *offs, int hrows, Expo x)
{
    if (gm == 1)
        launch<PREC, TILE, 1>(bt, grid, s, A, B, C, N, K, sa, sb, sc, scale, offs, hrows, x);
    else if (gm == 2)
        launch<PREC, TILE, 2>(bt, grid, s, A, B, C, N, K, sa, sb, sc, scale, offs, hrows, x);
    else
        launch<PREC, TILE, 0>(bt, grid, s, A, B, C, N, K, sa, sb, sc, scale, offs, hrows, x);
}

template <int PREC>
int gemm(bool bt, const float *A, const float *B, float *C, int M, int N, int K, int batch, long long sa,
         long long sb, long long sc, float scale, hipStream_t s, int gm, const size_t *offs, int hrows, Expo x)
{
    // 128 x 128 tiles where both dimensions allow, else 64 x 64
    const int tile = M % 128 == 0 && N % 128 == 0 ? 128 : 64;
    if (M % 64 || N % 64 || K % BK || batch < 1 || batch > 65535 || M / tile > 65535 || N / tile > 65535)
        return -1;
    const dim3 grid((unsigned)(N / tile), (unsigned)(M / tile), (unsigned)batch);
    if ((gm == 1 || (x.e && x.on_a)) && (hrows < 1 || M % hrows))
        return -1;
    if (tile == 128)
        launch_gm<PREC, 128>(gm, bt, grid, s, A, B, C, N, K, sa, sb, sc, scale, offs, hrows, x);
    else
        launch_gm<PREC, 64>(gm, bt, grid, s, A, B, C, N, K, sa, sb, sc, scale, offs, hrows, x);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int gemm_any(int prec, bool bt, const float *A, const float *B, float *C, int M, int N, int K, int batch,
             long long sa, long long sb, long long sc, float scale, hipStream_t s, int gm, const size_t *offs,
             int hrows, Expo x)
{
    if (prec == BDX_DENSE_F16)
        return gemm<BDX_DENSE_F16>(bt, A, B, C, M, N, K, batch, sa, sb, sc, scale, s, gm, offs, hrows, x);
    if (prec == BDX_DENSE_BF16)
        return gemm<BDX_DENSE_BF16>(bt, A, B, C, M, N, K, batch, sa, sb, sc, scale, s, gm, offs, hrows, x);
    return -1;
}

// Prescale exponents of the arrays, in two launches (graph-capturable: no memset node).
// k_clear zeroes one word per array; k_array_max raises word z to the bit pattern of
// max|x| over array z (non-negative floats order like their bit patterns; a NaN or an
// infinity gives a pattern >= that of infinity), each workgroup covering kMaxPart floats
// of one array (array z at in + offs[z], or in + z * hw).  The GEMMs turn a word into the
// exponent e with max|x| 2^-e in [0.5, 1) (frexp), and into 0 -- no prescale -- for an
// array of zeros or with a non-finite value.
constexpr int kMaxPart = 8192;

__global__ __launch_bounds__(256) void k_clear(unsigned *__restrict__ w, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        w[i] = 0u;
}

__global__ __launch_bounds__(256) void k_array_max(const float *__restrict__ in, size_t hw,
                                                   const size_t *__restrict__ offs, unsigned *__restrict__ bits)
{
    __shared__ unsigned part[4];
    const size_t z = blockIdx.y;
    const float4 *a = reinterpret_cast<const float4 *>(in + (offs ? offs[z] : z * hw));
    const size_t lo = (size_t)blockIdx.x * (kMaxPart / 4), hi = lo + kMaxPart / 4 < hw / 4 ? lo + kMaxPart / 4 : hw / 4;
    unsigned m = 0u;
    for (size_t i = lo + threadIdx.x; i < hi; i += 256) {
        const float4 v = a[i];
        m = max(m, max(max(__float_as_uint(fabsf(v.x)), __float_as_uint(fabsf(v.y))),
                       max(__float_as_uint(fabsf(v.z)), __float_as_uint(fabsf(v.w)))));
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
        m = max(m, (unsigned)__shfl_xor((int)m, d));
    if ((threadIdx.x & 63) == 0)
        part[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        atomicMax(bits + z, max(max(part[0], part[1]), max(part[2], part[3])));
}

int array_exponents(const float *in, size_t hw, size_t count, const size_t *offs, unsigned *bits, hipStream_t s)
{
    const size_t parts = (hw + kMaxPart - 1) / kMaxPart;
    if (count > 65535 || hw % 4)
        return -1;
    hipLaunchKernelGGL(k_clear, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, bits, count);
    hipLaunchKernelGGL(k_array_max, dim3((unsigned)parts, (unsigned)count), dim3(256), 0, s, in, hw, offs, bits);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace
}  // namespace bdx

// ---- static intermediate for calls that allocate nothing ----------------------------------
// One file-scope array per GPU (each device holds its own copy until the process ends).
// Host code never passes `scratch` itself to a kernel -- in host code the name is only a
// stand-in -- but the address hipGetSymbolAddress returns for it on that device.
static __device__ __align__(256) float scratch[BDX_STATIC_WORK_FLOATS];
namespace {
constexpr int kMaxDevices = 64;
float *scratch_on[kMaxDevices];   // per device: the address of its copy of `scratch`
std::mutex scratch_lock;
}  // namespace

extern "C" int bdx_dense_static_work(int device, float **ptr, size_t *floats)
{
    if (device < 0 || device >= kMaxDevices || !ptr || !floats)
        return -1;
    std::lock_guard<std::mutex> hold(scratch_lock);
    if (!scratch_on[device]) {
        int current = -1;
        if (hipGetDevice(&current) != hipSuccess)
            return -1;
        if (current != device && hipSetDevice(device) != hipSuccess)
            return -1;
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (current != device && hipSetDevice(current) != hipSuccess)
            return -1;
        if (e != hipSuccess || !p)
            return -1;
        scratch_on[device] = static_cast<float *>(p);
    }
    *ptr = scratch_on[device];
    *floats = BDX_STATIC_WORK_FLOATS;
    return 0;
}

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
    unsigned *expo = reinterpret_cast<unsigned *>(d_work + batch * hw);
    if (array_exponents(d_in, (size_t)hw, batch, nullptr, expo, s))
        return -1;
    // rows: T = X Q_w'^T (forward; B^T = Q_w') or T = Y Q_w' (inverse; B = Q_w'), each
    // array prescaled by 2^-e
    if (gemm_any(prec, mode == BDX_FORWARD, d_in, d_qw, d_work, (int)(batch * h), w, w, 1, 0, 0, 0, 1.0f, s, 0,
                 nullptr, h, Expo{expo, true}))
        return -1;
    // columns: Y = Q_h' T (forward) or X = Q_h'^T T (inverse), per array, times 2^e
    return gemm_any(prec, false, mode == BDX_FORWARD ? d_qh : d_qht, d_work, d_out, h, w, h, (int)batch, 0, hw,
                    hw, scale, s, 0, nullptr, 0, Expo{expo, false});
}

extern "C" int bdx_dense_transform_gather(const float *d_in, float *d_out, float *d_work, int h, int w, size_t count,
                                          const size_t *d_offsets, int mode, int prec, const float *d_qh,
                                          const float *d_qht, const float *d_qw, float scale, void *stream)
{
    using namespace bdx;
    if (count == 0)
        return 0;
    if (h < 64 || w < 64 || (mode != BDX_FORWARD && mode != BDX_INVERSE) || count > 65535 || !d_offsets)
        return -1;
    hipStream_t s = (hipStream_t)stream;
    const long long hw = (long long)h * w;
    unsigned *expo = reinterpret_cast<unsigned *>(d_work + count * hw);
    if (array_exponents(d_in, (size_t)hw, count, d_offsets, expo, s))
        return -1;
    // rows, gathered through the offsets: T = X Q_w'^T (forward) or T = Y Q_w' (inverse)
    if (gemm_any(prec, mode == BDX_FORWARD, d_in, d_qw, d_work, (int)(count * h), w, w, 1, 0, 0, 0, 1.0f, s, 1,
                 d_offsets, h, Expo{expo, true}))
        return -1;
    // columns, per array, scattered through the offsets
    return gemm_any(prec, false, mode == BDX_FORWARD ? d_qh : d_qht, d_work, d_out, h, w, h, (int)count, 0, hw, 0,
                    scale, s, 2, d_offsets, h, Expo{expo, false});
}
