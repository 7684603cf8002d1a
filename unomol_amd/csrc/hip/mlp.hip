// Dense kernels of the density-guess MLP: BF16-input / FP32-accumulate GEMMs on the
// gfx950 matrix cores (v_mfma_f32_16x16x32_bf16) with fused bias / SiLU epilogues, the
// masked-MSE loss, bias gradients, Adam and the BF16 weight refresh.
//
// GEMM: C[M x N] = op(A)[M x K] op(B)[K x N], op = identity or transpose, all operands
// row-major BF16.  M, N multiples of 64 and K a multiple of 32 (callers pad).  A block
// of 256 threads (4 waves) computes a 64 x 64 tile, each wave a 32 x 32 quadrant as 2 x 2
// MFMA tiles of 16 x 16; K advances 32 at a time through LDS.  Operand layout of the
// instruction (pinned by an on-device probe): lane l holds A[l%16][8(l/16)+e] and
// B[8(l/16)+e][l%16], e = 0..7; accumulator register r holds D[4(l/16)+r][l%16].
#include <hip/hip_runtime.h>
#include <cstdint>
#include "mlp_kernels.hpp"

namespace uno {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu_grad(float z) {
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

constexpr int kTM = 64, kTN = 64, kTK = 32, kPadK = 8;   // LDS rows padded by 8 bf16 (16 B)

template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K, const __bf16* __restrict__ A, int lda,
                                                    const __bf16* __restrict__ B, int ldb, MlpEpilogue ep) {
    __shared__ __attribute__((aligned(16))) __bf16 sA[kTM][kTK + kPadK];    // [i][k]
    __shared__ __attribute__((aligned(16))) __bf16 sB[kTN][kTK + kPadK];    // [j][k]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i0 = blockIdx.y * kTM, j0 = blockIdx.x * kTN;
    const int wi = (wave >> 1) * 32, wj = (wave & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += kTK) {
        // stage the 64 x 32 tiles: each thread moves 8 consecutive elements
        {
            if (!TA) {   // A[i][k] at A[i*lda + k]: 8 consecutive k
                const int r = tid >> 2, c = (tid & 3) * 8;
                const __bf16* src = A + (long)(i0 + r) * lda + k0 + c;
#pragma unroll
                for (int e = 0; e < 8; ++e) sA[r][c + e] = src[e];
            } else {     // A stored K x M: op(A)[i][k] = A[k*lda + i]: 8 consecutive i
                const int kk = tid >> 3, c = (tid & 7) * 8;
                const __bf16* src = A + (long)(k0 + kk) * lda + i0 + c;
#pragma unroll
                for (int e = 0; e < 8; ++e) sA[c + e][kk] = src[e];
            }
            if (!TB) {   // B stored K x N: op(B)[k][j] = B[k*ldb + j]: 8 consecutive j
                const int kk = tid >> 3, c = (tid & 7) * 8;
                const __bf16* src = B + (long)(k0 + kk) * ldb + j0 + c;
#pragma unroll
                for (int e = 0; e < 8; ++e) sB[c + e][kk] = src[e];
            } else {     // B stored N x K: op(B)[k][j] = B[j*ldb + k]: 8 consecutive k
                const int r = tid >> 2, c = (tid & 3) * 8;
                const __bf16* src = B + (long)(j0 + r) * ldb + k0 + c;
#pragma unroll
                for (int e = 0; e < 8; ++e) sB[r][c + e] = src[e];
            }
        }
        __syncthreads();
        bf16x8 av[2], bv[2];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int e = 0; e < 8; ++e) av[a][e] = sA[wi + a * 16 + (lane & 15)][8 * (lane >> 4) + e];
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int e = 0; e < 8; ++e) bv[b][e] = sB[wj + b * 16 + (lane & 15)][8 * (lane >> 4) + e];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a], bv[b], acc[a][b], 0, 0, 0);
        __syncthreads();
    }
    // epilogue
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + wi + a * 16 + 4 * (lane >> 4) + r;
                const int j = j0 + wj + b * 16 + (lane & 15);
                const long o = (long)i * ep.ldc + j;
                float v = acc[a][b][r];
                if constexpr (EPI == kEpiStore) {
                    if (ep.bias) v += ep.bias[j];
                    ep.out_f32[o] = v;
                } else if constexpr (EPI == kEpiBiasSilu) {
                    v += ep.bias[j];
                    if (ep.out_f32) ep.out_f32[o] = v;        // pre-activation kept for backward
                    ep.out_bf16[o] = (__bf16)silu(v);
                } else {   // kEpiSiluGrad: v is dH; dZ = dH silu'(Z)
                    const float dz = v * silu_grad(ep.aux_f32[o]);
                    ep.out_f32[o] = dz;
                    ep.out_bf16[o] = (__bf16)dz;
                }
            }
}

hipError_t gemm_bf16(bool ta, bool tb, int epi, int M, int N, int K, const __bf16* A, int lda, const __bf16* B,
                     int ldb, const MlpEpilogue& ep, hipStream_t st) {
    if (M <= 0 || N <= 0 || K <= 0 || M % kTM || N % kTN || K % kTK) return hipErrorInvalidValue;
    const dim3 grid(N / kTN, M / kTM), block(256);
#define UNO_G(TA_, TB_, E_)                                                                                \
    if (ta == TA_ && tb == TB_ && epi == E_) {                                                            \
        hipLaunchKernelGGL((k_gemm_bf16<TA_, TB_, E_>), grid, block, 0, st, M, N, K, A, lda, B, ldb, ep); \
        return hipGetLastError();                                                                          \
    }
    UNO_G(false, false, kEpiStore)
    UNO_G(false, false, kEpiBiasSilu)
    UNO_G(false, true, kEpiStore)
    UNO_G(false, true, kEpiSiluGrad)
    UNO_G(true, false, kEpiStore)
#undef UNO_G
    return hipErrorInvalidValue;
}

// rows idx[r] of a float matrix -> BF16 (and optionally f32) batch rows
__global__ void k_gather_rows(int nrow, int ncol, const int* __restrict__ idx, const float* __restrict__ src,
                              __bf16* __restrict__ dst_bf16, float* __restrict__ dst_f32) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nrow * ncol) return;
    const int r = (int)(t / ncol), c = (int)(t % ncol);
    const float v = idx[r] >= 0 ? src[(long)idx[r] * ncol + c] : 0.0f;
    if (dst_bf16) dst_bf16[t] = (__bf16)v;
    if (dst_f32) dst_f32[t] = v;
}

hipError_t gather_rows(int nrow, int ncol, const int* idx, const float* src, __bf16* dst_bf16, float* dst_f32,
                       hipStream_t st) {
    const long len = (long)nrow * ncol;
    hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, st, nrow, ncol, idx, src,
                       dst_bf16, dst_f32);
    return hipGetLastError();
}

// masked MSE: dY = 2 m (y - t) / norm (f32 and BF16), per-block partial sums of m (y - t)^2
__global__ void k_mse(long len, const float* __restrict__ y, const float* __restrict__ t, const float* __restrict__ m,
                      float inv_norm, float* __restrict__ dy, __bf16* __restrict__ dy_bf16, float* __restrict__ part) {
    __shared__ float red[256];
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float s = 0.0f;
    if (i < len) {
        const float d = m[i] * (y[i] - t[i]);
        s = d * (y[i] - t[i]);
        const float g = 2.0f * d * inv_norm;
        dy[i] = g;
        dy_bf16[i] = (__bf16)g;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

hipError_t mse_loss(long len, const float* y, const float* t, const float* m, float inv_norm, float* dy,
                    __bf16* dy_bf16, float* part, int* nblocks, hipStream_t st) {
    *nblocks = (int)((len + 255) / 256);
    hipLaunchKernelGGL(k_mse, dim3((unsigned)*nblocks), dim3(256), 0, st, len, y, t, m, inv_norm, dy, dy_bf16, part);
    return hipGetLastError();
}

// db[j] = sum_i dZ[i][j] (one thread per column, fixed order)
__global__ void k_colsum(int rows, int cols, const float* __restrict__ dz, float* __restrict__ db) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cols) return;
    float s = 0.0f;
    for (int i = 0; i < rows; ++i) s += dz[(long)i * cols + j];
    db[j] = s;
}

hipError_t col_sum(int rows, int cols, const float* dz, float* db, hipStream_t st) {
    hipLaunchKernelGGL(k_colsum, dim3((cols + 255) / 256), dim3(256), 0, st, rows, cols, dz, db);
    return hipGetLastError();
}

// Adam on FP32 master parameters; refreshes the BF16 copy of the weights
__global__ void k_adam(long n, float* __restrict__ p, float* __restrict__ mo, float* __restrict__ ve,
                       const float* __restrict__ g, float lr, float b1, float b2, float c1, float c2, float eps,
                       __bf16* __restrict__ pb) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float m = b1 * mo[i] + (1.0f - b1) * gi;
    const float v = b2 * ve[i] + (1.0f - b2) * gi * gi;
    mo[i] = m;
    ve[i] = v;
    const float pi = p[i] - lr * (m * c1) / (sqrtf(v * c2) + eps);
    p[i] = pi;
    if (pb) pb[i] = (__bf16)pi;
}

hipError_t adam_step(long n, float* p, float* m, float* v, const float* g, float lr, float b1, float b2, int t,
                     float eps, __bf16* pb, hipStream_t st) {
    const float c1 = 1.0f / (1.0f - powf(b1, (float)t)), c2 = 1.0f / (1.0f - powf(b2, (float)t));
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, p, m, v, g, lr, b1, b2, c1, c2,
                       eps, pb);
    return hipGetLastError();
}

__global__ void k_to_bf16(long n, const float* __restrict__ x, __bf16* __restrict__ y) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (__bf16)x[i];
}

hipError_t to_bf16(long n, const float* x, __bf16* y, hipStream_t st) {
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, x, y);
    return hipGetLastError();
}

}  // namespace uno
