// Host-callable launchers of the density-guess MLP kernels (mlp.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace uno {

enum MlpEpi : int {
    kEpiStore = 0,      // out_f32 = acc (+ bias)
    kEpiBiasSilu = 1,   // z = acc + bias: out_f32 = z (if set), out_bf16 = silu(z)
    kEpiSiluGrad = 2,   // acc = dH: dZ = dH silu'(aux_f32) -> out_f32 and out_bf16
};

struct MlpEpilogue {
    int ldc;
    const float* bias;
    const float* aux_f32;
    float* out_f32;
    __bf16* out_bf16;
};

// C = op(A) op(B), BF16 operands, FP32 accumulation; M, N multiples of 64, K of 32
hipError_t gemm_bf16(bool ta, bool tb, int epi, int M, int N, int K, const __bf16* A, int lda, const __bf16* B,
                     int ldb, const MlpEpilogue& ep, hipStream_t st);
// dst rows r = src rows idx[r] (idx < 0: zero row)
hipError_t gather_rows(int nrow, int ncol, const int* idx, const float* src, __bf16* dst_bf16, float* dst_f32,
                       hipStream_t st);
// masked squared error: dy = 2 m (y - t) inv_norm; part[b] = block partial of m (y - t)^2
hipError_t mse_loss(long len, const float* y, const float* t, const float* m, float inv_norm, float* dy,
                    __bf16* dy_bf16, float* part, int* nblocks, hipStream_t st);
hipError_t col_sum(int rows, int cols, const float* dz, float* db, hipStream_t st);
// Adam step t (1-based) on FP32 parameters; pb (optional) receives the BF16 copy
hipError_t adam_step(long n, float* p, float* m, float* v, const float* g, float lr, float b1, float b2, int t,
                     float eps, __bf16* pb, hipStream_t st);
hipError_t to_bf16(long n, const float* x, __bf16* y, hipStream_t st);

}  // namespace uno
