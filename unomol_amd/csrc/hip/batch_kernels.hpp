// Dense kernels of the batched SCF (batch.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace uno {

// one product C = alpha op(A) op(B) + beta C of a group (row-major operands)
struct GemmGroup {
    int m, n, k;
    int ta, tb;
    const double* A;
    const double* B;
    double* C;
    int lda, ldb, ldc;
};
// one 32 x 32 output tile of a group
struct GemmTile {
    int group, ti, tj;
};
hipError_t grouped_gemm(int ntiles, const GemmTile* tiles, const GemmGroup* groups, double alpha, double beta,
                        hipStream_t st);

// one symmetric matrix per block: A (n x n, destroyed), V (n x n work), evals [n], Z (n x n)
struct BatchMat {
    int n;
    double* A;
    double* V;
    double* evals;
    double* Z;
};
// n <= 256 per matrix; sweeps[b] (optional) receives the sweeps taken
hipError_t batched_jacobi(int nmat, const BatchMat* mats, int max_sweeps, int* sweeps, hipStream_t st);

}  // namespace uno
