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

// one molecule's n x n matrix in a store (and, for the shell maxima, its shells)
struct MatSlot {
    long long off;      // first element in every store
    int n;
    int nsh;            // shells of the molecule
    int shoff;          // first shell in the shell tables
    long long dshoff;   // first element of its nsh x nsh block maxima
};
hipError_t batch_dot(int nslot, const MatSlot* slots, const double* A, const double* B, double* out, bool acc,
                     hipStream_t st);
hipError_t batch_diffnorm(int nslot, const MatSlot* slots, const double* A, const double* B, double* out,
                          hipStream_t st);
hipError_t batch_sym(int nslot, const MatSlot* slots, const double* W, double* G, bool acc, hipStream_t st);
hipError_t batch_shell_max(int nslot, const MatSlot* slots, const int* sh_off, const int* sh_n, const double* D,
                           const double* D1, const double* D2, double* Dsh, hipStream_t st);
// F_k = sum_j coef[k ncoef + j] src[j]_k (src: device array of ncoef device store pointers)
hipError_t batch_lincomb(int nslot, const MatSlot* slots, int ncoef, const double* coef, const double* const* src,
                         double* F, hipStream_t st);

}  // namespace uno
