// Host-callable launchers of the device kernels (all asynchronous on the given stream).
#pragma once
#include <hip/hip_runtime.h>
#include "devtypes.hpp"

namespace uno {

// ERIs (eri.hip)
hipError_t eri_launch(int mode, int bra_class, int ket_class, const EriArgs& a, hipStream_t st);
int eri_class_ncomp(int bra_class, int ket_class);
hipError_t rys_eval_launch(int n, int npts, const double* x, double* u, double* w, hipStream_t st);

// dense helpers (linalg.hip); matrices are row-major n x n
hipError_t k_unpack(int n, const double* packed, double* dense, hipStream_t st);       // symmetric fill
hipError_t k_pack(int n, const double* dense, double* packed, hipStream_t st);         // lower triangle
hipError_t k_pack_sym_add(int n, const double* W, double* packed, hipStream_t st);     // packed += (W+W^T)_lower
hipError_t k_sym_w(int n, const double* W, double* G, hipStream_t st);                 // G = W + W^T
hipError_t k_axpby(long len, double a, const double* x, double b, const double* y, double* z, hipStream_t st);
hipError_t k_fill(long len, double v, double* x, hipStream_t st);
// out[0] += sum_ij A_ij B_ij (full matrices, deterministic single-block reduction)
hipError_t k_dot(long len, const double* A, const double* B, double* out, hipStream_t st);
// out[0] = sqrt(sum (A-B)^2) / n
hipError_t k_diffnorm(long len, int n, const double* A, const double* B, double* out, hipStream_t st);
// Dsh[s][t] = max |D_ij| over the function block of shells s, t
hipError_t k_shell_max(int n, int nsh, const int* sh_off, const int* sh_n, const double* D, const double* D1,
                       const double* D2, double* Dsh, hipStream_t st);
// out[0] = max |x|
hipError_t k_absmax(long len, const double* x, double* out, hipStream_t st);
// C = alpha op(A) op(B) + beta C  (FP64 MFMA 16x16x4)
hipError_t gemm_f64(bool ta, bool tb, int m, int n, int k, double alpha, const double* A, int lda, const double* B,
                    int ldb, double beta, double* C, int ldc, hipStream_t st);

// eigensolver pieces (linalg.hip)
struct EigWork {
    double* A;      // n x n working copy (destroyed)
    double* Q;      // n x n accumulated reflectors -> eigenvectors
    double* V;      // n x n reflector vectors (row k = v_k)
    double* tau;    // n
    double* p;      // n
    double* d;      // n  diagonal of T
    double* e;      // n  off-diagonal of T
    double* scal;   // 4 scalars
};
hipError_t householder_tridiag(int n, EigWork& w, hipStream_t st);
hipError_t accumulate_q(int n, EigWork& w, hipStream_t st);
// apply recorded QL sweeps to the rows of Q: sweeps[s] = {lo, hi, first rotation}
hipError_t apply_sweeps(int n, double* Q, int nsweep, const int* sweeps, const double* cs, hipStream_t st);
// Z[:, j] = Q[:, perm[j]]
hipError_t permute_cols(int n, const double* Q, const int* perm, double* Z, hipStream_t st);

}  // namespace uno
