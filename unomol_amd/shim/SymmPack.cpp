// Drop-in replacement of the reference's SymmPack.cpp (SymmPack.hpp:13-26): the two O(N^3)
// routines run on the GPU through libunomol_hip.so, the two O(N^2) ones stay on the host.
// Conventions as the reference's: packed lower triangle row-wise, rsp destroys its input and
// returns ascending eigenvalues with the eigenvectors in the columns of the row-major z,
// sp_trans works in place.
#include <cmath>
#include "SymmPack.hpp"
#include "unomol.h"

namespace unomol {
namespace SymmPack {

// sum over the full symmetric matrices of a_ij b_ij (SymmPack.cpp:7-19)
double TraceSymmPackProduct(double a[], double b[], int n) noexcept {
    double off = 0.0, diag = 0.0;
    long ij = 0;
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < i; ++j, ++ij) off += a[ij] * b[ij];
        diag += a[ij] * b[ij];
        ++ij;
    }
    return 2.0 * off + diag;
}

// sqrt(sum over the full matrices of (a - b)^2) / n (SymmPack.cpp:21-36)
double SymmPackDiffNorm(double a[], double b[], int n) noexcept {
    double off = 0.0, diag = 0.0;
    long ij = 0;
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < i; ++j, ++ij) {
            const double d = a[ij] - b[ij];
            off += d * d;
        }
        const double d = a[ij] - b[ij];
        diag += d * d;
        ++ij;
    }
    return std::sqrt(2.0 * off + diag) / n;
}

void rsp(const int n, double* __restrict__ a, double* __restrict__ z, double* __restrict__ evals,
         double* __restrict__ /*tmp*/) noexcept {
    if (uno_eigh_packed(n, a, z, evals) != UNO_OK) fatal_error(uno_last_error());
}

void sp_trans(int n, double* a, double* z, double* /*tmp*/) noexcept {
    if (uno_sp_trans(n, a, z) != UNO_OK) fatal_error(uno_last_error());
}

}  // namespace SymmPack
}  // namespace unomol
