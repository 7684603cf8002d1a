// One-electron matrices on the GPU (oneint.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace uno {

struct OneIntShell {
    int l, npr, cen, poff, off;   // angular momentum, primitives, centre, first primitive, first function
};

struct OneIntArgs {
    const OneIntShell* shells;
    const double* alpha;    // exponents
    const double* coef;     // normalised contraction coefficients
    const double* xyz;      // centres (bohr), 3 per centre
    const double* charge;   // nuclear charges
    int ncen, n;
    int skip;               // centre left out of the nuclear attraction (-1: none)
    double* S;              // dense n x n outputs
    double* T;
    double* H;
};

// pairs (shell of angular momentum la, shell of lb), la >= lb <= 2; every element of the
// three matrices belongs to exactly one pair's block or its transpose
hipError_t oneint_launch(int la, int lb, int npair, const int2* pairs, const OneIntArgs& a, hipStream_t st);

}  // namespace uno
