// Dense symmetric eigensolver for device-resident matrices (contract of the
// reference's SymmPack::rsp, SymmPack.cpp:272-288: eigenvalues ascending, eigenvectors
// in the columns of a row-major Z).  All steps run on the GPU:
//   1. Householder tridiagonalisation A = Q T Q^T (one reflector per column, one launch
//      per column: the previous rank-2 update fused with the next matrix-vector product);
//   2. eigenvalues of T by multisection on Sturm counts, one eigenvalue per wave;
//   3. eigenvectors of T by inverse iteration, one per thread (the vectors stored as the
//      columns of Yt, so a wave's accesses coalesce), then two-pass modified Gram-Schmidt
//      inside clusters of close eigenvalues (exact degeneracies included);
//   4. Z = Q Yt: the reflectors applied to Yt in compact-WY blocks of 32 (every block's
//      T Vb formed in one launch, then two MFMA GEMMs per block).
// The reference reduces with Householder and diagonalises T with implicit QL
// (SymmPack.cpp:67-269); the contract (not the steps) is what the callers rely on.
#include <algorithm>
#include <cmath>
#include "engine.hpp"
#include "../hip/kernels.hpp"

namespace uno {

struct EigenWS {
    int n = 0;
    DBuf<double> V, tau, p, d, e, scal, evals, Y, wk;
    DBuf<int> clusters;
    std::vector<double> hev;
    std::vector<int> hcl;
    hipError_t ensure(int nn) {
        if (nn <= n && V.p) return hipSuccess;
        n = nn;
        const size_t sq = (size_t)nn * nn;
        hipError_t r;
        if ((r = V.alloc(sq)) || (r = tau.alloc(nn)) || (r = p.alloc(nn)) ||
            (r = d.alloc(nn)) ||
            (r = e.alloc(nn)) || (r = scal.alloc(4)) || (r = evals.alloc(nn)) || (r = Y.alloc(sq)) ||
            (r = wk.alloc(5 * sq)) || (r = clusters.alloc(2 * (size_t)nn)))
            return r;
        return hipSuccess;
    }
};

static int eigh_impl(EigenWS& w, int n, double* dA, double* dZ, double* h_evals, hipStream_t st) {
    UNO_HIP(w.ensure(n));
    if (n <= 2) {
        // closed form for the trivial sizes
        double a[4] = {0, 0, 0, 0};
        UNO_HIP(hipMemcpyAsync(a, dA, (size_t)n * n * sizeof(double), hipMemcpyDeviceToHost, st));
        UNO_HIP(hipStreamSynchronize(st));
        double z[4] = {1, 0, 0, 1};
        if (n == 1) {
            h_evals[0] = a[0];
        } else {
            const double p = a[0], q = a[3], r = a[1];
            const double m = 0.5 * (p + q), dd = std::hypot(0.5 * (p - q), r);
            h_evals[0] = m - dd;
            h_evals[1] = m + dd;
            if (r != 0.0 || p > q) {
                // eigenvector of the smaller eigenvalue: (r, l0 - p) or (l0 - q, r)
                double x = r, y = h_evals[0] - p;
                if (std::fabs(x) + std::fabs(y) == 0.0) { x = h_evals[0] - q; y = r; }
                const double nrm = std::hypot(x, y);
                x /= nrm; y /= nrm;
                z[0] = x; z[2] = y; z[1] = -y; z[3] = x;
            }
        }
        UNO_HIP(hipMemcpyAsync(dZ, z, (size_t)n * n * sizeof(double), hipMemcpyHostToDevice, st));
        UNO_HIP(hipStreamSynchronize(st));
        return UNO_OK;
    }
    EigWork ew{dA, w.V.p, w.tau.p, w.p.p, w.d.p, w.e.p, w.scal.p, w.evals.p, w.Y.p, w.wk.p};
    UNO_HIP(householder_tridiag(n, ew, st));
    UNO_HIP(tridiag_eigen(n, ew, -1, nullptr, st));
    w.hev.resize(n);
    UNO_HIP(hipMemcpyAsync(w.hev.data(), w.evals.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
    UNO_HIP(hipStreamSynchronize(st));
    // clusters of eigenvalues closer than 1e-6 of the spectral scale get orthogonalised
    // together; farther pairs are separated by inverse iteration itself
    const double scale = std::max(1.0, std::max(std::fabs(w.hev.front()), std::fabs(w.hev.back())));
    w.hcl.clear();
    int start = 0;
    for (int j = 1; j <= n; ++j) {
        if (j == n || w.hev[j] - w.hev[j - 1] > 1e-6 * scale) {
            if (j - start > 1) {
                w.hcl.push_back(start);
                w.hcl.push_back(j);
            }
            start = j;
        }
    }
    const int ncl = (int)(w.hcl.size() / 2);
    if (ncl) UNO_HIP(w.clusters.upload(w.hcl.data(), w.hcl.size(), st));
    UNO_HIP(tridiag_eigen(n, ew, ncl, w.clusters.p, st));
    // Z = Q Yt: the eigenvectors of T back-transformed by the reflectors in place
    UNO_HIP(hipMemcpyAsync(dZ, w.Y.p, (size_t)n * n * sizeof(double), hipMemcpyDeviceToDevice, st));
    UNO_HIP(apply_q(n, ew, dZ, st));
    std::copy(w.hev.begin(), w.hev.end(), h_evals);
    return UNO_OK;
}

int eigh_device(uno_ctx* c, int n, double* dA, double* dZ, double* h_evals) {
    if (!c->eig) c->eig = new EigenWS();
    return eigh_impl(*c->eig, n, dA, dZ, h_evals, c->st);
}

int eigh_standalone(int n, double* dA, double* dZ, double* h_evals, hipStream_t st) {
    EigenWS w;
    const int rc = eigh_impl(w, n, dA, dZ, h_evals, st);
    (void)hipStreamSynchronize(st);
    return rc;
}

void delete_eigen(EigenWS* w) { delete w; }

}  // namespace uno

uno_ctx::~uno_ctx() {
    uno::delete_eigen(eig);
    if (stage_ev) (void)hipEventSynchronize(stage_ev);
    if (stage_pin) (void)hipHostFree(stage_pin);
    if (stage_ev) (void)hipEventDestroy(stage_ev);
    if (st) (void)hipStreamDestroy(st);
}
