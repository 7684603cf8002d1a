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

// device work buffers of one solve
struct EigBufs {
    double *V, *tau, *p, *d, *e, *scal, *evals, *Y, *wk;
    int* clusters;
};

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

// The solver proper on caller-provided work buffers (device: b; host: hev[n], hcl[n]).
static int eigh_core(int n, double* dA, double* dZ, double* h_evals, const EigBufs& b, double* hev, int* hcl,
                     hipStream_t st) {
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
    EigWork ew{dA, b.V, b.tau, b.p, b.d, b.e, b.scal, b.evals, b.Y, b.wk};
    UNO_HIP(householder_tridiag(n, ew, st));
    UNO_HIP(tridiag_eigen(n, ew, -1, nullptr, st));
    UNO_HIP(hipMemcpyAsync(hev, b.evals, n * sizeof(double), hipMemcpyDeviceToHost, st));
    UNO_HIP(hipStreamSynchronize(st));
    // clusters of eigenvalues closer than 1e-6 of the spectral scale get orthogonalised
    // together; farther pairs are separated by inverse iteration itself ({first, end} pairs,
    // at most n ints)
    const double scale = std::max(1.0, std::max(std::fabs(hev[0]), std::fabs(hev[n - 1])));
    int ncl = 0, start = 0;
    for (int j = 1; j <= n; ++j) {
        if (j == n || hev[j] - hev[j - 1] > 1e-6 * scale) {
            if (j - start > 1) {
                hcl[2 * ncl] = start;
                hcl[2 * ncl + 1] = j;
                ++ncl;
            }
            start = j;
        }
    }
    if (ncl) UNO_HIP(hipMemcpyAsync(b.clusters, hcl, 2 * (size_t)ncl * sizeof(int), hipMemcpyHostToDevice, st));
    UNO_HIP(tridiag_eigen(n, ew, ncl, b.clusters, st));
    // Z = Q Yt: the eigenvectors of T back-transformed by the reflectors in place
    UNO_HIP(hipMemcpyAsync(dZ, b.Y, (size_t)n * n * sizeof(double), hipMemcpyDeviceToDevice, st));
    UNO_HIP(apply_q(n, ew, dZ, st));
    std::copy(hev, hev + n, h_evals);
    return UNO_OK;
}

static int eigh_impl(EigenWS& w, int n, double* dA, double* dZ, double* h_evals, hipStream_t st) {
    UNO_HIP(w.ensure(n));
    w.hev.resize(std::max(n, 1));
    w.hcl.resize(std::max(n, 2));
    const EigBufs b{w.V.p, w.tau.p, w.p.p, w.d.p, w.e.p, w.scal.p, w.evals.p, w.Y.p, w.wk.p, w.clusters.p};
    return eigh_core(n, dA, dZ, h_evals, b, w.hev.data(), w.hcl.data(), st);
}

int eigh_device(uno_ctx* c, int n, double* dA, double* dZ, double* h_evals) {
    if (!c->eig) c->eig = new EigenWS();
    return eigh_impl(*c->eig, n, dA, dZ, h_evals, c->st);
}

size_t eigh_ws_doubles(int n) {
    const size_t sq = (size_t)n * n, pad = 32;
    // V, Y, wk (5 n^2), tau, p, d, e, evals, scal, clusters (n ints), each 256-byte aligned
    return 7 * sq + 5 * (n + pad) + (4 + pad) + (n / 2 + 1 + pad) + 10 * pad;
}

int eigh_in_workspace(int n, double* dA, double* dZ, double* h_evals, double* ws, double* hev, int* hcl,
                      hipStream_t st) {
    const size_t sq = (size_t)n * n;
    size_t off = 0;
    auto take = [&](size_t count) {
        double* p = ws + off;
        off += (count + 31) / 32 * 32;
        return p;
    };
    EigBufs b;
    b.V = take(sq);
    b.Y = take(sq);
    b.wk = take(5 * sq);
    b.tau = take(n);
    b.p = take(n);
    b.d = take(n);
    b.e = take(n);
    b.evals = take(n);
    b.scal = take(4);
    b.clusters = reinterpret_cast<int*>(take(n / 2 + 1));
    if (off > eigh_ws_doubles(n)) {
        set_error("eigensolver work space layout exceeds its size");
        return UNO_EINVAL;
    }
    return eigh_core(n, dA, dZ, h_evals, b, hev, hcl, st);
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
    if (st) (void)hipStreamSynchronize(st);
    for (hipEvent_t e : prof_ev) (void)hipEventDestroy(e);
    for (int w = 0; w + 1 < uno::kMaxLaunchStreams; ++w) {
        if (wst[w]) {
            (void)hipStreamSynchronize(wst[w]);
            (void)hipStreamDestroy(wst[w]);
        }
        if (wjoin[w]) (void)hipEventDestroy(wjoin[w]);
    }
    if (wfork) (void)hipEventDestroy(wfork);
    if (st) (void)hipStreamDestroy(st);
}
