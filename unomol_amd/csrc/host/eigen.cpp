// Dense symmetric eigensolver for device-resident matrices (contract of the
// reference's SymmPack::rsp, SymmPack.cpp:272-288: eigenvalues ascending, eigenvectors
// in the columns of a row-major Z).
//
//   1. Householder tridiagonalisation A = Q T Q^T on the GPU (one reflector per column;
//      symv + rank-2 update kernels), with Q accumulated on the GPU.
//   2. Implicit QL with Wilkinson-type shifts on the tridiagonal T on the host.  Its
//      scalar recurrence is a serial chain of O(N^2) dependent steps, which one CPU core
//      runs faster than one GPU lane; it only records the Givens rotations of each sweep
//      (the same algorithm as the reference's tql2, SymmPack.cpp:143-235, restated).
//   3. The recorded sweeps are applied to the rows of Q on the GPU, every row
//      independently (row-parallel, rows staged in LDS): Z = Q R.
//   4. Columns are permuted into ascending eigenvalue order on the GPU.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <numeric>
#include "engine.hpp"
#include "../hip/kernels.hpp"

namespace uno {

struct EigenWS {
    int n = 0;
    DBuf<double> A, Q, V, tau, p, d, e, scal, cs;
    DBuf<int> sweeps, perm;
    std::vector<double> hd, he, hcs;
    std::vector<int> hsw, hperm;
    hipError_t ensure(int nn) {
        if (nn <= n && A.p) return hipSuccess;
        n = nn;
        const size_t sq = (size_t)nn * nn;
        hipError_t r;
        if ((r = A.alloc(sq)) || (r = Q.alloc(sq)) || (r = V.alloc(sq)) || (r = tau.alloc(nn)) || (r = p.alloc(nn)) ||
            (r = d.alloc(nn)) || (r = e.alloc(nn)) || (r = scal.alloc(4)) || (r = perm.alloc(nn)))
            return r;
        return hipSuccess;
    }
};

// Implicit QL on (d, e), e[i] couples i and i+1.  Records sweeps {lo, hi, first} and
// rotation pairs (c, s).  Returns 0 or the 1-based index of an eigenvalue that did not
// converge within the iteration limit.
static int ql_record(int n, std::vector<double>& d, std::vector<double>& e, std::vector<int>& sw,
                     std::vector<double>& cs) {
    sw.clear();
    cs.clear();
    for (int l = 0; l < n; ++l) {
        int iter = 0, m;
        do {
            for (m = l; m < n - 1; ++m) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) + dd == dd) break;
            }
            if (m != l) {
                if (iter++ == 60) return l + 1;
                double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                double r = std::hypot(g, 1.0);
                g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? std::fabs(r) : -std::fabs(r)));
                double s = 1.0, c = 1.0, p = 0.0;
                const int first = (int)(cs.size() / 2);
                int i;
                bool underflow = false;
                for (i = m - 1; i >= l; --i) {
                    double f = s * e[i];
                    const double b = c * e[i];
                    e[i + 1] = (r = std::hypot(f, g));
                    if (r == 0.0) {
                        d[i + 1] -= p;
                        e[m] = 0.0;
                        underflow = true;
                        break;
                    }
                    s = f / r;
                    c = g / r;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2.0 * c * b;
                    d[i + 1] = g + (p = s * r);
                    g = c * r - b;
                    cs.push_back(c);
                    cs.push_back(s);
                }
                const int lo = underflow ? i + 1 : l;
                if (m > lo) {
                    sw.push_back(lo);
                    sw.push_back(m);
                    sw.push_back(first);
                }
                if (underflow && i >= l) continue;
                d[l] -= p;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
    }
    return 0;
}

static int eigh_impl(EigenWS& w, int n, double* dA, double* dZ, double* h_evals, hipStream_t st) {
    UNO_HIP(w.ensure(n));
    if (n == 1) {
        double a;
        UNO_HIP(hipMemcpyAsync(&a, dA, sizeof(double), hipMemcpyDeviceToHost, st));
        UNO_HIP(hipStreamSynchronize(st));
        h_evals[0] = a;
        const double one = 1.0;
        UNO_HIP(hipMemcpyAsync(dZ, &one, sizeof(double), hipMemcpyHostToDevice, st));
        UNO_HIP(hipStreamSynchronize(st));
        return UNO_OK;
    }
    EigWork ew{dA, w.Q.p, w.V.p, w.tau.p, w.p.p, w.d.p, w.e.p, w.scal.p};
    UNO_HIP(householder_tridiag(n, ew, st));
    UNO_HIP(accumulate_q(n, ew, st));
    w.hd.resize(n);
    w.he.resize(n);
    UNO_HIP(hipMemcpyAsync(w.hd.data(), w.d.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
    UNO_HIP(hipMemcpyAsync(w.he.data(), w.e.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
    UNO_HIP(hipStreamSynchronize(st));
    w.he[n - 1] = 0.0;
    const int bad = ql_record(n, w.hd, w.he, w.hsw, w.hcs);
    if (bad) {
        // same failure the reference reports (SymmPack.cpp:282-285)
        set_error("Eigenvalue " + std::to_string(bad) + " was not found (QL did not converge)");
        return UNO_EEIGEN;
    }
    const int nsweep = (int)(w.hsw.size() / 3);
    if (nsweep > 0) {
        UNO_HIP(w.sweeps.upload(w.hsw.data(), w.hsw.size(), st));
        UNO_HIP(w.cs.upload(w.hcs.data(), w.hcs.size(), st));
        UNO_HIP(apply_sweeps(n, w.Q.p, nsweep, w.sweeps.p, w.cs.p, st));
    }
    w.hperm.resize(n);
    std::iota(w.hperm.begin(), w.hperm.end(), 0);
    std::stable_sort(w.hperm.begin(), w.hperm.end(), [&](int a, int b) { return w.hd[a] < w.hd[b]; });
    for (int j = 0; j < n; ++j) h_evals[j] = w.hd[w.hperm[j]];
    UNO_HIP(w.perm.upload(w.hperm.data(), n, st));
    UNO_HIP(permute_cols(n, w.Q.p, w.perm.p, dZ, st));
    UNO_HIP(hipStreamSynchronize(st));
    return UNO_OK;
}

int eigh_device(uno_ctx* c, int n, double* dA, double* dZ, double* h_evals) {
    if (!c->eig) c->eig = new EigenWS();
    return eigh_impl(*c->eig, n, dA, dZ, h_evals, c->st);
}

int eigh_standalone(int n, double* dA, double* dZ, double* h_evals, hipStream_t st) {
    EigenWS w;
    return eigh_impl(w, n, dA, dZ, h_evals, st);
}

void delete_eigen(EigenWS* w) { delete w; }

}  // namespace uno

uno_ctx::~uno_ctx() {
    uno::delete_eigen(eig);
    if (st) (void)hipStreamDestroy(st);
}
