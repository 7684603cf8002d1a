// Pulay DIIS extrapolation coefficients, shared by the single-molecule and the batched SCF.
//
// The history is a ring of K slots, `filled` of them in use, `newest` the slot just written;
// B[i*K + j] = <e_i, e_j> for the error vectors e = X^T (F P S - S P F) X.  The coefficients
// minimise |sum_i c_i e_i| subject to sum_i c_i = 1.  Written in the differences from the
// newest vector (c_n = 1 - sum of the others), that is the least-squares problem
//     min | e_n + sum_i c_i (e_i - e_n) |,   normal equations  Q c = -r,
//     Q_ij = <e_i - e_n, e_j - e_n>,  r_i = <e_i - e_n, e_n>,
// whose solution is unique only while the differences are linearly independent.  They stop
// being so when the error space is smaller than the history (a molecule whose symmetry
// leaves one or two orbital rotations: H2's error vectors are all parallel) or when the
// iteration stalls; the bordered Lagrange system is then singular, and an elimination
// returns whichever member of the family of minimisers rounding selects.  One such member
// repeats the previous extrapolation exactly: the density then stops changing while the
// orbitals are not yet self-consistent, and a density-difference test reports convergence.
// So the oldest vectors are dropped until Q, scaled to unit diagonal, has Cholesky pivots
// above kMinPivot; with no usable pair the newest Fock matrix is taken alone.
#include <algorithm>
#include <cmath>
#include <vector>

#include "engine.hpp"

namespace uno {

namespace {
constexpr double kMinPivot = 1e-10;   // on the unit-diagonal Q: independence to ~1e-5 in angle
constexpr double kMinDiff = 1e-14;    // |e_i - e_n|^2 against max(|e_i|^2, |e_n|^2): rounding level
}  // namespace

int diis_coefficients(const double* B, int K, int filled, int newest, double* coef) {
    for (int i = 0; i < K; ++i) coef[i] = 0.0;
    coef[newest] = 1.0;
    const double bnn = B[newest * K + newest];
    for (int nd = filled; nd >= 2; --nd) {
        const int m = nd - 1;
        std::vector<int> s(m);
        for (int i = 0; i < m; ++i) s[i] = ((newest - 1 - i) % K + K) % K;   // newest first
        std::vector<double> Q((size_t)m * m), r(m), d(m);
        bool ok = true;
        for (int i = 0; i < m && ok; ++i) {
            const double bin = B[s[i] * K + newest];
            r[i] = bin - bnn;
            for (int j = 0; j < m; ++j)
                Q[(size_t)i * m + j] = B[s[i] * K + s[j]] - bin - B[newest * K + s[j]] + bnn;
            const double qii = Q[(size_t)i * m + i];
            if (!(qii > kMinDiff * std::max(B[s[i] * K + s[i]], bnn))) ok = false;
            else d[i] = std::sqrt(qii);
        }
        if (!ok) continue;
        // Cholesky of the unit-diagonal matrix R = D^-1 Q D^-1
        std::vector<double> L((size_t)m * m, 0.0);
        for (int j = 0; j < m && ok; ++j) {
            double v = 1.0;
            for (int k = 0; k < j; ++k) v -= L[(size_t)j * m + k] * L[(size_t)j * m + k];
            if (!(v > kMinPivot)) { ok = false; break; }
            const double ljj = std::sqrt(v);
            L[(size_t)j * m + j] = ljj;
            for (int i = j + 1; i < m; ++i) {
                double w = Q[(size_t)i * m + j] / (d[i] * d[j]);
                for (int k = 0; k < j; ++k) w -= L[(size_t)i * m + k] * L[(size_t)j * m + k];
                L[(size_t)i * m + j] = w / ljj;
            }
        }
        if (!ok) continue;
        // R z = -D^-1 r, c = D^-1 z
        std::vector<double> z(m);
        for (int i = 0; i < m; ++i) {
            double w = -r[i] / d[i];
            for (int k = 0; k < i; ++k) w -= L[(size_t)i * m + k] * z[k];
            z[i] = w / L[(size_t)i * m + i];
        }
        for (int i = m - 1; i >= 0; --i) {
            double w = z[i];
            for (int k = i + 1; k < m; ++k) w -= L[(size_t)k * m + i] * z[k];
            z[i] = w / L[(size_t)i * m + i];
        }
        double sum = 0.0;
        for (int i = 0; i < m; ++i) {
            coef[s[i]] = z[i] / d[i];
            sum += coef[s[i]];
        }
        coef[newest] = 1.0 - sum;
        return nd;
    }
    return 1;
}

}  // namespace uno

extern "C" int uno_diis_coefficients(const double* B, int K, int filled, int newest, double* coef, int* used) {
    using namespace uno;
    if (!B || !coef || !used || K <= 0 || filled < 1 || filled > K || newest < 0 || newest >= K) {
        set_error("bad argument");
        return UNO_EINVAL;
    }
    *used = diis_coefficients(B, K, filled, newest, coef);
    return UNO_OK;
}
