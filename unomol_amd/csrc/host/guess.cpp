// Learned initial density: atom-pair descriptors, local frames and the density blocks
// the network predicts (no reference counterpart: the reference's only guess is the
// core Hamiltonian, RHF.hpp:205-211).
//
// Frames.  Atom A's frame R_A has rows u1 = direction to its nearest atom, u2 = the
// direction to the next atom not collinear with u1, orthogonalised, u3 = u1 x u2 (ties
// in distance broken by atom index; lab axes stand in when neighbours are missing).
// A block P_AB is predicted in A's frame: the molecule rotated by Q = R_A.  Rotating
// a molecule by Q maps each shell's Cartesian functions to phi'_mu = sum_nu M_mu,nu(Q)
// phi_nu, with M_mu,nu = (N_mu / N_nu) [coefficient of y^nu in prod_a (Q^T y)_a^mu_a]
// (N the per-component normalisation), and the density to P' = M^T P M.  Hence
// P_loc,AB = M_A(Q)^T P_AB M_B(Q) and back P_AB = M_A(Q^T)^T P_loc,AB M_B(Q^T)
// (M(Q)^-1 = M(Q^T), a representation of the rotation group).
//
// Descriptor of the ordered pair A <= B (kGuessIn = 128 floats): element one-hots of A
// and B (H C N O F), same-atom flag, Gaussian radial basis of |AB|, the unit vector AB
// and B's frame axes in A's frame, and for A and for B a neighbour environment in A's
// frame: per class (H / heavy) and radial shell, the smooth count and the summed unit
// directions.  Target: the block in A's frame, entry (mu, nu) at 16 mu + nu.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>
#include "engine.hpp"

namespace uno {

constexpr int kGuessIn = kGuessInDim;
constexpr int kGuessOut = kGuessOutDim;
constexpr int kGuessMaxFun = 16;      // functions per atom the block layout holds
constexpr double kGuessCut = 10.0;    // bohr: pairs further apart get zero blocks
constexpr double kEnvCut = 8.0;       // bohr: neighbour environment radius

namespace {

int element_slot(double z) {
    const int iz = (int)std::lround(z);
    switch (iz) {
        case 1: return 0;
        case 6: return 1;
        case 7: return 2;
        case 8: return 3;
        case 9: return 4;
        default: return -1;
    }
}

struct Frames {
    std::vector<std::array<double, 9>> R;          // rows u1, u2, u3
    std::vector<std::vector<int>> shells;          // shells of each atom, deck order
    std::vector<int> nfun;                         // functions per atom
};

void normalize3(double* v) {
    const double s = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] /= s; v[1] /= s; v[2] /= s;
}

bool build_frames(const Basis& B, Frames& F, std::string& err) {
    const int na = B.ncen;
    F.R.assign(na, {});
    F.shells.assign(na, {});
    F.nfun.assign(na, 0);
    for (size_t s = 0; s < B.sh.size(); ++s) {
        if (B.sh[s].l > 2) {
            err = "the learned guess covers s, p and d shells only";
            return false;
        }
        F.shells[B.sh[s].cen].push_back((int)s);
        F.nfun[B.sh[s].cen] += ncart(B.sh[s].l);
    }
    for (int a = 0; a < na; ++a) {
        if (element_slot(B.q[a]) < 0) {
            err = "the learned guess covers H, C, N, O and F only (atom " + std::to_string(a) + ")";
            return false;
        }
        if (F.nfun[a] > kGuessMaxFun) {
            err = "the learned guess holds at most 16 functions per atom";
            return false;
        }
        // neighbours by (distance rounded to 1e-6 bohr, index)
        std::vector<std::pair<double, int>> nb;
        for (int b = 0; b < na; ++b) {
            if (b == a) continue;
            const double* p = &B.xyz[3 * a];
            const double* q = &B.xyz[3 * b];
            const double d = std::sqrt((q[0] - p[0]) * (q[0] - p[0]) + (q[1] - p[1]) * (q[1] - p[1]) +
                                       (q[2] - p[2]) * (q[2] - p[2]));
            nb.push_back({std::round(d * 1e6) * 1e-6, b});
        }
        std::sort(nb.begin(), nb.end());
        double u1[3] = {0, 0, 1}, u2[3] = {1, 0, 0};
        if (!nb.empty()) {
            const double* p = &B.xyz[3 * a];
            const double* q = &B.xyz[3 * nb[0].second];
            for (int k = 0; k < 3; ++k) u1[k] = q[k] - p[k];
            normalize3(u1);
            bool found = false;
            for (size_t i = 1; i < nb.size() && !found; ++i) {
                const double* r = &B.xyz[3 * nb[i].second];
                double v[3] = {r[0] - p[0], r[1] - p[1], r[2] - p[2]};
                const double d = v[0] * u1[0] + v[1] * u1[1] + v[2] * u1[2];
                for (int k = 0; k < 3; ++k) v[k] -= d * u1[k];
                if (std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) > 0.1) {
                    for (int k = 0; k < 3; ++k) u2[k] = v[k];
                    found = true;
                }
            }
            if (!found) {
                // lab axis least aligned with u1
                int best = 0;
                for (int k = 1; k < 3; ++k)
                    if (std::fabs(u1[k]) < std::fabs(u1[best])) best = k;
                u2[0] = u2[1] = u2[2] = 0.0;
                u2[best] = 1.0;
                const double d = u1[best];
                for (int k = 0; k < 3; ++k) u2[k] -= d * u1[k];
            }
        } else {
            u2[0] = 1.0; u2[1] = 0.0; u2[2] = 0.0;
        }
        normalize3(u2);
        const double u3[3] = {u1[1] * u2[2] - u1[2] * u2[1], u1[2] * u2[0] - u1[0] * u2[2],
                              u1[0] * u2[1] - u1[1] * u2[0]};
        for (int k = 0; k < 3; ++k) {
            F.R[a][k] = u1[k];
            F.R[a][3 + k] = u2[k];
            F.R[a][6 + k] = u3[k];
        }
    }
    return true;
}

}  // namespace

// M(Q) of one shell (ncart x ncart, row-major): phi'_mu = sum_nu M_mu,nu phi_nu for the
// molecule rotated by Q (positions r -> Q r)
void shell_rotation(int l, const double* Q, double* M) {
    const int nc = ncart(l);
    std::fill(M, M + nc * nc, 0.0);
    for (int mu = 0; mu < nc; ++mu) {
        int e[3];
        cart(l, mu, e[0], e[1], e[2]);
        // expand prod_a (sum_i Q_ia y_i)^e_a into monomials y^nu
        std::map<std::array<int, 3>, double> poly{{{0, 0, 0}, 1.0}};
        for (int a = 0; a < 3; ++a)
            for (int rep = 0; rep < e[a]; ++rep) {
                std::map<std::array<int, 3>, double> nx;
                for (const auto& t : poly)
                    for (int i = 0; i < 3; ++i) {
                        auto k = t.first;
                        ++k[i];
                        nx[k] += t.second * Q[i * 3 + a];
                    }
                poly.swap(nx);
            }
        const double nm = cart_norm(e[0], e[1], e[2]);
        for (int nu = 0; nu < nc; ++nu) {
            int f[3];
            cart(l, nu, f[0], f[1], f[2]);
            const auto it = poly.find({f[0], f[1], f[2]});
            if (it != poly.end()) M[mu * nc + nu] = nm / cart_norm(f[0], f[1], f[2]) * it->second;
        }
    }
}

namespace {

// block-diagonal M of all shells of an atom (n x n)
void atom_rotation(const Basis& B, const Frames& F, int atom, const double* Q, std::vector<double>& M) {
    const int n = F.nfun[atom];
    M.assign((size_t)n * n, 0.0);
    int o = 0;
    double sm[36];
    for (int s : F.shells[atom]) {
        const int l = B.sh[s].l, nc = ncart(l);
        shell_rotation(l, Q, sm);
        for (int i = 0; i < nc; ++i)
            for (int j = 0; j < nc; ++j) M[(size_t)(o + i) * n + o + j] = sm[i * nc + j];
        o += nc;
    }
}

// atom-ordered function indices of an atom
std::vector<int> atom_functions(const Basis& B, const Frames& F, int atom) {
    std::vector<int> f;
    for (int s : F.shells[atom])
        for (int k = 0; k < ncart(B.sh[s].l); ++k) f.push_back(B.off[s] + k);
    return f;
}

double dist(const Basis& B, int a, int b) {
    const double* p = &B.xyz[3 * a];
    const double* q = &B.xyz[3 * b];
    return std::sqrt((q[0] - p[0]) * (q[0] - p[0]) + (q[1] - p[1]) * (q[1] - p[1]) + (q[2] - p[2]) * (q[2] - p[2]));
}

// environment of atom x seen in frame R (40 floats)
void environment(const Basis& B, int x, const double* R, float* out) {
    static const double centre[5] = {1.0, 2.5, 4.0, 5.5, 7.0};
    const double w = 0.75;
    std::fill(out, out + 40, 0.0f);
    for (int y = 0; y < B.ncen; ++y) {
        if (y == x) continue;
        const double d = dist(B, x, y);
        if (d > kEnvCut) continue;
        const int cls = element_slot(B.q[y]) == 0 ? 0 : 1;
        double v[3];
        for (int k = 0; k < 3; ++k) v[k] = (B.xyz[3 * y + k] - B.xyz[3 * x + k]) / d;
        double vl[3];
        for (int k = 0; k < 3; ++k) vl[k] = R[3 * k] * v[0] + R[3 * k + 1] * v[1] + R[3 * k + 2] * v[2];
        const double fc = 0.5 * (std::cos(M_PI * d / kEnvCut) + 1.0);
        for (int b = 0; b < 5; ++b) {
            const double g = std::exp(-((d - centre[b]) / w) * ((d - centre[b]) / w)) * fc;
            float* o = out + (cls * 5 + b) * 4;
            o[0] += (float)g;
            for (int k = 0; k < 3; ++k) o[1 + k] += (float)(g * vl[k]);
        }
    }
}

void descriptor(const Basis& B, const Frames& F, int a, int b, float* x) {
    std::fill(x, x + kGuessIn, 0.0f);
    const double* R = F.R[a].data();
    x[element_slot(B.q[a])] = 1.0f;
    x[5 + element_slot(B.q[b])] = 1.0f;
    x[10] = a == b ? 1.0f : 0.0f;
    const double d = dist(B, a, b);
    for (int k = 0; k < 12; ++k) x[11 + k] = (float)std::exp(-((d - k) / 0.8) * ((d - k) / 0.8));
    if (a != b)
        for (int k = 0; k < 3; ++k) {
            double s = 0.0;
            for (int i = 0; i < 3; ++i) s += R[3 * k + i] * (B.xyz[3 * b + i] - B.xyz[3 * a + i]) / d;
            x[23 + k] = (float)s;
        }
    // B's frame axes in A's frame: (R_B R_A^T)
    const double* Rb = F.R[b].data();
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += Rb[3 * i + k] * R[3 * j + k];
            x[26 + 3 * i + j] = (float)s;
        }
    environment(B, a, R, x + 35);
    environment(B, b, R, x + 75);
}

void pairs_within_cut(const Basis& B, std::vector<std::pair<int, int>>& pairs) {
    pairs.clear();
    for (int a = 0; a < B.ncen; ++a)
        for (int b = a; b < B.ncen; ++b)
            if (dist(B, a, b) <= kGuessCut) pairs.push_back({a, b});
}

// Q^T of a row-major 3x3
void transpose3(const double* Q, double* T) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) T[i * 3 + j] = Q[j * 3 + i];
}

}  // namespace

int guess_features(const Basis& B, std::vector<float>& X, std::vector<int>& pairs, std::string& err) {
    Frames F;
    if (!build_frames(B, F, err)) return UNO_EINVAL;
    std::vector<std::pair<int, int>> pp;
    pairs_within_cut(B, pp);
    X.assign(pp.size() * kGuessIn, 0.0f);
    pairs.assign(2 * pp.size(), 0);
    // pairs are independent (each writes its own row of X)
    const long npp = (long)pp.size();
#pragma omp parallel for schedule(dynamic, 32)
    for (long k = 0; k < npp; ++k) {
        descriptor(B, F, pp[k].first, pp[k].second, &X[k * kGuessIn]);
        pairs[2 * k] = pp[k].first;
        pairs[2 * k + 1] = pp[k].second;
    }
    return UNO_OK;
}

// dense symmetric P (n x n, row-major) from predicted blocks in the pair frames
int guess_assemble(const Basis& B, const std::vector<int>& pairs, const float* Y, std::vector<double>& P,
                   std::string& err) {
    Frames F;
    if (!build_frames(B, F, err)) return UNO_EINVAL;
    const int n = B.norb;
    P.assign((size_t)n * n, 0.0);
    const long np = (long)(pairs.size() / 2);
    // the pairs' blocks of P are disjoint (a pair (a, b) with a < b writes the (a, b) and
    // (b, a) blocks, (a, a) the diagonal block), so the pairs run in parallel
#pragma omp parallel
    {
    std::vector<double> Ma, Mb, T;
#pragma omp for schedule(dynamic, 16)
    for (long k = 0; k < np; ++k) {
        const int a = pairs[2 * k], b = pairs[2 * k + 1];
        double Qt[9];
        transpose3(F.R[a].data(), Qt);
        atom_rotation(B, F, a, Qt, Ma);
        atom_rotation(B, F, b, Qt, Mb);
        const int na = F.nfun[a], nb = F.nfun[b];
        const float* blk = Y + k * kGuessOut;
        // P_AB = Ma^T L Mb with L the local block
        T.assign((size_t)na * nb, 0.0);
        for (int i = 0; i < na; ++i)
            for (int j = 0; j < nb; ++j) {
                double s = 0.0;
                for (int q = 0; q < nb; ++q) s += (double)blk[i * kGuessMaxFun + q] * Mb[(size_t)q * nb + j];
                T[(size_t)i * nb + j] = s;
            }
        const std::vector<int> fa = atom_functions(B, F, a), fb = atom_functions(B, F, b);
        for (int i = 0; i < na; ++i)
            for (int j = 0; j < nb; ++j) {
                double s = 0.0;
                for (int q = 0; q < na; ++q) s += Ma[(size_t)q * na + i] * T[(size_t)q * nb + j];
                if (a == b) {
                    P[(size_t)fa[i] * n + fb[j]] += 0.5 * s;
                    P[(size_t)fb[j] * n + fa[i]] += 0.5 * s;
                } else {
                    P[(size_t)fa[i] * n + fb[j]] = s;
                    P[(size_t)fb[j] * n + fa[i]] = s;
                }
            }
    }
    }
    return UNO_OK;
}

// descriptors, target blocks (in the pair frames) and masks from a density (dense n x n)
int guess_samples(const Basis& B, const double* P, std::vector<float>& X, std::vector<float>& Y,
                  std::vector<float>& Mk, std::string& err) {
    std::vector<int> pairs;
    if (int rc = guess_features(B, X, pairs, err)) return rc;
    Frames F;
    build_frames(B, F, err);
    const int n = B.norb;
    const size_t np = pairs.size() / 2;
    Y.assign(np * kGuessOut, 0.0f);
    Mk.assign(np * kGuessOut, 0.0f);
    std::vector<double> Ma, Mb, T;
    for (size_t k = 0; k < np; ++k) {
        const int a = pairs[2 * k], b = pairs[2 * k + 1];
        atom_rotation(B, F, a, F.R[a].data(), Ma);
        atom_rotation(B, F, b, F.R[a].data(), Mb);
        const int na = F.nfun[a], nb = F.nfun[b];
        const std::vector<int> fa = atom_functions(B, F, a), fb = atom_functions(B, F, b);
        // L = Ma^T P_AB Mb
        T.assign((size_t)na * nb, 0.0);
        for (int i = 0; i < na; ++i)
            for (int j = 0; j < nb; ++j) {
                double s = 0.0;
                for (int q = 0; q < nb; ++q) s += P[(size_t)fa[i] * n + fb[q]] * Mb[(size_t)q * nb + j];
                T[(size_t)i * nb + j] = s;
            }
        for (int i = 0; i < na; ++i)
            for (int j = 0; j < nb; ++j) {
                double s = 0.0;
                for (int q = 0; q < na; ++q) s += Ma[(size_t)q * na + i] * T[(size_t)q * nb + j];
                Y[k * kGuessOut + i * kGuessMaxFun + j] = (float)s;
                Mk[k * kGuessOut + i * kGuessMaxFun + j] = 1.0f;
            }
    }
    return UNO_OK;
}

}  // namespace uno
