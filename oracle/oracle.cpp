// TEST INFRASTRUCTURE ONLY -- CPU restatement of the reference path (the oracle).
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this
// library, as the checker.  The product (unomol_amd) never links or calls it.
//
// Each function cites the reference file:line whose behaviour it restates.  The code
// is this project's own; where an order, threshold or quirk has to match (canonical
// quartet order, primitive cut, Fgamma switch, packed layout) it is kept and named.
#include <cmath>
#include <cstring>
#include <cstdio>
#include <cfloat>
#include <vector>
#include <algorithm>
#include "oracle.h"

namespace {

// ----------------------------------------------------------------------------
// basis (Basis.hpp:56-75 normalisation; AuxFunctions.hpp:35-64 ordering/norms)
// ----------------------------------------------------------------------------
struct Sh {
    int l, npr, cen;
    std::vector<double> a, c;   // exponents, normalised coefficients
};

struct Bas {
    std::vector<Sh> sh;
    std::vector<double> xyz, q;
    std::vector<int> off;
    int norb = 0, ncen = 0, maxl = 0;
};

// Cartesian components of shell l: lx descending, then ly descending.
static void comps(int l, std::vector<int>& lx, std::vector<int>& ly, std::vector<int>& lz) {
    lx.clear(); ly.clear(); lz.clear();
    for (int x = l; x >= 0; --x)
        for (int y = l - x; y >= 0; --y) { lx.push_back(x); ly.push_back(y); lz.push_back(l - x - y); }
}

// Component normalisation 1/sqrt(df[lx] df[ly] df[lz]).  df follows the reference
// table (AuxFunctions.hpp:49-58): 1, 1, 3, 45, ... -- exact (2l-1)!! for l<=2,
// the reference's value for l>=3 is kept so f/g shells match it.
static double compnorm(int lx, int ly, int lz) {
    double df[16];
    df[0] = 1.0;
    double m = 1.0;
    for (int i = 1; i < 16; ++i) { df[i] = df[i - 1] * m; m *= (2 * i + 1); }
    return 1.0 / std::sqrt(df[lx] * df[ly] * df[lz]);
}

static Bas make_basis(const orc_basis* b) {
    Bas B;
    B.ncen = b->ncen;
    B.xyz.assign(b->xyz, b->xyz + 3 * b->ncen);
    B.q.assign(b->charge, b->charge + b->ncen);
    int off = 0;
    for (int s = 0; s < b->nshell; ++s) {
        Sh x;
        x.l = b->sh_l[s]; x.npr = b->sh_npr[s]; x.cen = b->sh_cen[s];
        x.a.assign(b->alpha + b->sh_poff[s], b->alpha + b->sh_poff[s] + x.npr);
        x.c.assign(b->coef + b->sh_poff[s], b->coef + b->sh_poff[s] + x.npr);
        // self-overlap of the x^l component is (2l-1)!!-scaled; renormalise the
        // contraction to unity (Basis.hpp:56-75)
        double lp = 1.5 + x.l, sum = 0.0;
        for (int i = 0; i < x.npr; ++i)
            for (int j = 0; j < x.npr; ++j)
                sum += x.c[i] * x.c[j] * std::pow(std::sqrt(x.a[i] * x.a[j]) / (x.a[i] + x.a[j]), lp);
        sum *= 2.0 * std::sqrt(2.0);
        sum = 1.0 / std::sqrt(sum);
        const double pi32 = std::pow(M_PI, 1.5);
        for (int i = 0; i < x.npr; ++i)
            x.c[i] = x.c[i] * sum * std::sqrt(std::pow(2.0 * x.a[i], lp) / pi32);
        B.off.push_back(off);
        off += (x.l + 1) * (x.l + 2) / 2;
        B.maxl = std::max(B.maxl, x.l);
        B.sh.push_back(x);
    }
    B.norb = off;
    return B;
}

static inline double dist2(const double* a, const double* b) {
    double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    return dx * dx + dy * dy + dz * dz;
}

// ----------------------------------------------------------------------------
// Rys ERIs (TwoElectronInts.cpp:420-509; Rys.hpp:85-143,173-212)
// ----------------------------------------------------------------------------
struct RysWork {
    // G[axis][root][i][j], i <= L12, j <= L34
    double G[3][9][9][9];
    double u[9], w[9];
    double binom[9][9];
    RysWork() {
        for (int i = 0; i < 9; ++i)
            for (int j = 0; j <= i; ++j) {
                double c = 1.0;
                for (int k = 0; k < j; ++k) c = c * (i - k) / (k + 1);
                binom[i][j] = c;
            }
    }
};

// 2-D recurrence table for one axis/root (Rys.hpp:194-212 recurrence relations)
static void vrr2d(double (*G)[9], int L12, int L34, double B00, double B1, double B1p,
                  double C, double Cp) {
    G[0][0] = 1.0;
    if (L12 > 0) G[1][0] = C;
    for (int i = 1; i < L12; ++i) G[i + 1][0] = i * B1 * G[i - 1][0] + C * G[i][0];
    for (int j = 0; j < L34; ++j) {
        for (int i = 0; i <= L12; ++i) {
            double v = Cp * G[i][j];
            if (j > 0) v += j * B1p * G[i][j - 1];
            if (i > 0) v += i * B00 * G[i - 1][j];
            G[i][j + 1] = v;
        }
    }
}

// binomial transfer to (l1,l2;l3,l4) on one axis (Rys.hpp:173-192)
static double hrr(const double (*G)[9], const RysWork& W, double ab, double cd,
                  int l1, int l2, int l3, int l4) {
    double s = 0.0, abp = 1.0;
    for (int i = 0; i <= l2; ++i) {
        double cdp = 1.0;
        for (int j = 0; j <= l4; ++j) {
            s += W.binom[l2][i] * W.binom[l4][j] * abp * cdp * G[l1 + l2 - i][l3 + l4 - j];
            cdp *= cd;
        }
        abp *= ab;
    }
    return s;
}

struct Quartet {
    const Sh *s1, *s2, *s3, *s4;
    const double *A, *B, *C, *D;
    bool same12, same34;
};

// contracted ERIs for one shell quartet, component list given by (i,j,k,l) local
// component indices and per-component norm factors
static void quartet_rys(const Quartet& Q, int ncomp, const int* cl, const double* norms,
                        double* out, RysWork& W) {
    const double SRterm = 2.0 * std::pow(M_PI, 2.5);
    const double thresh = 1.0e-12;     // primitive cut on the bare prefactor (:477-479)
    const Sh& s1 = *Q.s1; const Sh& s2 = *Q.s2; const Sh& s3 = *Q.s3; const Sh& s4 = *Q.s4;
    double ab[3], cd[3];
    for (int x = 0; x < 3; ++x) { ab[x] = Q.A[x] - Q.B[x]; cd[x] = Q.C[x] - Q.D[x]; }
    const double ab2 = dist2(Q.A, Q.B), cd2 = dist2(Q.C, Q.D);
    const int L12 = s1.l + s2.l, L34 = s3.l + s4.l;
    const int nr = (L12 + L34) / 2 + 1;
    std::vector<int> c1x, c1y, c1z, c2x, c2y, c2z, c3x, c3y, c3z, c4x, c4y, c4z;
    comps(s1.l, c1x, c1y, c1z); comps(s2.l, c2x, c2y, c2z);
    comps(s3.l, c3x, c3y, c3z); comps(s4.l, c4x, c4y, c4z);
    for (int c = 0; c < ncomp; ++c) out[c] = 0.0;
    for (int i = 0; i < s1.npr; ++i) {
        // same shell: lower triangle of primitive pairs, off-diagonal doubled (:444-449)
        int jend = Q.same12 ? i + 1 : s2.npr;
        for (int j = 0; j < jend; ++j) {
            double f12 = (Q.same12 && i != j) ? 2.0 : 1.0;
            double a = s1.a[i], b = s2.a[j], p = a + b, pi = 1.0 / p;
            double c12 = s1.c[i] * s2.c[j] * f12;
            double s12 = std::exp(-a * b * ab2 * pi);
            double P[3], PA[3];
            for (int x = 0; x < 3; ++x) { P[x] = (a * Q.A[x] + b * Q.B[x]) * pi; PA[x] = P[x] - Q.A[x]; }
            for (int k = 0; k < s3.npr; ++k) {
                int lend = Q.same34 ? k + 1 : s4.npr;
                for (int l = 0; l < lend; ++l) {
                    double f34 = (Q.same34 && k != l) ? 2.0 : 1.0;
                    double c = s3.a[k], d = s4.a[l], qq = c + d, qi = 1.0 / qq;
                    double c34 = s3.c[k] * s4.c[l] * f34;
                    double s34 = std::exp(-c * d * cd2 * qi);
                    double t = p + qq;
                    double sr = SRterm * s12 * s34 * pi * qi / std::sqrt(t);
                    if (sr < thresh) continue;
                    sr *= c12 * c34;
                    double QC[3], PQ[3];
                    for (int x = 0; x < 3; ++x) {
                        double Qx = (c * Q.C[x] + d * Q.D[x]) * qi;
                        QC[x] = Qx - Q.C[x];
                        PQ[x] = P[x] - Qx;
                    }
                    double rho = p * qq / t;
                    double X = rho * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
                    orc_rys(nr, X, W.u, W.w);
                    for (int r = 0; r < nr; ++r) {
                        double fff = W.u[r] / t;
                        double B00 = 0.5 * fff;
                        double B1 = (0.5 - B00 * qq) / p;
                        double B1p = (0.5 - B00 * p) / qq;
                        for (int x = 0; x < 3; ++x) {
                            double Cc = PA[x] - qq * PQ[x] * fff;
                            double Cp = QC[x] + p * PQ[x] * fff;
                            vrr2d(W.G[x][r], L12, L34, B00, B1, B1p, Cc, Cp);
                        }
                    }
                    for (int cc = 0; cc < ncomp; ++cc) {
                        int i1 = cl[4 * cc], i2 = cl[4 * cc + 1], i3 = cl[4 * cc + 2], i4 = cl[4 * cc + 3];
                        double s = 0.0;
                        for (int r = 0; r < nr; ++r) {
                            double ix = hrr(W.G[0][r], W, ab[0], cd[0], c1x[i1], c2x[i2], c3x[i3], c4x[i4]);
                            double iy = hrr(W.G[1][r], W, ab[1], cd[1], c1y[i1], c2y[i2], c3y[i3], c4y[i4]);
                            double iz = hrr(W.G[2][r], W, ab[2], cd[2], c1z[i1], c2z[i2], c3z[i3], c4z[i4]);
                            s += ix * iy * iz * W.w[r];
                        }
                        out[cc] += s * sr * norms[cc];
                    }
                }
            }
        }
    }
}

// canonical enumeration (TwoElectronInts.cpp:541-672)
template <class Sink>
static void eri_enumerate(const Bas& B, Sink sink) {
    RysWork W;
    const int ns = (int)B.sh.size();
    std::vector<int> cl;
    std::vector<double> nrm, vals;
    std::vector<int> fi;
    for (int ish = 0; ish < ns; ++ish)
        for (int jsh = 0; jsh <= ish; ++jsh)
            for (int ksh = 0; ksh <= ish; ++ksh)
                for (int lsh = 0; lsh <= ksh; ++lsh) {
                    const Sh* S[4] = {&B.sh[ish], &B.sh[jsh], &B.sh[ksh], &B.sh[lsh]};
                    int o[4] = {B.off[ish], B.off[jsh], B.off[ksh], B.off[lsh]};
                    int n[4];
                    for (int x = 0; x < 4; ++x) n[x] = (S[x]->l + 1) * (S[x]->l + 2) / 2;
                    // bra/ket are swapped to l1>=l2, l3>=l4 (:563-580, :604-621)
                    bool sw12 = S[0]->l < S[1]->l, sw34 = S[2]->l < S[3]->l;
                    cl.clear(); nrm.clear(); fi.clear();
                    for (int a = 0; a < n[0]; ++a) {
                        int ir = o[0] + a;
                        for (int b = 0; b < n[1]; ++b) {
                            int jr = o[1] + b;
                            if (jr > ir) break;
                            for (int c = 0; c < n[2]; ++c) {
                                int kr = o[2] + c;
                                if (kr > ir) break;
                                for (int d = 0; d < n[3]; ++d) {
                                    int lr = o[3] + d;
                                    if (lr > kr || (ir == kr && lr > jr)) break;
                                    int i1 = sw12 ? b : a, i2 = sw12 ? a : b;
                                    int i3 = sw34 ? d : c, i4 = sw34 ? c : d;
                                    cl.push_back(i1); cl.push_back(i2); cl.push_back(i3); cl.push_back(i4);
                                    std::vector<int> x1, y1, z1;
                                    double nf = 1.0;
                                    int comp[4] = {a, b, c, d};
                                    for (int s = 0; s < 4; ++s) {
                                        comps(S[s]->l, x1, y1, z1);
                                        nf *= compnorm(x1[comp[s]], y1[comp[s]], z1[comp[s]]);
                                    }
                                    nrm.push_back(nf);
                                    fi.push_back(ir); fi.push_back(jr); fi.push_back(kr); fi.push_back(lr);
                                }
                            }
                        }
                    }
                    int knt = (int)nrm.size();
                    if (!knt) continue;
                    Quartet Q;
                    const double* X[4];
                    for (int s = 0; s < 4; ++s) X[s] = &B.xyz[3 * S[s]->cen];
                    Q.s1 = sw12 ? S[1] : S[0]; Q.s2 = sw12 ? S[0] : S[1];
                    Q.A = sw12 ? X[1] : X[0];  Q.B = sw12 ? X[0] : X[1];
                    Q.s3 = sw34 ? S[3] : S[2]; Q.s4 = sw34 ? S[2] : S[3];
                    Q.C = sw34 ? X[3] : X[2];  Q.D = sw34 ? X[2] : X[3];
                    Q.same12 = (ish == jsh);
                    Q.same34 = (ksh == lsh);
                    vals.assign(knt, 0.0);
                    quartet_rys(Q, knt, cl.data(), nrm.data(), vals.data(), W);
                    for (int c = 0; c < knt; ++c)
                        sink(vals[c], fi[4 * c], fi[4 * c + 1], fi[4 * c + 2], fi[4 * c + 3]);
                }
}

// ----------------------------------------------------------------------------
// G replay (TwoElectronInts.cpp:699-820).  Packed index of the pair (p,q), p>=q.
// ----------------------------------------------------------------------------
static inline long pk(int p, int q) { return p >= q ? (long)p * (p + 1) / 2 + q : (long)q * (q + 1) / 2 + p; }

static void g_one_rhf(const double* P, double* G, double v, int i, int j, int k, int l) {
    // canonical i>=j, i>=k, k>=l, (i==k => j>=l); the six distinct images of the
    // function quartet contribute 2J - K with P the half density.
    long ij = pk(i, j), kl = pk(k, l), ik = pk(i, k), il = pk(i, l), jk = pk(j, k), jl = pk(j, l);
    double da = 2.0 * v * P[ij], db = 2.0 * v * P[kl];
    double sjl = v * P[ik], sjk = v * P[il], sik = v * P[jl], sil = v * P[jk];
    if (k != l) {
        db += db;
        G[ik] -= sik;
        if (i != j && j >= k) G[jk] -= sjk;
    }
    G[il] -= sil;
    G[ij] += db;
    if (i != j && j >= l) G[jl] -= sjl;
    if (ij != kl) {
        if (i != j) da += da;
        if (j <= k) {
            G[jk] -= sjk;
            if (i == k && i != j) G[ik] -= sik;
            if (k != l && j <= l) G[jl] -= sjl;
        }
        G[kl] += da;
    }
}

static void g_one_uhf(const double* PA, const double* PB, double* GA, double* GB, double v,
                      int i, int j, int k, int l) {
    long ij = pk(i, j), kl = pk(k, l), ik = pk(i, k), il = pk(i, l), jk = pk(j, k), jl = pk(j, l);
    double da = v * (PA[ij] + PB[ij]), db = v * (PA[kl] + PB[kl]);
    const double* Pp[2] = {PA, PB};
    double* Gp[2] = {GA, GB};
    double sjl[2], sjk[2], sik[2], sil[2];
    for (int s = 0; s < 2; ++s) {
        sjl[s] = v * Pp[s][ik]; sjk[s] = v * Pp[s][il];
        sik[s] = v * Pp[s][jl]; sil[s] = v * Pp[s][jk];
    }
    if (k != l) {
        db += db;
        for (int s = 0; s < 2; ++s) {
            Gp[s][ik] -= sik[s];
            if (i != j && j >= k) Gp[s][jk] -= sjk[s];
        }
    }
    for (int s = 0; s < 2; ++s) {
        Gp[s][il] -= sil[s];
        Gp[s][ij] += db;
        if (i != j && j >= l) Gp[s][jl] -= sjl[s];
    }
    if (ij != kl) {
        if (i != j) da += da;
        for (int s = 0; s < 2; ++s) {
            if (j <= k) {
                Gp[s][jk] -= sjk[s];
                if (i != j && i <= k) Gp[s][ik] -= sik[s];
                if (k != l && j <= l) Gp[s][jl] -= sjl[s];
            }
            Gp[s][kl] += da;
        }
    }
}

// ----------------------------------------------------------------------------
// dense symmetric eigensolver: cyclic Jacobi (contract of SymmPack.cpp:272-288:
// ascending eigenvalues, eigenvectors in columns of a row-major Z)
// ----------------------------------------------------------------------------
static int jacobi_eig(int n, std::vector<double>& A, std::vector<double>& Z, std::vector<double>& ev) {
    Z.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) Z[(size_t)i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                double v = A[(size_t)i * n + j] * A[(size_t)i * n + j];
                tot += v;
                if (i != j) off += v;
            }
        if (off <= 1e-34 * tot || off == 0.0) break;
        for (int p = 0; p < n - 1; ++p)
            for (int qx = p + 1; qx < n; ++qx) {
                double apq = A[(size_t)p * n + qx];
                if (apq == 0.0) continue;
                double app = A[(size_t)p * n + p], aqq = A[(size_t)qx * n + qx];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + qx];
                    A[(size_t)k * n + p] = c * akp - s * akq;
                    A[(size_t)k * n + qx] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    double apk = A[(size_t)p * n + k], aqk = A[(size_t)qx * n + k];
                    A[(size_t)p * n + k] = c * apk - s * aqk;
                    A[(size_t)qx * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    double zkp = Z[(size_t)k * n + p], zkq = Z[(size_t)k * n + qx];
                    Z[(size_t)k * n + p] = c * zkp - s * zkq;
                    Z[(size_t)k * n + qx] = s * zkp + c * zkq;
                }
            }
    }
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return A[(size_t)a * n + a] < A[(size_t)b * n + b]; });
    ev.resize(n);
    std::vector<double> Zs((size_t)n * n);
    for (int c = 0; c < n; ++c) {
        ev[c] = A[(size_t)idx[c] * n + idx[c]];
        for (int k = 0; k < n; ++k) Zs[(size_t)k * n + c] = Z[(size_t)k * n + idx[c]];
    }
    Z.swap(Zs);
    return 0;
}

static void unpack(int n, const double* a, std::vector<double>& A) {
    A.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) A[(size_t)i * n + j] = A[(size_t)j * n + i] = a[pk(i, j)];
}

// ----------------------------------------------------------------------------
// one-electron integrals via McMurchie-Davidson (OneElectronInts.cpp:6-124)
// ----------------------------------------------------------------------------
// Hermite expansion coefficients E[i][j][t] for one axis
struct Herm { double E[8][8][16]; };

static void hermite(Herm& H, int la, int lb, double p, double PA, double PB) {
    double o2p = 0.5 / p;
    std::memset(&H, 0, sizeof(Herm));
    H.E[0][0][0] = 1.0;
    for (int i = 0; i <= la; ++i)
        for (int j = 0; j <= lb; ++j) {
            if (i == 0 && j == 0) continue;
            for (int t = 0; t <= i + j; ++t) {
                double v;
                if (i > 0) {
                    v = PA * H.E[i - 1][j][t] + (t + 1 <= i + j - 1 ? (t + 1) * H.E[i - 1][j][t + 1] : 0.0);
                    if (t > 0) v += o2p * H.E[i - 1][j][t - 1];
                } else {
                    v = PB * H.E[i][j - 1][t] + (t + 1 <= i + j - 1 ? (t + 1) * H.E[i][j - 1][t + 1] : 0.0);
                    if (t > 0) v += o2p * H.E[i][j - 1][t - 1];
                }
                H.E[i][j][t] = v;
            }
        }
}

// Boys function with the reference's nuclear-attraction semantics
// (MD_Rfunction.hpp:2185-2206): pure large-t limit for t > 20; otherwise the
// series at m summed until a term falls below 1e-12, then downward recursion.
static void fgamma_ref(double* fm, double t, int m) {
    if (t > 20.0) {
        fm[0] = 0.88622692545275801365 / std::sqrt(t);
        for (int i = 1; i <= m; ++i) fm[i] = fm[i - 1] * (i - 0.5) / t;
        return;
    }
    double mph = m + 0.5, term = 0.5 / mph, sum = term;
    for (int i = 1; i <= 200; ++i) {
        term *= t / (mph + i);
        sum += term;
        if (term < 1e-12) break;
    }
    double et = std::exp(-t);
    fm[m] = sum * et;
    for (int i = m - 1; i >= 0; --i) fm[i] = (fm[i + 1] * 2.0 * t + et) / (2.0 * i + 1.0);
}

// R_{tuv} Hermite Coulomb integrals: R[t][u][v] = R^{(0)}_{tuv}
static void rtuv(double R[9][9][9], int L, double p, const double* PC, double pref) {
    double F[20];
    double T = p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
    fgamma_ref(F, T, L);
    static thread_local double Rn[20][9][9][9];
    for (int n = 0; n <= L; ++n) Rn[n][0][0][0] = pref * std::pow(-2.0 * p, n) * F[n];
    for (int n = L - 1; n >= 0; --n) {
        int lmax = L - n;
        for (int t = 0; t <= lmax; ++t)
            for (int u = 0; u <= lmax - t; ++u)
                for (int v = 0; v <= lmax - t - u; ++v) {
                    if (t + u + v == 0) continue;
                    double val;
                    if (t > 0) {
                        val = PC[0] * Rn[n + 1][t - 1][u][v];
                        if (t > 1) val += (t - 1) * Rn[n + 1][t - 2][u][v];
                    } else if (u > 0) {
                        val = PC[1] * Rn[n + 1][t][u - 1][v];
                        if (u > 1) val += (u - 1) * Rn[n + 1][t][u - 2][v];
                    } else {
                        val = PC[2] * Rn[n + 1][t][u][v - 1];
                        if (v > 1) val += (v - 1) * Rn[n + 1][t][u][v - 2];
                    }
                    Rn[n][t][u][v] = val;
                }
    }
    for (int t = 0; t <= L; ++t)
        for (int u = 0; u <= L - t; ++u)
            for (int v = 0; v <= L - t - u; ++v) R[t][u][v] = Rn[0][t][u][v];
}

static void oneint(const Bas& B, double* S, double* T, double* H, double* M) {
    const int ns = (int)B.sh.size();
    Herm Hx, Hy, Hz;
    double R[9][9][9];
    for (int is = 0; is < ns; ++is)
        for (int js = 0; js <= is; ++js) {
            const Sh& A = B.sh[is]; const Sh& Bb = B.sh[js];
            const double* RA = &B.xyz[3 * A.cen]; const double* RB = &B.xyz[3 * Bb.cen];
            std::vector<int> ax, ay, az, bx, by, bz;
            comps(A.l, ax, ay, az); comps(Bb.l, bx, by, bz);
            int na = (int)ax.size(), nb = (int)bx.size();
            std::vector<double> s(na * nb, 0.0), t(na * nb, 0.0), v(na * nb, 0.0), m(9 * na * nb, 0.0);
            double ab2 = dist2(RA, RB);
            for (int i = 0; i < A.npr; ++i)
                for (int j = 0; j < Bb.npr; ++j) {
                    double a = A.a[i], b = Bb.a[j], p = a + b;
                    double K = std::exp(-a * b / p * ab2);
                    double P[3];
                    for (int x = 0; x < 3; ++x) P[x] = (a * RA[x] + b * RB[x]) / p;
                    hermite(Hx, A.l + 1, Bb.l + 1, p, P[0] - RA[0], P[0] - RB[0]);
                    hermite(Hy, A.l + 1, Bb.l + 1, p, P[1] - RA[1], P[1] - RB[1]);
                    hermite(Hz, A.l + 1, Bb.l + 1, p, P[2] - RA[2], P[2] - RB[2]);
                    double s00 = std::pow(M_PI / p, 1.5) * K;
                    double cc = A.c[i] * Bb.c[j];
                    // nuclear attraction, all centres
                    int L = A.l + Bb.l;
                    double Vt[9][9][9];
                    std::memset(Vt, 0, sizeof(Vt));
                    for (int c = 0; c < B.ncen; ++c) {
                        double PC[3];
                        for (int x = 0; x < 3; ++x) PC[x] = P[x] - B.xyz[3 * c + x];
                        rtuv(R, L, p, PC, 2.0 * M_PI / p * K);
                        for (int tt = 0; tt <= L; ++tt)
                            for (int uu = 0; uu <= L - tt; ++uu)
                                for (int vv = 0; vv <= L - tt - uu; ++vv) Vt[tt][uu][vv] -= B.q[c] * R[tt][uu][vv];
                    }
                    for (int ia = 0; ia < na; ++ia)
                        for (int ib = 0; ib < nb; ++ib) {
                            int l1 = ax[ia], m1 = ay[ia], n1 = az[ia], l2 = bx[ib], m2 = by[ib], n2 = bz[ib];
                            double nf = compnorm(l1, m1, n1) * compnorm(l2, m2, n2) * cc;
                            double ex = Hx.E[l1][l2][0], ey = Hy.E[m1][m2][0], ez = Hz.E[n1][n2][0];
                            s[ia * nb + ib] += nf * s00 * ex * ey * ez;
                            auto kin = [&](const Herm& Hh, int k1, int k2) {
                                double r = 2.0 * a * b * Hh.E[k1 + 1][k2 + 1][0];
                                if (k1) r -= k1 * b * Hh.E[k1 - 1][k2 + 1][0];
                                if (k2) r -= k2 * a * Hh.E[k1 + 1][k2 - 1][0];
                                if (k1 && k2) r += 0.5 * k1 * k2 * Hh.E[k1 - 1][k2 - 1][0];
                                return r;
                            };
                            double tk = kin(Hx, l1, l2) * ey * ez + kin(Hy, m1, m2) * ex * ez + kin(Hz, n1, n2) * ex * ey;
                            t[ia * nb + ib] += nf * s00 * tk;
                            double sum = 0.0;
                            for (int tt = 0; tt <= l1 + l2; ++tt)
                                for (int uu = 0; uu <= m1 + m2; ++uu)
                                    for (int vv = 0; vv <= n1 + n2; ++vv)
                                        sum += Hx.E[l1][l2][tt] * Hy.E[m1][m2][uu] * Hz.E[n1][n2][vv] * Vt[tt][uu][vv];
                            v[ia * nb + ib] += nf * sum;
                            if (M) {
                                // moments about the origin (Moments.cpp:45-89)
                                double o2p = 0.5 / p, sv = nf * s00;
                                double ovl = sv * ex * ey * ez;
                                double e1x = Hx.E[l1][l2][1], e1y = Hy.E[m1][m2][1], e1z = Hz.E[n1][n2][1];
                                double e2x = Hx.E[l1][l2][2], e2y = Hy.E[m1][m2][2], e2z = Hz.E[n1][n2][2];
                                double* mm = &m[9 * (ia * nb + ib)];
                                mm[0] += ovl * P[0] + e1x * sv * ey * ez;
                                mm[1] += ovl * P[1] + e1y * sv * ex * ez;
                                mm[2] += ovl * P[2] + e1z * sv * ex * ey;
                                mm[3] += ovl * (P[0] * P[0] + o2p) + 2.0 * sv * ey * ez * (P[0] * e1x + e2x);
                                mm[4] += sv * ez * (e1x * e1y + e1x * ey * P[1] + ex * e1y * P[0]) + ovl * P[0] * P[1];
                                mm[5] += sv * ey * (e1x * e1z + e1x * ez * P[2] + ex * e1z * P[0]) + ovl * P[0] * P[2];
                                mm[6] += ovl * (P[1] * P[1] + o2p) + 2.0 * sv * ex * ez * (P[1] * e1y + e2y);
                                mm[7] += sv * ex * (e1y * e1z + e1y * ez * P[2] + ey * e1z * P[1]) + ovl * P[1] * P[2];
                                mm[8] += ovl * (P[2] * P[2] + o2p) + 2.0 * sv * ex * ey * (P[2] * e1z + e2z);
                            }
                        }
                }
            for (int ia = 0; ia < na; ++ia)
                for (int ib = 0; ib < nb; ++ib) {
                    int ir = B.off[is] + ia, jr = B.off[js] + ib;
                    if (jr > ir) continue;
                    long ij = pk(ir, jr);
                    if (S) S[ij] = s[ia * nb + ib];
                    if (T) T[ij] = t[ia * nb + ib];
                    if (H) H[ij] = t[ia * nb + ib] + v[ia * nb + ib];
                    if (M) for (int x = 0; x < 9; ++x) M[9 * ij + x] = m[9 * (ia * nb + ib) + x];
                }
        }
}

static double trace_pk(const double* a, const double* b, int n) {
    // sum_ij a_ij b_ij over the full symmetric matrix (SymmPack.cpp:7-18)
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < i; ++j) s += 2.0 * a[pk(i, j)] * b[pk(i, j)];
        s += a[pk(i, i)] * b[pk(i, i)];
    }
    return s;
}

static double diffnorm_pk(const double* a, const double* b, int n) {
    // sqrt(sum_ij (a-b)_ij^2)/n (SymmPack.cpp:20-36)
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < i; ++j) { double d = a[pk(i, j)] - b[pk(i, j)]; s += 2.0 * d * d; }
        double d = a[pk(i, i)] - b[pk(i, i)];
        s += d * d;
    }
    return std::sqrt(s) / n;
}

} // namespace

// ============================================================================
extern "C" {

int orc_norb(const orc_basis* b) { return make_basis(b).norb; }

int orc_eri_all(const orc_basis* b, long cap, double* val, int* ijkl, long* count) {
    Bas B = make_basis(b);
    long n = 0;
    eri_enumerate(B, [&](double v, int i, int j, int k, int l) {
        if (n < cap) {
            val[n] = v;
            ijkl[4 * n] = i; ijkl[4 * n + 1] = j; ijkl[4 * n + 2] = k; ijkl[4 * n + 3] = l;
        }
        ++n;
    });
    *count = n;
    return 0;
}

int orc_gmat_rhf(long n, const double* val, const int* f, int, const double* P, double* G) {
    for (long e = 0; e < n; ++e) g_one_rhf(P, G, val[e], f[4 * e], f[4 * e + 1], f[4 * e + 2], f[4 * e + 3]);
    return 0;
}

int orc_gmat_uhf(long n, const double* val, const int* f, int, const double* PA, const double* PB,
                 double* GA, double* GB) {
    for (long e = 0; e < n; ++e)
        g_one_uhf(PA, PB, GA, GB, val[e], f[4 * e], f[4 * e + 1], f[4 * e + 2], f[4 * e + 3]);
    return 0;
}

int orc_rsp(int n, double* a, double* z, double* evals) {
    std::vector<double> A, Z, ev;
    unpack(n, a, A);
    jacobi_eig(n, A, Z, ev);
    std::memcpy(z, Z.data(), sizeof(double) * n * n);
    std::memcpy(evals, ev.data(), sizeof(double) * n);
    for (long i = 0; i < (long)n * (n + 1) / 2; ++i) a[i] = 0.0;   // input destroyed
    return 0;
}

int orc_sp_trans(int n, double* a, const double* x) {
    std::vector<double> A;
    unpack(n, a, A);
    std::vector<double> T((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) {
            double aik = A[(size_t)i * n + k];
            for (int j = 0; j < n; ++j) T[(size_t)i * n + j] += aik * x[(size_t)k * n + j];
        }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = 0.0;
            for (int k = 0; k < n; ++k) s += x[(size_t)k * n + i] * T[(size_t)k * n + j];
            a[pk(i, j)] = s;
        }
    return 0;
}

int orc_oneint(const orc_basis* b, double* S, double* T, double* H) {
    Bas B = make_basis(b);
    oneint(B, S, T, H, nullptr);
    return 0;
}

int orc_moment_ints(const orc_basis* b, double* m) {
    Bas B = make_basis(b);
    oneint(B, nullptr, nullptr, nullptr, m);
    return 0;
}

int orc_scf(const orc_basis* b, int nelec, int maxits, double eps, int cflag, orc_scf_out* out,
            double* evals_out, double* P_out, double* mom_out) {
    Bas B = make_basis(b);
    const int n = B.norb;
    const long n2 = (long)n * (n + 1) / 2;
    const bool uhf = (nelec % 2) != 0;
    const int noA = uhf ? nelec - nelec / 2 : nelec / 2;
    const int noB = nelec / 2;
    // convergence floor (Basis.hpp:249-254)
    double xeps = DBL_EPSILON * n * n * 0.5;
    if (eps < xeps) eps = xeps;
    // ERIs once (conventional, like the reference store), keeping |v| > 1e-14
    std::vector<double> ev;
    std::vector<int> ef;
    eri_enumerate(B, [&](double v, int i, int j, int k, int l) {
        if (std::fabs(v) > 1e-14) { ev.push_back(v); ef.push_back(i); ef.push_back(j); ef.push_back(k); ef.push_back(l); }
    });
    std::vector<double> S(n2), T(n2), H(n2), M(9 * n2);
    oneint(B, S.data(), T.data(), H.data(), M.data());
    double nuc = 0.0;
    for (int i = 0; i < B.ncen; ++i)
        for (int j = i + 1; j < B.ncen; ++j)
            nuc += B.q[i] * B.q[j] / std::sqrt(dist2(&B.xyz[3 * i], &B.xyz[3 * j]));
    // X = U s^-1/2 with all vectors kept (RHF.hpp:214-233)
    std::vector<double> X((size_t)n * n), W((size_t)n * n), e(n), tmp(S);
    orc_rsp(n, tmp.data(), W.data(), e.data());
    for (int j = 0; j < n; ++j) {
        double f = e[j];
        if (std::fabs(f) < 1e-7 && f < 0.0) f = -f;
        if (f <= 0.0) return 2;
        e[j] = 1.0 / std::sqrt(f);
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) X[(size_t)i * n + j] = W[(size_t)i * n + j] * e[j];
    const int nsp = uhf ? 2 : 1;
    std::vector<double> P(nsp * n2), Pold(nsp * n2), G(nsp * n2), F(n2), C((size_t)n * n), Ev(nsp * n);
    auto diag_to_P = [&](const double* Fk, double* Pk, double* evk, int nocc, bool keepC) {
        std::vector<double> Fp(Fk, Fk + n2);
        orc_sp_trans(n, Fp.data(), X.data());
        orc_rsp(n, Fp.data(), W.data(), evk);
        // C = X W (RHF.hpp:178-190)
        std::vector<double> Cl((size_t)n * n, 0.0);
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < n; ++k) {
                double xik = X[(size_t)i * n + k];
                for (int j = 0; j < n; ++j) Cl[(size_t)i * n + j] += xik * W[(size_t)k * n + j];
            }
        for (int i = 0; i < n; ++i)
            for (int j = 0; j <= i; ++j) {
                double s = 0.0;
                for (int k = 0; k < nocc; ++k) s += Cl[(size_t)i * n + k] * Cl[(size_t)j * n + k];
                Pk[pk(i, j)] = s;
            }
        if (keepC) C = Cl;
    };
    // core-Hamiltonian guess (RHF.hpp:205-211; UHF.hpp:206-215 beta reuses alpha's C)
    diag_to_P(H.data(), P.data(), Ev.data(), noA, true);
    if (uhf) {
        for (int i = 0; i < n; ++i)
            for (int j = 0; j <= i; ++j) {
                double s = 0.0;
                for (int k = 0; k < noB; ++k) s += C[(size_t)i * n + k] * C[(size_t)j * n + k];
                P[n2 + pk(i, j)] = s;
            }
    }
    double energy = 0.0, eold = 0.0, ediff = 10.0, pdiff = 10.0;
    int iteration = 0;
    auto update = [&]() {
        std::fill(G.begin(), G.end(), 0.0);
        if (!uhf) {
            orc_gmat_rhf((long)ev.size(), ev.data(), ef.data(), n, P.data(), G.data());
            energy = 2.0 * trace_pk(P.data(), H.data(), n) + trace_pk(P.data(), G.data(), n);
        } else {
            orc_gmat_uhf((long)ev.size(), ev.data(), ef.data(), n, P.data(), P.data() + n2, G.data(), G.data() + n2);
            energy = trace_pk(P.data(), H.data(), n) + trace_pk(P.data() + n2, H.data(), n) +
                     0.5 * (trace_pk(P.data(), G.data(), n) + trace_pk(P.data() + n2, G.data() + n2, n));
        }
        ediff = energy - eold;
        eold = energy;
        pdiff = 0.0;
        for (int s = 0; s < nsp; ++s) {
            for (long i = 0; i < n2; ++i) F[i] = H[i] + G[s * n2 + i];
            std::copy(P.begin() + s * n2, P.begin() + (s + 1) * n2, Pold.begin() + s * n2);
            diag_to_P(F.data(), P.data() + s * n2, Ev.data() + s * n, s == 0 ? noA : noB, false);
            pdiff += diffnorm_pk(P.data() + s * n2, Pold.data() + s * n2, n);
        }
        ++iteration;
    };
    iteration = 0;
    eold = 0.0;
    update();
    double init_energy = energy + nuc;
    iteration = 1;
    if (uhf) eold = 0.0;   // UHF.hpp:157-158 resets eold before the second update
    update();
    auto converged = [&]() {
        if (cflag == 0) return ediff < eps;
        if (cflag == 1) return pdiff < eps;
        return pdiff < eps && ediff < eps;
    };
    while (iteration < maxits) {
        // damping only (RHF.hpp:536-569): keep the new P when the energy fell,
        // otherwise average it with the previous input density
        if (!(ediff < 0.0))
            for (long i = 0; i < nsp * n2; ++i) P[i] = 0.5 * (P[i] + Pold[i]);
        update();
        if (converged()) break;
    }
    double kin = uhf ? trace_pk(P.data(), T.data(), n) + trace_pk(P.data() + n2, T.data(), n)
                     : 2.0 * trace_pk(P.data(), T.data(), n);
    if (out) {
        out->energy = energy + nuc;
        out->init_energy = init_energy;
        out->ediff = ediff;
        out->pdiff = pdiff;
        out->nucrep = nuc;
        out->kinetic = kin;
        out->iterations = iteration;
        out->uhf = uhf ? 1 : 0;
    }
    if (evals_out) std::memcpy(evals_out, Ev.data(), sizeof(double) * nsp * n);
    if (P_out) std::memcpy(P_out, P.data(), sizeof(double) * nsp * n2);
    if (mom_out) {
        // AnalyzeMoments (Moments.cpp:189-270): packed P with x4 off-diagonal and x2
        // diagonal factors; UHF uses (PA+PB)/2
        double nM[9] = {0}, eM[9] = {0};
        for (int c = 0; c < B.ncen; ++c) {
            const double* r = &B.xyz[3 * c];
            double qc = B.q[c];
            double v[9] = {r[0], r[1], r[2], r[0] * r[0], r[0] * r[1], r[0] * r[2], r[1] * r[1], r[1] * r[2], r[2] * r[2]};
            for (int x = 0; x < 9; ++x) nM[x] += qc * v[x];
        }
        for (int i = 0; i < n; ++i)
            for (int j = 0; j <= i; ++j) {
                long ij = pk(i, j);
                double pij = uhf ? 0.5 * (P[ij] + P[n2 + ij]) : P[ij];
                double fct = (i == j) ? 2.0 : 4.0;
                for (int x = 0; x < 9; ++x) eM[x] += pij * fct * M[9 * ij + x];
            }
        for (int x = 0; x < 9; ++x) mom_out[x] = nM[x] - eM[x];
    }
    return 0;
}

} // extern "C"
