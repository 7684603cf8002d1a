// Host basis model and one-electron integrals (S, T, V, multipole moments).
//
// One-electron work is O(N^2 * ncen) and stays on the host: it runs once per SCF.
// Method: McMurchie-Davidson Hermite expansion (the reference's method,
// OneElectronInts.cpp:6-124, MD_Dfunction.hpp:39-72, MD_Rfunction.hpp:49-2182),
// written here from the recurrences.  The nuclear-attraction Boys function keeps the
// reference's semantics (MD_Rfunction.hpp:2185-2206): the pure large-t limit above
// t = 20, and below it the m-series summed until a term is < 1e-12 followed by
// downward recursion -- this is part of the energies the reference prints.
#include "basis.hpp"

#include <cmath>
#include <cstring>

namespace uno {

void cart(int l, int k, int& lx, int& ly, int& lz) {
    int idx = 0;
    for (int x = l; x >= 0; --x)
        for (int y = l - x; y >= 0; --y) {
            if (idx == k) {
                lx = x; ly = y; lz = l - x - y;
                return;
            }
            ++idx;
        }
    lx = ly = lz = 0;
}

double cart_norm(int lx, int ly, int lz) {
    // df(0)=1, df(k)=df(k-1)*(2k-3)!!-style table of the reference: 1, 1, 3, 45, 4725
    static const double df[5] = {1.0, 1.0, 3.0, 45.0, 4725.0};
    return 1.0 / std::sqrt(df[lx] * df[ly] * df[lz]);
}

bool make_basis(const uno_basis* b, Basis& B, std::string& err) {
    if (!b || b->nshell <= 0 || b->ncen <= 0) {
        err = "basis has no shells or centres";
        return false;
    }
    B = Basis();
    B.ncen = b->ncen;
    B.xyz.assign(b->xyz, b->xyz + 3 * b->ncen);
    B.q.assign(b->charge, b->charge + b->ncen);
    int off = 0;
    const double pi32 = std::pow(M_PI, 1.5);
    for (int s = 0; s < b->nshell; ++s) {
        Shell x;
        x.l = b->sh_l[s];
        x.npr = b->sh_npr[s];
        x.cen = b->sh_cen[s];
        if (x.l < 0 || x.l > 4 || x.npr <= 0 || x.cen < 0 || x.cen >= b->ncen) {
            err = "shell " + std::to_string(s) + " is malformed";
            return false;
        }
        x.a.assign(b->alpha + b->sh_poff[s], b->alpha + b->sh_poff[s] + x.npr);
        x.c.assign(b->coef + b->sh_poff[s], b->coef + b->sh_poff[s] + x.npr);
        const double lp = 1.5 + x.l;
        double ov = 0.0;
        for (int i = 0; i < x.npr; ++i)
            for (int j = 0; j < x.npr; ++j)
                ov += x.c[i] * x.c[j] * std::pow(std::sqrt(x.a[i] * x.a[j]) / (x.a[i] + x.a[j]), lp);
        ov *= 2.0 * std::sqrt(2.0);
        if (!(ov > 0.0)) {
            err = "shell " + std::to_string(s) + " has zero norm";
            return false;
        }
        const double sc = 1.0 / std::sqrt(ov);
        if (!b->normalized)
            for (int i = 0; i < x.npr; ++i) x.c[i] *= sc * std::sqrt(std::pow(2.0 * x.a[i], lp) / pi32);
        B.off.push_back(off);
        off += ncart(x.l);
        if (x.l > B.maxl) B.maxl = x.l;
        B.sh.push_back(std::move(x));
    }
    B.norb = off;
    return true;
}

namespace {

// E^{ij}_t for one axis, i <= la, j <= lb (la, lb <= 7)
struct Hermite1D {
    double e[8][8][16];
    void build(int la, int lb, double o2p, double pa, double pb) {
        std::memset(e, 0, sizeof(e));
        e[0][0][0] = 1.0;
        for (int j = 1; j <= lb; ++j)
            for (int t = 0; t <= j; ++t)
                e[0][j][t] = (t ? o2p * e[0][j - 1][t - 1] : 0.0) + pb * e[0][j - 1][t] + (t + 1) * e[0][j - 1][t + 1];
        for (int i = 1; i <= la; ++i)
            for (int j = 0; j <= lb; ++j)
                for (int t = 0; t <= i + j; ++t)
                    e[i][j][t] = (t ? o2p * e[i - 1][j][t - 1] : 0.0) + pa * e[i - 1][j][t] + (t + 1) * e[i - 1][j][t + 1];
    }
};

void boys_ref(double* f, double t, int m) {
    if (t > 20.0) {
        f[0] = 0.88622692545275801365 / std::sqrt(t);
        for (int i = 1; i <= m; ++i) f[i] = f[i - 1] * (i - 0.5) / t;
        return;
    }
    const double mh = m + 0.5;
    double term = 0.5 / mh, s = term;
    for (int i = 1; i <= 200; ++i) {
        term *= t / (mh + i);
        s += term;
        if (term < 1e-12) break;
    }
    const double et = std::exp(-t);
    f[m] = s * et;
    for (int i = m - 1; i >= 0; --i) f[i] = (2.0 * t * f[i + 1] + et) / (2.0 * i + 1.0);
}

// Hermite Coulomb R^{(0)}_{tuv}, t+u+v <= L (L <= 8), scaled by pref
void hermite_coulomb(double R[9][9][9], int L, double p, const double pc[3], double pref) {
    double f[20];
    boys_ref(f, p * (pc[0] * pc[0] + pc[1] * pc[1] + pc[2] * pc[2]), L);
    double w[9][9][9][9];   // [n][t][u][v]
    double mp = 1.0;
    for (int n = 0; n <= L; ++n) {
        w[n][0][0][0] = pref * mp * f[n];
        mp *= -2.0 * p;
    }
    for (int n = L - 1; n >= 0; --n) {
        const int lm = L - n;
        for (int t = 0; t <= lm; ++t)
            for (int u = 0; u <= lm - t; ++u)
                for (int v = 0; v <= lm - t - u; ++v) {
                    if (t + u + v == 0) continue;
                    double r;
                    if (t) r = pc[0] * w[n + 1][t - 1][u][v] + (t > 1 ? (t - 1) * w[n + 1][t - 2][u][v] : 0.0);
                    else if (u) r = pc[1] * w[n + 1][t][u - 1][v] + (u > 1 ? (u - 1) * w[n + 1][t][u - 2][v] : 0.0);
                    else r = pc[2] * w[n + 1][t][u][v - 1] + (v > 1 ? (v - 1) * w[n + 1][t][u][v - 2] : 0.0);
                    w[n][t][u][v] = r;
                }
    }
    for (int t = 0; t <= L; ++t)
        for (int u = 0; u <= L - t; ++u)
            for (int v = 0; v <= L - t - u; ++v) R[t][u][v] = w[0][t][u][v];
}

}  // namespace

void one_electron(const Basis& B, double* S, double* T, double* H, double* M) {
    const int ns = (int)B.sh.size();
    // shell pairs are independent and write disjoint packed elements: spread them over
    // the host cores (dynamic schedule, rows of the triangle differ in length)
#pragma omp parallel for schedule(dynamic, 1)
    for (int is = 0; is < ns; ++is)
        for (int js = 0; js <= is; ++js) {
            Hermite1D hx, hy, hz;
            double R[9][9][9];
            const Shell& A = B.sh[is];
            const Shell& Bs = B.sh[js];
            const double* ra = &B.xyz[3 * A.cen];
            const double* rb = &B.xyz[3 * Bs.cen];
            const int na = ncart(A.l), nb = ncart(Bs.l);
            std::vector<double> s(na * nb, 0.0), t(na * nb, 0.0), v(na * nb, 0.0), m(M ? 9 * na * nb : 0, 0.0);
            const double ab2 = (ra[0] - rb[0]) * (ra[0] - rb[0]) + (ra[1] - rb[1]) * (ra[1] - rb[1]) +
                               (ra[2] - rb[2]) * (ra[2] - rb[2]);
            int lx[15], ly[15], lz[15], mx[15], my[15], mz[15];
            for (int k = 0; k < na; ++k) cart(A.l, k, lx[k], ly[k], lz[k]);
            for (int k = 0; k < nb; ++k) cart(Bs.l, k, mx[k], my[k], mz[k]);
            const int L = A.l + Bs.l;
            for (int i = 0; i < A.npr; ++i)
                for (int j = 0; j < Bs.npr; ++j) {
                    const double a = A.a[i], b = Bs.a[j], p = a + b, ip = 1.0 / p;
                    const double K = std::exp(-a * b * ip * ab2);
                    double P[3];
                    for (int x = 0; x < 3; ++x) P[x] = (a * ra[x] + b * rb[x]) * ip;
                    hx.build(A.l + 1, Bs.l + 1, 0.5 * ip, P[0] - ra[0], P[0] - rb[0]);
                    hy.build(A.l + 1, Bs.l + 1, 0.5 * ip, P[1] - ra[1], P[1] - rb[1]);
                    hz.build(A.l + 1, Bs.l + 1, 0.5 * ip, P[2] - ra[2], P[2] - rb[2]);
                    const double s00 = std::pow(M_PI * ip, 1.5) * K;
                    const double cc = A.c[i] * Bs.c[j];
                    double V[9][9][9];
                    for (int x = 0; x <= L; ++x)
                        for (int y = 0; y <= L - x; ++y)
                            for (int z = 0; z <= L - x - y; ++z) V[x][y][z] = 0.0;
                    for (int c = 0; c < B.ncen; ++c) {
                        const double pc[3] = {P[0] - B.xyz[3 * c], P[1] - B.xyz[3 * c + 1], P[2] - B.xyz[3 * c + 2]};
                        hermite_coulomb(R, L, p, pc, 2.0 * M_PI * ip * K);
                        for (int x = 0; x <= L; ++x)
                            for (int y = 0; y <= L - x; ++y)
                                for (int z = 0; z <= L - x - y; ++z) V[x][y][z] -= B.q[c] * R[x][y][z];
                    }
                    for (int ia = 0; ia < na; ++ia)
                        for (int ib = 0; ib < nb; ++ib) {
                            const int l1 = lx[ia], m1 = ly[ia], n1 = lz[ia], l2 = mx[ib], m2 = my[ib], n2 = mz[ib];
                            const double nf = cart_norm(l1, m1, n1) * cart_norm(l2, m2, n2) * cc;
                            const double ex = hx.e[l1][l2][0], ey = hy.e[m1][m2][0], ez = hz.e[n1][n2][0];
                            const int o = ia * nb + ib;
                            s[o] += nf * s00 * ex * ey * ez;
                            // kinetic energy as (1/2)<grad a|grad b> per axis
                            auto kin = [&](const Hermite1D& h, int k1, int k2) {
                                double r = 2.0 * a * b * h.e[k1 + 1][k2 + 1][0];
                                if (k1) r -= k1 * b * h.e[k1 - 1][k2 + 1][0];
                                if (k2) r -= k2 * a * h.e[k1 + 1][k2 - 1][0];
                                if (k1 && k2) r += 0.5 * k1 * k2 * h.e[k1 - 1][k2 - 1][0];
                                return r;
                            };
                            t[o] += nf * s00 * (kin(hx, l1, l2) * ey * ez + kin(hy, m1, m2) * ex * ez + kin(hz, n1, n2) * ex * ey);
                            double vs = 0.0;
                            for (int x = 0; x <= l1 + l2; ++x)
                                for (int y = 0; y <= m1 + m2; ++y) {
                                    const double exy = hx.e[l1][l2][x] * hy.e[m1][m2][y];
                                    for (int z = 0; z <= n1 + n2; ++z) vs += exy * hz.e[n1][n2][z] * V[x][y][z];
                                }
                            v[o] += nf * vs;
                            if (M) {
                                // multipoles about the origin: x = P_x + (x - P_x)
                                const double sv = nf * s00, ov = sv * ex * ey * ez, o2p = 0.5 * ip;
                                const double x1 = hx.e[l1][l2][1], y1 = hy.e[m1][m2][1], z1 = hz.e[n1][n2][1];
                                const double x2 = hx.e[l1][l2][2], y2 = hy.e[m1][m2][2], z2 = hz.e[n1][n2][2];
                                double* q = &m[9 * o];
                                q[0] += ov * P[0] + sv * x1 * ey * ez;
                                q[1] += ov * P[1] + sv * y1 * ex * ez;
                                q[2] += ov * P[2] + sv * z1 * ex * ey;
                                q[3] += ov * (P[0] * P[0] + o2p) + 2.0 * sv * ey * ez * (P[0] * x1 + x2);
                                q[4] += sv * ez * (x1 * y1 + x1 * ey * P[1] + ex * y1 * P[0]) + ov * P[0] * P[1];
                                q[5] += sv * ey * (x1 * z1 + x1 * ez * P[2] + ex * z1 * P[0]) + ov * P[0] * P[2];
                                q[6] += ov * (P[1] * P[1] + o2p) + 2.0 * sv * ex * ez * (P[1] * y1 + y2);
                                q[7] += sv * ex * (y1 * z1 + y1 * ez * P[2] + ey * z1 * P[1]) + ov * P[1] * P[2];
                                q[8] += ov * (P[2] * P[2] + o2p) + 2.0 * sv * ex * ey * (P[2] * z1 + z2);
                            }
                        }
                }
            for (int ia = 0; ia < na; ++ia)
                for (int ib = 0; ib < nb; ++ib) {
                    const int ir = B.off[is] + ia, jr = B.off[js] + ib;
                    if (jr > ir) continue;
                    const long ij = pidx(ir, jr);
                    const int o = ia * nb + ib;
                    if (S) S[ij] = s[o];
                    if (T) T[ij] = t[o];
                    if (H) H[ij] = t[o] + v[o];
                    if (M) std::memcpy(M + 9 * ij, &m[9 * o], 9 * sizeof(double));
                }
        }
}

}  // namespace uno
