// One-electron matrices on the GPU: overlap S, kinetic T and core Hamiltonian H = T + V,
// one lane per shell pair, specialised per (la, lb) with la >= lb.
//
// Same method and arithmetic as the host code (McMurchie-Davidson Hermite expansion,
// the reference's OneElectronInts.cpp:6-124 and MD_Dfunction.hpp:39-72): per primitive
// pair the 1-D Hermite coefficients E^{ij}_t, S from E^{ij}_0, T as (1/2)<grad a|grad b>
// per axis, and V from the Hermite Coulomb integrals R_tuv summed over the nuclei.  The
// nuclear-attraction Boys function keeps the reference's semantics
// (MD_Rfunction.hpp:2185-2206): the pure large-t limit above t = 20; below it the m-series
// summed until a term is < 1e-12, then downward recursion.
// The "skip" centre (the DPM positron, OneElectronInts.cpp:43) is left out of V.
// Outputs are dense row-major N x N; each lane writes its block and its transpose.
#include <hip/hip_runtime.h>
#include "oneint_kernels.hpp"

namespace uno {

namespace {

__host__ __device__ constexpr int nc1(int l) { return (l + 1) * (l + 2) / 2; }
__host__ __device__ constexpr int cpow(int l, int k, int axis) {
    int idx = 0;
    for (int x = l; x >= 0; --x)
        for (int y = l - x; y >= 0; --y) {
            if (idx == k) return axis == 0 ? x : (axis == 1 ? y : l - x - y);
            ++idx;
        }
    return 0;
}
// per-component factor of the reference's table (AuxFunctions.hpp:49-64), l <= 2
__host__ __device__ constexpr double cnorm(int l, int k) {
    return (cpow(l, k, 0) == 2 || cpow(l, k, 1) == 2 || cpow(l, k, 2) == 2) ? 0.57735026918962576451 : 1.0;
}

__device__ void boys_ref(double* f, double t, int m) {
    if (t > 20.0) {
        f[0] = 0.88622692545275801365 / sqrt(t);
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
    const double et = exp(-t);
    f[m] = s * et;
    for (int i = m - 1; i >= 0; --i) f[i] = (2.0 * t * f[i + 1] + et) / (2.0 * i + 1.0);
}

// E^{ij}_t, i <= I, j <= J, t <= i + j
template <int I, int J>
struct Herm {
    double e[I + 1][J + 1][I + J + 1];
    __device__ __forceinline__ void build(double o2p, double pa, double pb) {
#pragma unroll
        for (int i = 0; i <= I; ++i)
#pragma unroll
            for (int j = 0; j <= J; ++j)
#pragma unroll
                for (int t = 0; t <= I + J; ++t) e[i][j][t] = 0.0;
        e[0][0][0] = 1.0;
#pragma unroll
        for (int j = 1; j <= J; ++j)
#pragma unroll
            for (int t = 0; t <= j; ++t)
                e[0][j][t] = (t ? o2p * e[0][j - 1][t - 1] : 0.0) + pb * e[0][j - 1][t] +
                             (t + 1 <= j - 1 ? (t + 1) * e[0][j - 1][t + 1] : 0.0);
#pragma unroll
        for (int i = 1; i <= I; ++i)
#pragma unroll
            for (int j = 0; j <= J; ++j)
#pragma unroll
                for (int t = 0; t <= i + j; ++t)
                    e[i][j][t] = (t ? o2p * e[i - 1][j][t - 1] : 0.0) + (t <= i - 1 + j ? pa * e[i - 1][j][t] : 0.0) +
                                 (t + 1 <= i - 1 + j ? (t + 1) * e[i - 1][j][t + 1] : 0.0);
    }
};

// lanes per shell pair: each sums the nuclear attraction over every kOiSplit-th centre,
// the partial sums then reduced over the group in a fixed order (the centre loop is the
// kernel's work; one lane per pair left most of the GPU idle)
constexpr int kOiSplit = 8;

template <int LA, int LB>
__global__ __launch_bounds__(128) void k_oneint(int npair, const int2* __restrict__ pairs, OneIntArgs a) {
    constexpr int NA = nc1(LA), NB = nc1(LB), L = LA + LB;
    const long gt = (long)blockIdx.x * blockDim.x + threadIdx.x;
    // groups of kOiSplit consecutive lanes: a group is wholly in or out of range
    const int g = (int)(gt / kOiSplit), sub = (int)(gt % kOiSplit);
    if (g >= npair) return;
    const int sa = pairs[g].x, sb = pairs[g].y;
    const OneIntShell A = a.shells[sa], B = a.shells[sb];
    const double* ra = a.xyz + 3 * A.cen;
    const double* rb = a.xyz + 3 * B.cen;
    const double ab2 = (ra[0] - rb[0]) * (ra[0] - rb[0]) + (ra[1] - rb[1]) * (ra[1] - rb[1]) +
                       (ra[2] - rb[2]) * (ra[2] - rb[2]);
    double s[NA * NB], t[NA * NB], v[NA * NB];
#pragma unroll
    for (int o = 0; o < NA * NB; ++o) s[o] = t[o] = v[o] = 0.0;
    for (int i = 0; i < A.npr; ++i)
        for (int j = 0; j < B.npr; ++j) {
            const double ea = a.alpha[A.poff + i], eb = a.alpha[B.poff + j];
            const double p = ea + eb, ip = 1.0 / p;
            const double K = exp(-ea * eb * ip * ab2);
            double P[3];
#pragma unroll
            for (int x = 0; x < 3; ++x) P[x] = (ea * ra[x] + eb * rb[x]) * ip;
            Herm<LA + 1, LB + 1> hx, hy, hz;
            hx.build(0.5 * ip, P[0] - ra[0], P[0] - rb[0]);
            hy.build(0.5 * ip, P[1] - ra[1], P[1] - rb[1]);
            hz.build(0.5 * ip, P[2] - ra[2], P[2] - rb[2]);
            const double s00 = pow(M_PI * ip, 1.5) * K;
            const double cc = a.coef[A.poff + i] * a.coef[B.poff + j];
            // V_tuv = -sum_c Z_c R_tuv(P - C), t + u + v <= L
            double V[L + 1][L + 1][L + 1];
#pragma unroll
            for (int x = 0; x <= L; ++x)
#pragma unroll
                for (int y = 0; y <= L; ++y)
#pragma unroll
                    for (int z = 0; z <= L; ++z) V[x][y][z] = 0.0;
            const double pref = 2.0 * M_PI * ip * K;
            for (int c = sub; c < a.ncen; c += kOiSplit) {
                if (c == a.skip) continue;   // OneElectronInts.cpp:43
                const double pc[3] = {P[0] - a.xyz[3 * c], P[1] - a.xyz[3 * c + 1], P[2] - a.xyz[3 * c + 2]};
                double f[L + 1];
                boys_ref(f, p * (pc[0] * pc[0] + pc[1] * pc[1] + pc[2] * pc[2]), L);
                double w[L + 1][L + 1][L + 1][L + 1];   // [n][t][u][v]
                double mp = 1.0;
#pragma unroll
                for (int n = 0; n <= L; ++n) {
                    w[n][0][0][0] = pref * mp * f[n];
                    mp *= -2.0 * p;
                }
#pragma unroll
                for (int n = L - 1; n >= 0; --n) {
#pragma unroll
                    for (int tt = 0; tt <= L - n; ++tt)
#pragma unroll
                        for (int u = 0; u <= L - n - tt; ++u)
#pragma unroll
                            for (int vv = 0; vv <= L - n - tt - u; ++vv) {
                                if (tt + u + vv == 0) continue;
                                double r;
                                if (tt) r = pc[0] * w[n + 1][tt - 1][u][vv] + (tt > 1 ? (tt - 1) * w[n + 1][tt - 2][u][vv] : 0.0);
                                else if (u) r = pc[1] * w[n + 1][tt][u - 1][vv] + (u > 1 ? (u - 1) * w[n + 1][tt][u - 2][vv] : 0.0);
                                else r = pc[2] * w[n + 1][tt][u][vv - 1] + (vv > 1 ? (vv - 1) * w[n + 1][tt][u][vv - 2] : 0.0);
                                w[n][tt][u][vv] = r;
                            }
                }
                const double zc = a.charge[c];
#pragma unroll
                for (int x = 0; x <= L; ++x)
#pragma unroll
                    for (int y = 0; y <= L - x; ++y)
#pragma unroll
                        for (int z = 0; z <= L - x - y; ++z) V[x][y][z] -= zc * w[0][x][y][z];
            }
#pragma unroll
            for (int ia = 0; ia < NA; ++ia)
#pragma unroll
                for (int ib = 0; ib < NB; ++ib) {
                    const int l1 = cpow(LA, ia, 0), m1 = cpow(LA, ia, 1), n1 = cpow(LA, ia, 2);
                    const int l2 = cpow(LB, ib, 0), m2 = cpow(LB, ib, 1), n2 = cpow(LB, ib, 2);
                    const double nf = cnorm(LA, ia) * cnorm(LB, ib) * cc;
                    const double ex = hx.e[l1][l2][0], ey = hy.e[m1][m2][0], ez = hz.e[n1][n2][0];
                    const int o = ia * NB + ib;
                    s[o] += nf * s00 * ex * ey * ez;
                    auto kin = [&](const Herm<LA + 1, LB + 1>& h, int k1, int k2) {
                        double r = 2.0 * ea * eb * h.e[k1 + 1][k2 + 1][0];
                        if (k1) r -= k1 * eb * h.e[k1 - 1][k2 + 1][0];
                        if (k2) r -= k2 * ea * h.e[k1 + 1][k2 - 1][0];
                        if (k1 && k2) r += 0.5 * k1 * k2 * h.e[k1 - 1][k2 - 1][0];
                        return r;
                    };
                    t[o] += nf * s00 * (kin(hx, l1, l2) * ey * ez + kin(hy, m1, m2) * ex * ez + kin(hz, n1, n2) * ex * ey);
                    double vs = 0.0;
#pragma unroll
                    for (int x = 0; x <= l1 + l2; ++x)
#pragma unroll
                        for (int y = 0; y <= m1 + m2; ++y) {
                            const double exy = hx.e[l1][l2][x] * hy.e[m1][m2][y];
#pragma unroll
                            for (int z = 0; z <= n1 + n2; ++z) vs += exy * hz.e[n1][n2][z] * V[x][y][z];
                        }
                    v[o] += nf * vs;
                }
        }
    // the group's partial nuclear sums (every lane of the group holds the total after)
#pragma unroll
    for (int o = 0; o < NA * NB; ++o)
#pragma unroll
        for (int m = 1; m < kOiSplit; m <<= 1) v[o] += __shfl_xor(v[o], m, kOiSplit);
    if (sub) return;
    const int n = a.n;
#pragma unroll
    for (int ia = 0; ia < NA; ++ia)
#pragma unroll
        for (int ib = 0; ib < NB; ++ib) {
            const int o = ia * NB + ib;
            const long r = A.off + ia, c = B.off + ib;
            a.S[r * n + c] = s[o];
            a.S[c * n + r] = s[o];
            a.T[r * n + c] = t[o];
            a.T[c * n + r] = t[o];
            a.H[r * n + c] = t[o] + v[o];
            a.H[c * n + r] = t[o] + v[o];
        }
}

}  // namespace

hipError_t oneint_launch(int la, int lb, int npair, const int2* pairs, const OneIntArgs& a, hipStream_t st) {
    if (npair <= 0) return hipSuccess;
    const dim3 grid((unsigned)(((long)npair * kOiSplit + 127) / 128)), block(128);
#define UNO_O(LA_, LB_)                                                                          \
    if (la == LA_ && lb == LB_) {                                                                \
        hipLaunchKernelGGL((k_oneint<LA_, LB_>), grid, block, 0, st, npair, pairs, a);           \
        return hipGetLastError();                                                                \
    }
    UNO_O(0, 0) UNO_O(1, 0) UNO_O(1, 1) UNO_O(2, 0) UNO_O(2, 1) UNO_O(2, 2)
#undef UNO_O
    return hipErrorInvalidValue;
}

}  // namespace uno
