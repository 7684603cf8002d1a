// Rys-quadrature shell-quartet ERIs for gfx950, specialised per (la,lb|lc,ld) class.
//
// Mathematics (the reference's, TwoElectronInts.cpp:420-509 and Rys.hpp:85-212):
//   (ab|cd) = sum_prims 2 pi^(5/2) K_ab K_cd / (p q sqrt(p+q)) sum_r w_r Ix Iy Iz,
// with the 2-D integrals from the Rys recurrence (B00, B1, B1', C, C') and the
// binomial transfer to (la,lb|lc,ld).  The primitive cut drops a primitive quartet
// whose bare prefactor 2 pi^(5/2) e_ab e_cd / (p q sqrt(p+q)) is below primcut, as the
// reference does with 1e-12 (TwoElectronInts.cpp:477-479).  The Rys roots/weights come
// from this project's own Chebyshev tables (rys_table.inc, gen_rys_table.py), evaluated in
// their power-basis re-expansion (rys_mono.inc, gen_rys_monomial.py).  Two roots use the
// reference-compatible table by default (rys_ref2.inc, gen_rys_ref2.py): the reference's
// own two-root quadrature, which departs from the exact one for 15 < x <= 33
// (Rys.cpp:615-624); EriArgs::rys2_exact selects the exact table instead.
//
// Two execution shapes:
//   * one lane per shell quartet (classes with <= 162 components): VRR, transfer and the
//     contracted component block live in registers, fully unrolled per class;
//   * one 64-lane wave per shell quartet (larger classes): lanes split the roots, the
//     (root, axis) recurrences, the transfer entries and the component assembly, with the
//     2-D tables staged in LDS.
// Both end in the same J/K digestion: the six density blocks the quartet touches are
// contracted with the ERI block and added to the unsymmetrised W = 4J' - K' (RHF) or
// Ws = 2J'[Pa+Pb] - K'[Ps] (UHF) with one FP64 atomic per output element; G = W + W^T.
#pragma once
#include <hip/hip_runtime.h>
#include "devtypes.hpp"

namespace uno {

#define UNO_RYS_QUAL static __device__ const
#include "../rys_table.inc"
#include "../rys_table_f32.inc"
#include "../rys_mono.inc"
#include "../rys_ref2.inc"
#undef UNO_RYS_QUAL

__host__ __device__ constexpr int ncart_c(int l) { return (l + 1) * (l + 2) / 2; }
__host__ __device__ constexpr int cart_pow(int l, int k, int axis) {
    int idx = 0;
    for (int x = l; x >= 0; --x)
        for (int y = l - x; y >= 0; --y) {
            if (idx == k) return axis == 0 ? x : (axis == 1 ? y : l - x - y);
            ++idx;
        }
    return 0;
}
// component normalisation for l <= 2: 1/sqrt(3) for the x^2, y^2, z^2 components of d
__host__ __device__ constexpr double cart_nrm(int l, int k) {
    return (cart_pow(l, k, 0) == 2 || cart_pow(l, k, 1) == 2 || cart_pow(l, k, 2) == 2) ? 0.57735026918962576451
                                                                                         : 1.0;
}
__host__ __device__ constexpr double binom_c(int n, int k) {
    double r = 1.0;
    for (int i = 0; i < k; ++i) r = r * (n - i) / (i + 1);
    return r;
}

constexpr double kTwoPi25 = 34.986836655249725;   // 2 pi^(5/2)

// 1/sqrt(s) for s in the normal range: the hardware estimate refined by two Newton steps
// in double (the estimate is good to about 2^-23 relative, each step squares the error);
// single precision uses the estimate as it is (1 ulp)
template <typename T>
__device__ __forceinline__ T rsqrt_nr(T s) {
    if constexpr (sizeof(T) == 4) {
        return __builtin_amdgcn_rsqf(s);
    } else {
        const T h = (T)0.5 * s;
        T y = __builtin_amdgcn_rsq(s);
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const T r = fma(-(h * y), y, (T)0.5);
            y = fma(y, r, y);
        }
        return y;
    }
}

constexpr int kRysAsym[6] = {0, 40, 48, 56, 62, 68};
// entries of the table of NR roots: intervals x 2 NR functions x coefficients
template <int NR, typename T>
__host__ __device__ constexpr int rys_table_size() {
    return kRysAsym[NR] * 2 * NR * (sizeof(T) == 4 ? UNO_RYS_NCOEF_F : UNO_RYS_NCOEF);
}
// the power-basis table (rys_mono.inc) of NR roots: per unit interval, 2 NR series of
// NCOEF coefficients each, roots first
template <int NR, typename T>
__device__ __forceinline__ const T* rys_table_global() {
    if constexpr (sizeof(T) == 4) return uno_rys_mono_f + uno_rys_off_f[NR];
    else return uno_rys_mono + uno_rys_off[NR];
}

// Rys roots u = t^2 and weights for argument x: Horner on the per-interval series in
// s = 2(x-k)-1 (degree 12 in double, degree 6 in float; the Chebyshev tables re-expanded
// in powers of s), Gauss-Hermite limit for large x.  tab is the table of NR roots (global
// memory or a copy in LDS).
template <int NR, typename T>
__device__ __forceinline__ void rys_roots(T x, T* u, T* w, const T* tab) {
    constexpr bool f32 = sizeof(T) == 4;
    if (x >= (T)kRysAsym[NR]) {
        const T isx = rsqrt_nr(x), ix = isx * isx;
        constexpr int o = NR * (NR - 1) / 2;
#pragma unroll
        for (int i = 0; i < NR; ++i) {
            if constexpr (f32) {
                u[i] = uno_rys_asym_u_f[o + i] * ix;
                w[i] = uno_rys_asym_w_f[o + i] * isx;
            } else {
                u[i] = uno_rys_asym_u[o + i] * ix;
                w[i] = uno_rys_asym_w[o + i] * isx;
            }
        }
        return;
    }
    constexpr int NC = f32 ? UNO_RYS_NCOEF_F : UNO_RYS_NCOEF;
    const int k = (int)x;
    const T s = (T)2 * (x - (T)k) - (T)1;
    const T* c = tab + k * (2 * NR * NC);
#pragma unroll
    for (int f = 0; f < 2 * NR; ++f) {
        const T* cf = c + f * NC;
        T v = cf[NC - 1];
#pragma unroll
        for (int m = NC - 2; m >= 0; --m) v = fma(v, s, cf[m]);
        if (f < NR) u[f] = v; else w[f - NR] = v;
    }
}
template <int NR, typename T>
__device__ __forceinline__ void rys_roots(T x, T* u, T* w) {
    rys_roots<NR, T>(x, u, w, rys_table_global<NR, T>());
}
// root r alone (0 <= r < NR), the same series as rys_roots: kernels that give each root a
// lane of its own
template <int NR, typename T>
__device__ __forceinline__ void rys_root_one(T x, int r, T& u, T& w, const T* tab) {
    constexpr bool f32 = sizeof(T) == 4;
    if (x >= (T)kRysAsym[NR]) {
        const T isx = rsqrt_nr(x), ix = isx * isx;
        constexpr int o = NR * (NR - 1) / 2;
        if constexpr (f32) {
            u = uno_rys_asym_u_f[o + r] * ix;
            w = uno_rys_asym_w_f[o + r] * isx;
        } else {
            u = uno_rys_asym_u[o + r] * ix;
            w = uno_rys_asym_w[o + r] * isx;
        }
        return;
    }
    constexpr int NC = f32 ? UNO_RYS_NCOEF_F : UNO_RYS_NCOEF;
    const int k = (int)x;
    const T s = (T)2 * (x - (T)k) - (T)1;
    const T* cu = tab + k * (2 * NR * NC) + r * NC;
    const T* cw = cu + NR * NC;
    T vu = cu[NC - 1], vw = cw[NC - 1];
#pragma unroll
    for (int m = NC - 2; m >= 0; --m) {
        vu = fma(vu, s, cu[m]);
        vw = fma(vw, s, cw[m]);
    }
    u = vu;
    w = vw;
}
// the table of NR roots a launch evaluates: the exact tables, except for two roots, where
// the reference-compatible table is the default (same layout and size)
template <int NR, typename T>
__device__ __forceinline__ const T* rys_table_of(const EriArgs& a) {
    if constexpr (NR == 2) {
        if (!a.rys2_exact) {
            if constexpr (sizeof(T) == 4) return uno_rys_ref2_mono_f;
            else return uno_rys_ref2_mono;
        }
    }
    return rys_table_global<NR, T>();
}

// One axis of one root: Rys 2-D recurrence G[e][f] (e <= LA+LB, f <= LC+LD) then the
// binomial transfer to I[a][b][c][d] (a <= LA, b <= LB, c <= LC, d <= LD), flattened.
template <int LA, int LB, int LC, int LD, typename T>
__device__ __forceinline__ void axis_table(T C, T Cp, T B00, T B1, T B1p, T ab, T cd, T* I) {
    constexpr int E = LA + LB, F = LC + LD;
    T G[E + 1][F + 1];
    G[0][0] = 1;
    if constexpr (E > 0) G[1][0] = C;
#pragma unroll
    for (int e = 1; e < E; ++e) G[e + 1][0] = (T)e * B1 * G[e - 1][0] + C * G[e][0];
#pragma unroll
    for (int f = 0; f < F; ++f)
#pragma unroll
        for (int e = 0; e <= E; ++e) {
            T v = Cp * G[e][f];
            if (f > 0) v = fma((T)f * B1p, G[e][f - 1], v);
            if (e > 0) v = fma((T)e * B00, G[e - 1][f], v);
            G[e][f + 1] = v;
        }
    // ket transfer: H[e][c][d] = sum_j C(d,j) cd^j G[e][c+d-j]
    T H[E + 1][LC + 1][LD + 1];
    T cdp[LD + 1];
    cdp[0] = 1;
#pragma unroll
    for (int j = 1; j <= LD; ++j) cdp[j] = cdp[j - 1] * cd;
#pragma unroll
    for (int e = 0; e <= E; ++e)
#pragma unroll
        for (int c = 0; c <= LC; ++c)
#pragma unroll
            for (int d = 0; d <= LD; ++d) {
                T v = 0;
#pragma unroll
                for (int j = 0; j <= d; ++j) v = fma((T)binom_c(d, j) * cdp[j], G[e][c + d - j], v);
                H[e][c][d] = v;
            }
    T abp[LB + 1];
    abp[0] = 1;
#pragma unroll
    for (int i = 1; i <= LB; ++i) abp[i] = abp[i - 1] * ab;
#pragma unroll
    for (int a = 0; a <= LA; ++a)
#pragma unroll
        for (int b = 0; b <= LB; ++b)
#pragma unroll
            for (int c = 0; c <= LC; ++c)
#pragma unroll
                for (int d = 0; d <= LD; ++d) {
                    T v = 0;
#pragma unroll
                    for (int i = 0; i <= b; ++i) v = fma((T)binom_c(b, i) * abp[i], H[a + b - i][c][d], v);
                    I[((a * (LB + 1) + b) * (LC + 1) + c) * (LD + 1) + d] = v;
                }
}

template <int LA, int LB, int LC, int LD>
struct ClassInfo {
    static constexpr int NA = ncart_c(LA), NB = ncart_c(LB), NC = ncart_c(LC), ND = ncart_c(LD);
    static constexpr int NCOMP = NA * NB * NC * ND;
    static constexpr int E = LA + LB, F = LC + LD, NR = (E + F) / 2 + 1;
    static constexpr int TBL = (LA + 1) * (LB + 1) * (LC + 1) * (LD + 1);
    __host__ __device__ static constexpr int tix(int i, int j, int k, int l, int axis) {
        return ((cart_pow(LA, i, axis) * (LB + 1) + cart_pow(LB, j, axis)) * (LC + 1) + cart_pow(LC, k, axis)) *
                   (LD + 1) +
               cart_pow(LD, l, axis);
    }
};

// The matrices one entry's quartets read and write: the launch's own, or for batched
// launches the entry's molecule's slice of every store
struct MolView {
    const double* D0;
    const double* D1;
    const double* D2;
    double* W0;
    double* W1;
    const double* dsh;
    int n, nsh;
};

__device__ __forceinline__ MolView mol_view(const EriArgs& a, int entry) {
    MolView v{a.D0, a.D1, a.D2, a.W0, a.W1, a.dsh, a.n, a.nsh};
    if (a.mols) {
        const MolDesc m = a.mols[a.e_mol[entry]];
        v.D0 = a.D0 + m.moff;
        v.D1 = a.D1 ? a.D1 + m.moff : nullptr;
        v.D2 = a.D2 ? a.D2 + m.moff : nullptr;
        v.W0 = a.W0 + m.moff;
        v.W1 = a.W1 ? a.W1 + m.moff : nullptr;
        v.dsh = a.dsh ? a.dsh + m.dshoff : nullptr;
        v.n = m.n;
        v.nsh = m.nsh;
    }
    return v;
}

// bra pair and first ket pair of entry i
__device__ __forceinline__ int entry_bra(const EriArgs& a, int i) { return a.e_bra ? a.e_bra[i] : a.bra_off + i; }
__device__ __forceinline__ int entry_ket(const EriArgs& a, int i) {
    return a.e_ket ? a.e_ket[i] : a.ket_off + (a.kstart ? a.kstart[i] : 0);
}

// Resolve the quartet handled by linear index g: (bra pair, ket pair), the entry, and the
// product of their Schwarz bounds (negative for explicit lists: no screening)
__device__ __forceinline__ bool quartet_of(const EriArgs& a, long long g, int& pab, int& pcd, double& q12,
                                           int& entry) {
    if (g >= a.nquart) return false;
    if (a.explicit_q) {
        const int2 e = a.explicit_q[g];
        pab = e.x;
        pcd = e.y;
        q12 = -1.0;
        entry = 0;
        return true;
    }
    // the block table brackets the bra of every quartet of this block
    const long long b = g / a.qpb;
    int lo = a.blk[b], hi = a.blk[b + 1];
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.prefix[mid] <= g) lo = mid; else hi = mid - 1;
    }
    const int jk = (int)(g - a.prefix[lo]);
    entry = lo;
    pab = entry_bra(a, lo);
    pcd = entry_ket(a, lo) + jk;
    q12 = a.pairQ[pab] * a.pairQ[pcd];
    return true;
}

__device__ __forceinline__ void atomic_add_f64(double* p, double v) { unsafeAtomicAdd(p, v); }

// Fixed-point accumulation: v = hi 2^-32 + lo 2^-72 + (rounding below 2^-73), hi and lo
// added to two int64 words.  Integer addition is associative, so the sum does not depend
// on the order in which the contributions arrive.  hi holds |W| up to 2^31; lo holds the
// remainders (|lo| <= 2^39 each) of up to 2^24 contributions per entry.
constexpr double kFixHiScale = 0x1p32;
constexpr double kFixLoScale = 0x1p72;
__device__ __forceinline__ void fixed_add(double* p, long long lo_off, double v) {
    const double h = rint(v * kFixHiScale);
    const double r = fma(-h, 1.0 / kFixHiScale, v);   // exact: |r| <= 2^-33, same grid as v
    const long long hi = (long long)h;
    const long long lo = (long long)rint(r * kFixLoScale);
    unsigned long long* w = reinterpret_cast<unsigned long long*>(p);
    if (hi) atomicAdd(w, (unsigned long long)hi);
    if (lo) atomicAdd(w + lo_off, (unsigned long long)lo);
}
// a Fock-matrix contribution of the wave-task and wave kernels
__device__ __forceinline__ void fock_add(const EriArgs& a, double* p, double v) {
    if (a.fix_lo) fixed_add(p, a.fix_lo, v);
    else unsafeAtomicAdd(p, v);
}

// density screen: Qab Qcd max(4|D_ab|, 4|D_cd|, |D_ac|, |D_ad|, |D_bc|, |D_bd|) < tau
__device__ __forceinline__ bool screened(const EriArgs& a, const MolView& v, const PairHdr& P1, const PairHdr& P2,
                                         double q12) {
    if (!v.dsh) return q12 * a.dmax_global < a.tau;
    const int ns = v.nsh;
    const double* m = v.dsh;
    const int s1 = P1.sh1, s2 = P1.sh2, s3 = P2.sh1, s4 = P2.sh2;
    double d = 4.0 * fmax(m[s1 * ns + s2], m[s3 * ns + s4]);
    d = fmax(d, fmax(fmax(m[s1 * ns + s3], m[s1 * ns + s4]), fmax(m[s2 * ns + s3], m[s2 * ns + s4])));
    return q12 * d < a.tau;
}

}  // namespace uno
