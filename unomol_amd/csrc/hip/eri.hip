// ERI + fused J/K kernels (one instantiation per quartet class and digest mode) and the
// host-side launcher.  See eri_core.hpp for the mathematics and execution shapes.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdio>
#include "eri_core.hpp"
#include "kernels.hpp"

namespace uno {

// ---------------------------------------------------------------------------
// contracted (un-normalised) ERI block of one shell quartet in one lane; returns the
// primitive quartets kept by the cut.  Ket primitives outside, bra primitives inside; the
// next bra primitive record is fetched while the current one is evaluated.  tab is the
// Rys table of this class's root count (global or an LDS copy).
// ---------------------------------------------------------------------------
template <int LA, int LB, int LC, int LD, typename T>
__device__ __forceinline__ unsigned eri_block(const EriArgs& a, const PairHdr& P1, const PairHdr& P2, T* e,
                                              const T* tab) {
    using CI = ClassInfo<LA, LB, LC, LD>;
    constexpr int NA = CI::NA, NB = CI::NB, NC = CI::NC, ND = CI::ND, NCOMP = CI::NCOMP, NR = CI::NR;
#pragma unroll
    for (int c = 0; c < NCOMP; ++c) e[c] = 0;
    unsigned nprimq = 0;
    const T ab[3] = {(T)P1.AB[0], (T)P1.AB[1], (T)P1.AB[2]};
    const T cd[3] = {(T)P2.AB[0], (T)P2.AB[1], (T)P2.AB[2]};
    const PrimPair* bp = a.prims + P1.prim0;
    for (int iq = 0; iq < P2.nprim; ++iq) {
        const PrimPair qq = a.prims[P2.prim0 + iq];
        const T q = (T)qq.p, iq_ = (T)qq.ip, sq = (T)qq.s, Kq = (T)qq.K;
        // coordinate differences in double, then rounded to the compute type
        const T QC[3] = {(T)(qq.P[0] - P2.A[0]), (T)(qq.P[1] - P2.A[1]), (T)(qq.P[2] - P2.A[2])};
        PrimPair nx;
        if (P1.nprim > 0) nx = bp[0];
        for (int ip = 0; ip < P1.nprim; ++ip) {
            const PrimPair pp = nx;
            if (ip + 1 < P1.nprim) nx = bp[ip + 1];
            const T p = (T)pp.p, ip_ = (T)pp.ip;
            // rt = (p+q)^(-1/2), it = 1/(p+q)
            const T rt = rsqrt_nr(p + q), it = rt * rt;
            const T bare = (T)kTwoPi25 * (T)pp.s * sq * ip_ * iq_ * rt;
            if (bare < (T)a.primcut) continue;
            ++nprimq;
            const T pref = (T)kTwoPi25 * (T)pp.K * Kq * ip_ * iq_ * rt;
            const T PA[3] = {(T)(pp.P[0] - P1.A[0]), (T)(pp.P[1] - P1.A[1]), (T)(pp.P[2] - P1.A[2])};
            const T PQ[3] = {(T)(pp.P[0] - qq.P[0]), (T)(pp.P[1] - qq.P[1]), (T)(pp.P[2] - qq.P[2])};
            const T x = p * q * it * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
            T u[NR], w[NR];
            rys_roots<NR, T>(x, u, w, tab);
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const T fff = u[r] * it;
                const T B00 = (T)0.5 * fff;
                const T B1 = ((T)0.5 - B00 * q) * ip_;
                const T B1p = ((T)0.5 - B00 * p) * iq_;
                T Ix[CI::TBL], Iy[CI::TBL], Iz[CI::TBL];
                axis_table<LA, LB, LC, LD, T>(PA[0] - q * PQ[0] * fff, QC[0] + p * PQ[0] * fff, B00, B1, B1p, ab[0],
                                              cd[0], Ix);
                axis_table<LA, LB, LC, LD, T>(PA[1] - q * PQ[1] * fff, QC[1] + p * PQ[1] * fff, B00, B1, B1p, ab[1],
                                              cd[1], Iy);
                axis_table<LA, LB, LC, LD, T>(PA[2] - q * PQ[2] * fff, QC[2] + p * PQ[2] * fff, B00, B1, B1p, ab[2],
                                              cd[2], Iz);
                const T wr = w[r] * pref;
#pragma unroll
                for (int i = 0; i < NA; ++i)
#pragma unroll
                    for (int j = 0; j < NB; ++j)
#pragma unroll
                        for (int k = 0; k < NC; ++k)
#pragma unroll
                            for (int l = 0; l < ND; ++l) {
                                const int c = ((i * NB + j) * NC + k) * ND + l;
                                e[c] = fma(wr * Ix[CI::tix(i, j, k, l, 0)], Iy[CI::tix(i, j, k, l, 1)] * Iz[CI::tix(i, j, k, l, 2)], e[c]);
                            }
            }
        }
    }
    return nprimq;
}

// normalisation (AuxFunctions.hpp:49-64) times a factor f
template <int LA, int LB, int LC, int LD, typename T>
__device__ __forceinline__ void normalise_block(T* e, T f) {
    using CI = ClassInfo<LA, LB, LC, LD>;
    constexpr int NA = CI::NA, NB = CI::NB, NC = CI::NC, ND = CI::ND;
#pragma unroll
    for (int i = 0; i < NA; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j)
#pragma unroll
            for (int k = 0; k < NC; ++k)
#pragma unroll
                for (int l = 0; l < ND; ++l)
                    e[((i * NB + j) * NC + k) * ND + l] *=
                        f * (T)(cart_nrm(LA, i) * cart_nrm(LB, j) * cart_nrm(LC, k) * cart_nrm(LD, l));
}

// ---------------------------------------------------------------------------
// one lane per shell quartet of an explicit or prefix-space list (block / Schwarz modes)
// ---------------------------------------------------------------------------
template <int LA, int LB, int LC, int LD, int MODE, typename T>
__global__ __launch_bounds__(256) void k_eri_lane(const EriArgs a) {
    using CI = ClassInfo<LA, LB, LC, LD>;
    constexpr int NA = CI::NA, NB = CI::NB, NC = CI::NC, ND = CI::ND, NCOMP = CI::NCOMP;
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    static_assert(MODE == kDigestBlock || MODE == kDigestSchwarz, "Fock builds use the wave-task kernels");
    int pab, pcd, entry;
    double q12;
    if (!quartet_of(a, g, pab, pcd, q12, entry)) return;
    const PairHdr P1 = a.pairs[pab];
    const PairHdr P2 = a.pairs[pcd];
    T e[NCOMP];
    const unsigned nprimq = eri_block<LA, LB, LC, LD, T>(a, P1, P2, e, rys_table_of<CI::NR, T>(a));
    if (a.counter) {
        unsigned long long* slot = a.counter + 2 * (blockIdx.x & (kCounterSlots - 1));
        atomicAdd(slot, (unsigned long long)nprimq);
        atomicAdd(slot + 1, 1ull);
    }
    normalise_block<LA, LB, LC, LD, T>(e, (T)1);
    if constexpr (MODE == kDigestBlock) {
#pragma unroll
        for (int c = 0; c < NCOMP; ++c) a.out[g * NCOMP + c] = (double)e[c];
    } else if constexpr (MODE == kDigestSchwarz) {
        double m = 0.0;
#pragma unroll
        for (int i = 0; i < NA; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                if constexpr (LA == LC && LB == LD) m = fmax(m, fabs((double)e[((i * NB + j) * NC + i) * ND + j]));
            }
        a.out[pab] = sqrt(m);
    }
}

// Rys tables of up to this many roots are staged in LDS by the Fock-build kernels
constexpr int kLdsTabMaxRoots = 3;
// LDS a row-merged block may take (rows, then the Rys table if it still fits): two blocks
// per compute unit
constexpr size_t kRowLdsBudget = 80 * 1024;
// classes the bra-group row kernels cover (components per shell quartet): groups of two
// bras up to 9, groups of four up to 3 (their component blocks and Coulomb sums for every
// bra of the group stay in registers)
constexpr int kRowsMbMaxComp2 = 9;
constexpr int kRowsMbMaxComp4 = 3;

// ---------------------------------------------------------------------------
// Fock builds: one wave per (bra pair, chunk of up to 64 kets).  The bra is the same in
// every lane, so its records are read at one address, the bra search is one per wave,
// and the Coulomb outputs of the bra are reduced across the wave before one atomic.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// ket Coulomb output (k, l) of ket pair P2 in the one-quartet-per-lane kernels: into the
// pair-major accumulator when the launch has one (consecutive kets of a run are
// consecutive there, so a wave's atomics share cache lines; the host adds it to W after the
// build), else straight into W (both spins)
template <int MODE>
__device__ __forceinline__ void jket_add(const EriArgs& a, const MolView& mv, const PairHdr& P2, int k, int l,
                                         int nd, double v) {
    const long o = (long)(P2.f1 + k) * mv.n + P2.f2 + l;
    fock_add(a, a.jp ? a.jp + P2.jo + k * nd + l : mv.W0 + o, v);
    if (MODE == kDigestUHF && !a.jp) fock_add(a, mv.W1 + o, v);
}

template <int LA, int LB, int LC, int LD, int MODE, typename T>
__device__ __forceinline__ void bra_task(const EriArgs& a, const long long t, const T* tab) {
    using CI = ClassInfo<LA, LB, LC, LD>;
    constexpr int NA = CI::NA, NB = CI::NB, NC = CI::NC, ND = CI::ND, NCOMP = CI::NCOMP;
    const int lane = threadIdx.x & 63;
    // bra of task t: the largest i with wpre[i] <= t, bracketed by the task block table
    int lo = a.tblk[t >> 6], hi = a.tblk[(t >> 6) + 1];
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.wpre[mid] <= t) lo = mid; else hi = mid - 1;
    }
    const int bra = __builtin_amdgcn_readfirstlane(lo);   // entry of this task
    if (a.mol_active && !a.mol_active[a.e_mol[bra]]) return;   // a molecule already converged
    const int chunk = __builtin_amdgcn_readfirstlane((int)(t - a.wpre[bra]));
    const long long cnt = a.prefix[bra + 1] - a.prefix[bra];
    const int jk = chunk * 64 + lane;
    const int pab = __builtin_amdgcn_readfirstlane(entry_bra(a, bra));
    const PairHdr P1 = a.pairs[pab];
    const int pcd = entry_ket(a, bra) + jk;
    const MolView mv = mol_view(a, bra);
    bool active = jk < cnt;
    PairHdr P2;
    if (active) {
        // the screen reads only the ket's shell indices; the whole record for survivors
        active = !screened(a, mv, P1, a.pairs[pcd], a.pairQ[pab] * a.pairQ[pcd]);
        if (active) P2 = a.pairs[pcd];
    }
    T e[NCOMP];
    unsigned nprimq = 0;
    if (active) {
        nprimq = eri_block<LA, LB, LC, LD, T>(a, P1, P2, e, tab);
        const T f = (P1.same ? (T)0.5 : (T)1) * (P2.same ? (T)0.5 : (T)1) * (pab == pcd ? (T)0.5 : (T)1);
        normalise_block<LA, LB, LC, LD, T>(e, f);
    } else {
#pragma unroll
        for (int c = 0; c < NCOMP; ++c) e[c] = 0;
    }
    if (a.counter) {
        // evaluated primitive and shell quartets of the wave, into one of 64 slot pairs
        int np = (int)nprimq;
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) np += __shfl_xor(np, m, 64);
        const unsigned long long na = __popcll(__ballot(active));
        if (lane == 0) {
            unsigned long long* slot = a.counter + 2 * (blockIdx.x & (kCounterSlots - 1));
            atomicAdd(slot, (unsigned long long)np);
            atomicAdd(slot + 1, na);
        }
    }
    const int n = mv.n;
    const int fa = P1.f1, fb = P1.f2;
    constexpr T jf = (MODE == kDigestRHF) ? 4 : 2;
    // Coulomb, bra outputs: W_ab += jf sum_cd e D_cd, summed over the wave
    {
        T dcd[NC * ND];
        if (active) {
            const int fc = P2.f1, fd = P2.f2;
#pragma unroll
            for (int k = 0; k < NC; ++k)
#pragma unroll
                for (int l = 0; l < ND; ++l) dcd[k * ND + l] = (T)mv.D0[(long)(fc + k) * n + fd + l];
        } else {
#pragma unroll
            for (int kl = 0; kl < NC * ND; ++kl) dcd[kl] = 0;
        }
#pragma unroll
        for (int i = 0; i < NA; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                T s = 0;
#pragma unroll
                for (int kl = 0; kl < NC * ND; ++kl) s = fma(e[(i * NB + j) * NC * ND + kl], dcd[kl], s);
                const double tot = wave_sum((double)(jf * s));
                if (lane == 0) {
                    fock_add(a, mv.W0 + (long)(fa + i) * n + fb + j, tot);
                    if (MODE == kDigestUHF) fock_add(a, mv.W1 + (long)(fa + i) * n + fb + j, tot);
                }
            }
    }
    if (!active) return;
    const int fc = P2.f1, fd = P2.f2;
    // Coulomb, ket outputs: W_cd += jf sum_ab e D_ab
    {
        T dab[NA * NB];
#pragma unroll
        for (int i = 0; i < NA; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j) dab[i * NB + j] = (T)mv.D0[(long)(fa + i) * n + fb + j];
#pragma unroll
        for (int kl = 0; kl < NC * ND; ++kl) {
            T s = 0;
#pragma unroll
            for (int ij = 0; ij < NA * NB; ++ij) s = fma(e[ij * NC * ND + kl], dab[ij], s);
            jket_add<MODE>(a, mv, P2, kl / ND, kl % ND, ND, (double)(jf * s));
        }
    }
    // exchange, once per spin density
#pragma unroll
    for (int sp = 0; sp < (MODE == kDigestUHF ? 2 : 1); ++sp) {
        const double* DK = (MODE == kDigestUHF) ? (sp == 0 ? mv.D1 : mv.D2) : mv.D0;
        double* W = (sp == 0) ? mv.W0 : mv.W1;
        T dbd[NB * ND], dbc[NB * NC], dad[NA * ND], dac[NA * NC];
#pragma unroll
        for (int j = 0; j < NB; ++j) {
#pragma unroll
            for (int l = 0; l < ND; ++l) dbd[j * ND + l] = (T)DK[(long)(fb + j) * n + fd + l];
#pragma unroll
            for (int k = 0; k < NC; ++k) dbc[j * NC + k] = (T)DK[(long)(fb + j) * n + fc + k];
        }
#pragma unroll
        for (int i = 0; i < NA; ++i) {
#pragma unroll
            for (int l = 0; l < ND; ++l) dad[i * ND + l] = (T)DK[(long)(fa + i) * n + fd + l];
#pragma unroll
            for (int k = 0; k < NC; ++k) dac[i * NC + k] = (T)DK[(long)(fa + i) * n + fc + k];
        }
#pragma unroll
        for (int i = 0; i < NA; ++i)
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                T s = 0;
#pragma unroll
                for (int j = 0; j < NB; ++j)
#pragma unroll
                    for (int l = 0; l < ND; ++l) s = fma(e[((i * NB + j) * NC + k) * ND + l], dbd[j * ND + l], s);
                fock_add(a, W + (long)(fa + i) * n + fc + k, (double)(-s));
            }
#pragma unroll
        for (int i = 0; i < NA; ++i)
#pragma unroll
            for (int l = 0; l < ND; ++l) {
                T s = 0;
#pragma unroll
                for (int j = 0; j < NB; ++j)
#pragma unroll
                    for (int k = 0; k < NC; ++k) s = fma(e[((i * NB + j) * NC + k) * ND + l], dbc[j * NC + k], s);
                fock_add(a, W + (long)(fa + i) * n + fd + l, (double)(-s));
            }
#pragma unroll
        for (int j = 0; j < NB; ++j)
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                T s = 0;
#pragma unroll
                for (int i = 0; i < NA; ++i)
#pragma unroll
                    for (int l = 0; l < ND; ++l) s = fma(e[((i * NB + j) * NC + k) * ND + l], dad[i * ND + l], s);
                fock_add(a, W + (long)(fb + j) * n + fc + k, (double)(-s));
            }
#pragma unroll
        for (int j = 0; j < NB; ++j)
#pragma unroll
            for (int l = 0; l < ND; ++l) {
                T s = 0;
#pragma unroll
                for (int i = 0; i < NA; ++i)
#pragma unroll
                    for (int k = 0; k < NC; ++k) s = fma(e[((i * NB + j) * NC + k) * ND + l], dac[i * NC + k], s);
                fock_add(a, W + (long)(fb + j) * n + fd + l, (double)(-s));
            }
    }}

template <int LA, int LB, int LC, int LD, int MODE, typename T>
__global__ __launch_bounds__(256) void k_eri_bra(const EriArgs a) {
    using CI = ClassInfo<LA, LB, LC, LD>;
    const int lane = threadIdx.x & 63;
    // the Rys table of this root count, copied to LDS by the block (up to 35 KB)
    constexpr int NR = CI::NR;
    constexpr bool kLdsTab = NR <= kLdsTabMaxRoots;
    __shared__ T tab_lds[kLdsTab ? rys_table_size<NR, T>() : 1];
    const T* tab = rys_table_of<NR, T>(a);
    if constexpr (kLdsTab) {
        for (int i = threadIdx.x; i < rys_table_size<NR, T>(); i += blockDim.x) tab_lds[i] = tab[i];
        __syncthreads();
        tab = tab_lds;
    }
    if (a.queue) {
        // work queue: each wave takes the next task until none is left (every wave ends
        // after at most one grab past the end)
        for (;;) {
            unsigned long long g = 0;
            if (lane == 0) g = atomicAdd(a.queue, 1ULL);
            const long long t = (long long)__shfl(g, 0, 64);
            if (t >= a.ntask) break;
            bra_task<LA, LB, LC, LD, MODE, T>(a, t, tab);
        }
        return;
    }
    const long long t = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (t >= a.ntask) return;   // whole wave
    bra_task<LA, LB, LC, LD, MODE, T>(a, t, tab);
}

// ---------------------------------------------------------------------------
// Fock builds, bras with many kets: one block per (bra, segment of rseg kets), kRowThreads kets
// per pass.  The exchange contributions all land in the bra's rows of W (rows fa.., fb..),
// so the block merges them in LDS (one row of n doubles per bra function and spin) and
// adds the non-zero entries at the end with row-contiguous atomics; the bra's Coulomb
// outputs are reduced over the block.  Only the ket's Coulomb outputs are added per
// quartet.  Dynamic LDS: (NA + NB) x spins x ncols doubles (ncols: the functions the ket
// run touches), or with FB false only the NA rows of the first bra shell (the second
// shell's rows then take their atomics directly); TL: the Rys table is copied to LDS as
// well (otherwise read from global memory).
// ---------------------------------------------------------------------------
template <int LA, int LB, int MODE, bool FB>
__host__ __device__ constexpr int row_count() {
    return (ncart_c(LA) + (FB ? ncart_c(LB) : 0)) * (MODE == kDigestUHF ? 2 : 1);
}
// rows of the staged exchange density (a.dlds): every bra function, per spin density
template <int LA, int LB, int MODE>
__host__ __device__ constexpr int dens_row_count() {
    return (ncart_c(LA) + ncart_c(LB)) * (MODE == kDigestUHF ? 2 : 1);
}

// threads of a row-merged block (512 was measured slower: fewer blocks for the bras with
// few kets, and the same per-wave time)
constexpr int kRowThreads = 256;
constexpr int kRowWaves = kRowThreads / 64;

template <int LA, int LB, int LC, int LD, int MODE, typename T, bool TL, bool FB>
__global__ __launch_bounds__(kRowThreads) void k_eri_rows(const EriArgs a) {
    using CI = ClassInfo<LA, LB, LC, LD>;
    constexpr int NA = CI::NA, NB = CI::NB, NC = CI::NC, ND = CI::ND, NCOMP = CI::NCOMP, NR = CI::NR;
    constexpr int NAB = NA + (FB ? NB : 0), NROW = row_count<LA, LB, MODE, FB>();
    constexpr int NSP = MODE == kDigestUHF ? 2 : 1;
    extern __shared__ double rows[];
    __shared__ T tab_lds[TL ? rys_table_size<NR, T>() : 1];
    __shared__ double jred[kRowWaves][NA * NB];
    __shared__ unsigned long long cred[kRowWaves][2];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // the block's bra (uniform): largest r with rpre[r] <= blockIdx.x
    const long long b = blockIdx.x;
    int lo = 0, hi = a.nrbra - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.rpre[mid] <= b) lo = mid; else hi = mid - 1;
    }
    const int entry = __builtin_amdgcn_readfirstlane(a.rbra[lo]);
    const long long seg = b - a.rpre[lo];
    const long long cnt = a.prefix[entry + 1] - a.prefix[entry];
    const long long k0 = seg * a.rseg, k1 = min(cnt, k0 + (long long)a.rseg);
    const MolView mv = mol_view(a, entry);
    const int n = mv.n, nc = a.ncols;
    const int pab = __builtin_amdgcn_readfirstlane(entry_bra(a, entry));
    const PairHdr P1 = a.pairs[pab];
    const int ket0 = entry_ket(a, entry);
    const int fa = P1.f1, fb = P1.f2;
    for (int i = threadIdx.x; i < NROW * nc; i += blockDim.x) rows[i] = 0.0;
    // with a.dlds, the exchange density rows of the bra (rows fa.., fb.. of each spin
    // density, at the same columns) follow the accumulator rows: the exchange digestion
    // then reads LDS instead of gathering from global memory
    constexpr int NDR = NA + NB;
    const bool dl = a.dlds != 0;
    double* dlr = rows + NROW * nc;
    if (dl) {
        for (int i = threadIdx.x; i < NDR * NSP * nc; i += blockDim.x) {
            const int r = i / nc, col = i - r * nc;
            const int sp = r / NDR, rr = r - sp * NDR;
            const double* DK = (MODE == kDigestUHF) ? (sp == 0 ? mv.D1 : mv.D2) : mv.D0;
            const int grow = rr < NA ? fa + rr : fb + rr - NA;
            dlr[i] = DK[(long)grow * n + a.colinv[col]];
        }
    }
    const T* tab = rys_table_of<NR, T>(a);
    if constexpr (TL) {
        for (int i = threadIdx.x; i < rys_table_size<NR, T>(); i += blockDim.x) tab_lds[i] = tab[i];
        tab = tab_lds;
    }
    __syncthreads();
    constexpr T jf = (MODE == kDigestRHF) ? 4 : 2;
    double jacc[NA * NB];
#pragma unroll
    for (int ij = 0; ij < NA * NB; ++ij) jacc[ij] = 0.0;
    T dab[NA * NB];
#pragma unroll
    for (int i = 0; i < NA; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) dab[i * NB + j] = (T)mv.D0[(long)(fa + i) * n + fb + j];
    unsigned long long nprim = 0, nshell = 0;
    for (long long base = k0; base < k1; base += blockDim.x) {
        const long long jk = base + threadIdx.x;
        bool active = jk < k1;
        PairHdr P2;
        const int pcd = ket0 + (int)jk;
        if (active) {
            // the screen reads only the ket's shell indices; the whole record for survivors
            active = !screened(a, mv, P1, a.pairs[pcd], a.pairQ[pab] * a.pairQ[pcd]);
            if (active) P2 = a.pairs[pcd];
        }
        if (!active) continue;
        T e[NCOMP];
        nprim += eri_block<LA, LB, LC, LD, T>(a, P1, P2, e, tab);
        ++nshell;
        const T f = (P1.same ? (T)0.5 : (T)1) * (P2.same ? (T)0.5 : (T)1) * (pab == pcd ? (T)0.5 : (T)1);
        normalise_block<LA, LB, LC, LD, T>(e, f);
        const int fc = P2.f1, fd = P2.f2;
        const int lc = a.colmap[fc], ld = a.colmap[fd];   // their LDS row columns
        // Coulomb, bra outputs: kept per thread
        {
            T dcd[NC * ND];
#pragma unroll
            for (int k = 0; k < NC; ++k)
#pragma unroll
                for (int l = 0; l < ND; ++l) dcd[k * ND + l] = (T)mv.D0[(long)(fc + k) * n + fd + l];
#pragma unroll
            for (int ij = 0; ij < NA * NB; ++ij) {
                T s = 0;
#pragma unroll
                for (int kl = 0; kl < NC * ND; ++kl) s = fma(e[ij * NC * ND + kl], dcd[kl], s);
                jacc[ij] += (double)(jf * s);
            }
        }
        // Coulomb, ket outputs: one atomic each
#pragma unroll
        for (int kl = 0; kl < NC * ND; ++kl) {
            T s = 0;
#pragma unroll
            for (int ij = 0; ij < NA * NB; ++ij) s = fma(e[ij * NC * ND + kl], dab[ij], s);
            jket_add<MODE>(a, mv, P2, kl / ND, kl % ND, ND, (double)(jf * s));
        }
        // exchange into the LDS rows: row i (a function i), row NA + j (b function j, FB)
#pragma unroll
        for (int sp = 0; sp < NSP; ++sp) {
            const double* DK = (MODE == kDigestUHF) ? (sp == 0 ? mv.D1 : mv.D2) : mv.D0;
            double* R = rows + sp * NAB * nc;
            double* WB = (MODE == kDigestUHF && sp) ? mv.W1 : mv.W0;
            T dbd[NB * ND], dbc[NB * NC], dad[NA * ND], dac[NA * NC];
            if (dl) {
                const double* DR = dlr + sp * NDR * nc;
#pragma unroll
                for (int j = 0; j < NB; ++j) {
#pragma unroll
                    for (int l = 0; l < ND; ++l) dbd[j * ND + l] = (T)DR[(NA + j) * nc + ld + l];
#pragma unroll
                    for (int k = 0; k < NC; ++k) dbc[j * NC + k] = (T)DR[(NA + j) * nc + lc + k];
                }
#pragma unroll
                for (int i = 0; i < NA; ++i) {
#pragma unroll
                    for (int l = 0; l < ND; ++l) dad[i * ND + l] = (T)DR[i * nc + ld + l];
#pragma unroll
                    for (int k = 0; k < NC; ++k) dac[i * NC + k] = (T)DR[i * nc + lc + k];
                }
            } else {
#pragma unroll
                for (int j = 0; j < NB; ++j) {
#pragma unroll
                    for (int l = 0; l < ND; ++l) dbd[j * ND + l] = (T)DK[(long)(fb + j) * n + fd + l];
#pragma unroll
                    for (int k = 0; k < NC; ++k) dbc[j * NC + k] = (T)DK[(long)(fb + j) * n + fc + k];
                }
#pragma unroll
                for (int i = 0; i < NA; ++i) {
#pragma unroll
                    for (int l = 0; l < ND; ++l) dad[i * ND + l] = (T)DK[(long)(fa + i) * n + fd + l];
#pragma unroll
                    for (int k = 0; k < NC; ++k) dac[i * NC + k] = (T)DK[(long)(fa + i) * n + fc + k];
                }
            }
#pragma unroll
            for (int i = 0; i < NA; ++i) {
#pragma unroll
                for (int k = 0; k < NC; ++k) {
                    T s = 0;
#pragma unroll
                    for (int j = 0; j < NB; ++j)
#pragma unroll
                        for (int l = 0; l < ND; ++l) s = fma(e[((i * NB + j) * NC + k) * ND + l], dbd[j * ND + l], s);
                    atomicAdd(R + i * nc + lc + k, (double)(-s));
                }
#pragma unroll
                for (int l = 0; l < ND; ++l) {
                    T s = 0;
#pragma unroll
                    for (int j = 0; j < NB; ++j)
#pragma unroll
                        for (int k = 0; k < NC; ++k) s = fma(e[((i * NB + j) * NC + k) * ND + l], dbc[j * NC + k], s);
                    atomicAdd(R + i * nc + ld + l, (double)(-s));
                }
            }
#pragma unroll
            for (int j = 0; j < NB; ++j) {
#pragma unroll
                for (int k = 0; k < NC; ++k) {
                    T s = 0;
#pragma unroll
                    for (int i = 0; i < NA; ++i)
#pragma unroll
                        for (int l = 0; l < ND; ++l) s = fma(e[((i * NB + j) * NC + k) * ND + l], dad[i * ND + l], s);
                    if constexpr (FB) atomicAdd(R + (NA + j) * nc + lc + k, (double)(-s));
                    else atomic_add_f64(WB + (long)(fb + j) * n + fc + k, (double)(-s));
                }
#pragma unroll
                for (int l = 0; l < ND; ++l) {
                    T s = 0;
#pragma unroll
                    for (int i = 0; i < NA; ++i)
#pragma unroll
                        for (int k = 0; k < NC; ++k) s = fma(e[((i * NB + j) * NC + k) * ND + l], dac[i * NC + k], s);
                    if constexpr (FB) atomicAdd(R + (NA + j) * nc + ld + l, (double)(-s));
                    else atomic_add_f64(WB + (long)(fb + j) * n + fd + l, (double)(-s));
                }
            }
        }
    }
    // bra Coulomb outputs: wave sums, then the four waves
#pragma unroll
    for (int ij = 0; ij < NA * NB; ++ij) {
        const double v = wave_sum(jacc[ij]);
        if (lane == 0) jred[wave][ij] = v;
    }
    if (a.counter) {
        int np = (int)nprim, ns = (int)nshell;
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
            np += __shfl_xor(np, m, 64);
            ns += __shfl_xor(ns, m, 64);
        }
        if (lane == 0) {
            cred[wave][0] = (unsigned long long)np;
            cred[wave][1] = (unsigned long long)ns;
        }
    }
    __syncthreads();
    if (threadIdx.x < NA * NB) {
        const int i = threadIdx.x / NB, j = threadIdx.x % NB;
        double v = 0.0;
#pragma unroll
        for (int w = 0; w < kRowWaves; ++w) v += jred[w][threadIdx.x];
        atomic_add_f64(mv.W0 + (long)(fa + i) * n + fb + j, v);
        if (MODE == kDigestUHF) atomic_add_f64(mv.W1 + (long)(fa + i) * n + fb + j, v);
    }
    if (a.counter && threadIdx.x == 0) {
        unsigned long long* slot = a.counter + 2 * (blockIdx.x & (kCounterSlots - 1));
        unsigned long long np = 0, ns = 0;
        for (int w = 0; w < kRowWaves; ++w) {
            np += cred[w][0];
            ns += cred[w][1];
        }
        atomicAdd(slot, np);
        atomicAdd(slot + 1, ns);
    }
    // merged exchange rows: consecutive threads take consecutive columns of a row
#pragma unroll
    for (int r = 0; r < NROW; ++r) {
        const int sp = r / NAB, rr = r % NAB;
        const int grow = rr < NA ? fa + rr : fb + rr - NA;
        double* W = (MODE == kDigestUHF && sp) ? mv.W1 : mv.W0;
        for (int col = threadIdx.x; col < nc; col += blockDim.x) {
            const double v = rows[r * nc + col];
            if (v != 0.0) atomic_add_f64(W + (long)grow * n + a.colinv[col], v);
        }
    }
}

// ---------------------------------------------------------------------------
// Fock builds, groups of bras with many kets: one block per (group of up to NBR bras of a
// space, segment of rseg kets).  The kets of an FP64 space are the same run for every bra,
// so each ket's pair record and primitive-pair records are loaded once for the whole group
// (ket primitives outside, the group's bras and their primitives inside) instead of once
// per bra: the row kernels above wait on those loads.  Digestion per bra as in k_eri_rows
// with every bra row merged in LDS (NBR x rows x ncols doubles) and the Rys table in LDS.
// ---------------------------------------------------------------------------
// one primitive quartet into the component block e (the body of eri_block's loops); false
// when the bare-prefactor cut drops it
template <int LA, int LB, int LC, int LD, typename T>
__device__ __forceinline__ bool prim_quartet(const EriArgs& a, const double* A1, const PrimPair& pp, const PrimPair& qq,
                                             const T* QC, const T* ab, const T* cd, const T* tab, T* e) {
    using CI = ClassInfo<LA, LB, LC, LD>;
    constexpr int NA = CI::NA, NB = CI::NB, NC = CI::NC, ND = CI::ND, NR = CI::NR;
    const T q = (T)qq.p, iq_ = (T)qq.ip;
    const T p = (T)pp.p, ip_ = (T)pp.ip;
    const T rt = rsqrt_nr(p + q), it = rt * rt;
    const T bare = (T)kTwoPi25 * (T)pp.s * (T)qq.s * ip_ * iq_ * rt;
    if (bare < (T)a.primcut) return false;
    const T pref = (T)kTwoPi25 * (T)pp.K * (T)qq.K * ip_ * iq_ * rt;
    const T PA[3] = {(T)(pp.P[0] - A1[0]), (T)(pp.P[1] - A1[1]), (T)(pp.P[2] - A1[2])};
    const T PQ[3] = {(T)(pp.P[0] - qq.P[0]), (T)(pp.P[1] - qq.P[1]), (T)(pp.P[2] - qq.P[2])};
    const T x = p * q * it * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
    T u[NR], w[NR];
    rys_roots<NR, T>(x, u, w, tab);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const T fff = u[r] * it;
        const T B00 = (T)0.5 * fff;
        const T B1 = ((T)0.5 - B00 * q) * ip_;
        const T B1p = ((T)0.5 - B00 * p) * iq_;
        T Ix[CI::TBL], Iy[CI::TBL], Iz[CI::TBL];
        axis_table<LA, LB, LC, LD, T>(PA[0] - q * PQ[0] * fff, QC[0] + p * PQ[0] * fff, B00, B1, B1p, ab[0], cd[0], Ix);
        axis_table<LA, LB, LC, LD, T>(PA[1] - q * PQ[1] * fff, QC[1] + p * PQ[1] * fff, B00, B1, B1p, ab[1], cd[1], Iy);
        axis_table<LA, LB, LC, LD, T>(PA[2] - q * PQ[2] * fff, QC[2] + p * PQ[2] * fff, B00, B1, B1p, ab[2], cd[2], Iz);
        const T wr = w[r] * pref;
#pragma unroll
        for (int i = 0; i < NA; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j)
#pragma unroll
                for (int k = 0; k < NC; ++k)
#pragma unroll
                    for (int l = 0; l < ND; ++l) {
                        const int c = ((i * NB + j) * NC + k) * ND + l;
                        e[c] = fma(wr * Ix[CI::tix(i, j, k, l, 0)], Iy[CI::tix(i, j, k, l, 1)] * Iz[CI::tix(i, j, k, l, 2)], e[c]);
                    }
    }
    return true;
}

template <int LA, int LB, int MODE>
__host__ __device__ constexpr int rows_mb_count() {
    return (ncart_c(LA) + ncart_c(LB)) * (MODE == kDigestUHF ? 2 : 1);
}

// waves per SIMD the bra-group kernels' registers must allow: the single-bra kernel's
// occupancy for the smallest classes, one wave less for the others
template <int LA, int LB, int LC, int LD, int NBR>
constexpr int rows_mb_min_waves() {
    constexpr int nc = ClassInfo<LA, LB, LC, LD>::NCOMP;
    return nc == 1 ? (NBR == 2 ? 5 : 4) : (nc == 3 ? (NBR == 2 ? 4 : 3) : 2);
}
template <int LA, int LB, int LC, int LD, int MODE, typename T, int NBR>
__global__ __launch_bounds__(kRowThreads, (rows_mb_min_waves<LA, LB, LC, LD, NBR>())) void k_eri_rows_mb(const EriArgs a) {
    using CI = ClassInfo<LA, LB, LC, LD>;
    constexpr int NA = CI::NA, NB = CI::NB, NC = CI::NC, ND = CI::ND, NCOMP = CI::NCOMP, NR = CI::NR;
    constexpr int NAB = NA + NB, NSP = MODE == kDigestUHF ? 2 : 1, NROW = NAB * NSP, NANB = NA * NB;
    extern __shared__ double rows[];   // bra b's row r at rows[(b * NROW + r) * nc]
    __shared__ T tab_lds[rys_table_size<NR, T>()];
    __shared__ double jred[kRowWaves][NBR * NANB];
    __shared__ unsigned long long cred[kRowWaves][2];
    __shared__ PairHdr sP1[NBR];
    __shared__ T sdab[NBR][NANB];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // the block's group (uniform): largest g with rpre[g] <= blockIdx.x
    const long long blk = blockIdx.x;
    const int ngrp = (a.nrbra + NBR - 1) / NBR;
    int lo = 0, hi = ngrp - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.rpre[mid] <= blk) lo = mid; else hi = mid - 1;
    }
    const int g0 = __builtin_amdgcn_readfirstlane(lo * NBR);
    const int nb = min(NBR, a.nrbra - g0);
    long long cnt[NBR];
    int pab[NBR];
    long long cmax = 0;
#pragma unroll
    for (int b = 0; b < NBR; ++b) {
        const int e = a.rbra[g0 + (b < nb ? b : 0)];
        cnt[b] = b < nb ? a.prefix[e + 1] - a.prefix[e] : 0;
        pab[b] = __builtin_amdgcn_readfirstlane(entry_bra(a, e));
        cmax = max(cmax, cnt[b]);
    }
    const int entry0 = a.rbra[g0];
    const long long seg = blk - a.rpre[lo];
    const long long k0 = seg * a.rseg, k1 = min(cmax, k0 + (long long)a.rseg);
    const MolView mv = mol_view(a, entry0);
    const int n = mv.n, nc = a.ncols;
    const int ket0 = entry_ket(a, entry0);
    if (threadIdx.x < nb) sP1[threadIdx.x] = a.pairs[entry_bra(a, a.rbra[g0 + threadIdx.x])];
    for (int i = threadIdx.x; i < NBR * NROW * nc; i += blockDim.x) rows[i] = 0.0;
    const T* tab = rys_table_of<NR, T>(a);
    for (int i = threadIdx.x; i < rys_table_size<NR, T>(); i += blockDim.x) tab_lds[i] = tab[i];
    tab = tab_lds;
    __syncthreads();
    for (int t = threadIdx.x; t < nb * NANB; t += blockDim.x) {
        const int b = t / NANB, ij = t - b * NANB;
        sdab[b][ij] = (T)mv.D0[(long)(sP1[b].f1 + ij / NB) * n + sP1[b].f2 + ij % NB];
    }
    __syncthreads();
    constexpr T jf = (MODE == kDigestRHF) ? 4 : 2;
    // the bra Coulomb sums: reduced over the wave per ket pass and bra (in converged control
    // flow), lane 0 keeping the wave's running totals (registers per bra, not per lane)
    for (int t = threadIdx.x; t < kRowWaves * NBR * NANB; t += blockDim.x) (&jred[0][0])[t] = 0.0;
    __syncthreads();
    unsigned long long nprim = 0, nshell = 0;
    for (long long base = k0; base < k1; base += blockDim.x) {
        const long long jk = base + threadIdx.x;
        bool act[NBR];
        bool any = false;
        // lanes past the segment read a valid record (the segment's first ket), unused: the
        // wave stays converged for its Coulomb sums
        const int pcd = ket0 + (int)(jk < k1 ? jk : k0);
        const PairHdr P2 = a.pairs[pcd];
        if (jk < k1) {
            const double qcd = a.pairQ[pcd];
#pragma unroll
            for (int b = 0; b < NBR; ++b) {
                act[b] = b < nb && jk < cnt[b] && !screened(a, mv, sP1[b], P2, a.pairQ[pab[b]] * qcd);
                any = any || act[b];
            }
        } else {
#pragma unroll
            for (int b = 0; b < NBR; ++b) act[b] = false;
        }
        if (__ballot(any) == 0ULL) continue;   // the whole wave: the Coulomb sums below are wave-wide
        T e[NBR][NCOMP];
#pragma unroll
        for (int b = 0; b < NBR; ++b)
#pragma unroll
            for (int c = 0; c < NCOMP; ++c) e[b][c] = 0;
        const T cd[3] = {(T)P2.AB[0], (T)P2.AB[1], (T)P2.AB[2]};
        // ket primitives outside: each record is loaded once for all the group's bras
        const int npq = any ? P2.nprim : 0;
        for (int iq = 0; iq < npq; ++iq) {
            const PrimPair qq = a.prims[P2.prim0 + iq];
            const T QC[3] = {(T)(qq.P[0] - P2.A[0]), (T)(qq.P[1] - P2.A[1]), (T)(qq.P[2] - P2.A[2])};
#pragma unroll
            for (int b = 0; b < NBR; ++b) {
                if (!act[b]) continue;
                const int p0 = __builtin_amdgcn_readfirstlane(sP1[b].prim0);
                const int np1 = __builtin_amdgcn_readfirstlane(sP1[b].nprim);
                const double A1[3] = {sP1[b].A[0], sP1[b].A[1], sP1[b].A[2]};
                const T ab[3] = {(T)sP1[b].AB[0], (T)sP1[b].AB[1], (T)sP1[b].AB[2]};
                for (int ip = 0; ip < np1; ++ip)
                    if (prim_quartet<LA, LB, LC, LD, T>(a, A1, a.prims[p0 + ip], qq, QC, ab, cd, tab, e[b])) ++nprim;
            }
        }
        const int fc = P2.f1, fd = P2.f2;
        const int lc = a.colmap[fc], ld = a.colmap[fd];
        T dcd[NC * ND];
#pragma unroll
        for (int k = 0; k < NC; ++k)
#pragma unroll
            for (int l = 0; l < ND; ++l) dcd[k * ND + l] = (T)mv.D0[(long)(fc + k) * n + fd + l];
#pragma unroll
        for (int b = 0; b < NBR; ++b) {
            // Coulomb, bra outputs: this pass's sum over the wave (every lane takes part)
            double jb[NANB];
#pragma unroll
            for (int ij = 0; ij < NANB; ++ij) jb[ij] = 0.0;
            if (act[b]) {
                const T f = (sP1[b].same ? (T)0.5 : (T)1) * (P2.same ? (T)0.5 : (T)1) * (pab[b] == pcd ? (T)0.5 : (T)1);
                normalise_block<LA, LB, LC, LD, T>(e[b], f);
#pragma unroll
                for (int ij = 0; ij < NANB; ++ij) {
                    T s = 0;
#pragma unroll
                    for (int kl = 0; kl < NC * ND; ++kl) s = fma(e[b][ij * NC * ND + kl], dcd[kl], s);
                    jb[ij] = (double)(jf * s);
                }
            }
#pragma unroll
            for (int ij = 0; ij < NANB; ++ij) {
                const double v = wave_sum(jb[ij]);
                if (lane == 0) jred[wave][b * NANB + ij] += v;
            }
            if (!act[b]) continue;
            ++nshell;
            const int fa = sP1[b].f1, fb = sP1[b].f2;
            // Coulomb, ket outputs: one atomic each
#pragma unroll
            for (int kl = 0; kl < NC * ND; ++kl) {
                T s = 0;
#pragma unroll
                for (int ij = 0; ij < NANB; ++ij) s = fma(e[b][ij * NC * ND + kl], sdab[b][ij], s);
                jket_add<MODE>(a, mv, P2, kl / ND, kl % ND, ND, (double)(jf * s));
            }
            // exchange into bra b's LDS rows: row i (a function i), row NA + j (b function j)
#pragma unroll
            for (int sp = 0; sp < NSP; ++sp) {
                const double* DK = (MODE == kDigestUHF) ? (sp == 0 ? mv.D1 : mv.D2) : mv.D0;
                double* R = rows + (b * NROW + sp * NAB) * nc;
                T dbd[NB * ND], dbc[NB * NC], dad[NA * ND], dac[NA * NC];
#pragma unroll
                for (int j = 0; j < NB; ++j) {
#pragma unroll
                    for (int l = 0; l < ND; ++l) dbd[j * ND + l] = (T)DK[(long)(fb + j) * n + fd + l];
#pragma unroll
                    for (int k = 0; k < NC; ++k) dbc[j * NC + k] = (T)DK[(long)(fb + j) * n + fc + k];
                }
#pragma unroll
                for (int i = 0; i < NA; ++i) {
#pragma unroll
                    for (int l = 0; l < ND; ++l) dad[i * ND + l] = (T)DK[(long)(fa + i) * n + fd + l];
#pragma unroll
                    for (int k = 0; k < NC; ++k) dac[i * NC + k] = (T)DK[(long)(fa + i) * n + fc + k];
                }
#pragma unroll
                for (int i = 0; i < NA; ++i) {
#pragma unroll
                    for (int k = 0; k < NC; ++k) {
                        T s = 0;
#pragma unroll
                        for (int j = 0; j < NB; ++j)
#pragma unroll
                            for (int l = 0; l < ND; ++l) s = fma(e[b][((i * NB + j) * NC + k) * ND + l], dbd[j * ND + l], s);
                        atomicAdd(R + i * nc + lc + k, (double)(-s));
                    }
#pragma unroll
                    for (int l = 0; l < ND; ++l) {
                        T s = 0;
#pragma unroll
                        for (int j = 0; j < NB; ++j)
#pragma unroll
                            for (int k = 0; k < NC; ++k) s = fma(e[b][((i * NB + j) * NC + k) * ND + l], dbc[j * NC + k], s);
                        atomicAdd(R + i * nc + ld + l, (double)(-s));
                    }
                }
#pragma unroll
                for (int j = 0; j < NB; ++j) {
#pragma unroll
                    for (int k = 0; k < NC; ++k) {
                        T s = 0;
#pragma unroll
                        for (int i = 0; i < NA; ++i)
#pragma unroll
                            for (int l = 0; l < ND; ++l) s = fma(e[b][((i * NB + j) * NC + k) * ND + l], dad[i * ND + l], s);
                        atomicAdd(R + (NA + j) * nc + lc + k, (double)(-s));
                    }
#pragma unroll
                    for (int l = 0; l < ND; ++l) {
                        T s = 0;
#pragma unroll
                        for (int i = 0; i < NA; ++i)
#pragma unroll
                            for (int k = 0; k < NC; ++k) s = fma(e[b][((i * NB + j) * NC + k) * ND + l], dac[i * NC + k], s);
                        atomicAdd(R + (NA + j) * nc + ld + l, (double)(-s));
                    }
                }
            }
        }
    }
    // bra Coulomb outputs: the waves' totals (jred), summed over the four waves below
    if (a.counter) {
        int np = (int)nprim, ns = (int)nshell;
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
            np += __shfl_xor(np, m, 64);
            ns += __shfl_xor(ns, m, 64);
        }
        if (lane == 0) {
            cred[wave][0] = (unsigned long long)np;
            cred[wave][1] = (unsigned long long)ns;
        }
    }
    __syncthreads();
    if (threadIdx.x < nb * NANB) {
        const int b = threadIdx.x / NANB, ij = threadIdx.x - b * NANB;
        double v = 0.0;
#pragma unroll
        for (int w = 0; w < kRowWaves; ++w) v += jred[w][threadIdx.x];
        const long o = (long)(sP1[b].f1 + ij / NB) * n + sP1[b].f2 + ij % NB;
        atomic_add_f64(mv.W0 + o, v);
        if (MODE == kDigestUHF) atomic_add_f64(mv.W1 + o, v);
    }
    if (a.counter && threadIdx.x == 0) {
        unsigned long long* slot = a.counter + 2 * (blockIdx.x & (kCounterSlots - 1));
        unsigned long long np = 0, ns = 0;
        for (int w = 0; w < kRowWaves; ++w) {
            np += cred[w][0];
            ns += cred[w][1];
        }
        atomicAdd(slot, np);
        atomicAdd(slot + 1, ns);
    }
    // merged exchange rows: consecutive threads take consecutive columns of a row
    for (int br = 0; br < nb * NROW; ++br) {
        const int b = br / NROW, r = br - b * NROW;
        const int sp = r / NAB, rr = r - sp * NAB;
        const int grow = rr < NA ? sP1[b].f1 + rr : sP1[b].f2 + rr - NA;
        double* W = (MODE == kDigestUHF && sp) ? mv.W1 : mv.W0;
        for (int col = threadIdx.x; col < nc; col += blockDim.x) {
            const double v = rows[br * nc + col];
            if (v != 0.0) atomic_add_f64(W + (long)grow * n + a.colinv[col], v);
        }
    }
}

template <typename V>
__device__ __forceinline__ V sel3(int ax, V a0, V a1, V a2) {
    return ax == 0 ? a0 : (ax == 1 ? a1 : a2);
}

// ---------------------------------------------------------------------------
// one 64-lane wave (one block) per shell quartet, already resolved (and screened) to its
// bra pair, ket pair, entry and bound
// ---------------------------------------------------------------------------
// One (root, axis) of the wave kernel: the 2-D recurrence G[e][f] scaled by wz (the weight
// times the prefactor on the z axis), the ket transfer H[e][c][d] = sum_j C(d,j) cd^j
// G[e][c+d-j] and the bra transfer I[a][b][c][d] = sum_i C(b,i) ab^i H[a+b-i][c][d], all
// unrolled at compile time, each I entry stored (to LDS) as it is formed
template <int LA, int LB, int LC, int LD, typename T>
__device__ __forceinline__ void wave_axis(T C, T Cp, T B00, T B1, T B1p, T ab, T cd, T wz, T* I) {
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
#pragma unroll
    for (int e = 0; e <= E; ++e)
#pragma unroll
        for (int f = 0; f <= F; ++f) G[e][f] *= wz;
    T cdp[LD + 1], abp[LB + 1];
    cdp[0] = 1;
#pragma unroll
    for (int j = 1; j <= LD; ++j) cdp[j] = cdp[j - 1] * cd;
    abp[0] = 1;
#pragma unroll
    for (int i = 1; i <= LB; ++i) abp[i] = abp[i - 1] * ab;
    T H[E + 1][LC + 1][LD + 1];
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

// per component c of a class (the wave kernel's order): its 2-D table index on each
// axis (x | y << 8 | z << 16) and the number of x^2, y^2, z^2 d functions among its four
// (<< 24), fixed at compile time
template <int LA, int LB, int LC, int LD>
struct WaveCompTab {
    int v[ClassInfo<LA, LB, LC, LD>::NCOMP];
    constexpr WaveCompTab() : v() {
        using CI = ClassInfo<LA, LB, LC, LD>;
        for (int c = 0; c < CI::NCOMP; ++c) {
            const int l = c % CI::ND, k = (c / CI::ND) % CI::NC, j = (c / (CI::ND * CI::NC)) % CI::NB,
                      i = c / (CI::ND * CI::NC * CI::NB);
            const int nd2 = (cart_nrm(LA, i) != 1.0) + (cart_nrm(LB, j) != 1.0) + (cart_nrm(LC, k) != 1.0) +
                            (cart_nrm(LD, l) != 1.0);
            v[c] = CI::tix(i, j, k, l, 0) | CI::tix(i, j, k, l, 1) << 8 | CI::tix(i, j, k, l, 2) << 16 | nd2 << 24;
        }
    }
};

template <int LA, int LB, int LC, int LD, int MODE, typename T>
__device__ __forceinline__ unsigned wave_one(const EriArgs& a, const long long g, const int pab, const int pcd,
                                             const int entry) {
    using CI = ClassInfo<LA, LB, LC, LD>;
    constexpr int NA = CI::NA, NB = CI::NB, NC = CI::NC, ND = CI::ND, NCOMP = CI::NCOMP, NR = CI::NR;
    constexpr int E = CI::E, F = CI::F, TBL = CI::TBL;
    constexpr int PER = (NCOMP + 63) / 64;
    __shared__ T sUW[2 * NR];
    __shared__ T sI[NR * 3][TBL];
    __shared__ T sE[NCOMP];
    const int lane = threadIdx.x;
    // the two pair records live in LDS (read where needed) rather than in 40 registers
    __shared__ PairHdr sP[2];
    static_assert(sizeof(PairHdr) % 4 == 0 && sizeof(PairHdr) / 4 <= 32, "pair record copy by lanes");
    constexpr int kHdrWords = sizeof(PairHdr) / 4;
    if (lane < kHdrWords) reinterpret_cast<int*>(&sP[0])[lane] = reinterpret_cast<const int*>(a.pairs + pab)[lane];
    else if (lane >= 32 && lane < 32 + kHdrWords)
        reinterpret_cast<int*>(&sP[1])[lane - 32] = reinterpret_cast<const int*>(a.pairs + pcd)[lane - 32];
    __syncthreads();
    const PairHdr& P1 = sP[0];
    const PairHdr& P2 = sP[1];
    const MolView mv = mol_view(a, entry);
    // each lane owns components lane + 64 m: their table indices from the class's table
    static constexpr WaveCompTab<LA, LB, LC, LD> kComp{};
    // (kept packed, one register per component, unpacked where used)
    int cix[PER];
    T acc[PER];
#pragma unroll
    for (int m = 0; m < PER; ++m) {
        acc[m] = 0;
        const int c = lane + 64 * m;
        cix[m] = kComp.v[c < NCOMP ? c : 0];
    }
    unsigned long long nprimq = 0;
    for (int ip = 0; ip < P1.nprim; ++ip) {
        const PrimPair pp = a.prims[P1.prim0 + ip];
        for (int iq = 0; iq < P2.nprim; ++iq) {
            const PrimPair qq = a.prims[P2.prim0 + iq];
            const T p = (T)pp.p, q = (T)qq.p, ip_ = (T)pp.ip, iq_ = (T)qq.ip;
            const T it = (T)1 / (p + q), rt = sqrt(it);
            const T bare = (T)kTwoPi25 * (T)pp.s * (T)qq.s * ip_ * iq_ * rt;
            if (bare < (T)a.primcut) continue;   // uniform across the block
            ++nprimq;
            const T pref = (T)kTwoPi25 * (T)pp.K * (T)qq.K * ip_ * iq_ * rt;
            const T PQ[3] = {(T)(pp.P[0] - qq.P[0]), (T)(pp.P[1] - qq.P[1]), (T)(pp.P[2] - qq.P[2])};
            const T x = p * q * it * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
            if (lane < NR) {
                // root lane
                T u, w;
                rys_root_one<NR, T>(x, lane, u, w, rys_table_of<NR, T>(a));
                sUW[lane] = u;
                sUW[NR + lane] = w * pref;
            }
            __syncthreads();
            if (lane < 3 * NR) {
                // (root, axis) = (lane / 3, lane % 3): recurrence and transfer in registers,
                // the transferred table into LDS
                const int r = lane / 3, ax = lane % 3;
                const T fff = sUW[r] * it;
                const T B00 = (T)0.5 * fff, B1 = ((T)0.5 - B00 * q) * ip_, B1p = ((T)0.5 - B00 * p) * iq_;
                // axis picked by selects, not by a dynamic index (which would put the pair
                // records in scratch memory)
                const T pqa = sel3(ax, PQ[0], PQ[1], PQ[2]);
                const T C = (T)sel3(ax, pp.P[0] - P1.A[0], pp.P[1] - P1.A[1], pp.P[2] - P1.A[2]) - q * pqa * fff;
                const T Cp = (T)sel3(ax, qq.P[0] - P2.A[0], qq.P[1] - P2.A[1], qq.P[2] - P2.A[2]) + p * pqa * fff;
                const T ab = (T)sel3(ax, P1.AB[0], P1.AB[1], P1.AB[2]), cd = (T)sel3(ax, P2.AB[0], P2.AB[1], P2.AB[2]);
                const T wz = (ax == 2) ? sUW[NR + r] : (T)1;   // weight x prefactor folded into z
                wave_axis<LA, LB, LC, LD, T>(C, Cp, B00, B1, B1p, ab, cd, wz, sI[lane]);
            }
            __syncthreads();
#pragma unroll
            for (int m = 0; m < PER; ++m) {
                if (lane + 64 * m < NCOMP) {
                    const int ix = cix[m] & 255, iy = (cix[m] >> 8) & 255, iz = (cix[m] >> 16) & 255;
                    T s = acc[m];
#pragma unroll
                    for (int r = 0; r < NR; ++r) s = fma(sI[3 * r][ix] * sI[3 * r + 1][iy], sI[3 * r + 2][iz], s);
                    acc[m] = s;
                }
            }
            __syncthreads();
        }
    }
    T f = 1;
    if (MODE <= kDigestUHF) f = (P1.same ? (T)0.5 : (T)1) * (P2.same ? (T)0.5 : (T)1) * (pab == pcd ? (T)0.5 : (T)1);
#pragma unroll
    for (int m = 0; m < PER; ++m) {
        const int c = lane + 64 * m;
        if (c < NCOMP) {
            // component normalisation: 1/sqrt(3) per x^2, y^2 or z^2 d function among the four
            T cn = 1;
            for (int q = kComp.v[c] >> 24; q > 0; --q) cn *= 0.57735026918962576451;
            sE[c] = acc[m] * f * cn;
        }
    }
    __syncthreads();
    if constexpr (MODE == kDigestBlock) {
        for (int c = lane; c < NCOMP; c += 64) a.out[g * NCOMP + c] = (double)sE[c];
        return (unsigned)nprimq;
    } else if constexpr (MODE == kDigestSchwarz) {
        if (lane == 0) {
            double m = 0.0;
            for (int i = 0; i < NA; ++i)
                for (int j = 0; j < NB; ++j) m = fmax(m, fabs((double)sE[((i * NB + j) * NC + i) * ND + j]));
            a.out[pab] = sqrt(m);
        }
        return (unsigned)nprimq;
    } else {
        const int n = mv.n;
        const int fa = P1.f1, fb = P1.f2, fc = P2.f1, fd = P2.f2;
        constexpr T jf = (MODE == kDigestRHF) ? 4 : 2;
        // Coulomb outputs: NA*NB (bra) + NC*ND (ket)
        for (int o = lane; o < NA * NB + NC * ND; o += 64) {
            T s = 0;
            if (o < NA * NB) {
                const int i = o / NB, j = o % NB;
                for (int k = 0; k < NC; ++k)
                    for (int l = 0; l < ND; ++l)
                        s = fma(sE[((i * NB + j) * NC + k) * ND + l], (T)mv.D0[(long)(fc + k) * n + fd + l], s);
                fock_add(a, mv.W0 + (long)(fa + i) * n + fb + j, (double)(jf * s));
                if (MODE == kDigestUHF) fock_add(a, mv.W1 + (long)(fa + i) * n + fb + j, (double)(jf * s));
            } else {
                const int kl = o - NA * NB, k = kl / ND, l = kl % ND;
                for (int i = 0; i < NA; ++i)
                    for (int j = 0; j < NB; ++j)
                        s = fma(sE[((i * NB + j) * NC + k) * ND + l], (T)mv.D0[(long)(fa + i) * n + fb + j], s);
                // straight into W (a wave's lanes already cover the ket block's contiguous rows)
                fock_add(a, mv.W0 + (long)(fc + k) * n + fd + l, (double)(jf * s));
                if (MODE == kDigestUHF) fock_add(a, mv.W1 + (long)(fc + k) * n + fd + l, (double)(jf * s));
            }
        }
        constexpr int NK = NA * NC + NA * ND + NB * NC + NB * ND;
        for (int sp = 0; sp < (MODE == kDigestUHF ? 2 : 1); ++sp) {
            const double* DK = (MODE == kDigestUHF) ? (sp == 0 ? mv.D1 : mv.D2) : mv.D0;
            double* W = sp == 0 ? mv.W0 : mv.W1;
            for (int o = lane; o < NK; o += 64) {
                T s = 0;
                if (o < NA * NC) {                       // (a,c) <- D_bd
                    const int i = o / NC, k = o % NC;
                    for (int j = 0; j < NB; ++j)
                        for (int l = 0; l < ND; ++l)
                            s = fma(sE[((i * NB + j) * NC + k) * ND + l], (T)DK[(long)(fb + j) * n + fd + l], s);
                    fock_add(a, W + (long)(fa + i) * n + fc + k, (double)(-s));
                } else if (o < NA * NC + NA * ND) {      // (a,d) <- D_bc
                    const int q = o - NA * NC, i = q / ND, l = q % ND;
                    for (int j = 0; j < NB; ++j)
                        for (int k = 0; k < NC; ++k)
                            s = fma(sE[((i * NB + j) * NC + k) * ND + l], (T)DK[(long)(fb + j) * n + fc + k], s);
                    fock_add(a, W + (long)(fa + i) * n + fd + l, (double)(-s));
                } else if (o < NA * NC + NA * ND + NB * NC) {   // (b,c) <- D_ad
                    const int q = o - NA * NC - NA * ND, j = q / NC, k = q % NC;
                    for (int i = 0; i < NA; ++i)
                        for (int l = 0; l < ND; ++l)
                            s = fma(sE[((i * NB + j) * NC + k) * ND + l], (T)DK[(long)(fa + i) * n + fd + l], s);
                    fock_add(a, W + (long)(fb + j) * n + fc + k, (double)(-s));
                } else {                                 // (b,d) <- D_ac
                    const int q = o - NA * NC - NA * ND - NB * NC, j = q / ND, l = q % ND;
                    for (int i = 0; i < NA; ++i)
                        for (int k = 0; k < NC; ++k)
                            s = fma(sE[((i * NB + j) * NC + k) * ND + l], (T)DK[(long)(fa + i) * n + fc + k], s);
                    fock_add(a, W + (long)(fb + j) * n + fd + l, (double)(-s));
                }
            }
        }
        return (unsigned)nprimq;
    }
}

// a block (one wave) takes qpb (<= 64) consecutive quartets of the launch: lane i resolves
// quartet i and applies the density screen (one round of loads for all of them), then the
// wave evaluates the survivors one after another.  qpb stays small: the blocks are many,
// so enough waves are in flight to cover each one's serial latency (64 per block was
// measured 2.5x slower)
// waves per SIMD the register allocation must allow: three up to 324 components ((dp|dp),
// (dd|pp) then fit without spilling, with the pair records in LDS), two up to 648, one for
// (dd|dd)
template <int LA, int LB, int LC, int LD>
constexpr int wave_min_waves() {
    return ClassInfo<LA, LB, LC, LD>::NCOMP <= 324 ? 3 : ClassInfo<LA, LB, LC, LD>::NCOMP <= 648 ? 2 : 1;
}
template <int LA, int LB, int LC, int LD, int MODE, typename T>
__global__ __launch_bounds__(64, (wave_min_waves<LA, LB, LC, LD>())) void k_eri_wave(const EriArgs a) {
    const int lane = threadIdx.x;
    unsigned long long nprim = 0, nshell = 0;
    // one quartet block per launched block, or (work queue) blocks taken one after another
    // until none is left (every block ends after at most one grab past the end)
    for (long long item = blockIdx.x;;) {
        if (a.queue) {
            unsigned long long g = 0;
            if (lane == 0) g = atomicAdd(a.queue, 1ULL);
            item = (long long)__shfl(g, 0, 64);
            if (item >= a.nqblk) break;
        }
        const long long g0 = item * a.qpb;
        int pab = 0, pcd = 0, entry = 0;
        double q12 = -1.0;
        bool live = lane < a.qpb && quartet_of(a, g0 + lane, pab, pcd, q12, entry);
        if (live && a.mol_active && !a.mol_active[a.e_mol[entry]]) live = false;
        if (MODE <= kDigestUHF && live && q12 >= 0.0)
            live = !screened(a, mol_view(a, entry), a.pairs[pab], a.pairs[pcd], q12);
        unsigned long long todo = __ballot(live);
        while (todo) {
            const int src = __builtin_ctzll(todo);
            todo &= todo - 1;
            const int qab = __builtin_amdgcn_readlane(pab, src), qcd = __builtin_amdgcn_readlane(pcd, src);
            const int qe = __builtin_amdgcn_readlane(entry, src);
            nprim += wave_one<LA, LB, LC, LD, MODE, T>(a, g0 + src, qab, qcd, qe);
            ++nshell;
            __syncthreads();
        }
        if (!a.queue) break;
    }
    if (a.counter && lane == 0 && nshell) {
        unsigned long long* slot = a.counter + 2 * (blockIdx.x & (kCounterSlots - 1));
        atomicAdd(slot, nprim);
        atomicAdd(slot + 1, nshell);
    }
}

// ---------------------------------------------------------------------------
// launcher: dispatch on the class and digest mode
// ---------------------------------------------------------------------------
// lane kernels: one quartet per lane, 256 per block; wave kernels: 8 per block (one wave)
// classes up to kLaneMaxComp components run one quartet per lane (the component block in
// registers); larger ones one quartet per wave
static constexpr int kLaneMaxComp = 162;
static constexpr int quartets_per_block(int ncomp) { return ncomp <= kLaneMaxComp ? 256 : 8; }
// work-queue launches: 256-thread blocks for the wave tasks (eight per CU on 256 CUs; the
// wave kernel's one-wave blocks four times as many)
static constexpr long long kQueueBlocks = 2048;

template <int LA, int LB, int LC, int LD, int MODE, typename T = double>
static hipError_t launch_cls(const EriArgs& a, hipStream_t st) {
    using CI = ClassInfo<LA, LB, LC, LD>;
    if (a.nquart <= 0) return hipSuccess;
    EriArgs b = a;
    b.qpb = quartets_per_block(CI::NCOMP);
    const long long nb = (a.nquart + b.qpb - 1) / b.qpb;
    if constexpr (CI::NCOMP <= kLaneMaxComp && MODE <= kDigestUHF) {
        // Fock builds: wave tasks (bra, chunk of 64 kets), four per block
        if (a.ntask <= 0 || !a.wpre || !a.tblk) return hipErrorInvalidValue;
        long long nt = (a.ntask + 3) / 4;
        if (a.queue) nt = std::min<long long>(nt, kQueueBlocks);
        hipLaunchKernelGGL((k_eri_bra<LA, LB, LC, LD, MODE, T>), dim3((unsigned)nt), dim3(256), 0, st, b);
    } else if constexpr (CI::NCOMP <= kLaneMaxComp) {
        hipLaunchKernelGGL((k_eri_lane<LA, LB, LC, LD, MODE, T>), dim3((unsigned)nb), dim3(256), 0, st, b);
    } else {
        long long grid = nb;
        if (a.queue) {
            b.nqblk = nb;
            grid = std::min<long long>(nb, 4 * kQueueBlocks);
        }
        hipLaunchKernelGGL((k_eri_wave<LA, LB, LC, LD, MODE, T>), dim3((unsigned)grid), dim3(64), 0, st, b);
    }
    return hipGetLastError();
}

// row-merged Fock-build launch of a lane class: one block per row task
template <int LA, int LB, int LC, int LD, int MODE, typename T, bool TL, bool FB>
static hipError_t launch_rows_v(const EriArgs& a, hipStream_t st) {
    const size_t dyn = (size_t)(row_count<LA, LB, MODE, FB>() + (a.dlds ? dens_row_count<LA, LB, MODE>() : 0)) *
                       a.ncols * sizeof(double);
    static size_t attr = 0;   // dynamic LDS limit set for this instantiation
    if (dyn > attr) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_eri_rows<LA, LB, LC, LD, MODE, T, TL, FB>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
        if (e != hipSuccess) return e;
        attr = dyn;
    }
    hipLaunchKernelGGL((k_eri_rows<LA, LB, LC, LD, MODE, T, TL, FB>), dim3((unsigned)a.nrtask), dim3(kRowThreads), dyn, st,
                       a);
    return hipGetLastError();
}

// bra-group launch (variants 4 and 5: groups of 2 and 4 bras)
template <int LA, int LB, int LC, int LD, int MODE, typename T, int NBR>
static hipError_t launch_rows_mb(const EriArgs& a, hipStream_t st) {
    const size_t dyn = (size_t)NBR * rows_mb_count<LA, LB, MODE>() * a.ncols * sizeof(double);
    static size_t attr = 0;   // dynamic LDS limit set for this instantiation
    if (dyn > attr) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_eri_rows_mb<LA, LB, LC, LD, MODE, T, NBR>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
        if (e != hipSuccess) return e;
        attr = dyn;
    }
    hipLaunchKernelGGL((k_eri_rows_mb<LA, LB, LC, LD, MODE, T, NBR>), dim3((unsigned)a.nrtask), dim3(kRowThreads), dyn,
                       st, a);
    return hipGetLastError();
}

// variant chosen by the host (eri_rows_variant): 1 all bra rows + table in LDS, 2 all bra
// rows, 3 the first shell's rows only
template <int LA, int LB, int LC, int LD, int MODE, typename T>
static hipError_t launch_rows(const EriArgs& a, int variant, hipStream_t st) {
    using CI = ClassInfo<LA, LB, LC, LD>;
    if constexpr (CI::NCOMP <= kLaneMaxComp && MODE <= kDigestUHF) {
        // the row kernels merge in LDS with FP64 atomics: no deterministic form
        if (a.nrtask <= 0 || !a.rbra || !a.rpre || a.rseg <= 0 || !a.colmap || !a.colinv || a.ncols <= 0 ||
            a.ncols > a.n || a.fix_lo)
            return hipErrorInvalidValue;
        switch (variant) {
            case 1: return launch_rows_v<LA, LB, LC, LD, MODE, T, true, true>(a, st);
            case 2: return launch_rows_v<LA, LB, LC, LD, MODE, T, false, true>(a, st);
            case 3: return launch_rows_v<LA, LB, LC, LD, MODE, T, false, false>(a, st);
            case 4:
            case 5:
                // bra groups: FP64 spaces (every bra's kets the same run) of the small classes
                if constexpr (sizeof(T) == 8 && CI::NCOMP <= kRowsMbMaxComp2) {
                    if (a.kstart || a.e_bra || a.dlds) return hipErrorInvalidValue;
                    if (variant == 4) return launch_rows_mb<LA, LB, LC, LD, MODE, T, 2>(a, st);
                    if constexpr (CI::NCOMP <= kRowsMbMaxComp4) return launch_rows_mb<LA, LB, LC, LD, MODE, T, 4>(a, st);
                }
                return hipErrorInvalidValue;
            default: return hipErrorInvalidValue;
        }
    } else {
        return hipErrorInvalidValue;
    }
}

template <int MODE, typename T>
static hipError_t launch_mode_rows(int cb, int ck, const EriArgs& a, int variant, hipStream_t st) {
#define UNO_Q(X, Y, LA, LB, LC, LD) \
    if (cb == X && ck == Y) return launch_rows<LA, LB, LC, LD, MODE, T>(a, variant, st);
    UNO_Q(0, 0, 0, 0, 0, 0)
    UNO_Q(1, 0, 1, 0, 0, 0) UNO_Q(1, 1, 1, 0, 1, 0)
    UNO_Q(2, 0, 1, 1, 0, 0) UNO_Q(2, 1, 1, 1, 1, 0) UNO_Q(2, 2, 1, 1, 1, 1)
    UNO_Q(3, 0, 2, 0, 0, 0) UNO_Q(3, 1, 2, 0, 1, 0) UNO_Q(3, 2, 2, 0, 1, 1) UNO_Q(3, 3, 2, 0, 2, 0)
    UNO_Q(4, 0, 2, 1, 0, 0) UNO_Q(4, 1, 2, 1, 1, 0) UNO_Q(4, 2, 2, 1, 1, 1) UNO_Q(4, 3, 2, 1, 2, 0)
    UNO_Q(5, 0, 2, 2, 0, 0) UNO_Q(5, 1, 2, 2, 1, 0)
#undef UNO_Q
    return hipErrorInvalidValue;
}

hipError_t eri_launch_rows(int mode, int cb, int ck, const EriArgs& a, int variant, hipStream_t st, bool fp32) {
    if (mode == kDigestRHF) return fp32 ? launch_mode_rows<kDigestRHF, float>(cb, ck, a, variant, st)
                                        : launch_mode_rows<kDigestRHF, double>(cb, ck, a, variant, st);
    if (mode == kDigestUHF) return fp32 ? launch_mode_rows<kDigestUHF, float>(cb, ck, a, variant, st)
                                        : launch_mode_rows<kDigestUHF, double>(cb, ck, a, variant, st);
    return hipErrorInvalidValue;
}

// the LDS a row-merged block of this variant needs at ncols columns, with or without the
// staged density rows (0: the variant does not apply)
static size_t rows_lds_bytes(int cb, int ck, int mode, int ncols, bool fp32, int variant, bool dlds) {
    static const int na[6] = {1, 3, 3, 6, 6, 6};          // functions of the first bra shell
    static const int nb[6] = {1, 1, 3, 1, 3, 6};          // and of the second
    static const int lsum[6] = {0, 1, 2, 2, 3, 4};
    const int nr = (lsum[cb] + lsum[ck]) / 2 + 1;
    const size_t tab = fp32 ? (size_t)kRysAsym[nr] * 2 * nr * UNO_RYS_NCOEF_F * sizeof(float)
                            : (size_t)kRysAsym[nr] * 2 * nr * UNO_RYS_NCOEF * sizeof(double);
    const size_t sp = mode == kDigestUHF ? 2 : 1, reserve = 1024;
    const size_t all = (size_t)(na[cb] + nb[cb]) * sp * ncols * sizeof(double),
                 first = (size_t)na[cb] * sp * ncols * sizeof(double);
    const size_t dens = dlds ? all : 0;
    static const int ncomp[6] = {1, 3, 9, 6, 18, 36};
    const int nc = ncomp[cb] * ncomp[ck];
    switch (variant) {
        case 1: return all + dens + tab + reserve;
        case 2: return all + dens + reserve;
        case 3: return first + dens + reserve;
        case 4: return (fp32 || dlds || nc > kRowsMbMaxComp2) ? 0 : 2 * all + tab + reserve;
        case 5: return (fp32 || dlds || nc > kRowsMbMaxComp4) ? 0 : 4 * all + tab + reserve;
        default: return 0;
    }
}

int eri_rows_group(int variant) { return variant == 4 ? 2 : variant == 5 ? 4 : 1; }

bool eri_rows_dlds_fits(int cb, int ck, int mode, int ncols, bool fp32, int variant) {
    const size_t b = rows_lds_bytes(cb, ck, mode, ncols, fp32, variant, true);
    return b > 0 && b <= kRowLdsBudget;
}

int eri_rows_variant(int cb, int ck, int mode, int ncols, bool fp32, int force) {
    if (!eri_uses_wave_tasks(cb, ck)) return 0;
    bool fits[6] = {true, false, false, false, false, false};
    for (int v = 1; v <= 5; ++v) {
        const size_t b = rows_lds_bytes(cb, ck, mode, ncols, fp32, v, false);
        fits[v] = b > 0 && b <= kRowLdsBudget;
    }
    if (force >= 1 && force <= 5 && fits[force]) return force;
    for (int v = 1; v <= 3; ++v)
        if (fits[v]) return v;
    return 0;
}

template <int MODE, typename T>
static hipError_t launch_mode(int cb, int ck, const EriArgs& a, hipStream_t st) {
    // pair classes: 0 ss, 1 ps, 2 pp, 3 ds, 4 dp, 5 dd
#define UNO_Q(X, Y, LA, LB, LC, LD) \
    if (cb == X && ck == Y) return launch_cls<LA, LB, LC, LD, MODE, T>(a, st);
    UNO_Q(0, 0, 0, 0, 0, 0)
    UNO_Q(1, 0, 1, 0, 0, 0) UNO_Q(1, 1, 1, 0, 1, 0)
    UNO_Q(2, 0, 1, 1, 0, 0) UNO_Q(2, 1, 1, 1, 1, 0) UNO_Q(2, 2, 1, 1, 1, 1)
    UNO_Q(3, 0, 2, 0, 0, 0) UNO_Q(3, 1, 2, 0, 1, 0) UNO_Q(3, 2, 2, 0, 1, 1) UNO_Q(3, 3, 2, 0, 2, 0)
    UNO_Q(4, 0, 2, 1, 0, 0) UNO_Q(4, 1, 2, 1, 1, 0) UNO_Q(4, 2, 2, 1, 1, 1) UNO_Q(4, 3, 2, 1, 2, 0)
    UNO_Q(4, 4, 2, 1, 2, 1)
    UNO_Q(5, 0, 2, 2, 0, 0) UNO_Q(5, 1, 2, 2, 1, 0) UNO_Q(5, 2, 2, 2, 1, 1) UNO_Q(5, 3, 2, 2, 2, 0)
    UNO_Q(5, 4, 2, 2, 2, 1) UNO_Q(5, 5, 2, 2, 2, 2)
#undef UNO_Q
    return hipErrorInvalidValue;
}

hipError_t eri_launch(int mode, int cb, int ck, const EriArgs& a, hipStream_t st, bool fp32) {
    if (fp32) {
        switch (mode) {
            case kDigestRHF: return launch_mode<kDigestRHF, float>(cb, ck, a, st);
            case kDigestUHF: return launch_mode<kDigestUHF, float>(cb, ck, a, st);
            default: return hipErrorInvalidValue;   // block and Schwarz modes are FP64 only
        }
    }
    switch (mode) {
        case kDigestRHF: return launch_mode<kDigestRHF, double>(cb, ck, a, st);
        case kDigestUHF: return launch_mode<kDigestUHF, double>(cb, ck, a, st);
        case kDigestBlock: return launch_mode<kDigestBlock, double>(cb, ck, a, st);
        default: break;
    }
    // Schwarz: diagonal classes only
    if (cb != ck) return hipErrorInvalidValue;
    switch (cb) {
        case 0: return launch_cls<0, 0, 0, 0, kDigestSchwarz>(a, st);
        case 1: return launch_cls<1, 0, 1, 0, kDigestSchwarz>(a, st);
        case 2: return launch_cls<1, 1, 1, 1, kDigestSchwarz>(a, st);
        case 3: return launch_cls<2, 0, 2, 0, kDigestSchwarz>(a, st);
        case 4: return launch_cls<2, 1, 2, 1, kDigestSchwarz>(a, st);
        case 5: return launch_cls<2, 2, 2, 2, kDigestSchwarz>(a, st);
    }
    return hipErrorInvalidValue;
}

int eri_class_ncomp(int cb, int ck) {
    static const int nab[6] = {1, 3, 9, 6, 18, 36};
    return nab[cb] * nab[ck];
}

int eri_quartets_per_block(int cb, int ck) { return quartets_per_block(eri_class_ncomp(cb, ck)); }

bool eri_uses_wave_tasks(int cb, int ck) { return eri_class_ncomp(cb, ck) <= kLaneMaxComp; }

// ---------------------------------------------------------------------------
// pair-major Coulomb accumulator into W
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_jp_scatter(int np, const PairHdr* __restrict__ pairs, long long njp,
                                                    const double* __restrict__ jp, double* W0, double* W1, int n) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const PairHdr h = pairs[p];
    const long long c1 = p + 1 < np ? (long long)pairs[p + 1].jo : njp;
    const int nc = (int)(c1 - h.jo);
    // block shape from its size: 1 ss, 3 ps, 9 pp, 6 ds, 18 dp, 36 dd
    const int nd = nc == 9 || nc == 18 ? 3 : nc == 36 ? 6 : 1;
    for (int kl = 0; kl < nc; ++kl) {
        const double v = jp[h.jo + kl];
        if (v == 0.0) continue;
        const long o = (long)(h.f1 + kl / nd) * n + h.f2 + kl % nd;
        W0[o] += v;
        if (W1) W1[o] += v;
    }
}

// the pair and primitive arrays reordered on the device: new pair k is old pair order[k],
// its primitive records moved to prim0[k] and its Coulomb block offset set to jo[k]
__global__ __launch_bounds__(256) void k_permute_pairs(int np, const int* __restrict__ order,
                                                       const int* __restrict__ prim0, const int* __restrict__ jo,
                                                       const PairHdr* __restrict__ oldp,
                                                       const PrimPair* __restrict__ oldq, PairHdr* __restrict__ newp,
                                                       PrimPair* __restrict__ newq) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np) return;
    PairHdr h = oldp[order[k]];
    const int p0 = h.prim0;
    h.prim0 = prim0[k];
    h.jo = jo[k];
    newp[k] = h;
    for (int q = 0; q < h.nprim; ++q) newq[h.prim0 + q] = oldq[p0 + q];
}

hipError_t permute_pairs(int np, const int* order, const int* prim0, const int* jo, const PairHdr* oldp,
                         const PrimPair* oldq, PairHdr* newp, PrimPair* newq, hipStream_t st) {
    if (np <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_permute_pairs, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, np, order, prim0, jo,
                       oldp, oldq, newp, newq);
    return hipGetLastError();
}

// fixed-point W (two int64 words per entry, fix_lo entries apart) to FP64, in place
__global__ __launch_bounds__(256) void k_fixed_to_f64(long long nn, double* W, long long lo_off) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    const long long* w = reinterpret_cast<const long long*>(W);
    const long long hi = w[i], lo = w[i + lo_off];
    W[i] = (double)hi * (1.0 / kFixHiScale) + (double)lo * (1.0 / kFixLoScale);
}

hipError_t fixed_to_f64(long long nn, double* W, long long lo_off, hipStream_t st) {
    if (nn <= 0) return hipSuccess;
    if (lo_off < nn) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_fixed_to_f64, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, nn, W, lo_off);
    return hipGetLastError();
}

hipError_t jp_scatter(int npairs, const PairHdr* pairs, long long njp, const double* jp, double* W0, double* W1, int n,
                      hipStream_t st) {
    if (npairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_jp_scatter, dim3((npairs + 255) / 256), dim3(256), 0, st, npairs, pairs, njp, jp, W0, W1, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Rys-table evaluation for the parity tests
// ---------------------------------------------------------------------------
template <int NR, typename T>
__global__ void k_rys_eval(int npts, const double* x, double* u, double* w, int ref2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    T uu[NR], ww[NR];
    EriArgs a{};
    a.rys2_exact = !ref2;
    rys_roots<NR, T>((T)x[i], uu, ww, rys_table_of<NR, T>(a));
    for (int r = 0; r < NR; ++r) { u[i * NR + r] = (double)uu[r]; w[i * NR + r] = (double)ww[r]; }
}

template <typename T>
static hipError_t rys_eval_t(int n, int npts, const double* x, double* u, double* w, hipStream_t st, int ref2) {
    const dim3 g((npts + 255) / 256), b(256);
    switch (n) {
        case 1: hipLaunchKernelGGL((k_rys_eval<1, T>), g, b, 0, st, npts, x, u, w, ref2); break;
        case 2: hipLaunchKernelGGL((k_rys_eval<2, T>), g, b, 0, st, npts, x, u, w, ref2); break;
        case 3: hipLaunchKernelGGL((k_rys_eval<3, T>), g, b, 0, st, npts, x, u, w, ref2); break;
        case 4: hipLaunchKernelGGL((k_rys_eval<4, T>), g, b, 0, st, npts, x, u, w, ref2); break;
        case 5: hipLaunchKernelGGL((k_rys_eval<5, T>), g, b, 0, st, npts, x, u, w, ref2); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t rys_eval_launch(int n, int npts, const double* x, double* u, double* w, hipStream_t st, bool fp32,
                           bool ref2) {
    return fp32 ? rys_eval_t<float>(n, npts, x, u, w, st, ref2) : rys_eval_t<double>(n, npts, x, u, w, st, ref2);
}

}  // namespace uno
