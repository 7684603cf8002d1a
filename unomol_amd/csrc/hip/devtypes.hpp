// Device-side data layout of the direct-SCF engine (shared by host builders and kernels).
//
// Shell pairs (a >= b by shell index) are stored with the higher-l shell first, so every
// quartet class (la,lb|lc,ld) has la >= lb and lc >= ld like the reference's swapped
// shell quartets (TwoElectronInts.cpp:563-580).  Each pair owns a contiguous run of
// primitive pairs.  Both records are 64 bytes so a lane fetches them as 4 x dwordx4.
#pragma once
#include <cstdint>

namespace uno {

struct alignas(16) PairHdr {
    double A[3];    // centre of the first (higher-l) shell
    double AB[3];   // A - B
    int f1, f2;     // first basis function of shell 1 and shell 2
    int prim0;      // first primitive pair
    int nprim;      // number of primitive pairs
    int same;       // both shells are the same shell (a == b)
    int sh1, sh2;   // shell indices (first, second)
    int jo;         // offset of the pair's function block in the pair-major Coulomb
                    // accumulator (runs of pairs are contiguous there)
};

struct alignas(16) PrimPair {
    double p, ip;   // exponent sum and its inverse
    double P[3];    // Gaussian product centre
    double K;       // c_i c_j (x2 for off-diagonal same-shell pairs) exp(-ab/p |AB|^2)
    double s;       // bare exp(-ab/p |AB|^2) for the primitive prefactor cut
    double pad;
};

// slot pairs of the evaluated-quartet counters of one launch (spread to avoid contention)
constexpr int kCounterSlots = 64;

// pair class index for (l1 >= l2), l <= 2: ss=0 ps=1 pp=2 ds=3 dp=4 dd=5
constexpr int pair_class(int l1, int l2) { return l1 * (l1 + 1) / 2 + l2; }
constexpr int kMaxL = 2;
constexpr int kNumPairClasses = 6;

// digest modes of the ERI kernels
enum DigestMode : int {
    kDigestRHF = 0,      // W += 4 J'[P] - K'[P]
    kDigestUHF = 1,      // Wa += 2 J'[Pa+Pb] - K'[Pa], Wb likewise
    kDigestBlock = 2,    // write the contracted, normalised ERI block
    kDigestSchwarz = 3,  // max |(ij|ij)| of a diagonal quartet
};

// one molecule of a batched launch: its dense matrices start at moff of every store,
// its shell-block density maxima at dshoff
struct MolDesc {
    long long moff;
    long long dshoff;
    int n, nsh;
};

struct EriArgs {
    const PairHdr* pairs;
    const PrimPair* prims;
    // implicit quartet space of one launch: bra pairs [bra_off, bra_off+nbra) and ket pairs
    // from ket_off, both runs of the pair array sorted by Schwarz bound (descending);
    // prefix[i] = first quartet of bra i, blk[b] = bra holding quartet b*qpb
    int bra_off, ket_off;
    const double* pairQ;      // Schwarz bound of every pair, in pair-array order
    const long long* prefix;
    const int* blk;
    const int* kstart;        // first ket of every bra's range (null: 0), FP32 ranges
                              // start after the bra's FP64 kets
    // bra-major launches: one wave per (bra, chunk of 64 kets); wpre[i] = first wave
    // task of bra i, tblk[b] = bra of task 64 b
    const long long* wpre;
    const int* tblk;
    long long ntask;
    // row-merged launches (bras with many kets): block task b belongs to the r-th listed
    // bra, rpre[r] <= b < rpre[r+1]; rbra[r] is its entry; a block covers rseg kets of it
    // and merges its exchange contributions to the bra's rows in LDS before adding them
    const int* rbra;
    const long long* rpre;
    int nrbra;
    long long nrtask;
    int rseg;
    // columns of those LDS rows: the functions the launch's ket run touches, in function
    // order (whole shells, so a shell's functions stay consecutive); colmap[f] = column of
    // function f (-1: untouched), colinv[c] = function of column c, ncols columns
    const int* colmap;
    const int* colinv;
    int ncols;
    // 1: the bra's exchange density rows (every bra function, each spin density, the same
    // columns) are staged in LDS after the accumulator rows
    int dlds;
    // batched launches: entry i (a bra with one ket range) names its bra pair, its first
    // ket pair and its molecule (null: the run offsets above and molecule view n/D/W)
    const int* e_bra;
    const int* e_ket;
    const int* e_mol;
    const MolDesc* mols;
    int qpb;                  // quartets per block of the launch
    int rys2_exact;           // 1: exact two-root Rys quadrature; 0: the reference's
                              // two-root quadrature (Rys.cpp:413-631), the default
    int nbra;
    long long nquart;
    const int2* explicit_q;   // explicit (bra, ket) list (block / Schwarz modes) or null
    int same_class;           // bra and ket lists are the same list
    double tau;               // quartet screening threshold on Qab Qcd Dmax
    double primcut;           // bare primitive prefactor cut (reference 1e-12)
    const double* dsh;        // [nsh*nsh] max |D| over each shell block (null = global bound only)
    int nsh;
    double dmax_global;
    int n;                    // basis functions
    const double* D0;         // RHF: P (half density); UHF: Pa + Pb (dense n x n)
    const double* D1;         // UHF: Pa
    const double* D2;         // UHF: Pb
    double* W0;               // RHF: W; UHF: Wa
    double* W1;               // UHF: Wb
    double* out;              // block / Schwarz output
    double* jp;               // pair-major accumulator of the ket Coulomb outputs (pair.jo)
                              // of the per-lane kernels; null: they go straight to W
    unsigned long long* counter;  // evaluated-quartet counters (may be null): kCounterSlots
                                  // (primitive, shell) pairs, summed by the host
    // work-queue launches (batched builds): a fixed grid whose waves (or one-wave blocks)
    // take the next work item from *queue (zeroed before the launch) until it passes the
    // item count (ntask wave tasks, or nqblk quartet blocks of the wave kernel); null: one
    // item per wave / block as launched
    unsigned long long* queue;
    long long nqblk;
    const int* mol_active;    // per molecule: 0 skips its items (null: every molecule)
    // deterministic accumulation (0: FP64 atomics): every W entry is two int64 words, the
    // word at the entry's own address and one fix_lo entries further on, holding the sum in
    // fixed point (integer additions give the same sum in any order); wave-task and
    // wave kernels only
    long long fix_lo;
};

}  // namespace uno
