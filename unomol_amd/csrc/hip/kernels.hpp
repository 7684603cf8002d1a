// Host-callable launchers of the device kernels (all asynchronous on the given stream).
#pragma once
#include <hip/hip_runtime.h>
#include "devtypes.hpp"

namespace uno {

// ERIs (eri.hip)
// fp32: evaluate the quartets in single precision (RHF/UHF digestion only), accumulating in FP64
hipError_t eri_launch(int mode, int bra_class, int ket_class, const EriArgs& a, hipStream_t st, bool fp32 = false);
int eri_class_ncomp(int bra_class, int ket_class);
// quartets one block of the class kernel walks (the block table granularity)
int eri_quartets_per_block(int bra_class, int ket_class);
// Fock-build launches of this class run as wave tasks (bra, up to 64 kets): the space needs
// wpre / tblk / ntask instead of the quartet block table
bool eri_uses_wave_tasks(int bra_class, int ket_class);
// row-merged Fock-build launch (bras with many kets) of a wave-task class: a.rbra / a.rpre /
// a.nrtask / a.rseg describe the block tasks, a.colmap / a.colinv / a.ncols the LDS row
// columns; variant from eri_rows_variant: 1 every bra row and the Rys table in LDS, 2 every
// bra row, 3 the first bra shell's rows only, 0 none fits the LDS budget at ncols columns
// (the class then runs as wave tasks only); force 1-3 picks that variant when it fits,
// force 4 or 5 the bra-group kernel (groups of 2 or 4 bras, every bra row and the table in
// LDS; FP64 spaces of the classes of at most 9 or 3 components)
hipError_t eri_launch_rows(int mode, int bra_class, int ket_class, const EriArgs& a, int variant, hipStream_t st,
                           bool fp32);
int eri_rows_variant(int bra_class, int ket_class, int mode, int ncols, bool fp32, int force);
// whether that variant still fits the LDS budget with the bra's exchange density rows
// staged as well (a.dlds = 1)
bool eri_rows_dlds_fits(int bra_class, int ket_class, int mode, int ncols, bool fp32, int variant);
// bras per row task of a variant (2 or 4 for the bra-group variants, else 1)
int eri_rows_group(int variant);
// W[f1 + k][f2 + l] += jp[jo + k nd + l] (and W1) for every pair of the array (pair blocks
// are disjoint, no atomics); the block size of a pair follows from the next pair's jo
hipError_t jp_scatter(int npairs, const PairHdr* pairs, long long njp, const double* jp, double* W0, double* W1, int n,
                      hipStream_t st);
// new pair k = old pair order[k], its primitives at prim0[k], Coulomb block at jo[k]
// (device arrays; new ones must not overlap the old)
hipError_t permute_pairs(int np, const int* order, const int* prim0, const int* jo, const PairHdr* oldp,
                         const PrimPair* oldq, PairHdr* newp, PrimPair* newq, hipStream_t st);
// deterministic builds: W accumulated in fixed point (entry i: int64 words i and i + lo_off)
// back to FP64 in place, first nn entries
hipError_t fixed_to_f64(long long nn, double* W, long long lo_off, hipStream_t st);
hipError_t rys_eval_launch(int n, int npts, const double* x, double* u, double* w, hipStream_t st, bool fp32 = false,
                           bool ref2 = false);

// dense helpers (linalg.hip); matrices are row-major n x n
hipError_t k_unpack(int n, const double* packed, double* dense, hipStream_t st);       // symmetric fill
hipError_t k_pack(int n, const double* dense, double* packed, hipStream_t st);         // lower triangle
hipError_t k_pack_sym_add(int n, const double* W, double* packed, hipStream_t st);     // packed += (W+W^T)_lower
// G = W + W^T, or G += W + W^T when accumulating
hipError_t k_sym_w(int n, const double* W, double* G, hipStream_t st, bool accumulate = false);
hipError_t k_axpby(long len, double a, const double* x, double b, const double* y, double* z, hipStream_t st);
hipError_t k_fill(long len, double v, double* x, hipStream_t st);
// X[i][j] = Z[i][j] f[j] (n x n, f on the device)
hipError_t k_scale_cols(int n, const double* Z, const double* f, double* X, hipStream_t st);
// out[0] += sum_ij A_ij B_ij (full matrices); deterministic two-pass reduction whose
// workspace ws holds kRedBlocks partial sums (one workspace per stream)
constexpr int kRedBlocks = 240;
hipError_t k_dot(long len, const double* A, const double* B, double* out, double* ws, hipStream_t st);
// out[0] = sqrt(sum (A-B)^2) / n
hipError_t k_diffnorm(long len, int n, const double* A, const double* B, double* out, double* ws, hipStream_t st);
// Dsh[s][t] = max |D_ij| over the function block of shells s, t
hipError_t k_shell_max(int n, int nsh, const int* sh_off, const int* sh_n, const double* D, const double* D1,
                       const double* D2, double* Dsh, hipStream_t st);
// out[0] = max |x|
hipError_t k_absmax(long len, const double* x, double* out, hipStream_t st);
// A_ii += alpha (n x n)
hipError_t k_shift_diag(int n, double alpha, double* A, hipStream_t st);
// purification: Gershgorin bounds of F (lohi: 2n work, bounds: 2 doubles) and
// X = (hi I - F) / (hi - lo); out[0] = tr(A)
hipError_t sp2_init(int n, const double* F, double* X, double* lohi, double* bounds, hipStream_t st);
hipError_t k_trace(int n, const double* A, double* out, hipStream_t st);
// launch-chained purification steps (linalg.hip): states is an array of step states
// (sp2_state_bytes() each); sp2_start fills state 0 and X0's tile sums (part0: 2 doubles per
// upper-triangle 64 x 64 tile); sp2_steps enqueues steps k0 .. k0+count-1; sp2_state_read
// waits for the stream and reads state k (cur = buffer holding X, done 1 = converged,
// 2 = failed, steps = steps taken, t1 = tr(X))
size_t sp2_state_bytes();
hipError_t sp2_start(int n, const double* X0, double* part0, void* st0, hipStream_t st);
hipError_t sp2_steps(int n, int nocc, double thr, void* states, int k0, int count, double* buf0, double* buf1,
                     double* part0, double* part1, hipStream_t st);
hipError_t sp2_state_read(const void* states, int k, int* cur, int* done, int* steps, double* t1, hipStream_t st);
// C = alpha op(A) op(B) + beta C  (FP64 MFMA 16x16x4)
hipError_t gemm_f64(bool ta, bool tb, int m, int n, int k, double alpha, const double* A, int lda, const double* B,
                    int ldb, double beta, double* C, int ldc, hipStream_t st);
// 1: this thread's gemm_f64 calls give bitwise-repeatable products (no split K, whose
// chunks are added with atomics); returns the previous setting
int gemm_deterministic(int on);
struct GemmDetScope {
    int prev;
    explicit GemmDetScope(bool on) : prev(gemm_deterministic(on ? 1 : 0)) {}
    ~GemmDetScope() { gemm_deterministic(prev); }
    GemmDetScope(const GemmDetScope&) = delete;
    GemmDetScope& operator=(const GemmDetScope&) = delete;
};

// eigensolver pieces (linalg.hip)
struct EigWork {
    double* A;      // n x n working copy (destroyed)
    double* V;      // n x n reflector vectors (row k = v_k, zero outside its support)
    double* tau;    // n
    double* p;      // n
    double* d;      // n  diagonal of T
    double* e;      // n  off-diagonal of T
    double* scal;   // 4 scalars
    double* evals;  // n  eigenvalues of T (ascending)
    double* Y;      // n x n eigenvectors of T, one per column (Yt)
    double* wk;     // 5 n^2 inverse-iteration work
};
hipError_t householder_tridiag(int n, EigWork& w, hipStream_t st);
// Z <- Q Z with Q = H_0 ... H_{n-3} from the reflectors in w.V / w.tau (uses w.wk)
hipError_t apply_q(int n, EigWork& w, double* Z, hipStream_t st);
// nclusters < 0: multisection for w.evals; otherwise inverse iteration into the columns of
// w.Y followed by Gram-Schmidt inside the given clusters ({first, end} pairs, device memory)
hipError_t tridiag_eigen(int n, EigWork& w, int nclusters, const int* clusters, hipStream_t st);

// streaming copy of `bytes` (read + write), reps times; total milliseconds in *ms
hipError_t copy_probe(size_t bytes, int reps, float* ms, hipStream_t st);

}  // namespace uno
