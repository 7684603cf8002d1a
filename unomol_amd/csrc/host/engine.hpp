// Internal engine state behind the C ABI (include/unomol.h).
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <map>
#include <tuple>
#include <vector>
#include "../../../include/unomol.h"
#include "../hip/devtypes.hpp"
#include "../hip/oneint_kernels.hpp"
#include "basis.hpp"

namespace uno {

void set_error(const std::string& msg);
const std::string& get_error();

#define UNO_HIP(expr)                                                                               \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) {                                                                     \
            ::uno::set_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " + __FILE__ + \
                             ":" + std::to_string(__LINE__));                                       \
            return UNO_EDEVICE;                                                                     \
        }                                                                                           \
    } while (0)

// device buffer owned by RAII
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t alloc(size_t cnt) {
        if (cnt <= n && p) return hipSuccess;
        release();
        n = cnt;
        return hipMalloc((void**)&p, (cnt ? cnt : 1) * sizeof(T));
    }
    hipError_t upload(const T* h, size_t cnt, hipStream_t st) {
        hipError_t e = alloc(cnt);
        if (e != hipSuccess) return e;
        return cnt ? hipMemcpyAsync(p, h, cnt * sizeof(T), hipMemcpyHostToDevice, st) : hipSuccess;
    }
};

// A run of the pair array: one angular pair class (ss ps pp ds dp dd) and one bucket of
// primitive-pair counts, sorted by Schwarz bound (descending).  Bucketing by primitive
// count keeps the loop trip counts of the lanes of a wave alike.
struct PairClassList {
    int lcls = 0, bucket = 0;
    int off = 0, count = 0;
    std::vector<double> Q;      // host copy of the bounds of the run
};

// One quartet space of a launch: per bra, a contiguous range of kets (prefix counts, the
// first ket of each range when it does not start at 0) and the block table.
struct QuartetSpace {
    // bras b0 .. b0 + prefix.size() - 2 of the bra run (the empty bras at either end are
    // trimmed); prefix[i] = first quartet of bra b0 + i
    int b0 = 0;
    std::vector<long long> prefix, wpre, rpre;
    std::vector<int> kstart, blk, tblk, rbra;
    DBuf<long long> d_prefix, d_wpre, d_rpre;
    DBuf<int> d_kstart, d_blk, d_tblk, d_rbra;
    long long nq = 0, ntask = 0, nrtask = 0;
};

// row-merged Fock-build tasks: bras with at least row_min_kets (default kRowMinKets) kets
// in a space run as blocks of kRowSeg kets whose exchange contributions are merged in LDS
// (the LDS variant per class and n from eri_rows_variant)
constexpr long long kRowMinKets = 192;
constexpr int kRowSeg = 1024;

struct TuneChoice {
    int variant;
    long long min_kets;
    int dlds = 0;   // the bra's exchange density rows staged in LDS as well
};

struct QuartetClass {
    int X, Y;                   // indices into the runs, X >= Y (bra run, ket run)
    QuartetSpace s64;           // FP64 quartets (all of them unless the build is mixed)
    QuartetSpace s32;           // mixed builds: quartets whose Schwarz bound is below fp32_thr
};

// profiling slots: kernel kind (0 the wave-task / per-quartet kernels, 1-3 the row-merged
// variants, 4-5 the bra-group row kernels), then precision (FP64, FP32), then the class
// pair (6 X + Y): slot = 72 kind + 36 prec + 6 X + Y
constexpr int kProfKinds = 6;
constexpr int kProfSlots = kProfKinds * 2 * kNumPairClasses * kNumPairClasses;

constexpr int kNumBuckets = 6;
// streams a Fock build's ERI launches may be spread over (HIP's default is four hardware
// queues per process)
constexpr int kMaxLaunchStreams = 4;
inline int prim_bucket(int nprim) {
    return nprim <= 1 ? 0 : nprim <= 3 ? 1 : nprim <= 6 ? 2 : nprim <= 9 ? 3 : nprim <= 18 ? 4 : 5;
}

struct EigenWS;

}  // namespace uno

struct uno_mlp;

struct uno_ctx {
    uno::Basis B;
    int n = 0, nsh = 0, start_shell = 0;
    hipStream_t st = nullptr;
    // options
    double tau = 1e-13, primcut = 1e-12, fp32_thr = 1e-6;
    // screening threshold of the mixed-precision builds (never below tau): their error
    // only steers the early iterations; the FP64 phase starts with a full rebuild at tau
    double tau_mixed = 3e-9;
    double mixed_switch = 3e-6;   // |dE| (Eh) ending uno_scf_run's mixed-precision phase
    int rebuild_every = 10;       // incremental builds between full rebuilds
    int precision = 0;
    // pair data
    std::vector<uno::PairHdr> pairs;
    std::vector<uno::PrimPair> prims;
    std::vector<double> pairQ;                  // Schwarz bound per pair
    std::vector<int> pair_of;                   // [a*nsh+b] (a>=b) -> pair index or -1
    uno::DBuf<uno::PairHdr> d_pairs;
    uno::DBuf<uno::PrimPair> d_prims;
    std::vector<uno::PairClassList> cls;   // runs ordered by (angular class, bucket)
    std::vector<uno::QuartetClass> qcls;
    uno::DBuf<double> d_pairQ;
    // per-shell function layout for the density screen
    uno::DBuf<int> d_shoff, d_shn;
    // device work matrices (n x n)
    uno::DBuf<double> D0, D1, D2, W0, W1, Dsh, pk, scal;
    uno::DBuf<double> JP;        // pair-major accumulator of the kets' Coulomb outputs
    uno::DBuf<double> red;       // partial sums of the two-pass reductions
    // LDS row columns of the row-merged launches, per pair run (ket side): run_ncols[r]
    // columns, d_colmap[r n + f] = column of function f or -1, d_colinv[r n + c] = function
    std::vector<int> run_ncols;
    uno::DBuf<int> d_colmap, d_colinv;
    std::vector<char> stage_h;   // a Fock build's index arrays, uploaded once per build
    char* stage_pin = nullptr;   // pinned copy of stage_h that the (asynchronous) upload reads
    size_t stage_pin_cap = 0;
    hipEvent_t stage_ev = nullptr;   // recorded after the upload: stage_pin is free again
    uno::DBuf<char> stage_d;
    long long njp = 0;
    uno::EigenWS* eig = nullptr;
    double last_quartets = 0.0, last_quartets32 = 0.0;
    // host milliseconds of the last build: density bound read-back, quartet spaces and
    // index arrays, staging upload, launch enqueue (the kernels run asynchronously)
    double last_ms[4] = {0.0, 0.0, 0.0, 0.0};
    // host milliseconds of the context build: pair records, Schwarz bounds (device),
    // sorted runs and their upload, the rest (columns, one-electron data, buffers)
    double ctx_ms[4] = {0.0, 0.0, 0.0, 0.0};
    // profiling pass: per quartet class (bra X, ket Y) kernel milliseconds and the
    // primitive / shell quartets the kernels evaluated
    int profile = 0;
    // Fock builds run bras with many kets as row-merged blocks (1) or all as wave tasks (0)
    int row_merge = 1;
    long long row_min_kets = uno::kRowMinKets;
    int row_variant = 0;     // row-merged LDS variant: 0 the best that fits, 1-3 forced
    int rys2_exact = 0;      // 1: exact two-root Rys quadrature; 0: the reference's (default)
    int purify = -1;         // SCF densities by purification: -1 from 256 functions, 0 never, 1 always
    int x_newton = -1;       // X = S^-1/2 by Newton-Schulz: -1 from kSp2MinN functions on, 0 never, 1 always
    int last_x_steps = -1;   // the last SCF's Newton-Schulz steps (-1: the canonical X)
    // 1: Fock builds accumulate in fixed point (bitwise the same G for the same input, in
    // any launch order) and run every bra as wave tasks (the row kernels merge with FP64
    // atomics); the SCF's dense products run without split K
    int deterministic = 0;
    // the ERI launches of a Fock build go round-robin over the context's stream and
    // launch_streams - 1 worker streams (default priority), joined back before the build's
    // last steps: the many small launches (FP32 tails, small quartet spaces) then run side
    // by side instead of each leaving most of the GPU idle
    int launch_streams = 3;
    hipStream_t wst[uno::kMaxLaunchStreams - 1] = {};
    hipEvent_t wfork = nullptr, wjoin[uno::kMaxLaunchStreams - 1] = {};
    int row_dens = -1;       // exchange density rows in LDS: -1 when they fit, 0 never, 1 forced where they fit
    // autotuned row-merge configuration per quartet space (run-pair index, precision, UHF,
    // mixed build): LDS variant (0 = wave tasks only) and the ket count from which a bra is
    // row-merged.  tuning = 1 while uno_autotune times the candidates.
    std::map<std::tuple<int, int, int, int>, uno::TuneChoice> tune;
    int tuning = 0;
    double tune_ms = 0.0;
    double prof_ms[uno::kProfSlots] = {};
    double prof_prim[uno::kProfSlots] = {};
    double prof_shell[uno::kProfSlots] = {};
    double prof_launch[uno::kProfSlots] = {};
    uno::DBuf<unsigned long long> d_counter;
    // profiled launches are bracketed by event pairs from a reusable pool (pair k: events
    // 2k, 2k+1; prof_evslot[k] its slot); builds leave them pending, and profile_collect
    // reads them (and the device counters) when the results are asked for, so that a
    // profiled SCF runs without a synchronisation per build
    std::vector<hipEvent_t> prof_ev;
    std::vector<int> prof_evslot;
    size_t prof_pending = 0;
    uno_mlp* guess_model = nullptr;   // not owned: the network of guess = 2
    int skip_center = -1;             // centre left out of the nuclear attraction (DPM positron)
    // one-electron data on the device: shells, primitives, centres and the shell pairs
    // of each (la >= lb) class
    uno::DBuf<uno::OneIntShell> oi_shells;
    uno::DBuf<double> oi_alpha, oi_coef, oi_xyz, oi_charge;
    uno::DBuf<int2> oi_pairs[uno::kNumPairClasses];
    int oi_npair[uno::kNumPairClasses] = {};
    ~uno_ctx();
};

namespace uno {

int ctx_build(uno_ctx* c, const uno_basis* b, int start_shell);
// quartet-space helpers of the Fock build (ctx.cpp)
void build_prefix(uno_ctx* c, QuartetClass& q, double dmax, double fp32_thr, double tau);
void build_tasks(QuartetSpace& s, long long row_min_kets = 0);
void build_rows(QuartetSpace& s, long long min_kets, int seg, int group = 1);
void build_blocks(QuartetSpace& s, int qpb);
EriArgs base_args(uno_ctx* c);
// keeps c's tuned configurations for later contexts of the same layout
void remember_tuning(uno_ctx* c);
// profiling: reset (and, when enabling, zero the device counters), or fold the pending
// event pairs and the counters into the context's per-slot totals
int profile_reset(uno_ctx* c, bool enable);
int profile_collect(uno_ctx* c);
// dense S, T, H = T + V (n x n, device) on the context's stream
int oneint_device(uno_ctx* c, double* S, double* T, double* H);
// Fock build from dense device densities already in c->D0 (and D1, D2 for UHF).
// Leaves W0 (and W1) = unsymmetrised accumulators; G = W + W^T.  mixed: quartets whose
// Schwarz bound Qab Qcd is below c->fp32_thr are evaluated in FP32.
int fock_build(uno_ctx* c, bool uhf, double dmax_hint, bool mixed = false);
int eri_emit(uno_ctx* c, long cap, double* val, int* ijkl, long* count);

// dense symmetric eigensolver on device data: A (destroyed) -> Z (eigenvectors in
// columns), evals ascending (host)
int eigh_device(uno_ctx* c, int n, double* dA, double* dZ, double* h_evals);
int eigh_standalone(int n, double* dA, double* dZ, double* h_evals, hipStream_t st);
// the same solve on one caller-provided device work space of eigh_ws_doubles(n) doubles
// (256-byte aligned) and host arrays hev[n], hcl[n]: no allocation
size_t eigh_ws_doubles(int n);

// Pulay DIIS coefficients over a ring of K slots (`filled` in use, `newest` just written)
// from B[i*K + j] = <e_i, e_j>; the oldest vectors are dropped while the differences from
// the newest are near linearly dependent.  Returns the number of vectors used (1: coef
// selects the newest alone).
int diis_coefficients(const double* B, int K, int filled, int newest, double* coef);
int eigh_in_workspace(int n, double* dA, double* dZ, double* h_evals, double* ws, double* hev, int* hcl,
                      hipStream_t st);

// dh (optional, host packed): added to the core Hamiltonian (model potentials)
int scf_run(uno_ctx* c, int nelec, const uno_scf_opts& o, uno_scf_result& r, double* evals, double* P, double* C,
            const double* dh = nullptr);
// -sum_k c_k (ab|g_k): the positron's six-term Gaussian charge at R against every basis
// function pair (GDPMInts.cpp:6-146), packed
int gdpm_matrix(uno_ctx* c, const double* R, std::vector<double>& dh);

// learned density guess (guess.cpp): descriptors of the atom pairs within the cutoff
// (kGuessIn floats each; pairs as (A, B), A <= B), target blocks and masks from a dense
// density (kGuessOut floats each), and a dense density from predicted blocks
constexpr int kGuessInDim = 128, kGuessOutDim = 256;
int guess_features(const Basis& B, std::vector<float>& X, std::vector<int>& pairs, std::string& err);
int guess_samples(const Basis& B, const double* P, std::vector<float>& X, std::vector<float>& Y,
                  std::vector<float>& M, std::string& err);
int guess_assemble(const Basis& B, const std::vector<int>& pairs, const float* Y, std::vector<double>& P,
                   std::string& err);
void shell_rotation(int l, const double* Q, double* M);

}  // namespace uno

extern "C" int uno_mlp_forward(uno_mlp* m, long n, const float* X, float* Y);

namespace uno {

}  // namespace uno
