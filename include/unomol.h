/*
 * unomol.h -- C ABI of libunomol_hip.so, the MI355X (gfx950) RHF/UHF engine.
 *
 * This is the drop-in boundary for the integrals + Fock build + diagonalisation path
 * of PatNichols/unomol.  Each entry point names the reference interface it replaces
 * (file:line in the reference tree).  Conventions kept from the reference:
 *   - packed symmetric storage, lower triangle row-wise: idx(i,j) = i(i+1)/2 + j, i>=j
 *     (TwoElectronInts.cpp:705-711);
 *   - P is the half density for RHF and the per-spin density for UHF;
 *   - dense matrices are row-major N x N; eigenvectors are columns; eigenvalues ascend
 *     (SymmPack.cpp:272-288);
 *   - Fock/G builds ACCUMULATE into caller-zeroed arrays (RHF.hpp:89-93);
 *   - rsp destroys its input, sp_trans works in place (SymmPack.hpp:13-26).
 * Every function returns an int status (UNO_OK == 0); uno_last_error() gives the
 * message of the last failure on the calling thread.  The reference calls exit() on
 * failure; a C++ shim maps nonzero status to its fatal_error() (Util.cpp:6-9).
 * A context is single-stream and not thread-safe, like the reference objects.
 */
#ifndef UNOMOL_H
#define UNOMOL_H
#ifdef __cplusplus
extern "C" {
#endif

#define UNO_OK        0
#define UNO_EINVAL    1   /* bad argument / malformed basis */
#define UNO_EDEVICE   2   /* HIP runtime error or no gfx950 device */
#define UNO_ENOMEM    3   /* device allocation failed */
#define UNO_ENOCONV   4   /* SCF reached maxits (results still returned) */
#define UNO_ESINGULAR 5   /* zero/negative overlap eigenvalue (RHF.hpp:222-223) */
#define UNO_EIO       6   /* file could not be read/written */
#define UNO_EEIGEN    7   /* tridiagonal QL failed to converge (SymmPack.cpp:282-285) */

typedef struct uno_ctx uno_ctx;

/* POD mirror of the reference Basis/Shell/Center (Basis.hpp:17-352).  With normalized 0
 * the coefficients are the raw deck values and the library normalises them as
 * Shell::normalize does (Basis.hpp:56-75); with normalized 1 they are taken as already
 * normalised, as a reference Basis holds them after its constructor (Basis.hpp:246). */
typedef struct uno_basis {
    int nshell, ncen;
    const int* sh_l;      /* [nshell] */
    const int* sh_npr;    /* [nshell] */
    const int* sh_cen;    /* [nshell] */
    const int* sh_poff;   /* [nshell] offset into alpha/coef */
    const double* alpha;
    const double* coef;
    const double* xyz;    /* [3*ncen] bohr */
    const double* charge; /* [ncen] */
    int normalized;       /* 0: raw deck coefficients, 1: already normalised */
} uno_basis;

const char* uno_last_error(void);

/* Replaces TwoElectronInts::TwoElectronInts(const Basis&, int start_shell, const string&)
 * (TwoElectronInts.hpp:83-88) and its calculate() (TwoElectronInts.cpp:511-697): instead
 * of storing ERIs it uploads the shell-pair data and computes Schwarz bounds on the
 * device (direct SCF).  start_shell restricts bra shells to index >= start_shell as the
 * reference's partial quartet set does (TwoElectronInts.cpp:541). */
int uno_ctx_create(const uno_basis* basis, int start_shell, uno_ctx** out);
/* Replaces TwoElectronInts::recalculate(const Basis&) (TwoElectronInts.hpp:94-96). */
int uno_ctx_rebuild(uno_ctx* ctx, const uno_basis* basis);
int uno_ctx_destroy(uno_ctx* ctx);
/* Tuning knobs (no reference counterpart): "schwarz" (quartet screening threshold on
 * Qab*Qcd*max|D|, default 1e-13), "schwarz_mixed" (the threshold of uno_scf_run's
 * mixed-precision phase, never below "schwarz"; default 3e-9: that phase ends with a
 * full FP64 rebuild at "schwarz"), "mixed_switch" (|dE| in Eh ending that phase; default
 * 3e-6), "x_newton" (X = S^-1/2 by Newton-Schulz instead of the canonical U s^-1/2: -1,
 * the default, from 256 functions on; 0 never; 1 always; the canonical X with its checks
 * takes over when the iteration does not converge; uno_ctx_get "x_newton_steps" reads the
 * last SCF's step count, -1 for the canonical X), "rebuild_every" (incremental builds
 * between full rebuilds; default 10), "deterministic" (1: Fock builds accumulate J and K
 * in 128-bit fixed point with integer atomics, every bra as wave tasks, and uno_scf_run's
 * dense products run without split K, so the same input gives bitwise the same G, energy,
 * orbital energies and density in every run; default 0, FP64 atomics),
 * "launch_streams" (1-4: the streams a Fock build's ERI launches are spread over, the
 * context's own and up to three workers at the default priority, joined before the build
 * returns; default 3, the fastest of 1-4 at 64 waters),
 * "fp32_threshold" (quartets whose bound is below it
 * run in FP32 when "precision" is 1; default 1e-6), "precision" (0 = FP64, 1 = mixed, for the
 * builds of uno_fock_rhf/uno_fock_uhf; uno_scf_run takes its own option),
 * "prim_cut" (bare primitive prefactor cut, default 1e-12 as TwoElectronInts.cpp:428),
 * "skip_center" (a centre left out of the nuclear attraction, as the reference leaves out
 * the DPM positron, OneElectronInts.cpp:43; default -1, none), "row_merge" (1: bras with
 * many kets run as blocks that merge their exchange contributions in LDS; 0: every bra
 * as per-wave tasks; default 1), "row_min_kets" (the ket count from which a bra is
 * row-merged; default 192), "row_variant" (LDS layout of the row-merged blocks: 1 every
 * bra row and the Rys table, 2 every bra row, 3 the first bra shell's rows; 0, the
 * default, takes the first that fits the LDS budget; a forced one that does not fit
 * falls back to that choice), "row_density_lds" (the bra's exchange density rows staged
 * in LDS beside the merged rows: -1, the default, wherever they fit the budget and the
 * autotuner does not prefer otherwise; 0 never; 1 wherever they fit), "rys2_exact" (0, the
 * default: the two-root classes use the reference's own two-root quadrature,
 * Rys.cpp:413-631, which for 15 < x <= 33 departs from the exact quadrature by up to
 * 3.5e-4 relative, so that ERIs match the reference to 1e-12; 1: the exact quadrature),
 * "purify" (uno_scf_run's densities between Fock builds: 1 by trace-correcting
 * purification of X^T F X with FP64 MFMA GEMMs, the eigensolve only for the final orbital
 * energies and vectors or where purification does not settle; 0 always the eigensolve;
 * -1, the default, purification from 256 functions). */
int uno_ctx_set(uno_ctx* ctx, const char* key, double value);
/* Reads the knobs above and, read-only: "last_quartets" (shell quartets evaluated by the
 * last Fock build), "last_quartets32" (of which in FP32), "pairs" (significant shell
 * pairs), "pair_runs" (runs of the pair array: one per angular pair class and
 * primitive-count bucket, at most 36), "quartet_launches" (ERI launches per Fock build),
 * and the context build's host milliseconds: "ctx_ms_pairs" (pair and primitive records
 * and their upload), "ctx_ms_schwarz" (Schwarz bounds on the device), "ctx_ms_runs"
 * (sorted runs, applied on the device) and "ctx_ms_rest". */
int uno_ctx_get(uno_ctx* ctx, const char* key, double* value);
int uno_ctx_norb(const uno_ctx* ctx);

/* Replaces TwoElectronInts::formGmatrix(const double* P, double* G)
 * (TwoElectronInts.cpp:822-844): G += 2J[P] - K[P], packed host arrays. */
int uno_fock_rhf(uno_ctx* ctx, const double* P, double* G);
/* Replaces TwoElectronInts::formGmatrix(PA, PB, GA, GB) (TwoElectronInts.cpp:846-869):
 * Gs += J[PA+PB] - K[Ps]. */
int uno_fock_uhf(uno_ctx* ctx, const double* PA, const double* PB, double* GA, double* GB);

/* Replaces SymmPack::rsp(n, a, z, evals, tmp) (SymmPack.cpp:272-288). */
int uno_eigh_packed(int n, double* a, double* z, double* evals);
/* No counterpart in the reference: uno_eigh_packed for callers that cannot allocate device
 * memory (code under stream capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, results and errors as uno_eigh_packed, for n <= 256; its device work
 * space is a fixed 5.1 MiB array per GPU inside the library (no hipMalloc/hipFree, no
 * stream creation: it runs on the null stream).  Larger n returns UNO_EINVAL.  Calls that
 * overlap on one GPU share that array and race; it is not re-entrant. */
int uno_eigh_packed_noalloc(int n, double* a, double* z, double* evals);
/* Replaces SymmPack::sp_trans(n, a, z, tmp) (SymmPack.cpp:294-348): a <- X^T a X. */
int uno_sp_trans(int n, double* a, const double* x);

/* SCF options/results for the device-resident loop.  With accel 0 and guess 0 the loop
 * is the reference's (RHF.hpp:114-176 / UHF.hpp:137-177: two updates, then damping by
 * averaging when the energy rose, signed convergence test). */
typedef struct uno_scf_opts {
    int maxits;          /* maximum iterations (deck line 2) */
    double eps;          /* convergence threshold (raised to DBL_EPSILON*N^2/2 as Basis.hpp:249-254) */
    int cflag;           /* 0 energy, 1 density, 2 both (RHF.hpp:390-402); 3 |dE| < eps */
    int accel;           /* 0 reference damping, 2 DIIS (Pulay) */
    int guess;           /* 0 core Hamiltonian, 1 caller's P (PMATRIX.DAT), 2 learned (the
                            network set by uno_ctx_set_guess) */
    int precision;       /* 0 FP64, 1 mixed: FP32 for quartets with Schwarz bound below
                            "fp32_threshold" until |dE| < 1e-5, then FP64 */
    int fp64_final;      /* mixed: number of final iterations in full FP64 (default 2: the
                            last two energies are both FP64 energies) */
    int incremental;     /* 1: build F from the density change (Delta-D) */
    int verbose;         /* 1: per-iteration lines on stderr like the reference */
    double field[3];     /* uniform electric field (au): H -= field . dipole integrals
                            (FField.cpp:6-20) */
    int single_start;    /* 1: one update before the loop and no UHF energy reset, as the
                            reference's finite-field and DPM SCFs (RHF.hpp:246-255); 2: the
                            same with eps used as given (no DBL_EPSILON N^2/2 floor) */
    int guess_check;     /* guess 2: 1 = also form the core-Hamiltonian guess (RHF.hpp:205-211)
                            and keep whichever has the lower iteration-0 energy (two extra
                            Fock builds); 0 = keep the learned guess unless it fails its own
                            checks (electron count off by more than 10 %, purification
                            diverging, non-finite values) */
} uno_scf_opts;

typedef struct uno_scf_result {
    double energy;        /* total energy of the last input density (reference semantics) */
    double init_energy;   /* iteration-0 total energy */
    double ediff, pdiff;  /* last signed energy change and density difference norm */
    double nucrep, kinetic, electronic;
    int iterations, converged, uhf, nfock;
    double wall_s, fock_s, diag_s;   /* seconds: whole SCF, Fock builds, eigen/transforms */
    double quartets;      /* shell quartets evaluated over all Fock builds */
    double quartets32;    /* of which evaluated in FP32 (mixed precision) */
    int nfull;            /* full (not incremental) Fock builds */
    double guess_s;       /* seconds spent on the learned guess (guess = 2) */
    double wall99_s;      /* batched runs: seconds until 99 % of the molecules had stopped
                             (converged or at maxits); single runs: wall_s */
    double setup_s;       /* seconds before the first update: buffers, one-electron
                             matrices, X, the guess */
    double diis_s;        /* seconds in the DIIS extrapolation (accel = 2) */
    int guess_fallback;   /* guess 2: 1 when the core-Hamiltonian guess replaced the learned one */
    double guess_energy;  /* guess 2: iteration-0 energy of the learned guess when guess_check
                             evaluated it (else 0) */
    double guess_energy_coreh;  /* the same for the core-Hamiltonian guess */
    int purified;         /* densities formed by purification instead of an eigensolve */
    int purify_steps;     /* purification steps (one GEMM each) over all of them */
} uno_scf_result;

void uno_scf_defaults(uno_scf_opts* opts);
/* Runs the SCF for nelec electrons (odd => UHF, Unomol.cc:13).
 * evals: [N] RHF, [2N] UHF (alpha then beta).  P: packed [N(N+1)/2] RHF or [2*...] UHF;
 * read as the guess when opts->guess == 1, always written with the final density.
 * C: optional [N*N] (RHF) or [2*N*N] (UHF) MO coefficients (columns). */
int uno_scf_run(uno_ctx* ctx, int nelec, const uno_scf_opts* opts, uno_scf_result* res,
                double* evals, double* P, double* C);

/* Debug/parity: every canonical ERI (i>=j, i>=k, k>=l, i==k => j>=l) in the reference's
 * order (TwoElectronInts.cpp:541-672), evaluated on the device.  cap==0 returns the
 * count only. */
int uno_eri_emit(uno_ctx* ctx, long cap, double* val, int* ijkl, long* count);
/* Debug/parity: Rys roots u=t^2 and weights from the device tables (n<=5). */
int uno_rys_eval(int n, int npts, const double* x, double* u, double* w);
/* Debug/parity: the same from the single-precision tables of the FP32 quartets
 * (x rounded to float; results widened to double). */
int uno_rys_eval_f32(int n, int npts, const double* x, double* u, double* w);
/* Debug/parity: two roots from the reference-compatible table that the ERI kernels use
 * by default (the reference's Rys::root2, Rys.cpp:413-631, including its 15 < x <= 33
 * behaviour); fp32 != 0 selects the single-precision table. */
int uno_rys_eval_ref2(int fp32, int npts, const double* x, double* u, double* w);
/* Debug/parity: Schwarz bounds sqrt(max|(ab|ab)|) per shell pair a>=b, packed order. */
int uno_schwarz(uno_ctx* ctx, double* q);

/* Autotuner (no reference counterpart): times every row-merge configuration (LDS variant,
 * the ket count from which a bra is row-merged, or wave tasks only) of every quartet space
 * of one Fock build on the given density (PB null: RHF P; else UHF Pa, Pb; packed, as
 * uno_fock_rhf/uhf) and keeps the fastest per space, for this context and for every later
 * context of the same basis layout in this process.  mixed = 1 tunes the mixed-precision
 * build's spaces.  best_ms (optional) receives the tuned build's summed kernel time.
 * G is not produced; the next Fock build starts from scratch. */
int uno_autotune(uno_ctx* ctx, const double* P, const double* PB, int mixed, double* best_ms);

/* Profiling of the ERI/Fock kernels (no reference counterpart; the reference only
 * prints stage timers, TwoElectronInts.cpp:690-695).  enable=1 resets and turns on a
 * profiling mode in which every Fock build records, per quartet class (bra pair class
 * X, ket pair class Y, index 6*X+Y; pair classes ss ps pp ds dp dd), the kernel time
 * (HIP events, ms), launches, and the primitive and shell quartets evaluated.
 * enable=0 turns it off (and clears the stats); stats are read with enable=-1, which
 * waits for the context's stream: builds record their events and counters without
 * synchronising, and the read folds them in.  Arrays hold 432 entries,
 * index 72*kind + 36*prec + 6*X + Y: kind 0 the per-wave-task (or per-quartet) kernels,
 * kinds 1-3 the row-merged kernels (1: every bra row and the Rys table in LDS, 2: every
 * bra row, 3: the first bra shell's rows), kinds 4-5 the bra-group row kernels (groups
 * of 2 and 4 bras); prec 0 FP64, 1 the FP32 launches of mixed builds. */
int uno_profile(uno_ctx* ctx, int enable, double* ms, double* launches, double* prim_quartets,
                double* shell_quartets);

/* Measurement helper (no reference counterpart): a plain device copy of `bytes`
 * (read once, written once) repeated reps times; *ms gets the total kernel time.  Used
 * to calibrate hardware traffic counters and as an HBM bandwidth reference. */
int uno_copy_probe(unsigned long long bytes, int reps, double* ms);

/* One-electron matrices (host, McMurchie-Davidson; OneElectronInts.cpp:127-202):
 * packed S, T, H = T + V. */
int uno_oneint(const uno_basis* basis, double* S, double* T, double* H);
/* The same S, T, H computed on the device from a context's basis (the SCF uses these). */
int uno_oneint_device(uno_ctx* ctx, double* S, double* T, double* H);
/* Raw dipole/quadrupole integrals about the origin, packed, 9 per element
 * (dx,dy,dz,qxx,qxy,qxz,qyy,qyz,qzz) as written to RMOM.DAT (Moments.cpp:95-186). */
int uno_moment_ints(const uno_basis* basis, double* m);

/* Deck tools.  kind: "water" (monomer), "water_cluster" (n waters), "molecule" (pool
 * index n: H2 HF H2O NH3 CH4 N2 CO CO2 C2H2, seeded distortion).  Writes patin.dat text. */
int uno_deck_generate(const char* kind, int n, unsigned long long seed, const char* basis,
                      int charge, double eps, int maxits, const char* path);

/* Drop-in run of the reference program (Unomol.cc:8-26) on the deck "patin.dat" in dir:
 * writes short.gs.out, scfout.gs.out, PMATRIX.DAT, intsout.out, MOMINTS.DAT, RMOM.DAT,
 * moments.out, mol_dipmom.out there.  opts may be NULL (reference semantics). */
int uno_run_deck(const char* dir, const uno_scf_opts* opts, uno_scf_result* res);

/* ---- Learned initial density (no reference counterpart; the reference's guess is
 * the core Hamiltonian, RHF.hpp:205-211).  A multilayer perceptron maps atom-pair
 * descriptors to density blocks in the first atom's local frame; it is trained and
 * evaluated on the GPU with BF16-input / FP32-accumulate matrix-core kernels.
 * Pairs (A <= B) within 10 bohr; descriptors have 128 floats, blocks 256 (entry
 * (mu, nu) of the atoms' Cartesian functions at 16 mu + nu); H, C, N, O, F with s, p, d
 * shells and at most 16 functions per atom. */
typedef struct uno_mlp uno_mlp;

/* in_dim, width, out_dim: positive multiples of 64; hidden_layers SiLU layers of width
 * width and a linear output layer.  Weights LeCun-normal from seed, biases zero. */
int uno_mlp_create(int in_dim, int width, int hidden_layers, int out_dim, unsigned long long seed,
                   uno_mlp** out);
int uno_mlp_destroy(uno_mlp* m);
int uno_mlp_nparams(const uno_mlp* m, long* n);
/* ndims in: capacity of dims; out: number of layer widths (input, hidden..., output) */
int uno_mlp_dims(const uno_mlp* m, int* ndims, int* dims);
/* FP32 parameters, per layer W [in x out] row-major then b [out] */
int uno_mlp_get_params(uno_mlp* m, float* params);
int uno_mlp_set_params(uno_mlp* m, const float* params);
/* Y [n x out] = network(X [n x in]), host arrays */
int uno_mlp_forward(uno_mlp* m, long n, const float* X, float* Y);
/* masked mean squared error sum M (Y - T)^2 / sum M over n <= 65536 rows and its
 * gradient with respect to every parameter (same order as the parameters) */
int uno_mlp_loss_grad(uno_mlp* m, long n, const float* X, const float* T, const float* M, float* loss,
                      float* grad);
/* Adam (beta 0.9/0.999) with a cosine-decayed rate from lr over `steps` minibatches of
 * `batch` rows (a multiple of 64) drawn by epoch shuffles from seed; *loss = masked MSE
 * of the final network over all n rows */
int uno_mlp_train(uno_mlp* m, long n, const float* X, const float* T, const float* M, int steps, int batch,
                  float lr, unsigned long long seed, float* loss);
int uno_mlp_save(uno_mlp* m, const char* path);
int uno_mlp_load(const char* path, uno_mlp** out);

/* Descriptors [count x 128] and atom pairs [count x 2] of a basis (cap 0: count only). */
int uno_guess_features(const uno_basis* basis, long cap, float* X, int* pairs, long* count);
/* Training samples from a packed (half) density: descriptors, target blocks in the
 * pair frames and their masks [count x 256] (cap 0: count only). */
int uno_guess_samples(const uno_basis* basis, const double* P, long cap, float* X, float* Y, float* M,
                      long* count);
/* Packed density from blocks Y [count x 256] of the pairs [count x 2]. */
int uno_guess_assemble(const uno_basis* basis, long count, const int* pairs, const float* Y, double* P);
/* Network used by uno_scf_run with guess = 2 (not owned; NULL clears). */
int uno_ctx_set_guess(uno_ctx* ctx, uno_mlp* model);
/* Pulay DIIS coefficients (accel = 2; no reference counterpart: the reference damps).
 * B[i*K + j] = <e_i, e_j> over a ring of K slots, `filled` in use, `newest` just written;
 * coef[K] minimises |sum c_i e_i| with sum c_i = 1 over the most recent vectors whose
 * differences from the newest are linearly independent (older ones are dropped until they
 * are); *used = the number of vectors kept (1: coef selects the newest alone). */
int uno_diis_coefficients(const double* B, int K, int filled, int newest, double* coef, int* used);

/* Debug/parity: the (ncart x ncart) map of one shell's Cartesian functions under a
 * rotation Q of the molecule (row-major 3x3; l <= 2): phi'_mu = sum_nu M_mu,nu phi_nu. */
int uno_shell_rotation(int l, const double* Q, double* M);

/* ---- Batched SCF (no reference counterpart: the reference runs one molecule per
 * process).  Many molecules of uneven size (each at most 256 functions), closed-shell
 * and open-shell mixed (odd electron counts run UHF), iterated in lockstep on the
 * device: one ERI launch per class pair across all molecules, grouped GEMMs and a
 * batched Jacobi eigensolver for the dense steps, Pulay DIIS per molecule.  FP64 full
 * builds from the core-Hamiltonian guess; opts->cflag / eps / maxits as for uno_scf_run.
 * The Fock build's quartet lists are built once per run against a density bound (twice each
 * molecule's largest shell-block density; rebuilt if a density outgrows it) with the
 * costliest bras first, the launches skip converged molecules by a device flag (the wave
 * classes take their work from a device queue), and the whole build is captured once as a
 * HIP graph and replayed every iteration. */
typedef struct uno_batch uno_batch;
int uno_batch_create(int nmol, const uno_basis* bases, const int* nelec, uno_batch** out);
int uno_batch_destroy(uno_batch* batch);
int uno_batch_nmol(const uno_batch* batch);
/* Options of the batched Fock build: "graph" (1: capture and replay, 0: plain launches),
 * "queue" (0: one block per task; 1: the one-quartet-per-wave classes take their quartet
 * blocks from a device work queue; 2: the lane classes' 64-ket tasks as well),
 * "launch_streams" (1-4: streams the list launches are spread over, forked and joined
 * inside the build, parallel branches of the captured graph; default 3).  uno_batch_get
 * also reads the last run's "space_builds", "graph_builds", "graph_replays" and
 * "launches_per_build". */
int uno_batch_set(uno_batch* batch, const char* key, double value);
int uno_batch_get(const uno_batch* batch, const char* key, double* value);
/* The final density of molecule m after uno_batch_run, dense row-major N x N: the RHF half
 * density, or alpha then beta (2 N x N) for a UHF molecule. */
int uno_batch_density(const uno_batch* batch, int m, double* P);
/* res[nmol] per molecule (wall_s, fock_s and quartets are the batch totals);
 * energies (optional) [nmol] total energies. */
int uno_batch_run(uno_batch* batch, const uno_scf_opts* opts, uno_scf_result* res, double* energies);

#ifdef __cplusplus
}
#endif
#endif
