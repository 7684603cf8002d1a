// Device-resident RHF/UHF SCF loop.
//
// Loop semantics are the reference's (RHF.hpp:87-176, UHF.hpp:101-177):
//   update(): G from the *input* density, E = 2tr(PH) + tr(PG) (RHF) or
//             tr(PaH)+tr(PbH)+(tr(PaGa)+tr(PbGb))/2 (UHF), ediff = E - Eold,
//             F = H + G, F' = X^T F X, diagonalise, C = X W, P = C_occ C_occ^T,
//             pdiff = |P - P_in| / n (per spin, summed for UHF);
//   two updates before the loop (UHF zeroes Eold before the second), then
//   { converger; update; stop if converged } with the signed test ediff < eps
//   (and/or pdiff < eps) and damping P = (P + P_in)/2 whenever the energy rose.
// All matrices stay in HBM as dense row-major N x N FP64; the host only reads a few
// scalars per iteration and the tridiagonal for the QL step.
// accel = 2 replaces the damping with Pulay DIIS on F (error X^T(FDS - SDF)X).
//
// Incremental builds (incremental = 1): G is linear in the density, so after a full
// build G[P_k] = G[P_ref] + G[P_k - P_ref]; the change is small late in the SCF and the
// density screen then drops most quartets.  A full build is redone every rebuild_every
// (context knob, default 10) incremental builds and whenever the precision changes.
// Mixed precision (precision = 1): Fock builds evaluate quartets with Schwarz bound
// Qab Qcd below fp32_threshold in FP32 (accumulating in FP64) until |dE| falls below
// the context's mixed_switch (default 3e-6 Eh) or the convergence test passes; then a
// full FP64 build starts the final
// phase, and convergence is accepted only after fp64_final FP64 iterations.
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <future>
#include "engine.hpp"
#include "../hip/kernels.hpp"

namespace uno {

namespace {
using clk = std::chrono::steady_clock;
double secs(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }

}  // namespace

struct ScfBufs {
    DBuf<double> H, T, S, X, F, Fp, Tm, Cw, G[2], C[2], Pold[2], diisF, diisE, diisRow, Etmp, P[2], Pref[2];
};

constexpr int kSp2MinN = 256;           // purification instead of the eigensolve from this size (purify = -1)
constexpr int kSp2MaxIt = 100;
constexpr double kSp2Idem = 1e-8;

int scf_run(uno_ctx* c, int nelec, const uno_scf_opts& o, uno_scf_result& r, double* evals_out, double* P_io,
            double* C_out, const double* dh_extra) {
    const auto t_start = clk::now();
    // a deterministic context: the dense products repeat bitwise as well
    const GemmDetScope gemm_det(c->deterministic != 0);
    std::memset(&r, 0, sizeof(r));
    const int n = c->n;
    const long nn = (long)n * n;
    const long n2 = (long)n * (n + 1) / 2;
    const bool uhf = (nelec % 2) != 0;
    const int ns = uhf ? 2 : 1;
    const int nocc[2] = {uhf ? nelec - nelec / 2 : nelec / 2, nelec / 2};
    if (nelec <= 0 || nocc[0] > n) {
        set_error("electron count does not fit the basis");
        return UNO_EINVAL;
    }
    if (o.guess == 2 && !c->guess_model) {
        set_error("guess 2 needs a density-guess network (uno_ctx_set_guess)");
        return UNO_EINVAL;
    }
    if (o.precision < 0 || o.precision > 1 || o.incremental < 0 || o.incremental > 1 || o.fp64_final < 0) {
        set_error("bad precision / incremental / fp64_final option");
        return UNO_EINVAL;
    }
    hipStream_t st = c->st;
    double eps = o.eps;
    const double xeps = DBL_EPSILON * n * n * 0.5;   // Basis.hpp:249-254
    if (eps < xeps && o.single_start != 2) eps = xeps;
    ScfBufs b;
    for (DBuf<double>* p : {&b.H, &b.T, &b.S, &b.X, &b.F, &b.Fp, &b.Tm, &b.Cw, &b.G[0], &b.G[1], &b.C[0], &b.C[1],
                            &b.Pold[0], &b.Pold[1], &b.Etmp, &b.P[0], &b.P[1]})
        UNO_HIP(p->alloc(nn));
    if (o.incremental)
        for (int s = 0; s < ns; ++s) UNO_HIP(b.Pref[s].alloc(nn));
    // The learned guess's host work (pair descriptors, the network on its own stream, the
    // blocks assembled into a dense symmetric matrix) runs on a worker thread while this
    // thread builds the one-electron matrices and X on the device.  The future is declared
    // after everything the task touches, so an early return waits for it first.
    struct GuessJob {
        std::vector<double> D;
        std::string err;
        long np = 0;
        double feat_s = 0.0, net_s = 0.0, asm_s = 0.0;
    } gj;
    std::future<int> guess_job;
    if (o.guess == 2)
        guess_job = std::async(std::launch::async, [&gj, c, n]() -> int {
            const auto g0 = clk::now();
            std::vector<float> X, Y;
            std::vector<int> pairs;
            if (int rc = guess_features(c->B, X, pairs, gj.err)) return rc;
            gj.feat_s = secs(g0, clk::now());
            gj.np = (long)pairs.size() / 2;
            Y.assign((size_t)gj.np * kGuessOutDim, 0.0f);
            if (gj.np > 0)
                if (int rc = uno_mlp_forward(c->guess_model, gj.np, X.data(), Y.data())) {
                    gj.err = get_error();
                    return rc;
                }
            gj.net_s = secs(g0, clk::now()) - gj.feat_s;
            if (int rc = guess_assemble(c->B, pairs, Y.data(), gj.D, gj.err)) return rc;
            std::vector<double>& D = gj.D;
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < i; ++j) {
                    const double v = 0.5 * (D[(size_t)i * n + j] + D[(size_t)j * n + i]);
                    D[(size_t)i * n + j] = D[(size_t)j * n + i] = v;
                }
            gj.asm_s = secs(g0, clk::now()) - gj.feat_s - gj.net_s;
            return UNO_OK;
        });
    // one-electron matrices on the device, once
    if (int rc = oneint_device(c, b.S.p, b.T.p, b.H.p)) return rc;
    if (o.field[0] != 0.0 || o.field[1] != 0.0 || o.field[2] != 0.0) {
        // finite field: H -= E . d with the raw dipole integrals (FField.cpp:6-20)
        std::vector<double> M(9 * (size_t)n2), dh(n2);
        one_electron(c->B, nullptr, nullptr, nullptr, M.data());
        for (long k = 0; k < n2; ++k)
            dh[k] = -(o.field[0] * M[9 * k] + o.field[1] * M[9 * k + 1] + o.field[2] * M[9 * k + 2]);
        UNO_HIP(c->pk.upload(dh.data(), n2, st));
        UNO_HIP(k_unpack(n, c->pk.p, b.Tm.p, st));
        UNO_HIP(k_axpby(nn, 1.0, b.H.p, 1.0, b.Tm.p, b.H.p, st));
    }
    if (dh_extra) {
        UNO_HIP(c->pk.upload(dh_extra, n2, st));
        UNO_HIP(k_unpack(n, c->pk.p, b.Tm.p, st));
        UNO_HIP(k_axpby(nn, 1.0, b.H.p, 1.0, b.Tm.p, b.H.p, st));
    }
    double nuc = 0.0;
    for (int i = 0; i < c->B.ncen; ++i)
        for (int j = i + 1; j < c->B.ncen; ++j) {
            const double* a = &c->B.xyz[3 * i];
            const double* bb = &c->B.xyz[3 * j];
            const double dx = a[0] - bb[0], dy = a[1] - bb[1], dz = a[2] - bb[2];
            nuc += c->B.q[i] * c->B.q[j] / std::sqrt(dx * dx + dy * dy + dz * dz);
        }
    // (verbose 2: wait for the one-electron kernels, so that the time printed below
    // separates them from X)
    if (o.verbose > 1) UNO_HIP(hipStreamSynchronize(st));
    const double s_oneint = secs(t_start, clk::now());
    // X with X^T S X = I.  From kSp2MinN functions on (knob x_newton): the Lowdin
    // S^-1/2 by the coupled Newton-Schulz iteration on S/c (c the Gershgorin bound of S, so
    // the spectrum lies in (0, 1] when S is positive definite): Y0 = S/c, Z0 = I,
    // T = (3I - ZY)/2, Y <- YT, Z <- TZ; Z -> (S/c)^-1/2 quadratically once |I - ZY| < 1,
    // three FP64 MFMA GEMMs per step instead of an eigensolve.  Every quantity the SCF
    // returns (energies, densities, orbital energies, MO coefficients up to the sign and
    // rotations within degenerate levels) is the same for any X with X^T S X = I.  When the
    // iteration does not reach its residual (S singular or indefinite, or too ill-
    // conditioned), the canonical X of the reference below takes over, with its checks.
    bool x_done = false;
    c->last_x_steps = -1;
    if (c->x_newton == 1 || (c->x_newton < 0 && n >= kSp2MinN)) {
        DBuf<double> nsw;
        UNO_HIP(nsw.alloc(2 * (size_t)n + 8));
        UNO_HIP(sp2_init(n, b.S.p, b.Tm.p, nsw.p, nsw.p + 2 * (size_t)n, st));   // Gershgorin bounds of S
        double bnd[2] = {0.0, 0.0};
        UNO_HIP(hipMemcpyAsync(bnd, nsw.p + 2 * (size_t)n, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
        UNO_HIP(hipStreamSynchronize(st));
        const double cs = bnd[1];
        if (cs > 0.0 && std::isfinite(cs)) {
            double* Y = b.Tm.p;
            double* Z = b.Cw.p;
            double* Tt = b.Fp.p;
            double* Yn = b.Etmp.p;
            double* Zn = b.F.p;
            UNO_HIP(k_axpby(nn, 1.0 / cs, b.S.p, 0.0, b.S.p, Y, st));
            UNO_HIP(k_fill(nn, 0.0, Z, st));
            UNO_HIP(k_shift_diag(n, 1.0, Z, st));
            double res_prev = 1e300;
            for (int it = 0; it < 80; ++it) {
                UNO_HIP(gemm_f64(false, false, n, n, n, -0.5, Z, n, Y, n, 0.0, Tt, n, st));   // -ZY/2
                UNO_HIP(k_shift_diag(n, 1.5, Tt, st));                                       // T = (3I - ZY)/2
                // |I - ZY| = 2 |T - I|
                UNO_HIP(k_shift_diag(n, -1.0, Tt, st));
                UNO_HIP(k_absmax(nn, Tt, c->scal.p + 3, st));
                UNO_HIP(k_shift_diag(n, 1.0, Tt, st));
                double res = 0.0;
                UNO_HIP(hipMemcpyAsync(&res, c->scal.p + 3, sizeof(double), hipMemcpyDeviceToHost, st));
                UNO_HIP(hipStreamSynchronize(st));
                res *= 2.0;
                // the residual stays just below 1 for the first steps, while the small
                // eigenvalues of S/c grow by about 1.5x per step; only a blow-up (S
                // indefinite) ends the iteration early
                if (!std::isfinite(res) || res > 1e3) break;
                // converged: at rounding level, or no longer falling once below 1e-10
                if (res < 1e-13 || (res < 1e-10 && res > 0.5 * res_prev)) {
                    UNO_HIP(k_axpby(nn, 1.0 / std::sqrt(cs), Z, 0.0, Z, b.X.p, st));
                    x_done = true;
                    c->last_x_steps = it;
                    if (o.verbose > 1) std::fprintf(stderr, "[scf] X by Newton-Schulz: %d steps, |I - ZY| %.1e\n", it, res);
                    break;
                }
                res_prev = res;
                UNO_HIP(gemm_f64(false, false, n, n, n, 1.0, Y, n, Tt, n, 0.0, Yn, n, st));   // Y T
                UNO_HIP(gemm_f64(false, false, n, n, n, 1.0, Tt, n, Z, n, 0.0, Zn, n, st));   // T Z
                std::swap(Y, Yn);
                std::swap(Z, Zn);
            }
        }
    }
    // X = U s^-1/2 keeping every vector (RHF.hpp:214-233)
    if (!x_done) {
        std::vector<double> ev(n);
        UNO_HIP(hipMemcpyAsync(b.Tm.p, b.S.p, nn * sizeof(double), hipMemcpyDeviceToDevice, st));
        const auto tx = clk::now();
        if (int rc = eigh_device(c, n, b.Tm.p, b.Cw.p, ev.data())) return rc;
        if (o.verbose > 1) std::fprintf(stderr, "[scf] eigensolve of S %.2f ms\n", 1e3 * secs(tx, clk::now()));
        for (int j = 0; j < n; ++j) {
            double f = ev[j];
            if (std::fabs(f) < 1e-7 && f < 0.0) f = -f;
            if (f == 0.0) { set_error("Zero Eigenvalue in Overlap Matrix"); return UNO_ESINGULAR; }
            if (f < 0.0) { set_error("Negative Eigenvalue in Overlap Matrix"); return UNO_ESINGULAR; }
            ev[j] = 1.0 / std::sqrt(f);
        }
        // column j of U scaled by s_j^-1/2 on the device
        UNO_HIP(c->pk.upload(ev.data(), n, st));
        UNO_HIP(k_scale_cols(n, b.Cw.p, c->pk.p, b.X.p, st));
    }
    if (o.verbose > 1)
        std::fprintf(stderr, "[scf] setup: buffers and one-electron matrices %.2f ms, X %.2f ms\n", 1e3 * s_oneint,
                     1e3 * (secs(t_start, clk::now()) - s_oneint));

    double* P[2] = {b.P[0].p, b.P[1].p};
    std::vector<double> evals(2 * (size_t)n, 0.0);
    double diag_s = 0.0, fock_s = 0.0, diis_s = 0.0;

    // F (device, n x n) -> P[s] with nocc electrons; keeps C[s] and the eigenvalues
    // (write_P false: only C[s] and the eigenvalues)
    auto diag_to_P = [&](const double* Fd, int s, int no, bool write_P = true) -> int {
        const auto t0 = clk::now();
        UNO_HIP(gemm_f64(false, false, n, n, n, 1.0, Fd, n, b.X.p, n, 0.0, b.Tm.p, n, st));       // F X
        UNO_HIP(gemm_f64(true, false, n, n, n, 1.0, b.X.p, n, b.Tm.p, n, 0.0, b.Fp.p, n, st));   // X^T F X
        const auto te = clk::now();
        if (int rc = eigh_device(c, n, b.Fp.p, b.Cw.p, evals.data() + (size_t)s * n)) return rc;
        if (o.verbose > 1) std::fprintf(stderr, "[scf] eigensolve %.2f ms\n", 1e3 * secs(te, clk::now()));
        UNO_HIP(gemm_f64(false, false, n, n, n, 1.0, b.X.p, n, b.Cw.p, n, 0.0, b.C[s].p, n, st));  // C = X W
        if (write_P)
            UNO_HIP(gemm_f64(false, true, n, n, no, 1.0, b.C[s].p, n, b.C[s].p, n, 0.0, P[s], n, st)); // C_occ C_occ^T
        diag_s += secs(t0, clk::now());
        return UNO_OK;
    };

    // F (device) -> P[s] by purification, without an eigensolve: F' = X^T F X; from
    // X_0 = (hi I - F') / (hi - lo) (Gershgorin bounds of F'), X <- X^2 or 2X - X^2,
    // whichever trace lies nearer nocc (Niklasson's trace-correcting recursion), until
    // tr(X) - tr(X^2) = sum l(1-l) over X's eigenvalues is below kSp2Idem, then two more
    // steps (each squares the distance of every eigenvalue to 0 or 1); P = X_S P' X_S^T.
    // The result is the projector onto F's nocc lowest eigenvectors, the density the
    // eigensolve gives.  ok = false (P untouched) when the recursion does not settle
    // within kSp2MaxIt steps or ends with the wrong electron count (a level degenerate at
    // the occupation boundary): the caller then diagonalises.  Each step is one launch
    // (hip/linalg.hip, k_sp2_step: the symmetric MFMA product on the upper-triangle tiles
    // with the choice and the next traces fused in); the steps pass their state from launch
    // to launch on the device and the host reads it only after a batch of launches (the
    // previous density's step count plus a margin; steps after convergence return at once).
    DBuf<double> sp2w;
    const bool use_sp2 = c->purify == 1 || (c->purify < 0 && n >= kSp2MinN);
    const int sp2_nb = (n + 63) / 64, sp2_tiles = sp2_nb * (sp2_nb + 1) / 2;
    const size_t sp2_state_d = (sp2_state_bytes() * (kSp2MaxIt + 2) + 7) / 8;
    double *sp2_lohi = nullptr, *sp2_bounds = nullptr, *sp2_part[2] = {nullptr, nullptr};
    void* sp2_states = nullptr;
    if (use_sp2) {
        UNO_HIP(sp2w.alloc(2 * (size_t)n + 4 + 4 * (size_t)sp2_tiles + sp2_state_d));
        sp2_lohi = sp2w.p;
        sp2_bounds = sp2w.p + 2 * (size_t)n;
        sp2_part[0] = sp2_bounds + 4;
        sp2_part[1] = sp2_part[0] + 2 * (size_t)sp2_tiles;
        sp2_states = sp2_part[1] + 2 * (size_t)sp2_tiles;
    }
    bool last_sp2[2] = {false, false};
    int sp2_budget = 64;   // steps enqueued before the first look at the state; then last count + 6
    auto density_sp2 = [&](const double* Fd, int s, int no, bool& ok) -> int {
        ok = false;
        const auto t0 = clk::now();
        UNO_HIP(gemm_f64(false, false, n, n, n, 1.0, Fd, n, b.X.p, n, 0.0, b.Tm.p, n, st));       // F X
        UNO_HIP(gemm_f64(true, false, n, n, n, 1.0, b.X.p, n, b.Tm.p, n, 0.0, b.Fp.p, n, st));   // X^T F X
        UNO_HIP(sp2_init(n, b.Fp.p, b.Cw.p, sp2_lohi, sp2_bounds, st));
        UNO_HIP(sp2_start(n, b.Cw.p, sp2_part[0], sp2_states, st));
        int k = 0, cur = 0, done = 0, steps = 0, budget = sp2_budget;
        double t1 = 0.0;
        while (k < kSp2MaxIt) {
            const int cnt = std::min(budget, kSp2MaxIt - k);
            UNO_HIP(sp2_steps(n, no, kSp2Idem, sp2_states, k, cnt, b.Cw.p, b.Tm.p, sp2_part[0], sp2_part[1], st));
            k += cnt;
            UNO_HIP(sp2_state_read(sp2_states, k, &cur, &done, &steps, &t1, st));
            if (done) break;
            budget = 16;
        }
        r.purify_steps += steps;
        ok = done == 1 && std::fabs(t1 - no) < 1e-6;
        if (ok) {
            sp2_budget = steps + 6;
            double* Xk = cur ? b.Tm.p : b.Cw.p;
            double* T = cur ? b.Cw.p : b.Tm.p;
            UNO_HIP(gemm_f64(false, false, n, n, n, 1.0, b.X.p, n, Xk, n, 0.0, T, n, st));    // X_S P'
            UNO_HIP(gemm_f64(false, true, n, n, n, 1.0, T, n, b.X.p, n, 0.0, P[s], n, st));   // (X_S P') X_S^T
            ++r.purified;
        }
        diag_s += secs(t0, clk::now());
        return UNO_OK;
    };
    // the density of spin s from F: purification where enabled, else (or when it does not
    // settle) the eigensolve
    auto density_of = [&](const double* Fd, int s, int no) -> int {
        bool ok = false;
        if (use_sp2)
            if (int rc = density_sp2(Fd, s, no, ok)) return rc;
        last_sp2[s] = ok;
        if (!ok) return diag_to_P(Fd, s, no);
        return UNO_OK;
    };

    // core-Hamiltonian guess (RHF.hpp:205-211; UHF.hpp:206-215: beta reuses alpha's orbitals)
    auto core_guess = [&]() -> int {
        if (int rc = diag_to_P(b.H.p, 0, nocc[0])) return rc;
        if (uhf) {
            UNO_HIP(gemm_f64(false, true, n, n, nocc[1], 1.0, b.C[0].p, n, b.C[0].p, n, 0.0, P[1], n, st));
            std::memcpy(evals.data() + n, evals.data(), n * sizeof(double));
        }
        return UNO_OK;
    };
    // iteration-0 electronic energy of the densities in P (one full Fock build in the
    // precision of the first SCF build), without touching the SCF's state
    auto energy_of = [&](double& e) -> int {
        double* Din[2] = {uhf ? c->D1.p : c->D0.p, c->D2.p};
        for (int s = 0; s < ns; ++s) UNO_HIP(hipMemcpyAsync(Din[s], P[s], nn * sizeof(double), hipMemcpyDeviceToDevice, st));
        if (uhf) UNO_HIP(k_axpby(nn, 1.0, c->D1.p, 1.0, c->D2.p, c->D0.p, st));
        const bool mixed = o.precision == 1;
        const double tau0 = c->tau;
        if (mixed) c->tau = std::max(tau0, c->tau_mixed);
        const int rc = fock_build(c, uhf, -1.0, mixed);
        c->tau = tau0;
        if (rc) return rc;
        double h[2];
        UNO_HIP(k_fill(2, 0.0, c->scal.p, st));
        for (int s = 0; s < ns; ++s) {
            UNO_HIP(k_sym_w(n, s == 0 ? c->W0.p : c->W1.p, b.G[s].p, st, false));
            UNO_HIP(k_dot(nn, P[s], b.H.p, c->scal.p, c->red.p, st));
            UNO_HIP(k_dot(nn, P[s], b.G[s].p, c->scal.p + 1, c->red.p, st));
        }
        UNO_HIP(hipMemcpyAsync(h, c->scal.p, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
        UNO_HIP(hipStreamSynchronize(st));
        e = uhf ? h[0] + 0.5 * h[1] : 2.0 * h[0] + h[1];
        return UNO_OK;
    };
    // tr(P S) of spin s
    auto trace_ps = [&](int s, double& t) -> int {
        UNO_HIP(k_fill(1, 0.0, c->scal.p + 3, st));
        UNO_HIP(k_dot(nn, P[s], b.S.p, c->scal.p + 3, c->red.p, st));
        UNO_HIP(hipMemcpyAsync(&t, c->scal.p + 3, sizeof(double), hipMemcpyDeviceToHost, st));
        UNO_HIP(hipStreamSynchronize(st));
        return UNO_OK;
    };

    // guess
    if (o.guess == 2) {
        // learned: pair descriptors (host) -> network (GPU) -> blocks assembled in the lab
        // frame, symmetrised; each spin scaled to its electron count (tr(P S) = nocc); the
        // closed-shell density then McWeeny-purified, P <- 3 PSP - 2 PSPSP, twice, and
        // scaled again.  A prediction whose electron count is off by more than 10 %, or
        // whose purification moves away from idempotency, or that is not finite, is
        // replaced by the core-Hamiltonian guess; with guess_check the core-H guess is
        // formed as well and the lower iteration-0 energy wins.
        const auto t0 = clk::now();
        if (int rc = guess_job.get()) {
            set_error(gj.err);
            return rc;
        }
        std::vector<double>& D = gj.D;
        if (o.verbose > 1)
            std::fprintf(stderr, "[scf] guess (worker thread): %ld pairs, descriptors %.2f ms, network %.2f ms, "
                                 "assembly %.2f ms; waited %.2f ms for it\n", gj.np, 1e3 * gj.feat_s, 1e3 * gj.net_s,
                         1e3 * gj.asm_s, 1e3 * secs(t0, clk::now()));
        bool ok = true;
        for (double v : D) ok = ok && std::isfinite(v);
        for (int s = 0; s < ns && ok; ++s) {
            UNO_HIP(hipMemcpyAsync(P[s], D.data(), nn * sizeof(double), hipMemcpyHostToDevice, st));
            double t = 0.0;
            if (int rc = trace_ps(s, t)) return rc;
            const double want = nocc[s];
            if (!(std::fabs(t - 0.5 * nelec) <= 0.1 * 0.5 * nelec) || !(t > 0.0)) { ok = false; break; }
            UNO_HIP(k_axpby(nn, want / t, P[s], 0.0, P[s], P[s], st));
            if (!uhf) {
                double idem[2] = {0.0, 0.0};
                for (int it = 0; it < 2; ++it) {
                    UNO_HIP(gemm_f64(false, false, n, n, n, 1.0, b.S.p, n, P[s], n, 0.0, b.Tm.p, n, st));     // S P
                    UNO_HIP(gemm_f64(false, false, n, n, n, 1.0, P[s], n, b.Tm.p, n, 0.0, b.Fp.p, n, st));    // P S P
                    UNO_HIP(gemm_f64(false, false, n, n, n, 1.0, b.Fp.p, n, b.Tm.p, n, 0.0, b.Cw.p, n, st));  // P S P S P
                    UNO_HIP(k_axpby(nn, 1.0, b.Fp.p, -1.0, P[s], b.Tm.p, st));                              // PSP - P
                    UNO_HIP(k_absmax(nn, b.Tm.p, c->scal.p + 3, st));
                    UNO_HIP(hipMemcpyAsync(&idem[it], c->scal.p + 3, sizeof(double), hipMemcpyDeviceToHost, st));
                    UNO_HIP(k_axpby(nn, 3.0, b.Fp.p, -2.0, b.Cw.p, P[s], st));
                }
                UNO_HIP(hipStreamSynchronize(st));
                if (!(idem[1] <= idem[0]) || !(idem[1] < 0.25)) ok = false;
                if (ok) {
                    if (int rc = trace_ps(s, t)) return rc;
                    if (!(t > 0.0)) ok = false;
                    else UNO_HIP(k_axpby(nn, want / t, P[s], 0.0, P[s], P[s], st));
                }
            }
        }
        if (ok && o.guess_check) {
            double el = 0.0, ec = 0.0;
            if (int rc = energy_of(el)) return rc;
            for (int s = 0; s < ns; ++s)
                UNO_HIP(hipMemcpyAsync(b.Pold[s].p, P[s], nn * sizeof(double), hipMemcpyDeviceToDevice, st));
            if (int rc = core_guess()) return rc;
            if (int rc = energy_of(ec)) return rc;
            r.guess_energy = el + nuc;
            r.guess_energy_coreh = ec + nuc;
            if (el <= ec)
                for (int s = 0; s < ns; ++s)
                    UNO_HIP(hipMemcpyAsync(P[s], b.Pold[s].p, nn * sizeof(double), hipMemcpyDeviceToDevice, st));
            else r.guess_fallback = 1;
        } else if (!ok) {
            r.guess_fallback = 1;
            if (int rc = core_guess()) return rc;
        }
        r.guess_s = secs(t0, clk::now());
    } else if (o.guess == 1 && P_io) {
        for (int s = 0; s < ns; ++s) {
            UNO_HIP(c->pk.upload(P_io + s * n2, n2, st));
            UNO_HIP(k_unpack(n, c->pk.p, P[s], st));
        }
    } else {
        if (int rc = core_guess()) return rc;
    }

    // DIIS state
    const int kDiis = 8;
    std::vector<double> Bm;
    int ndiis = 0, diis_next = 0;
    if (o.accel == 2) {
        UNO_HIP(b.diisF.alloc((size_t)kDiis * ns * nn));
        UNO_HIP(b.diisE.alloc((size_t)kDiis * ns * nn));
        Bm.assign(kDiis * kDiis, 0.0);
    }

    double energy = 0.0, eold = 0.0, ediff = 10.0, pdiff = 10.0;
    int iteration = 0;
    double hs[4];
    // build control: full vs incremental, mixed vs FP64 phase
    bool need_full = true;
    int nincr = 0;
    int phase = o.precision == 1 ? 0 : 1;   // 0 mixed, 1 FP64
    int fp64_updates = 0;
    auto update = [&]() -> int {
        const auto t0 = clk::now();
        const bool full = need_full || !o.incremental || nincr >= c->rebuild_every;
        const bool mixed = phase == 0;
        // densities of this build: P, or its change since the last build (G is linear in P)
        double* Din[2] = {uhf ? c->D1.p : c->D0.p, c->D2.p};
        for (int s = 0; s < ns; ++s) {
            if (full) UNO_HIP(hipMemcpyAsync(Din[s], P[s], nn * sizeof(double), hipMemcpyDeviceToDevice, st));
            else UNO_HIP(k_axpby(nn, 1.0, P[s], -1.0, b.Pref[s].p, Din[s], st));
        }
        if (uhf) UNO_HIP(k_axpby(nn, 1.0, c->D1.p, 1.0, c->D2.p, c->D0.p, st));
        {
            // the mixed phase screens at tau_mixed (its error only steers the early
            // iterations; the FP64 phase starts with a full rebuild at the context's tau)
            const double tau0 = c->tau;
            if (mixed) c->tau = std::max(tau0, c->tau_mixed);
            const int rc = fock_build(c, uhf, -1.0, mixed);
            c->tau = tau0;
            if (rc) return rc;
        }
        for (int s = 0; s < ns; ++s) {
            UNO_HIP(k_sym_w(n, s == 0 ? c->W0.p : c->W1.p, b.G[s].p, st, !full));
            if (o.incremental)
                UNO_HIP(hipMemcpyAsync(b.Pref[s].p, P[s], nn * sizeof(double), hipMemcpyDeviceToDevice, st));
        }
        if (full) { need_full = false; nincr = 0; ++r.nfull; } else ++nincr;
        if (!mixed) ++fp64_updates;
        r.quartets32 += c->last_quartets32;
        UNO_HIP(k_fill(4, 0.0, c->scal.p, st));
        for (int s = 0; s < ns; ++s) {
            UNO_HIP(k_dot(nn, P[s], b.H.p, c->scal.p, c->red.p, st));
            UNO_HIP(k_dot(nn, P[s], b.G[s].p, c->scal.p + 1, c->red.p, st));
        }
        UNO_HIP(hipMemcpyAsync(hs, c->scal.p, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
        UNO_HIP(hipStreamSynchronize(st));
        const double build_s = secs(t0, clk::now());
        fock_s += build_s;
        if (o.verbose > 1)
            std::fprintf(stderr,
                         "[scf] build %d: %s %s, %.2f ms (host: bound %.2f, spaces %.2f, staging %.2f, enqueue %.2f), "
                         "%.3e shell quartets (%.3e FP32)\n",
                         r.nfock + 1, full ? "full" : "incremental", mixed ? "mixed" : "FP64", 1e3 * build_s, c->last_ms[0],
                         c->last_ms[1], c->last_ms[2], c->last_ms[3], c->last_quartets, c->last_quartets32);
        r.quartets += c->last_quartets;
        ++r.nfock;
        energy = uhf ? hs[0] + 0.5 * hs[1] : 2.0 * hs[0] + hs[1];
        ediff = energy - eold;
        eold = energy;
        // F per spin (+ DIIS extrapolation)
        double* Fs[2] = {b.F.p, b.Etmp.p};
        for (int s = 0; s < ns; ++s) UNO_HIP(k_axpby(nn, 1.0, b.H.p, 1.0, b.G[s].p, Fs[s], st));
        if (o.accel == 2) {
            const auto td = clk::now();
            const int slot = diis_next;
            diis_next = (diis_next + 1) % kDiis;
            if (ndiis < kDiis) ++ndiis;
            for (int s = 0; s < ns; ++s) {
                double* Fh = b.diisF.p + ((size_t)slot * ns + s) * nn;
                double* Eh = b.diisE.p + ((size_t)slot * ns + s) * nn;
                UNO_HIP(hipMemcpyAsync(Fh, Fs[s], nn * sizeof(double), hipMemcpyDeviceToDevice, st));
                // e = F P S - S P F, then X^T e X
                UNO_HIP(gemm_f64(false, false, n, n, n, 1.0, Fs[s], n, P[s], n, 0.0, b.Tm.p, n, st));
                UNO_HIP(gemm_f64(false, false, n, n, n, 1.0, b.Tm.p, n, b.S.p, n, 0.0, b.Fp.p, n, st));
                UNO_HIP(k_sym_w(n, b.Fp.p, b.Cw.p, st));   // FPS + (FPS)^T = FPS + SPF
                UNO_HIP(k_axpby(nn, 2.0, b.Fp.p, -1.0, b.Cw.p, b.Fp.p, st));   // FPS - SPF
                UNO_HIP(gemm_f64(false, false, n, n, n, 1.0, b.Fp.p, n, b.X.p, n, 0.0, b.Tm.p, n, st));
                UNO_HIP(gemm_f64(true, false, n, n, n, 1.0, b.X.p, n, b.Tm.p, n, 0.0, Eh, n, st));
            }
            // B matrix row/column of the new slot
            std::vector<double> row(kDiis, 0.0);
            DBuf<double>& dr = b.diisRow;
            UNO_HIP(dr.alloc(kDiis));
            UNO_HIP(k_fill(kDiis, 0.0, dr.p, st));
            for (int j = 0; j < ndiis; ++j)
                for (int s = 0; s < ns; ++s)
                    UNO_HIP(k_dot(nn, b.diisE.p + ((size_t)slot * ns + s) * nn, b.diisE.p + ((size_t)j * ns + s) * nn,
                                  dr.p + j, c->red.p, st));
            UNO_HIP(hipMemcpyAsync(row.data(), dr.p, kDiis * sizeof(double), hipMemcpyDeviceToHost, st));
            UNO_HIP(hipStreamSynchronize(st));
            for (int j = 0; j < ndiis; ++j) Bm[slot * kDiis + j] = Bm[j * kDiis + slot] = row[j];
            if (ndiis >= 2) {
                double cf[kDiis];
                if (diis_coefficients(Bm.data(), kDiis, ndiis, slot, cf) >= 2) {
                    for (int s = 0; s < ns; ++s) {
                        UNO_HIP(k_fill(nn, 0.0, Fs[s], st));
                        for (int i = 0; i < kDiis; ++i)
                            if (cf[i] != 0.0)
                                UNO_HIP(k_axpby(nn, 1.0, Fs[s], cf[i], b.diisF.p + ((size_t)i * ns + s) * nn, Fs[s], st));
                    }
                }
            }
            diis_s += secs(td, clk::now());
        }
        pdiff = 0.0;
        for (int s = 0; s < ns; ++s) {
            UNO_HIP(hipMemcpyAsync(b.Pold[s].p, P[s], nn * sizeof(double), hipMemcpyDeviceToDevice, st));
            // spin-beta Fock lives in Etmp; copy to F's slot is unnecessary: diag reads Fs[s]
            if (int rc = density_of(Fs[s], s, nocc[s])) return rc;
            UNO_HIP(k_diffnorm(nn, n, P[s], b.Pold[s].p, c->scal.p + 2 + s, c->red.p, st));
        }
        UNO_HIP(hipMemcpyAsync(hs + 2, c->scal.p + 2, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
        UNO_HIP(hipStreamSynchronize(st));
        pdiff = hs[2] + (uhf ? hs[3] : 0.0);
        ++iteration;
        if (o.verbose)
            std::fprintf(stderr, "Iteration     =  %6d\nEnergy        =  %25.16le\nDelta Energy  =  %25.15le\n"
                                 "Delta Density =  %25.15le\n\n",
                         iteration, energy + nuc, ediff, pdiff);
        return UNO_OK;
    };
    auto test = [&]() {
        if (o.cflag == 0) return ediff < eps;
        if (o.cflag == 1) return pdiff < eps;
        if (o.cflag == 3) return std::fabs(ediff) < eps;   // |dE| < eps (no reference counterpart)
        return pdiff < eps && ediff < eps;
    };
    auto converged = [&]() {
        const bool pass = test();
        if (phase == 0) {
            // leave the mixed phase: the next build is a full FP64 one
            if (pass || std::fabs(ediff) < c->mixed_switch) {
                phase = 1;
                need_full = true;
            }
            return false;
        }
        return pass && fp64_updates >= (o.precision == 1 ? std::max(o.fp64_final, 1) : 0);
    };
    iteration = 0;
    eold = 0.0;
    r.setup_s = secs(t_start, clk::now());
    if (int rc = update()) return rc;
    const double init_energy = energy + nuc;
    if (!o.single_start) {
        iteration = 1;
        if (uhf) eold = 0.0;   // UHF.hpp:157-158
        if (int rc = update()) return rc;
    }
    bool conv = false;
    while (iteration < o.maxits) {
        if (o.accel != 2 && !(ediff < 0.0))
            for (int s = 0; s < ns; ++s) UNO_HIP(k_axpby(nn, 0.5, P[s], 0.5, b.Pold[s].p, P[s], st));
        if (int rc = update()) return rc;
        if (converged()) { conv = true; break; }
    }
    // orbital energies and vectors of the last Fock matrix where its density came from
    // purification (the density itself is not touched)
    {
        double* Flast[2] = {b.F.p, b.Etmp.p};
        for (int s = 0; s < ns; ++s)
            if (last_sp2[s])
                if (int rc = diag_to_P(Flast[s], s, nocc[s], false)) return rc;
    }
    // kinetic energy and outputs
    UNO_HIP(k_fill(1, 0.0, c->scal.p, st));
    for (int s = 0; s < ns; ++s) UNO_HIP(k_dot(nn, P[s], b.T.p, c->scal.p, c->red.p, st));
    double kin = 0.0;
    UNO_HIP(hipMemcpyAsync(&kin, c->scal.p, sizeof(double), hipMemcpyDeviceToHost, st));
    UNO_HIP(hipStreamSynchronize(st));
    if (!uhf) kin *= 2.0;
    if (evals_out) std::memcpy(evals_out, evals.data(), sizeof(double) * ns * n);
    if (P_io)
        for (int s = 0; s < ns; ++s) {
            UNO_HIP(k_pack(n, P[s], c->pk.p, st));
            UNO_HIP(hipMemcpyAsync(P_io + s * n2, c->pk.p, n2 * sizeof(double), hipMemcpyDeviceToHost, st));
        }
    if (C_out)
        for (int s = 0; s < ns; ++s)
            UNO_HIP(hipMemcpyAsync(C_out + s * nn, b.C[s].p, nn * sizeof(double), hipMemcpyDeviceToHost, st));
    UNO_HIP(hipStreamSynchronize(st));
    r.energy = energy + nuc;
    r.init_energy = init_energy;
    r.ediff = ediff;
    r.pdiff = pdiff;
    r.nucrep = nuc;
    r.kinetic = kin;
    r.electronic = energy;
    r.iterations = iteration;
    r.converged = conv ? 1 : 0;
    r.uhf = uhf ? 1 : 0;
    r.fock_s = fock_s;
    r.diag_s = diag_s;
    r.diis_s = diis_s;
    r.wall_s = secs(t_start, clk::now());
    r.wall99_s = r.wall_s;
    return conv ? UNO_OK : UNO_ENOCONV;
}

}  // namespace uno
