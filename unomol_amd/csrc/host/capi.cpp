// C ABI (include/unomol.h) and the drop-in deck runner with the reference's outputs.
#include <cmath>
#include <fstream>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>
#include "deck.hpp"
#include "engine.hpp"
#include "../hip/kernels.hpp"

using namespace uno;

namespace {

void basis_view(const Deck& d, uno_basis& b) {
    b.nshell = d.nshell;
    b.ncen = d.ncen;
    b.sh_l = d.sh_l.data();
    b.sh_npr = d.sh_npr.data();
    b.sh_cen = d.sh_cen.data();
    b.sh_poff = d.sh_poff.data();
    b.alpha = d.alpha.data();
    b.coef = d.coef.data();
    b.xyz = d.xyz.data();
    b.charge = d.charge.data();
    b.normalized = 0;
}

// a <- X^T a X on the host-facing boundary, computed on the device
int sp_trans_dev(int n, double* a, const double* x, hipStream_t st) {
    const long nn = (long)n * n, n2 = (long)n * (n + 1) / 2;
    DBuf<double> pk, A, X, T;
    UNO_HIP(pk.upload(a, n2, st));
    UNO_HIP(A.alloc(nn));
    UNO_HIP(T.alloc(nn));
    UNO_HIP(X.upload(x, nn, st));
    UNO_HIP(k_unpack(n, pk.p, A.p, st));
    UNO_HIP(gemm_f64(false, false, n, n, n, 1.0, A.p, n, X.p, n, 0.0, T.p, n, st));
    UNO_HIP(gemm_f64(true, false, n, n, n, 1.0, X.p, n, T.p, n, 0.0, A.p, n, st));
    UNO_HIP(k_pack(n, A.p, pk.p, st));
    UNO_HIP(hipMemcpyAsync(a, pk.p, n2 * sizeof(double), hipMemcpyDeviceToHost, st));
    UNO_HIP(hipStreamSynchronize(st));
    return UNO_OK;
}

FILE* create(const std::string& dir, const char* name) { return std::fopen((dir + "/" + name).c_str(), "w"); }

const char* kRule = "xxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxx\n";

// Writes the reference's output files (RHF.hpp:404-533, UHF.hpp:399-485, Moments.cpp:95-469)
int write_outputs(const std::string& dir, const Deck& d, const Basis& B, const uno_scf_result& r,
                  const std::vector<double>& S, const std::vector<double>& T, const std::vector<double>& H,
                  const std::vector<double>& M, const std::vector<double>& P, const std::vector<double>& ev,
                  const std::vector<double>& C, int maxits) {
    const int n = B.norb;
    const long n2 = (long)n * (n + 1) / 2;
    const bool uhf = r.uhf != 0;
    const int noA = uhf ? d.nelec - d.nelec / 2 : d.nelec / 2, noB = d.nelec / 2;
    FILE* f = create(dir, "intsout.out");
    if (!f) { set_error("could not create intsout.out"); return UNO_EIO; }
    std::fprintf(f, "  UnoMol alpha version \n If you are using this for serious research you are insane!\n\n");
    std::fprintf(f, "Number of Shells   = %12u\nNumber of Orbitals = %12u\nNumber of Centers  = %12u\n", d.nshell, n,
                 d.ncen);
    std::fprintf(f, "Max L value        = %12u\nMax SCF Iterations = %12u\nNumber of Electrons= %12u\n", B.maxl, maxits,
                 uhf ? d.nelec : 2 * (d.nelec / 2));
    std::fprintf(f, " Centers \nCharge     Position(x,y,z  in bohr)\n");
    for (int i = 0; i < d.ncen; ++i)
        std::fprintf(f, "%15.10lf %15.10lf %15.10lf %15.10lf\n", d.charge[i], d.xyz[3 * i], d.xyz[3 * i + 1],
                     d.xyz[3 * i + 2]);
    std::fprintf(f, "\n Shells\n");
    for (size_t s = 0; s < B.sh.size(); ++s) {
        std::fprintf(f, "%3d %3d %3d\n", B.sh[s].npr, B.sh[s].l, B.sh[s].cen);
        for (int p = 0; p < B.sh[s].npr; ++p) std::fprintf(f, " %20.10lf %20.10lf\n", B.sh[s].a[p], B.sh[s].c[p]);
    }
    const char* names[3] = {"Overlap Matrix \n", "Kinetic Energy Matrix \n", "Core Hamiltonian Matrix \n"};
    const std::vector<double>* mats[3] = {&S, &T, &H};
    for (int m = 0; m < 3; ++m) {
        std::fprintf(f, m ? "\n%s" : "\n%s", names[m]);
        long ij = 0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j <= i; ++j, ++ij) std::fprintf(f, "%7u %7u %24.16le\n", i, j, (*mats[m])[ij]);
    }
    std::fclose(f);
    // raw and scaled moment integral files (80-byte MomInts records, Structs.hpp:17-20)
    struct MomRec { double v[9]; unsigned int ijr; unsigned int pad; };
    FILE* fm = create(dir, "MOMINTS.DAT");
    FILE* fr = create(dir, "RMOM.DAT");
    if (!fm || !fr) { set_error("could not create MOMINTS.DAT/RMOM.DAT"); return UNO_EIO; }
    for (size_t is = 0; is < B.sh.size(); ++is)
        for (size_t js = 0; js <= is; ++js)
            for (int a = 0; a < ncart(B.sh[is].l); ++a)
                for (int bb = 0; bb < ncart(B.sh[js].l); ++bb) {
                    const int ir = B.off[is] + a, jr = B.off[js] + bb;
                    if (jr > ir) break;
                    MomRec rec{};
                    rec.ijr = (unsigned)pidx(ir, jr);
                    std::memcpy(rec.v, &M[9 * rec.ijr], 9 * sizeof(double));
                    std::fwrite(&rec, sizeof rec, 1, fr);
                    const double fct = (ir == jr) ? 2.0 : 4.0;
                    for (double& v : rec.v) v *= fct;
                    std::fwrite(&rec, sizeof rec, 1, fm);
                }
    std::fclose(fm);
    std::fclose(fr);
    // PMATRIX.DAT: raw packed doubles (alpha then beta for UHF)
    FILE* fp = create(dir, "PMATRIX.DAT");
    if (!fp) { set_error("could not create PMATRIX.DAT"); return UNO_EIO; }
    std::fwrite(P.data(), sizeof(double), (uhf ? 2 : 1) * n2, fp);
    std::fclose(fp);
    f = create(dir, "short.gs.out");
    std::fprintf(f, "%25.16le\n%25.16le\n%25.16le\n", r.init_energy, r.energy, r.ediff);
    std::fclose(f);
    f = create(dir, "scfout.gs.out");
    const double virial = std::fabs((r.energy - r.kinetic) / r.kinetic / 2.0);
    if (maxits <= r.iterations) std::fprintf(f, "WARNING CONVERGENCE _NOT_ REACHED!\n");
    std::fprintf(f, "Final Iteration              = %12u\n", r.iterations);
    std::fprintf(f, "Hartree Fock Energy          = %25.15le Hartree\n", r.energy);
    std::fprintf(f, "Electronic Energy            = %25.15le Hartree\n", r.electronic);
    std::fprintf(f, "Nuclear Rep. Energy          = %25.15le Hartree\n", r.nucrep);
    std::fprintf(f, "Kinetic Energy               = %25.15le Hartree\n", r.kinetic);
    std::fprintf(f, "Difference in Final Energies = %25.16le Hartree\n", r.ediff);
    std::fprintf(f, "Pmatrix diff norm            = %25.16le\n", r.pdiff);
    std::fprintf(f, "virial                       = %25.16le\n", virial);
    std::fprintf(f, "%s", kRule);
    if (!uhf) {
        std::fprintf(f, "               Orbital Energies\nOrbital Energy          Occupancy\n");
        for (int i = 0; i < n; ++i) std::fprintf(f, "%7u %25.16le %12u\n", i + 1, ev[i], i < noA ? 2 : 0);
    } else {
        std::fprintf(f, "        Alpha Orbital Energies\nOrbital Energy          Occupancy\n");
        for (int i = 0; i < n; ++i) std::fprintf(f, "%7u %25.16le %12u\n", i + 1, ev[i], i < noA ? 1 : 0);
        std::fprintf(f, "%s", kRule);
        std::fprintf(f, "        Beta Orbital Energies\nOrbital Energy          Occupancy\n");
        for (int i = 0; i < n; ++i) std::fprintf(f, "%7u %25.16le %12u\n", i + 1, ev[n + i], i < noB ? 1 : 0);
    }
    std::fprintf(f, "%s", kRule);
    // Mulliken populations: diag of P S (x2 for the RHF half density)
    const int ns = uhf ? 2 : 1;
    std::vector<double> pop((size_t)ns * n, 0.0);
    std::vector<double> ovl((size_t)ns * n * n, 0.0);
    for (int s = 0; s < ns; ++s)
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                double sum = 0.0;
                for (int k = 0; k < n; ++k) sum += P[s * n2 + pidx(i, k)] * S[pidx(k, j)];
                ovl[((size_t)s * n + i) * n + j] = uhf ? sum : 2.0 * sum;
            }
    std::vector<double> net(d.charge);
    for (size_t is = 0; is < B.sh.size(); ++is)
        for (int a = 0; a < ncart(B.sh[is].l); ++a) {
            const int ir = B.off[is] + a;
            for (int s = 0; s < ns; ++s) net[B.sh[is].cen] -= ovl[((size_t)s * n + ir) * n + ir];
        }
    std::fprintf(f, "%s           Mulliken Populations \n", kRule);
    if (!uhf) {
        std::fprintf(f, "Orbital  Net Population\n");
        for (int i = 0; i < n; ++i) std::fprintf(f, "%7u %25.16le\n", i + 1, ovl[(size_t)i * n + i]);
    } else {
        std::fprintf(f, "Orbital  Alpha Population          Beta Population\n");
        for (int i = 0; i < n; ++i)
            std::fprintf(f, "%7u %25.16le %25.16le\n", i + 1, ovl[(size_t)i * n + i], ovl[((size_t)n + i) * n + i]);
    }
    std::fprintf(f, "%sMulliken Atomic Charges \nOrbital   Nuclear Charge   Net Charge \n", kRule);
    for (int i = 0; i < d.ncen; ++i) std::fprintf(f, "%7u %15.10lf %15.10lf \n", i + 1, d.charge[i], net[i]);
    if (d.prt_flag[2]) {
        // Mulliken overlap populations (RHF.hpp:647-657, UHF.hpp:742-753)
        FILE* fo = create(dir, "overlaps.out");
        if (!fo) { std::fclose(f); set_error("could not create overlaps.out"); return UNO_EIO; }
        std::fprintf(fo, "   MULLIKEN OVERLAP POPULATIONS\n");
        std::fprintf(fo, uhf ? "  i      j        Alpha value                 Beta Value\n" : "  i      j                   value \n");
        for (int i = 0; i < n; ++i)
            for (int j = 0; j <= i; ++j) {
                if (uhf)
                    std::fprintf(fo, "%5d %5d  %24.16le  %24.16le\n", i, j, ovl[(size_t)i * n + j],
                                 ovl[((size_t)n + i) * n + j]);
                else
                    std::fprintf(fo, "%5d %5d  %24.16le\n", i, j, ovl[(size_t)i * n + j]);
            }
        std::fclose(fo);
    }
    if (d.prt_flag[0]) {
        for (int s = 0; s < ns; ++s) {
            std::fprintf(f, "xxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxx\n");
            std::fprintf(f, uhf ? (s ? "              Beta Fock Eigenvectors\n\n" : "              Alpha Fock Eigenvectors\n\n")
                                : "              Fock Eigenvectors\n\n");
            const int no = s ? noB : noA;
            for (int i = 0; i < n; ++i) {
                std::fprintf(f, "Column %5d Occupancy= %2d EigenEnergy= %25.15le\n", i, i < no ? (uhf ? 1 : 2) : 0,
                             ev[(size_t)s * n + i]);
                const double* Cm = &C[(size_t)s * n * n + i];
                int j = 0;
                for (; j + 4 <= n; j += 4)
                    std::fprintf(f, "%18.10le %18.10le %18.10le %18.10le\n", Cm[(size_t)j * n], Cm[(size_t)(j + 1) * n],
                                 Cm[(size_t)(j + 2) * n], Cm[(size_t)(j + 3) * n]);
                for (; j < n; ++j) std::fprintf(f, "%18.10le ", Cm[(size_t)j * n]);
                std::fprintf(f, "\n");
            }
        }
    }
    std::fclose(f);
    // multipole analysis (Moments.cpp:189-363): electronic part from packed P with
    // x4 off-diagonal / x2 diagonal factors; UHF uses (Pa + Pb)/2
    double nM[9] = {0}, eM[9] = {0}, tM[9];
    for (int c = 0; c < d.ncen; ++c) {
        const double* rr = &d.xyz[3 * c];
        const double v[9] = {rr[0], rr[1], rr[2], rr[0] * rr[0], rr[0] * rr[1], rr[0] * rr[2], rr[1] * rr[1],
                             rr[1] * rr[2], rr[2] * rr[2]};
        for (int x = 0; x < 9; ++x) nM[x] += d.charge[c] * v[x];
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
            const long ij = pidx(i, j);
            const double pij = uhf ? 0.5 * (P[ij] + P[n2 + ij]) : P[ij];
            const double fct = (i == j) ? 2.0 : 4.0;
            for (int x = 0; x < 9; ++x) eM[x] += pij * fct * M[9 * ij + x];
        }
    for (int x = 0; x < 9; ++x) tM[x] = nM[x] - eM[x];
    const double q00 = tM[8] - 0.5 * (tM[3] + tM[6]);
    f = create(dir, "moments.out");
    std::fprintf(f, " MULTIPOLE MOMENT ANALYSIS \n units in bohr - hartree atomic units \n\n DIPOLE MOMENTS \n\n");
    std::fprintf(f, "          Total         Electronic        Nuclear \n");
    const char* dl[3] = {" x", " y", " z"};
    for (int x = 0; x < 3; ++x) std::fprintf(f, "%s  %15.7le  %15.7le  %15.7le \n", dl[x], tM[x], eM[x], nM[x]);
    std::fprintf(f, "\n QUADRUPOLE MOMENTS \n%s\n", uhf ? " " : "");
    std::fprintf(f, "          Total         Electronic        Nuclear \n");
    const char* ql[6] = {" xx", " xy", " xz", " yy", " yz", " zz"};
    for (int x = 0; x < 6; ++x) std::fprintf(f, "%s  %15.7le  %15.7le  %15.7le \n", ql[x], tM[3 + x], eM[3 + x], nM[3 + x]);
    std::fprintf(f, "\n Dipole moment     = %25.15le \n Quadrupole moment = %25.15le \n", tM[2], q00);
    std::fclose(f);
    // MO transition dipoles (Moments.cpp:365-469): C^T mu C per spin
    f = create(dir, "mol_dipmom.out");
    std::fprintf(f, " MULTIPOLE MOMENT ANALYSIS \n units in bohr - hartree atomic units \n\n");
    for (int s = 0; s < ns; ++s) {
        std::fprintf(f, uhf ? (s ? " BETA MO TRANSISITION DIPOLE MOMENTS \n\n" : " ALPHA MO TRANSISITION DIPOLE MOMENTS \n\n")
                            : " MO TRANSISITION DIPOLE MOMENTS \n\n");
        std::fprintf(f, " orbital 1       orbital2      dx-dy-dz\n");
        std::vector<double> dm[3];
        for (int x = 0; x < 3; ++x) {
            dm[x].resize(n2);
            for (long ij = 0; ij < n2; ++ij) dm[x][ij] = M[9 * ij + x];
            std::vector<double> Cs(C.begin() + (size_t)s * n * n, C.begin() + (size_t)(s + 1) * n * n);
            std::vector<double> a(n * (size_t)n, 0.0), t(n * (size_t)n, 0.0);
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j) a[(size_t)i * n + j] = dm[x][pidx(i, j)];
            for (int i = 0; i < n; ++i)
                for (int k = 0; k < n; ++k) {
                    const double aik = a[(size_t)i * n + k];
                    for (int j = 0; j < n; ++j) t[(size_t)i * n + j] += aik * Cs[(size_t)k * n + j];
                }
            for (int i = 0; i < n; ++i)
                for (int j = 0; j <= i; ++j) {
                    double sum = 0.0;
                    for (int k = 0; k < n; ++k) sum += Cs[(size_t)k * n + i] * t[(size_t)k * n + j];
                    dm[x][pidx(i, j)] = sum;
                }
        }
        long ij = 0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j <= i; ++j, ++ij)
                std::fprintf(f, " %12d %12d %15.6le %15.6le %15.6le\n", i, j, dm[0][ij], dm[1][ij], dm[2][ij]);
    }
    std::fclose(f);
    return UNO_OK;
}

}  // namespace

extern "C" {

const char* uno_last_error(void) { return get_error().c_str(); }

int uno_ctx_create(const uno_basis* basis, int start_shell, uno_ctx** out) {
    if (!basis || !out) { set_error("null argument"); return UNO_EINVAL; }
    uno_ctx* c = new (std::nothrow) uno_ctx();
    if (!c) { set_error("out of host memory"); return UNO_ENOMEM; }
    const int rc = ctx_build(c, basis, start_shell);
    if (rc) { delete c; *out = nullptr; return rc; }
    *out = c;
    return UNO_OK;
}

int uno_ctx_rebuild(uno_ctx* c, const uno_basis* basis) {
    if (!c || !basis) { set_error("null argument"); return UNO_EINVAL; }
    return ctx_build(c, basis, c->start_shell);
}

int uno_ctx_destroy(uno_ctx* c) {
    delete c;
    return UNO_OK;
}

int uno_ctx_norb(const uno_ctx* c) { return c ? c->n : 0; }

int uno_ctx_set(uno_ctx* c, const char* key, double v) {
    if (!c || !key) { set_error("null argument"); return UNO_EINVAL; }
    std::string k(key);
    if (k == "schwarz") c->tau = v;
    else if (k == "prim_cut") c->primcut = v;
    else if (k == "skip_center") c->skip_center = (int)v;
    else if (k == "fp32_threshold") c->fp32_thr = v;
    else if (k == "schwarz_mixed") c->tau_mixed = v;
    else if (k == "mixed_switch") c->mixed_switch = v;
    else if (k == "rebuild_every") c->rebuild_every = std::max(1, (int)v);
    else if (k == "precision") c->precision = (int)v;
    else if (k == "row_merge") c->row_merge = (int)v;
    else if (k == "row_min_kets") c->row_min_kets = (long long)v;
    else if (k == "row_variant") c->row_variant = (int)v;
    else if (k == "row_density_lds") c->row_dens = (int)v;
    else if (k == "rys2_exact") c->rys2_exact = (int)v;
    else if (k == "purify") c->purify = (int)v;
    else if (k == "x_newton") c->x_newton = (int)v;
    else if (k == "deterministic") c->deterministic = v != 0.0;
    else if (k == "launch_streams") c->launch_streams = std::max(1, std::min(kMaxLaunchStreams, (int)v));
    else { set_error("unknown option " + k); return UNO_EINVAL; }
    return UNO_OK;
}

int uno_ctx_get(uno_ctx* c, const char* key, double* v) {
    if (!c || !key || !v) { set_error("null argument"); return UNO_EINVAL; }
    std::string k(key);
    if (k == "schwarz") *v = c->tau;
    else if (k == "prim_cut") *v = c->primcut;
    else if (k == "skip_center") *v = c->skip_center;
    else if (k == "fp32_threshold") *v = c->fp32_thr;
    else if (k == "schwarz_mixed") *v = c->tau_mixed;
    else if (k == "mixed_switch") *v = c->mixed_switch;
    else if (k == "rebuild_every") *v = c->rebuild_every;
    else if (k == "precision") *v = c->precision;
    else if (k == "row_merge") *v = c->row_merge;
    else if (k == "row_min_kets") *v = (double)c->row_min_kets;
    else if (k == "row_variant") *v = c->row_variant;
    else if (k == "row_density_lds") *v = c->row_dens;
    else if (k == "rys2_exact") *v = c->rys2_exact;
    else if (k == "purify") *v = c->purify;
    else if (k == "x_newton") *v = c->x_newton;
    else if (k == "x_newton_steps") *v = c->last_x_steps;
    else if (k == "deterministic") *v = c->deterministic;
    else if (k == "launch_streams") *v = c->launch_streams;
    else if (k == "ctx_ms_pairs") *v = c->ctx_ms[0];
    else if (k == "ctx_ms_schwarz") *v = c->ctx_ms[1];
    else if (k == "ctx_ms_runs") *v = c->ctx_ms[2];
    else if (k == "ctx_ms_rest") *v = c->ctx_ms[3];
    else if (k == "last_quartets") *v = c->last_quartets;
    else if (k == "last_quartets32") *v = c->last_quartets32;
    else if (k == "pair_runs") *v = (double)c->cls.size();
    else if (k == "quartet_launches") *v = (double)c->qcls.size();
    else if (k == "pairs") *v = (double)c->pairs.size();
    else { set_error("unknown option " + k); return UNO_EINVAL; }
    return UNO_OK;
}

int uno_fock_rhf(uno_ctx* c, const double* P, double* G) {
    if (!c || !P || !G) { set_error("null argument"); return UNO_EINVAL; }
    const int n = c->n;
    const long n2 = (long)n * (n + 1) / 2;
    hipStream_t st = c->st;
    UNO_HIP(c->pk.upload(P, n2, st));
    UNO_HIP(k_unpack(n, c->pk.p, c->D0.p, st));
    if (int rc = fock_build(c, false, -1.0, c->precision == 1)) return rc;
    UNO_HIP(c->pk.upload(G, n2, st));
    UNO_HIP(k_pack_sym_add(n, c->W0.p, c->pk.p, st));
    UNO_HIP(hipMemcpyAsync(G, c->pk.p, n2 * sizeof(double), hipMemcpyDeviceToHost, st));
    UNO_HIP(hipStreamSynchronize(st));
    return UNO_OK;
}

int uno_fock_uhf(uno_ctx* c, const double* PA, const double* PB, double* GA, double* GB) {
    if (!c || !PA || !PB || !GA || !GB) { set_error("null argument"); return UNO_EINVAL; }
    const int n = c->n;
    const long n2 = (long)n * (n + 1) / 2, nn = (long)n * n;
    hipStream_t st = c->st;
    UNO_HIP(c->pk.upload(PA, n2, st));
    UNO_HIP(k_unpack(n, c->pk.p, c->D1.p, st));
    UNO_HIP(c->pk.upload(PB, n2, st));
    UNO_HIP(k_unpack(n, c->pk.p, c->D2.p, st));
    UNO_HIP(k_axpby(nn, 1.0, c->D1.p, 1.0, c->D2.p, c->D0.p, st));
    if (int rc = fock_build(c, true, -1.0, c->precision == 1)) return rc;
    double* Gs[2] = {GA, GB};
    double* Ws[2] = {c->W0.p, c->W1.p};
    for (int s = 0; s < 2; ++s) {
        UNO_HIP(c->pk.upload(Gs[s], n2, st));
        UNO_HIP(k_pack_sym_add(n, Ws[s], c->pk.p, st));
        UNO_HIP(hipMemcpyAsync(Gs[s], c->pk.p, n2 * sizeof(double), hipMemcpyDeviceToHost, st));
        UNO_HIP(hipStreamSynchronize(st));
    }
    return UNO_OK;
}

int uno_autotune(uno_ctx* c, const double* PA, const double* PB, int mixed, double* best_ms) {
    if (!c || !PA) { set_error("null argument"); return UNO_EINVAL; }
    const int n = c->n;
    const long n2 = (long)n * (n + 1) / 2, nn = (long)n * n;
    hipStream_t st = c->st;
    const bool uhf = PB != nullptr;
    UNO_HIP(c->pk.upload(PA, n2, st));
    UNO_HIP(k_unpack(n, c->pk.p, uhf ? c->D1.p : c->D0.p, st));
    if (uhf) {
        UNO_HIP(c->pk.upload(PB, n2, st));
        UNO_HIP(k_unpack(n, c->pk.p, c->D2.p, st));
        UNO_HIP(k_axpby(nn, 1.0, c->D1.p, 1.0, c->D2.p, c->D0.p, st));
    }
    c->tuning = 1;
    c->tune_ms = 0.0;
    const int rc = fock_build(c, uhf, -1.0, mixed != 0);
    c->tuning = 0;
    if (rc) return rc;
    UNO_HIP(hipStreamSynchronize(st));
    remember_tuning(c);
    if (best_ms) *best_ms = c->tune_ms;
    return UNO_OK;
}

int uno_eigh_packed(int n, double* a, double* z, double* evals) {
    if (n <= 0 || !a || !z || !evals) { set_error("bad argument"); return UNO_EINVAL; }
    hipStream_t st;
    UNO_HIP(hipStreamCreate(&st));
    const long nn = (long)n * n, n2 = (long)n * (n + 1) / 2;
    int rc;
    {
        DBuf<double> pk, A, Z;
        UNO_HIP(pk.upload(a, n2, st));
        UNO_HIP(A.alloc(nn));
        UNO_HIP(Z.alloc(nn));
        UNO_HIP(k_unpack(n, pk.p, A.p, st));
        rc = eigh_standalone(n, A.p, Z.p, evals, st);
        if (!rc) {
            UNO_HIP(hipMemcpyAsync(z, Z.p, nn * sizeof(double), hipMemcpyDeviceToHost, st));
            UNO_HIP(hipStreamSynchronize(st));
            std::memset(a, 0, n2 * sizeof(double));   // input destroyed like rsp
        }
    }
    (void)hipStreamDestroy(st);
    return rc;
}

int uno_sp_trans(int n, double* a, const double* x) {
    if (n <= 0 || !a || !x) { set_error("bad argument"); return UNO_EINVAL; }
    hipStream_t st;
    UNO_HIP(hipStreamCreate(&st));
    const int rc = sp_trans_dev(n, a, x, st);
    (void)hipStreamDestroy(st);
    return rc;
}

void uno_scf_defaults(uno_scf_opts* o) {
    std::memset(o, 0, sizeof(*o));
    o->maxits = 300;
    o->eps = 1e-10;
    o->cflag = 2;
    o->accel = 0;
    o->guess = 0;
    o->precision = 0;
    o->fp64_final = 2;
    o->incremental = 0;
    o->verbose = 0;
}

int uno_scf_run(uno_ctx* c, int nelec, const uno_scf_opts* opts, uno_scf_result* res, double* evals, double* P,
                double* C) {
    if (!c || !res) { set_error("null argument"); return UNO_EINVAL; }
    uno_scf_opts o;
    if (opts) o = *opts; else uno_scf_defaults(&o);
    return scf_run(c, nelec, o, *res, evals, P, C);
}

int uno_eri_emit(uno_ctx* c, long cap, double* val, int* ijkl, long* count) {
    if (!c || !count) { set_error("null argument"); return UNO_EINVAL; }
    return eri_emit(c, cap, val, ijkl, count);
}

static int rys_eval_on(hipStream_t st, int n, int npts, const double* x, double* u, double* w, bool fp32,
                       bool ref2) {
    DBuf<double> dx, du, dw;
    UNO_HIP(dx.upload(x, npts, st));
    UNO_HIP(du.alloc((size_t)npts * n));
    UNO_HIP(dw.alloc((size_t)npts * n));
    UNO_HIP(rys_eval_launch(n, npts, dx.p, du.p, dw.p, st, fp32, ref2));
    UNO_HIP(hipMemcpyAsync(u, du.p, (size_t)npts * n * sizeof(double), hipMemcpyDeviceToHost, st));
    UNO_HIP(hipMemcpyAsync(w, dw.p, (size_t)npts * n * sizeof(double), hipMemcpyDeviceToHost, st));
    UNO_HIP(hipStreamSynchronize(st));
    return UNO_OK;
}

static int rys_eval_impl(int n, int npts, const double* x, double* u, double* w, bool fp32, bool ref2 = false) {
    if (n < 1 || n > 5 || npts <= 0 || !x || !u || !w) { set_error("bad argument"); return UNO_EINVAL; }
    hipStream_t st;
    UNO_HIP(hipStreamCreate(&st));
    const int rc = rys_eval_on(st, n, npts, x, u, w, fp32, ref2);
    (void)hipStreamDestroy(st);
    return rc;
}

int uno_rys_eval(int n, int npts, const double* x, double* u, double* w) {
    return rys_eval_impl(n, npts, x, u, w, false);
}

int uno_rys_eval_f32(int n, int npts, const double* x, double* u, double* w) {
    return rys_eval_impl(n, npts, x, u, w, true);
}

int uno_rys_eval_ref2(int fp32, int npts, const double* x, double* u, double* w) {
    return rys_eval_impl(2, npts, x, u, w, fp32 != 0, true);
}

int uno_schwarz(uno_ctx* c, double* q) {
    if (!c || !q) { set_error("null argument"); return UNO_EINVAL; }
    // packed shell-pair order a >= b; pairs dropped by the primitive cut report 0
    for (int a = 0; a < c->nsh; ++a)
        for (int b = 0; b <= a; ++b) {
            const int p = c->pair_of[(size_t)a * c->nsh + b];
            q[pidx(a, b)] = p >= 0 ? c->pairQ[p] : 0.0;
        }
    return UNO_OK;
}

int uno_profile(uno_ctx* c, int enable, double* ms, double* launches, double* prim, double* shell) {
    if (!c) { set_error("null argument"); return UNO_EINVAL; }
    const int ncls = kProfSlots;
    if (enable >= 0) {
        c->profile = enable;
        if (int rc = profile_reset(c, enable != 0)) return rc;
    } else if (int rc = profile_collect(c)) {
        return rc;
    }
    for (int i = 0; i < ncls; ++i) {
        if (ms) ms[i] = c->prof_ms[i];
        if (launches) launches[i] = c->prof_launch[i];
        if (prim) prim[i] = c->prof_prim[i];
        if (shell) shell[i] = c->prof_shell[i];
    }
    return UNO_OK;
}

int uno_oneint_device(uno_ctx* c, double* S, double* T, double* H) {
    if (!c || !S || !T || !H) { set_error("null argument"); return UNO_EINVAL; }
    const int n = c->n;
    const size_t nn = (size_t)n * n;
    DBuf<double> dS, dT, dH;
    UNO_HIP(dS.alloc(nn));
    UNO_HIP(dT.alloc(nn));
    UNO_HIP(dH.alloc(nn));
    if (int rc = oneint_device(c, dS.p, dT.p, dH.p)) return rc;
    double* out[3] = {S, T, H};
    double* src[3] = {dS.p, dT.p, dH.p};
    for (int k = 0; k < 3; ++k) {
        UNO_HIP(k_pack(n, src[k], c->pk.p, c->st));
        UNO_HIP(hipMemcpyAsync(out[k], c->pk.p, (size_t)n * (n + 1) / 2 * sizeof(double), hipMemcpyDeviceToHost,
                               c->st));
        UNO_HIP(hipStreamSynchronize(c->st));
    }
    return UNO_OK;
}

int uno_shell_rotation(int l, const double* Q, double* M) {
    if (l < 0 || l > 2 || !Q || !M) { set_error("bad argument (0 <= l <= 2)"); return UNO_EINVAL; }
    shell_rotation(l, Q, M);
    return UNO_OK;
}

int uno_guess_features(const uno_basis* b, long cap, float* X, int* pairs, long* count) {
    if (!b || !count) { set_error("null argument"); return UNO_EINVAL; }
    Basis B;
    std::string err;
    if (!make_basis(b, B, err)) { set_error(err); return UNO_EINVAL; }
    std::vector<float> x;
    std::vector<int> pr;
    if (int rc = guess_features(B, x, pr, err)) { set_error(err); return rc; }
    *count = (long)pr.size() / 2;
    if (cap == 0) return UNO_OK;
    if (cap < *count || !X || !pairs) { set_error("capacity too small"); return UNO_EINVAL; }
    std::memcpy(X, x.data(), x.size() * sizeof(float));
    std::memcpy(pairs, pr.data(), pr.size() * sizeof(int));
    return UNO_OK;
}

int uno_guess_samples(const uno_basis* b, const double* P, long cap, float* X, float* Y, float* M, long* count) {
    if (!b || !P || !count) { set_error("null argument"); return UNO_EINVAL; }
    Basis B;
    std::string err;
    if (!make_basis(b, B, err)) { set_error(err); return UNO_EINVAL; }
    const int n = B.norb;
    std::vector<double> D((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) D[(size_t)i * n + j] = P[pidx(i, j)];
    std::vector<float> x, y, m;
    if (int rc = guess_samples(B, D.data(), x, y, m, err)) { set_error(err); return rc; }
    *count = (long)x.size() / kGuessInDim;
    if (cap == 0) return UNO_OK;
    if (cap < *count || !X || !Y || !M) { set_error("capacity too small"); return UNO_EINVAL; }
    std::memcpy(X, x.data(), x.size() * sizeof(float));
    std::memcpy(Y, y.data(), y.size() * sizeof(float));
    std::memcpy(M, m.data(), m.size() * sizeof(float));
    return UNO_OK;
}

int uno_guess_assemble(const uno_basis* b, long count, const int* pairs, const float* Y, double* P) {
    if (!b || !pairs || !Y || !P || count < 0) { set_error("null argument"); return UNO_EINVAL; }
    Basis B;
    std::string err;
    if (!make_basis(b, B, err)) { set_error(err); return UNO_EINVAL; }
    for (long k = 0; k < 2 * count; ++k)
        if (pairs[k] < 0 || pairs[k] >= B.ncen) { set_error("atom index out of range"); return UNO_EINVAL; }
    std::vector<int> pr(pairs, pairs + 2 * count);
    std::vector<double> D;
    if (int rc = guess_assemble(B, pr, Y, D, err)) { set_error(err); return rc; }
    const int n = B.norb;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) P[pidx(i, j)] = D[(size_t)i * n + j];
    return UNO_OK;
}

int uno_ctx_set_guess(uno_ctx* c, uno_mlp* model) {
    if (!c) { set_error("null argument"); return UNO_EINVAL; }
    c->guess_model = model;
    return UNO_OK;
}

int uno_copy_probe(unsigned long long bytes, int reps, double* ms) {
    if (!ms || reps <= 0 || bytes < 32) { set_error("bad argument"); return UNO_EINVAL; }
    float t = 0.f;
    UNO_HIP(copy_probe((size_t)bytes, reps, &t, nullptr));
    *ms = t;
    return UNO_OK;
}

int uno_oneint(const uno_basis* basis, double* S, double* T, double* H) {
    Basis B;
    std::string err;
    if (!make_basis(basis, B, err)) { set_error(err); return UNO_EINVAL; }
    one_electron(B, S, T, H, nullptr);
    return UNO_OK;
}

int uno_moment_ints(const uno_basis* basis, double* m) {
    Basis B;
    std::string err;
    if (!make_basis(basis, B, err)) { set_error(err); return UNO_EINVAL; }
    one_electron(B, nullptr, nullptr, nullptr, m);
    return UNO_OK;
}

int uno_deck_generate(const char* kind, int n, unsigned long long seed, const char* basis, int charge, double eps,
                      int maxits, const char* path) {
    if (!kind || !basis || !path) { set_error("null argument"); return UNO_EINVAL; }
    std::vector<Atom> atoms;
    std::string k(kind);
    if (k == "water") water_monomer(atoms);
    else if (k == "water_cluster") {
        if (n <= 0) { set_error("water_cluster needs n > 0"); return UNO_EINVAL; }
        water_cluster(n, seed, atoms);
    } else if (k == "molecule") small_molecule(n, seed, atoms);
    else { set_error("unknown deck kind " + k); return UNO_EINVAL; }
    Deck d;
    std::string err;
    if (!build_deck(atoms, basis, charge, d, err)) { set_error(err); return UNO_EINVAL; }
    d.eps = eps;
    d.maxits = maxits;
    if (!write_deck(path, d, err)) { set_error(err); return UNO_EIO; }
    return UNO_OK;
}

// DPM polarization-potential scan (RHF.hpp:292-388, UHF.hpp:293-383): the basis is
// augmented by the positron centre (charge +1) carrying the shells of posin.bas; for every
// point of pos.grid.dat the positron's six-Gaussian model potential (GDPMInts.cpp) is
// added to H and an SCF (eps 1e-12, one update before the loop) starts from the
// ground-state density.  vpol = E_final - E_init, alfa = -2 vpol r^4 and the static term
// vstat = E_init - E_gs (at the first point the reference adds the nuclear repulsion once
// more, RHF.hpp:336; kept) go to vpol.out and spol.out.
static int run_dpm(const std::string& dir, const Deck& d, const uno_scf_opts& o, const uno_scf_result& gs,
                   const std::vector<double>& Pgs) {
    std::ifstream ab((dir + "/posin.bas").c_str());
    if (!ab) { set_error("could not open posin.bas"); return UNO_EIO; }
    int xsh = 0, xno = 0, xmaxl = 0;
    ab >> xsh >> xno >> xmaxl;
    Deck a = d;
    a.ncen = d.ncen + 1;
    a.charge.push_back(1.0);
    a.xyz.insert(a.xyz.end(), {0.0, 0.0, 0.0});
    for (int k = 0; k < xsh; ++k) {
        int npr = 0, l = 0;
        if (!(ab >> npr >> l) || npr <= 0 || l < 0) { set_error("malformed posin.bas"); return UNO_EIO; }
        a.sh_l.push_back(l);
        a.sh_npr.push_back(npr);
        a.sh_cen.push_back(d.ncen);
        a.sh_poff.push_back((int)a.alpha.size());
        for (int q = 0; q < npr; ++q) {
            double al = 0.0, co = 0.0;
            if (!(ab >> al >> co)) { set_error("malformed posin.bas"); return UNO_EIO; }
            a.alpha.push_back(al);
            a.coef.push_back(co);
        }
        a.norb += ncart(l);
    }
    a.nshell = d.nshell + xsh;
    std::ifstream gr((dir + "/pos.grid.dat").c_str());
    if (!gr) { set_error("could not open pos.grid.dat"); return UNO_EIO; }
    int npts = 0;
    gr >> npts;
    const bool uhf = d.nelec % 2;
    const int ns = uhf ? 2 : 1;
    const long n2g = (long)(a.norb - xno) * (a.norb - xno + 1) / 2;
    const long n2 = (long)a.norb * (a.norb + 1) / 2;
    FILE* vout = std::fopen((dir + "/vpol.out").c_str(), "w");
    FILE* sout = std::fopen((dir + "/spol.out").c_str(), "w");
    if (!vout || !sout) {
        if (vout) std::fclose(vout);
        if (sout) std::fclose(sout);
        set_error("could not create vpol.out / spol.out");
        return UNO_EIO;
    }
    int rc = UNO_OK;
    for (int i = 0; i < npts && rc == UNO_OK; ++i) {
        double R[3];
        if (!(gr >> R[0] >> R[1] >> R[2])) { set_error("malformed pos.grid.dat"); rc = UNO_EIO; break; }
        for (int x = 0; x < 3; ++x) a.xyz[3 * d.ncen + x] = R[x];
        uno_basis b;
        basis_view(a, b);
        uno_ctx* c = nullptr;
        if ((rc = uno_ctx_create(&b, 0, &c))) break;
        c->skip_center = d.ncen;   // the positron's charge is not in V (OneElectronInts.cpp:43)
        std::vector<double> dh;
        if ((rc = gdpm_matrix(c, R, dh))) { uno_ctx_destroy(c); break; }
        // the ground-state density padded with zeros for the positron-centre functions
        std::vector<double> P(ns * n2, 0.0);
        for (int s = 0; s < ns; ++s) std::copy(Pgs.begin() + s * n2g, Pgs.begin() + (s + 1) * n2g, P.begin() + s * n2);
        uno_scf_opts so = o;
        so.eps = 1e-12;   // set after the deck's floor (RHF.hpp:303): no floor here
        so.guess = 1;
        so.single_start = 2;
        so.field[0] = so.field[1] = so.field[2] = 0.0;
        std::vector<double> ev(ns * (size_t)a.norb), C(ns * (size_t)a.norb * a.norb);
        uno_scf_result rs{};
        const int src = scf_run(c, d.nelec, so, rs, ev.data(), P.data(), C.data(), dh.data());
        uno_ctx_destroy(c);
        if (src && src != UNO_ENOCONV) { rc = src; break; }
        const double vpol = rs.energy - rs.init_energy;
        const double vstat = rs.init_energy - gs.energy + (i == 0 ? rs.nucrep : 0.0);
        const double r2 = R[0] * R[0] + R[1] * R[1] + R[2] * R[2];
        const double alfa = -2.0 * vpol * r2 * r2;
        if (uhf)
            std::fprintf(vout, "%15.10lf %15.10lf %15.10lf %25.15le %25.15le\n", R[0], R[1], R[2], vpol, alfa);
        else
            std::fprintf(vout, "%15.10lf %15.10lf %15.10lf %25.15le %25.15le %25.15le %25.15le\n", R[0], R[1], R[2],
                         vpol, alfa, vstat, vstat + vpol);
        std::fprintf(sout, "%3d %20.10le %20.10le %20.10le %20.10le %25.15le\n", i, gs.energy, rs.init_energy,
                     rs.energy, vstat, rs.ediff);
        std::fflush(vout);
        std::fflush(sout);
    }
    std::fclose(vout);
    std::fclose(sout);
    return rc;
}

int uno_run_deck(const char* dir_c, const uno_scf_opts* opts, uno_scf_result* res) {
    const std::string dir = dir_c ? dir_c : ".";
    Deck d;
    std::string err;
    if (!read_deck(dir + "/patin.dat", d, err)) { set_error(err); return UNO_EIO; }
    uno_basis b;
    basis_view(d, b);
    uno_ctx* c = nullptr;
    if (int rc = uno_ctx_create(&b, 0, &c)) return rc;
    uno_scf_opts o;
    if (opts) o = *opts; else uno_scf_defaults(&o);
    o.maxits = d.maxits;
    o.eps = d.eps;
    o.cflag = d.scf_flag[0];
    const int n = c->n;
    const long n2 = (long)n * (n + 1) / 2;
    const bool uhf = d.nelec % 2;
    std::vector<double> P((uhf ? 2 : 1) * n2), ev((uhf ? 2 : 1) * (size_t)n), C((uhf ? 2 : 1) * (size_t)n * n);
    if (d.scf_flag[2]) {
        // restart from PMATRIX.DAT (RHF.hpp:120-123)
        FILE* f = std::fopen((dir + "/PMATRIX.DAT").c_str(), "rb");
        if (!f || std::fread(P.data(), sizeof(double), P.size(), f) != P.size()) {
            if (f) std::fclose(f);
            uno_ctx_destroy(c);
            set_error("could not read PMATRIX.DAT");
            return UNO_EIO;
        }
        std::fclose(f);
        o.guess = 1;
    }
    uno_scf_result r{};
    int rc = uno_scf_run(c, d.nelec, &o, &r, ev.data(), P.data(), C.data());
    if (rc && rc != UNO_ENOCONV) { uno_ctx_destroy(c); return rc; }
    std::vector<double> S(n2), T(n2), H(n2), M(9 * n2);
    one_electron(c->B, S.data(), T.data(), H.data(), M.data());
    int wrc = write_outputs(dir, d, c->B, r, S, T, H, M, P, ev, C, d.maxits);
    if (!wrc && d.int_flag[1]) {
        // finite-field analysis (RHF.hpp:235-271, UHF.hpp:239-272): a field of 5e-3 au along
        // each axis, SCF from the ground-state density with one update before the loop
        const double emag = 5.0e-3;
        double polnrg[3], alf[3];
        for (int ix = 0; ix < 3 && !wrc; ++ix) {
            uno_scf_opts of = o;
            of.field[0] = of.field[1] = of.field[2] = 0.0;
            of.field[ix] = emag;
            of.guess = 1;
            of.single_start = 1;
            std::vector<double> Pf(P), evf(ev.size()), Cf(C.size());
            uno_scf_result rf{};
            const int frc = uno_scf_run(c, d.nelec, &of, &rf, evf.data(), Pf.data(), Cf.data());
            if (frc && frc != UNO_ENOCONV) { wrc = frc; break; }
            polnrg[ix] = rf.energy - r.energy;
            alf[ix] = -polnrg[ix] * 2.0 / emag / emag;
        }
        if (!wrc) {
            FILE* f = create(dir, "finitefield.out");
            if (!f) { set_error("could not create finitefield.out"); wrc = UNO_EIO; }
            else {
                std::fprintf(f, "        Unomol Finite Field Analysis\n");
                std::fprintf(f, "        Electric field magnitude = %20.12le\n", emag);
                std::fprintf(f, "     alfa                pol energy\n");
                std::fprintf(f, " x   %20.12le %20.12le\n", alf[0], polnrg[0]);
                std::fprintf(f, " y   %20.12le %20.12le\n", alf[1], polnrg[1]);
                std::fprintf(f, " z   %20.12le %20.12le\n", alf[2], polnrg[2]);
                if (!uhf) {
                    std::fprintf(f, " alf isotropic = %20.12le\n", (alf[0] + alf[1] + alf[2]) / 3.0);
                    std::fprintf(f, " alf aniso     = %20.12le\n", (2.0 * alf[2] - alf[1] - alf[0]) / 3.0);
                }
                std::fclose(f);
            }
        }
    }
    if (!wrc && d.int_flag[0]) wrc = run_dpm(dir, d, o, r, P);
    if (res) *res = r;
    uno_ctx_destroy(c);
    return wrc ? wrc : rc;
}

}  // extern "C"
