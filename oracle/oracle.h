/*
 * oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference's (PatNichols/unomol) integrals + Fock build +
 * diagonalisation + SCF path.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker: the product library
 * (unomol_amd) never links, loads or calls it.
 *
 * Every entry point returns 0 on success.  Basis input is the raw deck content
 * (un-normalised contraction coefficients exactly as written in patin.dat); the
 * oracle normalises them itself (reference Basis.hpp:56-75).
 *
 * Parity pins (see tests/): reference goldens test/scfout.dat.*, short.dat.*,
 * moments.dat.*; reference ERIs / one-electron matrices written by oracle/_ref;
 * Boys/Rys moment identities against mpmath; closed-form (ss|ss); brute-force
 * symmetric G; numpy eigh.
 */
#ifndef UNO_ORACLE_H
#define UNO_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_basis {
    int nshell, ncen;
    const int* sh_l;        /* [nshell] angular momentum */
    const int* sh_npr;      /* [nshell] primitives */
    const int* sh_cen;      /* [nshell] centre index */
    const int* sh_poff;     /* [nshell] offset into alpha/coef */
    const double* alpha;    /* raw exponents */
    const double* coef;     /* raw (un-normalised) coefficients */
    const double* xyz;      /* [ncen*3] bohr */
    const double* charge;   /* [ncen] */
} orc_basis;

/* Rys roots u_i = t_i^2 (ascending) and weights, from quad-precision Boys moments
 * and Golub-Welsch (n <= 9). */
int orc_rys(int n, double x, double* u, double* w);
/* Boys F_0..F_m(x) in quad precision, rounded. */
int orc_boys(int m, double x, double* f);

/* number of basis functions */
int orc_norb(const orc_basis* b);

/* Unique ERIs in the reference's canonical order (TwoElectronInts.cpp:511-697).
 * With cap==0 only the count is returned in *count.  No |val| drop is applied;
 * the reference keeps |val|>1e-14 (TwoElectronInts.cpp:667). */
int orc_eri_all(const orc_basis* b, long cap, double* val, int* ijkl, long* count);

/* G-matrix replay of a unique-ERI list (TwoElectronInts.cpp:699-820).
 * Accumulates into G (packed lower, caller-zeroed). */
int orc_gmat_rhf(long n, const double* val, const int* ijkl, int norb,
                 const double* P, double* G);
int orc_gmat_uhf(long n, const double* val, const int* ijkl, int norb,
                 const double* PA, const double* PB, double* GA, double* GB);

/* Packed symmetric eigendecomposition (SymmPack.cpp:272-288 contract):
 * a destroyed, evals ascending, z row-major with eigenvectors in columns. */
int orc_rsp(int n, double* a, double* z, double* evals);
/* a <- X^T a X in place (packed a, dense row-major X)  (SymmPack.cpp:294-348) */
int orc_sp_trans(int n, double* a, const double* x);

/* One-electron S, T, H=T+V packed (OneElectronInts.cpp:127-202) */
int orc_oneint(const orc_basis* b, double* S, double* T, double* H);

/* Dipole/quadrupole integrals about the origin, packed, raw (RMOM.DAT content,
 * Moments.cpp:5-93): m[9*ij + {dx,dy,dz,qxx,qxy,qxz,qyy,qyz,qzz}] */
int orc_moment_ints(const orc_basis* b, double* m);

typedef struct orc_scf_out {
    double energy;       /* final total energy (electronic + nuclear) */
    double init_energy;  /* iteration-0 total energy */
    double ediff, pdiff, nucrep, kinetic;
    int iterations;
    int uhf;
} orc_scf_out;

/* Full SCF in the reference's loop semantics (RHF.hpp:114-176, UHF.hpp:137-177),
 * core-Hamiltonian guess.  evals: [norb] (RHF) or [2*norb] (UHF a then b);
 * P: packed [no2] (RHF half density) or [2*no2] (UHF a then b).
 * moments: [9] total (tM) multipoles, dx..qzz. */
int orc_scf(const orc_basis* b, int nelec, int maxits, double eps, int cflag,
            orc_scf_out* out, double* evals, double* P, double* moments);

#ifdef __cplusplus
}
#endif
#endif
