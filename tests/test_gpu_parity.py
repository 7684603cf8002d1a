"""GPU parity tests: the HIP path (libunomol_hip.so, called through the C ABI) against
the oracle and the reference's own outputs, on the same inputs."""
import os
import sys

import numpy as np
import pytest

import oracle_ffi as O
import refdata as R
import unomol_amd
from unomol_amd import _native as N
from unomol_amd.deck import read_deck

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu-test] " + msg, flush=True)
    sys.stdout.flush()


def _pack(M):
    n = M.shape[0]
    return np.ascontiguousarray(M[np.tril_indices(n)])


def _unpack(p, n):
    M = np.zeros((n, n))
    M[np.tril_indices(n)] = p
    return M + np.tril(M, -1).T


def _tints(deck_file, schwarz=None):
    b = unomol_amd.Basis(deck_file)
    t = unomol_amd.TwoElectronInts(b)
    if schwarz is not None:
        t.set("schwarz", schwarz)
    return b, t


def test_rys_tables_vs_mpmath(golden):
    say("rys tables vs mpmath")
    z = np.load(os.path.join(golden, "rys_mp.npz"))
    for n in range(1, 6):
        u, w = unomol_amd.rys_eval(n, z["x"])
        np.testing.assert_allclose(u, z["u"][n, :, :n], rtol=2e-15)
        np.testing.assert_allclose(w, z["w"][n, :, :n], rtol=2e-15)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
def test_rys_moment_identity(n):
    say("rys moment identity n=%d" % n)
    rng = np.random.default_rng(10 + n)
    x = np.concatenate([rng.uniform(0, 1, 300), rng.uniform(0, 80, 2000), rng.uniform(80, 1e4, 100)])
    u, w = unomol_amd.rys_eval(n, x)
    for i in range(0, len(x), 7):
        f = O.boys(2 * n - 1, x[i])
        mom = np.array([np.sum(w[i] * u[i] ** k) for k in range(2 * n)])
        np.testing.assert_allclose(mom, f, rtol=1e-13)


def test_rys_ref2_table_vs_reference(golden):
    """The two-root table the ERI kernels use by default against the reference's own
    two-root quadrature (Rys::root2, sampled from the reference build): 4e-15 relative
    where it is the reference's fit (15 < x <= 40, including its 15 < x <= 33 departure
    from the exact rule), within the reference's own fit error (5e-13) elsewhere, where
    the table is the exact rule; the float table to float rounding.  x = 15 itself is
    left out: the reference takes its x <= 15 branch there, the table its [15, 16) interval."""
    say("rys reference-compatible 2-root table vs the reference's root2")
    z = np.load(os.path.join(golden, "rys2_ref.npz"))
    keep = z["x"] != 15.0
    x, r, w = z["x"][keep], z["r"][keep], z["w"][keep]
    u_ref = r / (1 + r)
    inside = (x > 15.0) & (x <= 40.0)
    u, ww = unomol_amd.rys_eval_ref2(x)
    for got, ref in ((u, u_ref), (ww, w)):
        rel = np.abs(got - ref) / np.abs(ref)
        assert rel[inside].max() < 4e-15
        assert rel[~inside].max() < 5e-13
    uf, wf = unomol_amd.rys_eval_ref2(x, fp32=True)
    np.testing.assert_allclose(uf, u_ref, rtol=2e-6)
    np.testing.assert_allclose(wf, w, rtol=2e-6)
    # the exact table departs from the reference there by far more than 1e-12
    ue, _ = unomol_amd.rys_eval(2, x[inside])
    assert (np.abs(ue - u_ref[inside]) / u_ref[inside]).max() > 1e-6


@pytest.mark.parametrize("tag", ["A_water_631g", "water_631gs"])
def test_eri_vs_reference(golden, tag):
    """Every canonical ERI of the deck against the reference's own ERI code: 1e-12
    absolute in every class with the default (reference-compatible) two-root quadrature;
    with the exact quadrature (rys2_exact = 1), against the reference with its two-root
    fit replaced by the exact rule."""
    say("eri emit vs reference " + tag)
    z = np.load(os.path.join(golden, "eri_%s.npz" % tag))
    b, t = _tints(os.path.join(golden, "decks", tag + ".dat"))
    v, f = t.eri_emit()
    assert (f == z["ijkl"]).all()
    assert np.abs(v - z["ref"]).max() < 1e-12
    t.set("rys2_exact", 1)
    v2, f2 = t.eri_emit()
    assert (f2 == z["ijkl"]).all()
    assert np.abs(v2 - z["ref_x2"]).max() < 1e-12
    # the two modes differ only in the two-root classes
    fl = np.concatenate([[l] * ((l + 1) * (l + 2) // 2) for l in b.deck["sh_l"]])
    L = fl[f].sum(axis=1)
    two = (L == 2) | (L == 3)
    assert np.abs(v2 - v)[~two].max() == 0.0


@pytest.mark.parametrize("tag", ["631.h2o", "631.ch4", "631.co", "w2_631gs"])
def test_eri_vs_oracle(tag):
    say("eri emit vs oracle " + tag)
    path = R.deck_path(tag)
    b, t = _tints(path)
    v, f = t.eri_emit()
    ov, of = O.eri_all(read_deck(path))
    assert (f == of).all()
    assert np.abs(v - ov).max() < 1e-12


def _oracle_g(deck, P, PB=None):
    v, f = O.eri_all(deck)
    n = deck["norb"]
    if PB is None:
        return O.gmat_rhf(v, f, n, P)
    return O.gmat_uhf(v, f, n, P, PB)


@pytest.mark.parametrize("tag", ["A_water_631g", "631.h2o", "631.nh3", "w2_631gs"])
def test_fock_rhf_vs_oracle(tag):
    say("fock rhf vs oracle " + tag)
    path = R.deck_path(tag)
    deck = read_deck(path)
    n = deck["norb"]
    rng = np.random.default_rng(1)
    A = rng.standard_normal((n, n)) * 0.1
    P = _pack(A + A.T)
    b, t = _tints(path, schwarz=1e-16)
    G = np.zeros_like(P)
    t.formGmatrix(P, G)
    Gref = _oracle_g(deck, P)
    assert np.abs(G - Gref).max() < 1e-11
    # accumulation semantics: a second call adds
    t.formGmatrix(P, G)
    assert np.abs(G - 2 * Gref).max() < 2e-11


@pytest.mark.parametrize("tag", ["water_631gs_cation", "631.hf"])
def test_fock_uhf_vs_oracle(tag):
    say("fock uhf vs oracle " + tag)
    path = R.deck_path(tag)
    deck = read_deck(path)
    n = deck["norb"]
    rng = np.random.default_rng(2)
    A = rng.standard_normal((n, n)) * 0.1
    B = rng.standard_normal((n, n)) * 0.1
    PA, PB = _pack(A + A.T), _pack(B + B.T)
    b, t = _tints(path, schwarz=1e-16)
    GA, GB = np.zeros_like(PA), np.zeros_like(PB)
    t.formGmatrix(PA, PB, GA, GB)
    RA, RB = _oracle_g(deck, PA, PB)
    assert np.abs(GA - RA).max() < 1e-11
    assert np.abs(GB - RB).max() < 1e-11


def test_pair_runs_grouped(tmp_path):
    """The pair array of a water cluster splits into at most one run per angular pair
    class and primitive-count bucket, and every run pair (X >= Y) is one ERI launch."""
    say("pair runs of a 16-water cluster")
    deck = str(tmp_path / "w16.dat")
    unomol_amd.generate_deck(deck, "water_cluster", 16, 5, "6-31G*", 0, 1e-8, 50)
    b = unomol_amd.Basis(deck)
    t = unomol_amd.TwoElectronInts(b)
    runs = int(t.get("pair_runs"))
    assert 1 <= runs <= 36
    assert int(t.get("quartet_launches")) == runs * (runs + 1) // 2
    assert int(t.get("pairs")) > 0


@pytest.mark.parametrize("n", [1, 2, 3, 13, 64, 200, 777])
def test_eigh_vs_numpy(n):
    say("eigh n=%d" % n)
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    A = A + A.T
    if n >= 13:
        # a degenerate cluster like symmetry-equivalent orbitals
        Qm, _ = np.linalg.qr(rng.standard_normal((n, n)))
        ev = np.sort(rng.uniform(-20, 5, n))
        ev[3:6] = ev[3]
        A = (Qm * ev) @ Qm.T
        A = 0.5 * (A + A.T)
    a = _pack(A)
    z = np.zeros(n * n)
    e = np.zeros(n)
    unomol_amd.SymmPack.rsp(n, a, z, e)
    Z = z.reshape(n, n)
    ref = np.linalg.eigvalsh(A)
    scale = max(1.0, np.abs(ref).max())
    assert np.abs(e - ref).max() < 1e-13 * scale * max(1, n / 10)
    assert np.all(np.diff(e) >= 0)
    assert np.abs(A @ Z - Z * e).max() < 1e-12 * scale * max(1, n / 10)
    assert np.abs(Z.T @ Z - np.eye(n)).max() < 1e-12 * max(1, n / 10)


def test_sp_trans_vs_numpy():
    say("sp_trans")
    rng = np.random.default_rng(7)
    for n in (1, 5, 70, 131):
        A = rng.standard_normal((n, n))
        A = A + A.T
        X = rng.standard_normal((n, n))
        a = _pack(A)
        unomol_amd.SymmPack.sp_trans(n, a, X)
        assert np.abs(_unpack(a, n) - X.T @ A @ X).max() < 1e-11 * n


SCF_TAGS = [t for t in R.RHF_TAGS if t not in R.DEGENERATE_TAGS] + R.UHF_TAGS + R.GEN_TAGS


@pytest.mark.parametrize("tag", SCF_TAGS)
def test_scf_vs_reference(tmp_path, tag):
    """The drop-in run (patin.dat in, the reference's output files out) against the
    reference rebuilt from its sources on the same deck: initial and final energies and
    every orbital energy to 1e-9 Eh, moments to 1e-7."""
    say("scf drop-in run " + tag)
    deck = open(R.deck_path(tag)).read()
    (tmp_path / "patin.dat").write_text(deck)
    r = unomol_amd.run_deck(tmp_path)
    assert r.converged
    src = os.path.join(R.GOLDEN, "ref_run", tag)
    ref = R.read_scfout(os.path.join(src, "scfout.gs.out"))
    short = R.read_short(os.path.join(src, "short.gs.out"))
    got = R.read_short(str(tmp_path / "short.gs.out"))
    mine = R.read_scfout(str(tmp_path / "scfout.gs.out"))
    assert abs(got["energy"] - short["energy"]) < 1e-9
    assert abs(got["init"] - short["init"]) < 1e-9
    np.testing.assert_allclose(mine["evals"], ref["evals"], atol=1e-9)
    if len(ref["evals_b"]):
        np.testing.assert_allclose(mine["evals_b"], ref["evals_b"], atol=1e-9)
    mref = R.read_moments(os.path.join(src, "moments.out"))
    mgot = R.read_moments(str(tmp_path / "moments.out"))
    R.assert_moments(tag, [mgot[k] for k in R.MOM_KEYS], mref)
    # the reference's published goldens of its own test decks (test/short.dat.*)
    pub = os.path.join(R.GOLDEN, "ref_test", "short.dat." + tag)
    if os.path.exists(pub):
        p = R.read_short(pub)
        assert abs(got["energy"] - p["energy"]) < 1e-9
        assert abs(got["init"] - p["init"]) < 1e-9
    R.assert_output_files(tmp_path, read_deck(R.deck_path(tag))["norb"], uhf=bool(len(ref["evals_b"])))


@pytest.mark.parametrize("tag", ["631.co", "w2_631gs", "water_631gs_cation"])
def test_scf_exact_roots_vs_reference_x2(tag):
    """The opt-in exact two-root quadrature (rys2_exact = 1) against the reference with
    its two-root fit replaced by the exact rule (oracle/ref_root2_exact.cpp)."""
    say("scf with exact 2-root quadrature " + tag)
    path = R.deck_path(tag)
    b, t = _tints(path)
    t.set("rys2_exact", 1)
    cls = unomol_amd.UnRestrictedHartreeFock if b.deck["nelec"] % 2 else unomol_amd.RestrictedHartreeFock
    hf = cls(b, t)
    e = hf.findEnergy()
    assert hf.result.converged
    ref = R.read_short(os.path.join(R.GOLDEN, "ref_run_x2", tag, "short.gs.out"))
    assert abs(e - ref["energy"]) < 1e-9
    assert abs(hf.result.init_energy - ref["init"]) < 1e-9


@pytest.mark.parametrize("tag,uhf", [("B_w16_631gs", False), ("B_w16_631gs_cation", True), ("631.h2o", False)])
def test_scf_purification_matches_eigensolve(tag, uhf):
    """SCF densities by trace-correcting purification (FP64 MFMA GEMMs, no eigensolve
    between Fock builds) against the eigensolve path: the same energy (1e-10 Eh), orbital
    energies (1e-9 Eh, from the final eigensolve) and density (1e-9), with every iteration's
    density purified."""
    say("scf purification vs eigensolve " + tag)
    path = os.path.join(R.GOLDEN, "decks", tag + ".dat") if tag.startswith("B_") else R.deck_path(tag)
    b, t = _tints(path)
    cls = unomol_amd.UnRestrictedHartreeFock if uhf else unomol_amd.RestrictedHartreeFock
    t.set("purify", 0)
    ref = cls(b, t)
    e0 = ref.findEnergy(accel=2, eps=1e-10, cflag=3, maxits=100)
    assert ref.result.converged and ref.result.purified == 0
    t.set("purify", 1)
    hf = cls(b, t)
    e1 = hf.findEnergy(accel=2, eps=1e-10, cflag=3, maxits=100)
    r = hf.result
    say("purification: %d densities purified in %d iterations, |dE| %.2e, diag %.3f s vs %.3f s" %
        (r.purified, r.iterations, abs(e1 - e0), r.diag_s, ref.result.diag_s))
    assert r.converged
    assert r.purified >= r.iterations * (2 if uhf else 1) - 2
    assert abs(e1 - e0) < 1e-10
    np.testing.assert_allclose(hf.evals, ref.evals, atol=1e-9)
    assert np.abs(hf.P - ref.P).max() < 1e-9
    t.set("purify", -1)


@pytest.mark.parametrize("tag,uhf", [("B_w16_631gs", False), ("B_w16_631gs_cation", True), ("631.h2o", False)])
def test_scf_newton_schulz_x_matches_canonical(tag, uhf):
    """X = S^-1/2 by Newton-Schulz (x_newton = 1) against the reference's canonical
    X = U s^-1/2 (x_newton = 0): the same energy (1e-10 Eh), orbital energies and density
    (1e-9); the reference damping path, as the reference runs it."""
    say("scf newton-schulz X vs canonical X " + tag)
    path = os.path.join(R.GOLDEN, "decks", tag + ".dat") if tag.startswith("B_") else R.deck_path(tag)
    b, t = _tints(path)
    cls = unomol_amd.UnRestrictedHartreeFock if uhf else unomol_amd.RestrictedHartreeFock
    out = []
    for xn in (0, 1):
        t.set("x_newton", xn)
        hf = cls(b, t)
        e = hf.findEnergy(accel=2, eps=1e-10, cflag=3, maxits=100)
        assert hf.result.converged
        steps = t.get("x_newton_steps")
        assert (steps > 0) if xn else (steps == -1), (xn, steps)
        out.append((e, hf.evals.copy(), hf.P.copy()))
    t.set("x_newton", -1)
    (e0, ev0, P0), (e1, ev1, P1) = out
    say("newton-schulz X: |dE| %.2e, max |d eps| %.2e, max |dP| %.2e" %
        (abs(e1 - e0), np.abs(ev1 - ev0).max(), np.abs(P1 - P0).max()))
    assert abs(e1 - e0) < 1e-10
    np.testing.assert_allclose(ev1, ev0, atol=1e-9)
    assert np.abs(P1 - P0).max() < 1e-9


def test_scf_diis_same_energy(tmp_path):
    say("scf diis water 6-31G*")
    path = R.deck_path("w2_631gs")
    b, t = _tints(path)
    e_ref = unomol_amd.RestrictedHartreeFock(b, t).findEnergy()
    hf = unomol_amd.RestrictedHartreeFock(b, t)
    e = hf.findEnergy(accel=2, eps=1e-11)
    assert hf.result.converged
    assert abs(e - e_ref) < 1e-9
    assert hf.result.iterations < 40


def test_rys_fp32_tables_vs_mpmath(golden):
    say("fp32 rys tables vs mpmath")
    z = np.load(os.path.join(golden, "rys_mp.npz"))
    for n in range(1, 6):
        u, w = unomol_amd.rys_eval(n, z["x"], fp32=True)
        np.testing.assert_allclose(u, z["u"][n, :, :n], rtol=2e-6)
        np.testing.assert_allclose(w, z["w"][n, :, :n], rtol=2e-6)


def test_fock_fp32_quartets_vs_fp64():
    """FP32 quartet kernels against the FP64 ones (pinned to the oracle above): every
    quartet in FP32 stays within single-precision error of the FP64 G, and the mixed
    build (FP32 only below the Schwarz bound 1e-6) within 1e-9."""
    say("fock fp32 / mixed vs fp64 w2_631gs")
    path = R.deck_path("w2_631gs")
    deck = read_deck(path)
    n = deck["norb"]
    rng = np.random.default_rng(3)
    A = rng.standard_normal((n, n)) * 0.05
    P = _pack(A + A.T + np.eye(n))
    b, t = _tints(path)
    G64 = np.zeros_like(P)
    t.formGmatrix(P, G64)
    assert t.get("last_quartets32") == 0
    t.set("precision", 1)
    t.set("fp32_threshold", 1e30)      # every quartet in FP32
    Gall = np.zeros_like(P)
    t.formGmatrix(P, Gall)
    assert t.get("last_quartets32") == t.get("last_quartets")
    assert np.abs(Gall - G64).max() < 2e-6 * np.abs(G64).max()
    t.set("fp32_threshold", 1e-6)
    Gmix = np.zeros_like(P)
    t.formGmatrix(P, Gmix)
    assert 0 < t.get("last_quartets32") < t.get("last_quartets")
    assert np.abs(Gmix - G64).max() < 1e-9


@pytest.mark.parametrize("tag", ["631.h2o", "w2_631gs"])
def test_fock_rows_rhf_vs_oracle(tag):
    """Every bra through the row-merged kernel (exchange merged in LDS rows, then added
    row-contiguously) against the oracle's brute-force G."""
    say("fock rhf, all bras row-merged, vs oracle " + tag)
    path = R.deck_path(tag)
    deck = read_deck(path)
    n = deck["norb"]
    rng = np.random.default_rng(11)
    A = rng.standard_normal((n, n)) * 0.1
    P = _pack(A + A.T)
    b, t = _tints(path, schwarz=1e-16)
    t.set("row_min_kets", 1)
    G = np.zeros_like(P)
    t.formGmatrix(P, G)
    Gref = _oracle_g(deck, P)
    assert np.abs(G - Gref).max() < 1e-11
    t.formGmatrix(P, G)
    assert np.abs(G - 2 * Gref).max() < 2e-11


@pytest.mark.parametrize("uhf", [False, True])
@pytest.mark.parametrize("variant", [4, 5])
def test_fock_rows_bra_groups_vs_oracle(uhf, variant):
    """Every bra row-merged in groups of 2 (variant 4) or 4 (variant 5) bras that load each
    ket's records once for the whole group, against the oracle's brute-force G (RHF on the
    two-water deck, UHF on the water cation); a second build accumulates."""
    tag = "water_631gs_cation" if uhf else "w2_631gs"
    say("fock %s, bra groups (variant %d), vs oracle %s" % ("uhf" if uhf else "rhf", variant, tag))
    path = R.deck_path(tag)
    deck = read_deck(path)
    n = deck["norb"]
    rng = np.random.default_rng(15 + variant)
    A = rng.standard_normal((n, n)) * 0.1
    B = rng.standard_normal((n, n)) * 0.1
    PA, PB = _pack(A + A.T), _pack(B + B.T)
    b, t = _tints(path, schwarz=1e-16)
    t.set("row_min_kets", 1)
    t.set("row_variant", variant)
    if uhf:
        GA, GB = np.zeros_like(PA), np.zeros_like(PB)
        t.formGmatrix(PA, PB, GA, GB)
        RA, RB = _oracle_g(deck, PA, PB)
        assert np.abs(GA - RA).max() < 1e-11 and np.abs(GB - RB).max() < 1e-11
    else:
        G = np.zeros_like(PA)
        t.formGmatrix(PA, G)
        Gref = _oracle_g(deck, PA)
        assert np.abs(G - Gref).max() < 1e-11
        t.formGmatrix(PA, G)
        assert np.abs(G - 2 * Gref).max() < 2e-11


def test_fock_rows_uhf_vs_oracle():
    say("fock uhf, all bras row-merged, vs oracle water_631gs_cation")
    path = R.deck_path("water_631gs_cation")
    deck = read_deck(path)
    n = deck["norb"]
    rng = np.random.default_rng(12)
    A = rng.standard_normal((n, n)) * 0.1
    B = rng.standard_normal((n, n)) * 0.1
    PA, PB = _pack(A + A.T), _pack(B + B.T)
    b, t = _tints(path, schwarz=1e-16)
    t.set("row_min_kets", 1)
    GA, GB = np.zeros_like(PA), np.zeros_like(PB)
    t.formGmatrix(PA, PB, GA, GB)
    RA, RB = _oracle_g(deck, PA, PB)
    assert np.abs(GA - RA).max() < 1e-11
    assert np.abs(GB - RB).max() < 1e-11


@pytest.mark.parametrize("uhf", [False, True])
def test_fock_rows_vs_waves_config_b(uhf):
    """16-water 6-31G* (304 functions): the default split (bras with >= 192 kets
    row-merged) against every bra as wave tasks, FP64 and all-FP32 quartets."""
    say("fock rows vs waves, config B deck, uhf=%s" % uhf)
    path = os.path.join(R.GOLDEN, "decks", "B_w16_631gs.dat")
    deck = read_deck(path)
    n = deck["norb"]
    rng = np.random.default_rng(13)
    A = rng.standard_normal((n, n)) * 0.02
    Bm = rng.standard_normal((n, n)) * 0.02
    PA, PB = _pack(A + A.T + np.eye(n)), _pack(Bm + Bm.T + np.eye(n))
    b, t = _tints(path)

    def build():
        if uhf:
            GA, GB = np.zeros_like(PA), np.zeros_like(PB)
            t.formGmatrix(PA, PB, GA, GB)
            return np.concatenate([GA, GB])
        G = np.zeros_like(PA)
        t.formGmatrix(PA, G)
        return G

    for prec, thr, tol in ((0, 1e-6, 1e-11), (1, 1e30, 1e-5)):
        t.set("precision", prec)
        t.set("fp32_threshold", thr)
        t.set("row_merge", 0)
        Gwave = build()
        scale = np.abs(Gwave).max()
        t.set("row_merge", 1)
        # LDS variants: rows + Rys table, every bra row, the first bra shell's rows; each
        # with the bra's exchange density rows read from global memory and staged in LDS;
        # then groups of 2 and 4 bras sharing their kets (FP64 spaces of the small classes;
        # the others keep the default variant)
        for variant in (1, 2, 3, 4, 5):
            for dens in (0, 1):
                t.set("row_variant", variant)
                t.set("row_density_lds", dens)
                Grow = build()
                assert np.abs(Grow - Gwave).max() < tol * max(scale, 1.0), (prec, variant, dens,
                                                                             np.abs(Grow - Gwave).max(), scale)
        t.set("row_variant", 0)
        t.set("row_density_lds", -1)


@pytest.mark.parametrize("uhf", [False, True])
def test_fock_deterministic_bitwise(uhf):
    """deterministic = 1: the Fock build accumulates in fixed point (two int64 words per
    entry, integer atomics), so repeated builds of the same density give bitwise the same
    G whatever order the waves run in.  Small deck: against the oracle's brute-force G
    (1e-11).  16-water 6-31G* deck: three builds bitwise equal, FP64 and mixed precision,
    and within 1e-11 (FP64) / 1e-5 (all-FP32 quartets) of the default build's scale."""
    say("fock deterministic, uhf=%s" % uhf)
    tag = "water_631gs_cation" if uhf else "631.h2o"
    path = R.deck_path(tag)
    deck = read_deck(path)
    n = deck["norb"]
    rng = np.random.default_rng(21)
    A = rng.standard_normal((n, n)) * 0.1
    Bm = rng.standard_normal((n, n)) * 0.1
    PA, PB = _pack(A + A.T), _pack(Bm + Bm.T)
    b, t = _tints(path, schwarz=1e-16)
    t.set("deterministic", 1)
    assert t.get("deterministic") == 1

    def build(t, PA, PB):
        if uhf:
            GA, GB = np.zeros_like(PA), np.zeros_like(PB)
            t.formGmatrix(PA, PB, GA, GB)
            return np.concatenate([GA, GB])
        G = np.zeros_like(PA)
        t.formGmatrix(PA, G)
        return G

    G1 = build(t, PA, PB)
    ref = _oracle_g(deck, PA, PB) if uhf else _oracle_g(deck, PA)
    ref = np.concatenate(ref) if uhf else ref
    assert np.abs(G1 - ref).max() < 1e-11
    assert np.array_equal(build(t, PA, PB), G1)

    path = os.path.join(R.GOLDEN, "decks", "B_w16_631gs.dat")
    n = read_deck(path)["norb"]
    rng = np.random.default_rng(22)
    A = rng.standard_normal((n, n)) * 0.02
    Bm = rng.standard_normal((n, n)) * 0.02
    PA, PB = _pack(A + A.T + np.eye(n)), _pack(Bm + Bm.T + np.eye(n))
    b, t = _tints(path)
    for prec, thr, tol in ((0, 1e-6, 1e-11), (1, 1e30, 1e-5)):
        t.set("precision", prec)
        t.set("fp32_threshold", thr)
        t.set("deterministic", 0)
        Gdef = [build(t, PA, PB) for _ in range(3)]
        scale = max(np.abs(Gdef[0]).max(), 1.0)
        t.set("deterministic", 1)
        Gdet = [build(t, PA, PB) for _ in range(3)]
        say("precision %d: default builds differing from the first: %d of 2 (max %.1e); "
            "deterministic vs default %.1e" %
            (prec, sum(not np.array_equal(g, Gdef[0]) for g in Gdef[1:]),
             max(np.abs(g - Gdef[0]).max() for g in Gdef[1:]), np.abs(Gdet[0] - Gdef[0]).max()))
        assert all(np.array_equal(g, Gdet[0]) for g in Gdet[1:])
        assert np.abs(Gdet[0] - Gdef[0]).max() < tol * scale


def test_scf_deterministic_bitwise():
    """A whole SCF with deterministic Fock builds (16-water 6-31G*, mixed precision,
    incremental builds, purification, DIIS) run twice: bitwise the same energy, orbital
    energies and density; and the default SCF's energy to 1e-10 Eh."""
    say("scf deterministic, 16 waters")
    path = os.path.join(R.GOLDEN, "decks", "B_w16_631gs.dat")
    b, t = _tints(path)
    kw = dict(accel=2, eps=1e-8, cflag=3, maxits=100, precision=1, incremental=1)
    e_def = unomol_amd.RestrictedHartreeFock(b, t).findEnergy(**kw)
    t.set("deterministic", 1)
    out = []
    for _ in range(2):
        hf = unomol_amd.RestrictedHartreeFock(b, t)
        e = hf.findEnergy(**kw)
        assert hf.result.converged
        out.append((e, hf.evals.copy(), hf.P.copy(), hf.result.iterations))
    (e0, ev0, P0, it0), (e1, ev1, P1, it1) = out
    say("deterministic SCF: %.12f (%d its) twice; default %.12f" % (e0, it0, e_def))
    assert e0 == e1 and it0 == it1
    assert np.array_equal(ev0, ev1) and np.array_equal(P0, P1)
    assert abs(e0 - e_def) < 1e-10


@pytest.mark.parametrize("uhf", [False, True])
def test_fock_launch_streams_same_matrix(uhf):
    """The ERI launches of a build spread over three streams (the default) against one
    stream (16-water 6-31G*, FP64 and mixed builds): the same G to FP64 rounding of the
    accumulation order, and bitwise the same in deterministic mode, whose fixed-point sums
    are converted only after every stream's launches have finished."""
    say("fock launch streams 3 vs 1, uhf=%s" % uhf)
    path = os.path.join(R.GOLDEN, "decks", "B_w16_631gs.dat")
    n = read_deck(path)["norb"]
    rng = np.random.default_rng(29)
    A = rng.standard_normal((n, n)) * 0.02
    Bm = rng.standard_normal((n, n)) * 0.02
    PA, PB = _pack(A + A.T + np.eye(n)), _pack(Bm + Bm.T + np.eye(n))
    b, t = _tints(path)

    def build():
        if uhf:
            GA, GB = np.zeros_like(PA), np.zeros_like(PB)
            t.formGmatrix(PA, PB, GA, GB)
            return np.concatenate([GA, GB])
        G = np.zeros_like(PA)
        t.formGmatrix(PA, G)
        return G

    assert t.get("launch_streams") == 3
    for prec in (0, 1):
        t.set("precision", prec)
        for det in (0, 1):
            t.set("deterministic", det)
            t.set("launch_streams", 1)
            G1 = build()
            t.set("launch_streams", 3)
            G3 = build()
            scale = max(np.abs(G1).max(), 1.0)
            say("precision %d deterministic %d: max |dG| %.2e" % (prec, det, np.abs(G3 - G1).max()))
            if det:
                assert np.array_equal(G3, G1)
            else:
                assert np.abs(G3 - G1).max() < 1e-12 * scale
    t.set("deterministic", 0)
    t.set("precision", 0)


def test_autotune_keeps_fock_matrix():
    """The autotuner times the row-merge configurations of every quartet space and keeps the
    fastest: the Fock matrix built afterwards (RHF and UHF, FP64 and mixed) equals the
    default configuration's, and a new context of the same deck picks the choices up."""
    say("autotune, config B deck")
    path = os.path.join(R.GOLDEN, "decks", "B_w16_631gs.dat")
    deck = read_deck(path)
    n = deck["norb"]
    rng = np.random.default_rng(14)
    A = rng.standard_normal((n, n)) * 0.02
    Bm = rng.standard_normal((n, n)) * 0.02
    PA, PB = _pack(A + A.T + np.eye(n)), _pack(Bm + Bm.T + np.eye(n))
    b, t = _tints(path)
    G0 = np.zeros_like(PA)
    t.formGmatrix(PA, G0)
    GA0, GB0 = np.zeros_like(PA), np.zeros_like(PB)
    t.formGmatrix(PA, PB, GA0, GB0)
    assert t.autotune(PA) > 0.0
    assert t.autotune(PA, PB) > 0.0
    G1 = np.zeros_like(PA)
    t.formGmatrix(PA, G1)
    GA1, GB1 = np.zeros_like(PA), np.zeros_like(PB)
    t.formGmatrix(PA, PB, GA1, GB1)
    assert np.abs(G1 - G0).max() < 1e-11 * max(1.0, np.abs(G0).max())
    assert np.abs(GA1 - GA0).max() < 1e-11 * max(1.0, np.abs(GA0).max())
    assert np.abs(GB1 - GB0).max() < 1e-11 * max(1.0, np.abs(GB0).max())
    t.set("precision", 1)
    Gm0 = np.zeros_like(PA)
    t.formGmatrix(PA, Gm0)
    assert t.autotune(PA, mixed=True) > 0.0
    Gm1 = np.zeros_like(PA)
    t.formGmatrix(PA, Gm1)
    assert np.abs(Gm1 - Gm0).max() < 1e-11 * max(1.0, np.abs(Gm0).max())
    # a second context of the same deck builds the same G with the remembered choices
    b2, t2 = _tints(path)
    G2 = np.zeros_like(PA)
    t2.formGmatrix(PA, G2)
    assert np.abs(G2 - G0).max() < 1e-11 * max(1.0, np.abs(G0).max())


def test_scf_incremental_same_energy():
    say("scf incremental builds w2_631gs")
    path = R.deck_path("w2_631gs")
    b, t = _tints(path)
    e_full = unomol_amd.RestrictedHartreeFock(b, t).findEnergy(accel=2, eps=1e-11, cflag=3)
    hf = unomol_amd.RestrictedHartreeFock(b, t)
    e = hf.findEnergy(accel=2, eps=1e-11, cflag=3, incremental=1)
    assert hf.result.converged
    assert hf.result.nfull < hf.result.nfock
    assert abs(e - e_full) < 1e-9


def test_scf_mixed_precision_energy(tmp_path):
    """Mixed precision + incremental builds on a 4-water 6-31G* cluster: the final energy
    is within 1e-8 Eh of the plain FP64 SCF, with FP32 quartets used on the way."""
    say("scf mixed precision 4-water 6-31G*")
    deck = str(tmp_path / "w4.dat")
    unomol_amd.generate_deck(deck, "water_cluster", 4, 4001, "6-31G*", 0, 1e-9, 100)
    b = unomol_amd.Basis(deck)
    t = unomol_amd.TwoElectronInts(b)
    ref = unomol_amd.RestrictedHartreeFock(b, t)
    e64 = ref.findEnergy(accel=2, eps=1e-9, cflag=3)
    assert ref.result.converged
    hf = unomol_amd.RestrictedHartreeFock(b, t)
    e = hf.findEnergy(accel=2, eps=1e-9, cflag=3, precision=1, incremental=1)
    r = hf.result
    assert r.converged
    assert r.quartets32 > 0
    assert abs(e - e64) < 1e-8
    ev64 = ref.evals.copy()
    np.testing.assert_allclose(hf.evals, ev64, atol=1e-6)
    # the mixed phase screens at its own threshold (default 3e-9); at the context's own
    # threshold (1e-13) it evaluates more quartets and lands on the same energy
    assert t.get("schwarz_mixed") == 3e-9
    t.set("schwarz_mixed", 1e-13)
    hf2 = unomol_amd.RestrictedHartreeFock(b, t)
    e2 = hf2.findEnergy(accel=2, eps=1e-9, cflag=3, precision=1, incremental=1)
    assert hf2.result.converged
    assert hf2.result.quartets > r.quartets
    assert abs(e2 - e64) < 1e-8
    t.set("schwarz_mixed", 3e-9)


def test_scf_mixed_precision_energy_uhf(tmp_path):
    """The same for the open-shell path: a 4-water 6-31G* cation (UHF doublet), mixed
    precision + incremental builds + the mixed-phase screening against the plain FP64 UHF
    SCF: the final energy within 1e-8 Eh, FP32 quartets used on the way."""
    say("scf mixed precision 4-water 6-31G* cation (UHF)")
    deck = str(tmp_path / "w4p.dat")
    unomol_amd.generate_deck(deck, "water_cluster", 4, 4001, "6-31G*", 1, 1e-9, 150)
    b = unomol_amd.Basis(deck)
    assert b.deck["nelec"] % 2 == 1
    t = unomol_amd.TwoElectronInts(b)
    ref = unomol_amd.UnRestrictedHartreeFock(b, t)
    e64 = ref.findEnergy(accel=2, eps=1e-9, cflag=3)
    assert ref.result.converged
    hf = unomol_amd.UnRestrictedHartreeFock(b, t)
    e = hf.findEnergy(accel=2, eps=1e-9, cflag=3, precision=1, incremental=1)
    assert hf.result.converged
    assert hf.result.quartets32 > 0
    assert abs(e - e64) < 1e-8, (e, e64)


B_TAGS = ["B_w16_631gs", "B_w16_631gs_cation"]


@pytest.mark.parametrize("tag", B_TAGS)
def test_config_b_trajectory_vs_reference(tmp_path, tag):
    """16-water 6-31G* cluster (304 functions), RHF and the +1 UHF doublet: three SCF
    iterations from the core-Hamiltonian guess with the reference's damping, against the
    reference rebuilt from its sources on the same deck."""
    say("config B trajectory " + tag)
    src = os.path.join(R.GOLDEN, "ref_run", tag)
    (tmp_path / "patin.dat").write_text(open(os.path.join(src, "patin.dat")).read())
    unomol_amd.run_deck(tmp_path)
    ref = R.read_short(os.path.join(src, "short.gs.out"))
    got = R.read_short(str(tmp_path / "short.gs.out"))
    assert abs(got["init"] - ref["init"]) < 1e-9
    assert abs(got["energy"] - ref["energy"]) < 1e-9
    assert abs(got["ediff"] - ref["ediff"]) < 1e-9
    mref = R.read_scfout(os.path.join(src, "scfout.gs.out"))
    mine = R.read_scfout(str(tmp_path / "scfout.gs.out"))
    np.testing.assert_allclose(mine["evals"], mref["evals"], atol=1e-9)


def _mlp_oracle():
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
    import mlp_oracle
    return mlp_oracle


def test_mlp_forward_and_gradient_vs_oracle():
    """BF16 matrix-core forward/backward of the guess network against the numpy oracle
    with the same bf16 rounding points (float32 accumulation order is the only
    difference)."""
    say("mlp forward / gradient vs oracle")
    MO = _mlp_oracle()
    net = unomol_amd.MLP(128, 128, 3, 256, seed=1)
    rng = np.random.default_rng(8)
    X = rng.standard_normal((200, 128)).astype(np.float32)
    T = (0.1 * rng.standard_normal((200, 256))).astype(np.float32)
    M = (rng.uniform(size=(200, 256)) > 0.4).astype(np.float32)
    p = net.get_params()
    Y = net.forward(X)
    Yo = MO.forward(p, net.dims, X)[0]
    assert np.abs(Y - Yo).max() < 2e-2 * np.abs(Yo).max()
    assert np.linalg.norm(Y - Yo) < 2e-3 * np.linalg.norm(Yo)
    loss, g = net.loss_grad(X, T, M)
    lo, go = MO.loss_grad(p, net.dims, X, T, M)
    assert abs(loss - lo) < 1e-3 * abs(lo)
    assert np.linalg.norm(g - go) < 1e-2 * np.linalg.norm(go)
    o = 0
    for l in range(len(net.dims) - 1):
        nw = net.dims[l] * net.dims[l + 1]
        gw, gwo = g[o:o + nw], go[o:o + nw]
        assert np.linalg.norm(gw - gwo) < 2e-2 * np.linalg.norm(gwo)
        o += nw + net.dims[l + 1]


def test_mlp_training_reduces_loss():
    say("mlp training on a smooth synthetic map")
    rng = np.random.default_rng(9)
    X = rng.standard_normal((4096, 128)).astype(np.float32)
    R = rng.standard_normal((128, 256)).astype(np.float32) / np.sqrt(128)
    T = np.tanh(X @ R).astype(np.float32)
    M = np.ones_like(T)
    net = unomol_amd.MLP(128, 128, 3, 256, seed=2)
    l0, _ = net.loss_grad(X[:1024], T[:1024], M[:1024])
    loss = net.train(X, T, M, steps=400, batch=512, lr=3e-3, seed=3)
    assert loss < 0.25 * l0


def test_learned_guess_fewer_iterations(tmp_path):
    """A network trained on densities of small clusters (2-6 waters) converges a held-out
    8-water cluster to the core-Hamiltonian result (1e-8 Eh) in fewer iterations."""
    say("learned guess: training set")
    from unomol_amd import guess_train as GT
    X, Y, M = GT.make_training_set(n_clusters=24, min_w=2, max_w=6, seed=500)
    say("learned guess: training on %d samples" % len(X))
    net, loss = GT.train_guess(X, Y, M, steps=800, batch=512)
    deck = str(tmp_path / "w8.dat")
    unomol_amd.generate_deck(deck, "water_cluster", 8, 8123, "6-31G*", 0, 1e-9, 100)
    b = unomol_amd.Basis(deck)
    t = unomol_amd.TwoElectronInts(b)
    ref = unomol_amd.RestrictedHartreeFock(b, t)
    e0 = ref.findEnergy(accel=2, eps=1e-9, cflag=3)
    unomol_amd.set_guess_model(t, net)
    hf = unomol_amd.RestrictedHartreeFock(b, t)
    e1 = hf.findEnergy(accel=2, eps=1e-9, cflag=3, guess=2)
    say("learned guess: %d vs %d iterations (core-H), loss %.2e" % (hf.result.iterations, ref.result.iterations,
                                                                   loss))
    assert ref.result.converged and hf.result.converged
    assert abs(e1 - e0) < 1e-8
    assert hf.result.iterations < ref.result.iterations
    # the guess itself is far closer than the core Hamiltonian's
    assert abs(hf.result.init_energy - e0) < 0.2 * abs(ref.result.init_energy - e0)


@pytest.mark.parametrize("tag", ["A_water_631g", "water_631gs", "631.ch4", "w2_631gs"])
def test_oneint_device_vs_reference(golden, tag):
    """Device S, T, H against the host restatement (itself pinned to the reference's
    intsout matrices at 1e-13) and, where the fixture holds them, the reference's own."""
    say("one-electron matrices on the device " + tag)
    path = R.deck_path(tag)
    b, t = _tints(path)
    S, T, H = t.oneint()
    hS, hT, hH = unomol_amd.oneint(read_deck(path))
    for got, ref in ((S, hS), (T, hT), (H, hH)):
        assert np.abs(got - ref).max() < 1e-13 * max(1.0, np.abs(ref).max())
    npz = os.path.join(golden, "eri_%s.npz" % tag)
    if os.path.exists(npz):
        z = np.load(npz)
        assert np.abs(S - z["S"]).max() < 1e-13
        assert np.abs(T - z["T"]).max() < 1e-12
        assert np.abs(H - z["H"]).max() < 1e-12


# SF6 in its TZ2P basis: the core-Hamiltonian HOMO is one of a triply degenerate set
# (orbitals 34-36 at -25.74293 Eh), so the guess density, and with it the iteration-0
# energy, depends on the rotation the eigensolver picks inside that set; the converged
# energy, orbital energies and polarizabilities do not.
FF_DEGENERATE_CORE_H = ("tz2p.sf6",)
FF_TAGS = ["n.dhdz", "d6s3p.h2", "dh95.h2", "dh95.c2h2", "dh95.co2", "tzp.ch4", "tz2p.sf6"]


@pytest.mark.parametrize("tag", FF_TAGS)
def test_finite_field_decks_vs_reference(tmp_path, tag):
    """Reference decks with the finite-field analysis (int_flag[1] = 1) and the nitrogen
    atom: ground state and, where requested, the three field SCFs and polarizabilities
    against the reference rebuilt from its sources."""
    say("finite-field deck " + tag)
    src = os.path.join(R.GOLDEN, "ref_run", tag)
    (tmp_path / "patin.dat").write_text(open(os.path.join(R.GOLDEN, "ref_test", "patin.dat." + tag)).read())
    r = unomol_amd.run_deck(tmp_path)
    ref = R.read_short(os.path.join(src, "short.gs.out"))
    got = R.read_short(str(tmp_path / "short.gs.out"))
    mref = R.read_scfout(os.path.join(src, "scfout.gs.out"))
    mine = R.read_scfout(str(tmp_path / "scfout.gs.out"))
    if tag not in FF_DEGENERATE_CORE_H:
        assert abs(got["init"] - ref["init"]) < 1e-9
    if not mref["converged"]:
        # the reference's damped SCF does not meet its test on this deck within 299
        # iterations: either it sits on the rounding floor of the solution (|dE| ~ 1e-13,
        # the density change just above the eps floor), where this engine may or may not
        # pass the same test but holds the same energy, or it still oscillates, and then
        # so does the same iteration here
        if abs(ref["ediff"]) < 1e-9:
            assert abs(got["energy"] - ref["energy"]) < 1e-9
        else:
            assert not r.converged and mine["iterations"] == mref["iterations"]
        return
    assert r.converged
    assert abs(got["energy"] - ref["energy"]) < 1e-9
    np.testing.assert_allclose(mine["evals"], mref["evals"], atol=1e-9)
    ffref = os.path.join(src, "finitefield.out")
    if os.path.exists(ffref):
        a = R.read_finitefield(ffref)
        b = R.read_finitefield(str(tmp_path / "finitefield.out"))
        assert b["field"] == a["field"]
        np.testing.assert_allclose(b["polnrg"], a["polnrg"], atol=2e-9)
        np.testing.assert_allclose(b["alpha"], a["alpha"], atol=2e-9 * 2 / a["field"] ** 2)


DPM_TAGS = ["dpm_water", "dpm_water_cation"]


@pytest.mark.parametrize("tag", DPM_TAGS)
def test_dpm_scan_vs_reference(tmp_path, tag):
    """DPM polarization-potential scan (int_flag[0] = 1): positron basis on a moving centre,
    the six-Gaussian model potential and one SCF per grid point, against the reference."""
    say("dpm scan " + tag)
    src = os.path.join(R.GOLDEN, "ref_run", tag)
    for fn in ("patin.dat", "posin.bas", "pos.grid.dat"):
        (tmp_path / fn).write_text(open(os.path.join(src, fn)).read())
    unomol_amd.run_deck(tmp_path)
    ref = np.loadtxt(os.path.join(src, "vpol.out"), ndmin=2)
    got = np.loadtxt(str(tmp_path / "vpol.out"), ndmin=2)
    assert got.shape == ref.shape
    np.testing.assert_allclose(got[:, :3], ref[:, :3], atol=1e-10)
    np.testing.assert_allclose(got[:, 3], ref[:, 3], atol=1e-9)                 # vpol
    r4 = (ref[:, :3] ** 2).sum(axis=1) ** 2
    np.testing.assert_allclose(got[:, 4], ref[:, 4], atol=1e-9 * 2 * r4.max())  # alfa = -2 vpol r^4
    if ref.shape[1] > 5:
        np.testing.assert_allclose(got[:, 5:], ref[:, 5:], atol=1e-9)          # vstat, vstat + vpol
    sref = np.loadtxt(os.path.join(src, "spol.out"), ndmin=2)
    sgot = np.loadtxt(str(tmp_path / "spol.out"), ndmin=2)
    # energies print with 10 digits after the point in e-format: compare to that precision
    np.testing.assert_allclose(sgot[:, 1:5], sref[:, 1:5], rtol=2e-10, atol=1e-9)


def _batch_decks(tmp_path, count, seed):
    """A mixed pool like the batched workload: distorted small molecules and 1-3-water
    clusters in 6-31G / 6-31G*, every other one a +1 cation (UHF)."""
    rng = np.random.default_rng(seed)
    paths = []
    for i in range(count):
        basis = "6-31G*" if rng.uniform() < 0.5 else "6-31G"
        charge = i % 2
        p = tmp_path / ("m%d.dat" % i)
        if rng.uniform() < 0.6:
            unomol_amd.generate_deck(p, "molecule", int(rng.integers(0, 9)), int(seed * 1000 + i), basis, charge,
                                     1e-10, 200)
        else:
            unomol_amd.generate_deck(p, "water_cluster", int(rng.integers(1, 4)), int(seed * 1000 + i), basis,
                                     charge, 1e-10, 200)
        paths.append(str(p))
    return paths


def test_batched_scf_matches_single(tmp_path):
    """Batched lockstep SCF (one launch per class pair over all molecules, grouped GEMMs,
    batched Jacobi) against the single-molecule SCF of each molecule."""
    say("batched scf: 12 mixed molecules")
    paths = _batch_decks(tmp_path, 12, 7)
    batch = unomol_amd.Batch(paths)
    e, res = batch.run(maxits=100, eps=1e-10, cflag=3, accel=2)
    for i, p in enumerate(paths):
        b = unomol_amd.Basis(p)
        t = unomol_amd.TwoElectronInts(b)
        cls = unomol_amd.UnRestrictedHartreeFock if b.deck["nelec"] % 2 else unomol_amd.RestrictedHartreeFock
        hf = cls(b, t)
        e1 = hf.findEnergy(maxits=100, eps=1e-10, cflag=3, accel=2)
        assert res[i].converged and hf.result.converged
        assert abs(e[i] - e1) < 1e-8, (i, p, e[i], e1)
    assert batch.get("graph_builds") >= 1 and batch.get("graph_replays") >= batch.get("graph_builds")


def test_batched_graph_and_queue_match_plain_launches(tmp_path):
    """The batched Fock build captured as a HIP graph and replayed, with work-queue
    launches, against plain launches of the same static quartet lists (no graph, no queue):
    the same energies, iteration counts (to one) and densities on every molecule; the graph replays
    every build of a run and the second run of the same batch (a fresh capture) agrees with
    the first.  The group launches run on three streams (the default; a captured graph then
    has parallel branches) and, in the last two runs, on one."""
    say("batched graph/queue vs plain: 16 mixed molecules")
    paths = _batch_decks(tmp_path, 16, 11)
    runs = []
    for graph, queue, streams in ((1, 1, 3), (1, 1, 3), (0, 0, 3), (0, 2, 3), (1, 1, 1), (0, 0, 1)):
        batch = unomol_amd.Batch(paths) if not runs or graph == 0 and queue == 0 else batch
        batch.set("graph", graph)
        batch.set("queue", queue)
        batch.set("launch_streams", streams)
        assert batch.get("launch_streams") == streams
        e, res = batch.run(maxits=100, eps=1e-10, cflag=2, accel=2)
        say("  graph %d queue %d streams %d: builds %d replays %d spaces %d launches/build %d" %
            (graph, queue, streams, batch.get("graph_builds"), batch.get("graph_replays"), batch.get("space_builds"),
             batch.get("launches_per_build")))
        assert all(r.converged for r in res)
        if graph:
            assert batch.get("graph_builds") == batch.get("space_builds")
            assert batch.get("graph_replays") >= max(r.iterations for r in res)
        else:
            assert batch.get("graph_builds") == 0
        runs.append((e, [r.iterations for r in res], [batch.density(i) for i in range(len(paths))]))
    e0, it0, P0 = runs[0]
    for e, it, P in runs[1:]:
        assert np.abs(e - e0).max() < 1e-9
        # atomics add in a different order from run to run: a molecule at its threshold may
        # take one iteration more or fewer
        assert max(abs(a - b) for a, b in zip(it, it0)) <= 1
        assert max(np.abs(a - b).max() for a, b in zip(P, P0)) < 1e-7


# (the open-shell atoms b.dhdz / f.dhdz are left out: their unpaired electron sits along
# an arbitrary axis of a degenerate p level, so P, the vectors and the overlaps are fixed
# only up to a rotation of the atom)
OUT_TAGS = ["631.h2o", "431.co", "water_631gs_cation"]


def _levels(e, tol=1e-5):
    """Groups of orbitals whose energies lie within tol of a neighbour (degenerate levels:
    only the subspace they span is fixed, not the vectors)."""
    groups, cur = [], [0]
    for i in range(1, len(e)):
        if e[i] - e[i - 1] < tol:
            cur.append(i)
        else:
            groups.append(cur)
            cur = [i]
    groups.append(cur)
    return groups


@pytest.mark.parametrize("tag", OUT_TAGS)
def test_output_files_vs_reference(tmp_path, tag):
    """Every file the drop-in run writes against the reference's own files on the same deck
    (printing switched on: eigenvectors and Mulliken overlap populations), then a restart
    from the reference's PMATRIX.DAT (scf_flag[2] = 1, RHF.hpp:120-123)."""
    say("output files vs reference " + tag)
    src = os.path.join(R.GOLDEN, "ref_out", tag, "print")
    (tmp_path / "patin.dat").write_text(open(os.path.join(src, "patin.dat")).read())
    r = unomol_amd.run_deck(tmp_path)
    assert r.converged
    n = read_deck(str(tmp_path / "patin.dat"))["norb"]
    n2 = n * (n + 1) // 2
    uhf = bool(r.uhf)
    g = lambda fn: str(tmp_path / fn)              # noqa: E731
    s = lambda fn: os.path.join(src, fn)            # noqa: E731
    # intsout.out: the header and shell dump as text, S T H to 1e-13
    hg, mg = R.read_intsout(g("intsout.out"))
    hs, ms = R.read_intsout(s("intsout.out"))
    assert hg == hs
    for a, b in zip(mg, ms):
        assert a.shape == b.shape == (n2,)
        assert np.abs(a - b).max() < 1e-13 * max(1.0, np.abs(b).max())
    # MOMINTS.DAT / RMOM.DAT: 80-byte records, same order and packed indices, values 1e-12
    for fn in ("MOMINTS.DAT", "RMOM.DAT"):
        a, b = R.read_momints(g(fn)), R.read_momints(s(fn))
        assert os.path.getsize(g(fn)) == os.path.getsize(s(fn)) == 80 * n2
        assert (a["ijr"] == b["ijr"]).all()
        assert np.abs(a["v"] - b["v"]).max() < 1e-12 * max(1.0, np.abs(b["v"]).max())
    # PMATRIX.DAT: packed doubles, alpha then beta
    pg, ps = np.fromfile(g("PMATRIX.DAT")), np.fromfile(s("PMATRIX.DAT"))
    assert pg.shape == ps.shape == ((2 if uhf else 1) * n2,)
    assert np.abs(pg - ps).max() < 1e-8
    # scfout: energies, Mulliken populations and charges, printed eigenvectors up to sign
    ms_, mg_ = R.read_scfout(s("scfout.gs.out")), R.read_scfout(g("scfout.gs.out"))
    np.testing.assert_allclose(mg_["evals"], ms_["evals"], atol=1e-9)
    popg, chg = R.read_mulliken(g("scfout.gs.out"))
    pops, chs = R.read_mulliken(s("scfout.gs.out"))
    assert popg.shape == pops.shape
    np.testing.assert_allclose(popg, pops, atol=1e-8)
    np.testing.assert_allclose(chg, chs, atol=2e-9)
    # printed eigenvectors: the projector onto every level's vectors (sign- and
    # rotation-free), to the printed 10 digits
    vg, vs = R.read_eigvecs(g("scfout.gs.out"), n), R.read_eigvecs(s("scfout.gs.out"), n)
    assert len(vg) == len(vs) == (2 if uhf else 1)
    single_all = []
    for (eg, cg), (es, cs) in zip(vg, vs):
        assert cg.shape == cs.shape == (n, n)
        np.testing.assert_allclose(eg, es, atol=1e-9)
        single = np.zeros(n, dtype=bool)
        for lev in _levels(es):
            Pg, Ps = cg[:, lev] @ cg[:, lev].T, cs[:, lev] @ cs[:, lev].T
            assert np.abs(Pg - Ps).max() < 1e-7, lev
            if len(lev) == 1:
                single[lev[0]] = True
        single_all.append(single)
    # overlaps.out: Mulliken overlap populations per packed pair (and spin)
    og, os_ = R.read_overlaps(g("overlaps.out")), R.read_overlaps(s("overlaps.out"))
    assert og.shape == os_.shape == (n2, 2 if uhf else 1)
    assert np.abs(og - os_).max() < 1e-8
    # mol_dipmom.out: MO transition dipoles (printed to 7 digits) up to the vectors' signs,
    # for pairs of non-degenerate orbitals
    dg, ds = R.read_mol_dipmom(g("mol_dipmom.out")), R.read_mol_dipmom(s("mol_dipmom.out"))
    assert len(dg) == len(ds) == (2 if uhf else 1)
    for bg, bs, keep in zip(dg, ds, single_all):
        assert bg.keys() == bs.keys()
        for (i, j), v in bs.items():
            if keep[i] and keep[j]:
                np.testing.assert_allclose(np.abs(bg[(i, j)]), np.abs(v), rtol=2e-6, atol=2e-8)
    # restart from the reference's converged density
    say("restart from PMATRIX.DAT " + tag)
    rs = os.path.join(R.GOLDEN, "ref_out", tag, "restart")
    rd = tmp_path / "restart"
    rd.mkdir()
    (rd / "patin.dat").write_text(open(os.path.join(rs, "patin.dat")).read())
    (rd / "PMATRIX.DAT").write_bytes(open(os.path.join(rs, "PMATRIX.in"), "rb").read())
    r2 = unomol_amd.run_deck(rd)
    assert r2.converged
    a, b = R.read_short(str(rd / "short.gs.out")), R.read_short(os.path.join(rs, "short.gs.out"))
    assert abs(a["init"] - b["init"]) < 1e-9
    assert abs(a["energy"] - b["energy"]) < 1e-9
    assert r2.iterations <= R.read_scfout(os.path.join(rs, "scfout.gs.out"))["iterations"] + 2


def test_error_paths(tmp_path):
    """Failures the reference turns into exit() come back as status codes with the
    reference's message: f shells (l = 3) rejected by the device context (the ERI kernels
    cover s, p, d), and a run that reaches maxits returns UNO_ENOCONV with its results and
    writes the reference's banner (RHF.hpp:413) into scfout.gs.out."""
    say("error paths")
    deck = open(R.deck_path("631.h2o")).read().split("\n")
    # an f shell on the oxygen: one primitive, l = 3
    hdr = deck[0].split()
    nsh, norb, ncen = int(hdr[0]), int(hdr[1]), int(hdr[2])
    deck[0] = "%7d%7d%7d  3" % (nsh + 1, norb + 10, ncen)
    fdeck = "\n".join(deck).rstrip("\n") + "\n  1  3  1\n  0.8000000000   1.0000000000\n"
    b = unomol_amd.Basis(text=fdeck)
    with pytest.raises(N.UnoError) as ei:
        unomol_amd.TwoElectronInts(b)
    assert ei.value.code == N.UNO_EINVAL
    # non-convergence: two iterations on a deck that needs ~45
    lines = open(R.deck_path("631.h2o")).read().split("\n")
    lines[1] = "     10      2"
    (tmp_path / "patin.dat").write_text("\n".join(lines))
    r = unomol_amd.run_deck(tmp_path)
    assert not r.converged and r.iterations == 2
    text = (tmp_path / "scfout.gs.out").read_text()
    assert text.startswith("WARNING CONVERGENCE _NOT_ REACHED!")
    assert R.read_scfout(str(tmp_path / "scfout.gs.out"))["converged"] is False


SHIM_EXE = os.path.join(R.GOLDEN, "..", "..", "oracle", "_ref", "Unomol_shim")


@pytest.mark.skipif(not os.path.exists(SHIM_EXE), reason="the shim program is built where the reference is mounted")
@pytest.mark.parametrize("tag", ["631.h2o", "water_631gs_cation"])
def test_reference_drivers_on_shim_vs_reference(tmp_path, tag):
    """The reference's own program (its unchanged Unomol.cc / RHF.hpp / UHF.hpp) built on
    the drop-in shim, i.e. every formGmatrix / rsp / sp_trans call going through the C ABI
    to the GPU, against the reference's outputs on the same deck: energies and orbital
    energies to 1e-9 Eh."""
    import subprocess
    say("reference drivers on the shim " + tag)
    (tmp_path / "patin.dat").write_text(open(R.deck_path(tag)).read())
    subprocess.run([os.path.abspath(SHIM_EXE)], cwd=str(tmp_path), check=True, timeout=300,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    src = os.path.join(R.GOLDEN, "ref_run", tag)
    a = R.read_short(str(tmp_path / "short.gs.out"))
    b = R.read_short(os.path.join(src, "short.gs.out"))
    assert abs(a["init"] - b["init"]) < 1e-9
    assert abs(a["energy"] - b["energy"]) < 1e-9
    ma, mb = R.read_scfout(str(tmp_path / "scfout.gs.out")), R.read_scfout(os.path.join(src, "scfout.gs.out"))
    np.testing.assert_allclose(ma["evals"], mb["evals"], atol=1e-9)
