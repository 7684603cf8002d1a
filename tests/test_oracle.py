"""CPU tests that pin the oracle (the CPU restatement in oracle/) to the reference and
to what the mathematics fixes.  No GPU is used."""
import math
import os
import sys

import numpy as np
import pytest

import oracle_ffi as O
import refdata as R
from unomol_amd.deck import read_deck

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def rysmp(golden):
    return np.load(os.path.join(golden, "rys_mp.npz"))


@pytest.fixture
def exact_rys2():
    """The oracle's exact two-root quadrature for one test (the default is the reference's)."""
    O.set_rys2_mode(False)
    yield
    O.set_rys2_mode(True)


def test_oracle_boys_vs_mpmath(rysmp):
    for i, x in enumerate(rysmp["x"]):
        f = O.boys(11, x)
        np.testing.assert_allclose(f, rysmp["boys"][i], rtol=2e-15, atol=0)


def test_oracle_rys_vs_mpmath(rysmp, exact_rys2):
    for n in range(1, 6):
        for i, x in enumerate(rysmp["x"]):
            u, w = O.rys(n, x)
            np.testing.assert_allclose(u, rysmp["u"][n, i, :n], rtol=4e-15)
            np.testing.assert_allclose(w, rysmp["w"][n, i, :n], rtol=4e-15)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7])
def test_oracle_rys_moment_identity(n, exact_rys2):
    # sum_i w_i u_i^k = F_k(x) for k < 2n: fixes roots and weights completely
    rng = np.random.default_rng(n)
    for x in np.concatenate([rng.uniform(0, 80, 20), [0.0, 1e-6, 15.0, 33.0]]):
        u, w = O.rys(n, x)
        f = O.boys(2 * n - 1, x)
        mom = np.array([np.sum(w * u ** k) for k in range(2 * n)])
        np.testing.assert_allclose(mom, f, rtol=5e-14)


def test_oracle_boys_closed_form():
    for x in [1e-8, 0.3, 2.0, 17.0, 45.0]:
        f0 = 0.5 * math.sqrt(math.pi / x) * math.erf(math.sqrt(x))
        assert abs(O.boys(0, x)[0] - f0) < 2e-16 * max(1.0, f0) * 4


def test_oracle_rys2_reference_form(golden):
    """The oracle's reference-compatible two-root quadrature (its six coefficients fitted
    to the reference's outputs on 15 < x <= 22) against every committed sample of the
    reference's root2 on (15, 40], to 1e-14 relative; the fit recovers a form whose
    large-x limit is the 4-point Gauss-Hermite rule (checked by the exact rule at x = 40,
    where the reference's e^-x terms are 4e-17)."""
    z = np.load(O.RYS2_NPZ)
    x, r, w = z["x"], z["r"], z["w"]
    m = (x > 15.0) & (x <= 40.0)
    worst = 0.0
    for xi, ri, wi in zip(x[m], r[m], w[m]):
        u, ww = O.rys(2, float(xi))
        worst = max(worst, np.max(np.abs(u - ri / (1 + ri)) / (ri / (1 + ri))), np.max(np.abs(ww - wi) / wi))
    assert worst < 1e-14
    # at 16 the reference departs from the exact rule by ~1e-4: the mode matters
    O.set_rys2_mode(False)
    ue, we = O.rys(2, 16.0)
    ue40, we40 = O.rys(2, 40.0)
    O.set_rys2_mode(True)
    ur, wr = O.rys(2, 16.0)
    ur40, wr40 = O.rys(2, 40.0)
    assert np.max(np.abs(ur - ue) / ue) > 1e-5
    np.testing.assert_allclose(ur40, ue40, rtol=1e-14)
    np.testing.assert_allclose(wr40, we40, rtol=1e-14)


def _horner_table(path, name, k, nc):
    import re
    text = open(path).read()
    body = re.search(r"%s\[(\d+)\] = \{(.*?)\};" % name, text, re.S).group(2)
    vals = np.array([float(v.rstrip("f")) for v in body.replace("\n", "").split(",") if v.strip()])
    return vals.reshape(-1, 4, nc)


def test_product_rys2_table_vs_reference_samples(golden):
    """The product's reference-compatible two-root table (unomol_amd/csrc/rys_ref2.inc),
    evaluated here the way the kernels do (Horner in s = 2(x-k)-1 on unit intervals),
    against the reference's own root2 samples: 2e-15 relative on (15, 40) in double,
    2e-6 in float; outside, the exact table agrees with the reference's fit to 5e-13.
    x = 15 itself is left out: there the reference takes its x <= 15 branch and the table
    its [15, 16) interval (3.5e-4 apart at that one argument)."""
    path = os.path.join(ROOT, "unomol_amd", "csrc", "rys_ref2.inc")
    z = np.load(O.RYS2_NPZ)
    x, r, w = z["x"], z["r"], z["w"]
    m = (x < 48.0) & (x != 15.0)
    x, r, w = x[m], r[m], w[m]
    ref = np.concatenate([r / (1 + r), w], axis=1)
    for name, nc, tol_in in (("uno_rys_ref2_mono", 13, 2e-15), ("uno_rys_ref2_mono_f", 7, 2e-6)):
        tab = _horner_table(path, name, 0, nc)
        assert tab.shape[0] == 48
        k = np.floor(x).astype(int)
        s = 2 * (x - k) - 1
        got = np.zeros_like(ref)
        for f in range(4):
            c = tab[k, f, :]
            v = c[:, nc - 1].copy()
            for j in range(nc - 2, -1, -1):
                v = v * s + c[:, j]
            got[:, f] = v
        rel = np.abs(got - ref) / np.abs(ref)
        inside = (x > 15.0) & (x < 40.0)
        assert rel[inside].max() < tol_in, name
        assert rel[~inside].max() < max(5e-13, tol_in), name


@pytest.mark.parametrize("tag", ["A_water_631g", "water_631gs"])
def test_oracle_eri_vs_reference(golden, tag):
    """Every canonical ERI against the reference's own ERI code at 1e-12 in every class;
    with the exact two-root rule, against the reference with its fit replaced by it."""
    z = np.load(os.path.join(golden, "eri_%s.npz" % tag))
    d = read_deck(os.path.join(golden, "decks", tag + ".dat"))
    v, f = O.eri_all(d)
    assert (f == z["ijkl"]).all(), "canonical order differs from the reference"
    assert np.abs(v - z["ref"]).max() < 1e-12
    O.set_rys2_mode(False)
    try:
        v2, f2 = O.eri_all(d)
    finally:
        O.set_rys2_mode(True)
    assert np.abs(v2 - z["ref_x2"]).max() < 1e-12
    # the reference's two-root fit moves the ERIs of the 2-root classes by up to ~1e-9
    L = _fun_l(d)[f].sum(axis=1)
    two = (L == 2) | (L == 3)
    assert np.abs(v2 - v)[~two].max() == 0.0
    assert np.abs(v2 - v)[two].max() > 1e-12


def _fun_l(d):
    fl = []
    for l in d["sh_l"]:
        fl += [l] * ((l + 1) * (l + 2) // 2)
    return np.array(fl)


def test_oracle_ssss_closed_form():
    # two normalised 1-primitive s functions on two centres: (ab|ab) closed form
    a, b, R = 0.8, 1.3, 1.7
    d = {"sh_l": np.array([0, 0], np.int32), "sh_npr": np.array([1, 1], np.int32),
         "sh_cen": np.array([0, 1], np.int32), "sh_poff": np.array([0, 1], np.int32),
         "alpha": np.array([a, b]), "coef": np.array([1.0, 1.0]),
         "xyz": np.array([[0, 0, 0], [0, 0, R]], float), "charge": np.array([1.0, 1.0])}
    v, f = O.eri_all(d)
    na, nb = (2 * a / math.pi) ** 0.75, (2 * b / math.pi) ** 0.75
    p = a + b
    K = math.exp(-a * b / p * R * R)
    # (10|10): bra and ket are both the a-b product
    T = 0.0
    exact = 2 * math.pi ** 2.5 / (p * p * math.sqrt(2 * p)) * K * K * (na * nb) ** 2 * O.boys(0, T)[0]
    idx = [i for i in range(len(v)) if tuple(f[i]) == (1, 0, 1, 0)][0]
    assert abs(v[idx] - exact) < 1e-14
    # (00|11): Coulomb between the two charge clouds
    T = (p / 2) * 0 + (2 * a * 2 * b / (2 * a + 2 * b)) * R * R
    exact = 2 * math.pi ** 2.5 / (2 * a * 2 * b * math.sqrt(2 * a + 2 * b)) * na ** 2 * nb ** 2 * O.boys(0, T)[0]
    idx = [i for i in range(len(v)) if tuple(f[i]) == (1, 1, 0, 0)][0]
    assert abs(v[idx] - exact) < 1e-14


def _full_eri(v, f, n):
    E = np.zeros((n, n, n, n))
    for val, (i, j, k, l) in zip(v, f):
        for (a, b, c, d) in ((i, j, k, l), (j, i, k, l), (i, j, l, k), (j, i, l, k),
                             (k, l, i, j), (l, k, i, j), (k, l, j, i), (l, k, j, i)):
            E[a, b, c, d] = val
    return E


def _pack(M):
    n = M.shape[0]
    return np.array([M[i, j] for i in range(n) for j in range(i + 1)])


def _unpack(p, n):
    M = np.zeros((n, n))
    k = 0
    for i in range(n):
        for j in range(i + 1):
            M[i, j] = M[j, i] = p[k]
            k += 1
    return M


def test_oracle_gmatrix_bruteforce(golden):
    d = read_deck(os.path.join(golden, "decks", "A_water_631g.dat"))
    n = d["norb"]
    v, f = O.eri_all(d)
    E = _full_eri(v, f, n)
    rng = np.random.default_rng(0)
    A = rng.standard_normal((n, n))
    P = A + A.T
    G = O.gmat_rhf(v, f, n, _pack(P))
    Gref = 2 * np.einsum("ijkl,kl->ij", E, P) - np.einsum("ikjl,kl->ij", E, P)
    np.testing.assert_allclose(_unpack(G, n), Gref, atol=1e-12)
    B = rng.standard_normal((n, n))
    PA, PB = P, B + B.T
    GA, GB = O.gmat_uhf(v, f, n, _pack(PA), _pack(PB))
    J = np.einsum("ijkl,kl->ij", E, PA + PB)
    np.testing.assert_allclose(_unpack(GA, n), J - np.einsum("ikjl,kl->ij", E, PA), atol=1e-12)
    np.testing.assert_allclose(_unpack(GB, n), J - np.einsum("ikjl,kl->ij", E, PB), atol=1e-12)


def test_oracle_rsp_and_sp_trans():
    rng = np.random.default_rng(3)
    for n in (1, 2, 7, 30):
        A = rng.standard_normal((n, n))
        A = A + A.T
        e, Z = O.rsp(n, _pack(A))
        np.testing.assert_allclose(e, np.linalg.eigvalsh(A), atol=1e-12)
        assert np.all(np.diff(e) >= 0)
        np.testing.assert_allclose(A @ Z, Z * e, atol=1e-11)
        np.testing.assert_allclose(Z.T @ Z, np.eye(n), atol=1e-12)
        X = rng.standard_normal((n, n))
        np.testing.assert_allclose(_unpack(O.sp_trans(n, _pack(A), X), n), X.T @ A @ X, atol=1e-11)


@pytest.mark.parametrize("tag", ["A_water_631g", "water_631gs"])
def test_oracle_oneint_vs_reference(golden, tag):
    z = np.load(os.path.join(golden, "eri_%s.npz" % tag))
    d = read_deck(os.path.join(golden, "decks", tag + ".dat"))
    S, T, H = O.oneint(d)
    for a, b in ((S, z["S"]), (T, z["T"]), (H, z["H"])):
        assert np.abs(a - b).max() < 1e-13


SCF_TAGS = [t for t in R.RHF_TAGS if t not in R.DEGENERATE_TAGS] + R.UHF_TAGS + \
    ["A_water_631g", "water_631gs", "water_631gs_cation"]


@pytest.mark.parametrize("tag", SCF_TAGS)
def test_oracle_scf_vs_reference_run(tag):
    """The oracle's SCF against the reference rebuilt from its sources: initial and final
    energies and every orbital energy to 1e-9 Eh, moments to 1e-7; and the reference's
    published goldens of its own decks (test/short.dat.*) to 1e-9 Eh."""
    d = read_deck(R.deck_path(tag))
    out, ev, P, mom = O.scf(d)
    ref = R.read_scfout(os.path.join(R.GOLDEN, "ref_run", tag, "scfout.gs.out"))
    short = R.read_short(os.path.join(R.GOLDEN, "ref_run", tag, "short.gs.out"))
    assert abs(out.energy - short["energy"]) < 1e-9
    assert abs(out.init_energy - short["init"]) < 1e-9
    n = d["norb"]
    np.testing.assert_allclose(ev[:n], ref["evals"], atol=1e-9)
    if out.uhf:
        np.testing.assert_allclose(ev[n:], ref["evals_b"], atol=1e-9)
    pub = os.path.join(R.GOLDEN, "ref_test", "short.dat." + tag)
    if os.path.exists(pub):
        p = R.read_short(pub)
        assert abs(out.energy - p["energy"]) < 1e-9
    mref = R.read_moments(os.path.join(R.GOLDEN, "ref_run", tag, "moments.out"))
    R.assert_moments(tag, mom, mref)


@pytest.mark.parametrize("tag", R.RHF_TAGS)
def test_reference_goldens_reproduced_by_rebuild(tag):
    """The committed regression goldens (test/short.dat.*) and the rebuilt reference agree."""
    a = R.read_short(os.path.join(R.GOLDEN, "ref_test", "short.dat." + tag))
    b = R.read_short(os.path.join(R.GOLDEN, "ref_run", tag, "short.gs.out"))
    assert abs(a["energy"] - b["energy"]) < 2e-12


@pytest.mark.parametrize("tag", R.DEGENERATE_TAGS)
def test_oracle_scf_degenerate_start(tag):
    """N2/STO-3G: the damped SCF from the core guess passes through iterations where a
    degenerate pi pair straddles the occupation boundary, so which solution it reaches
    depends on the eigensolver's arbitrary basis of that subspace.  Only self-consistency
    is required here; the reference's ground-state energy is a lower bound."""
    d = read_deck(R.deck_path(tag))
    out, ev, P, mom = O.scf(d)
    short = R.read_short(os.path.join(R.GOLDEN, "ref_run", tag, "short.gs.out"))
    assert out.iterations < d["maxits"]
    assert out.energy > short["energy"] - 1e-9


def test_mlp_oracle_gradient_finite_differences():
    """The MLP oracle's backward pass (float64 mode) against central differences of its
    own loss: pins the formulas the GPU gradient test relies on."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import mlp_oracle as MO
    rng = np.random.default_rng(0)
    dims = [6, 5, 5, 4]
    npar = sum(dims[l] * dims[l + 1] + dims[l + 1] for l in range(len(dims) - 1))
    p = rng.standard_normal(npar) * 0.5
    X = rng.standard_normal((7, 6))
    T = rng.standard_normal((7, 4))
    M = (rng.uniform(size=(7, 4)) > 0.3).astype(np.float64)
    _, g = MO.loss_grad(p, dims, X, T, M, bf16=False)
    h = 1e-6
    for k in range(0, npar, 3):
        e = np.zeros(npar)
        e[k] = h
        lp, _ = MO.loss_grad(p + e, dims, X, T, M, bf16=False)
        lm, _ = MO.loss_grad(p - e, dims, X, T, M, bf16=False)
        assert abs((lp - lm) / (2 * h) - g[k]) < 1e-7 * max(1.0, abs(g[k]))


def test_bf16_rounding_ties_to_even():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import mlp_oracle as MO
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e-3], dtype=np.float32)
    r = MO.bf16_round(x)
    np.testing.assert_array_equal(r[:4], np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0], dtype=np.float32))
    assert r[4] == np.float32(-2.5)
    assert abs(r[5] - 3.0e-3) <= 3.0e-3 * 2 ** -8
