"""CPU tests of the product library: it loads, exports every C-ABI symbol declared in
include/unomol.h, and its host-side pieces (deck I/O and generator, one-electron and
moment integrals) match the reference.  No device call is made here."""
import os
import re

import numpy as np
import pytest

import unomol_amd
from unomol_amd import _native as N
from unomol_amd.deck import read_deck

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_header_symbols():
    hdr = open(os.path.join(ROOT, "include", "unomol.h")).read()
    names = set(re.findall(r"\b(uno_[a-z_0-9]+)\s*\(", hdr))
    lib = N.lib()
    missing = [s for s in sorted(names) if not hasattr(lib, s)]
    assert not missing, missing
    assert names == set(N.EXPORTS)


@pytest.mark.parametrize("tag", ["A_water_631g", "water_631gs"])
def test_host_oneint_vs_reference(golden, tag):
    z = np.load(os.path.join(golden, "eri_%s.npz" % tag))
    d = read_deck(os.path.join(golden, "decks", tag + ".dat"))
    S, T, H = unomol_amd.oneint(d)
    assert np.abs(S - z["S"]).max() < 1e-13
    assert np.abs(T - z["T"]).max() < 1e-13
    assert np.abs(H - z["H"]).max() < 1e-13


def test_host_moment_ints_vs_oracle(golden):
    import oracle_ffi as O
    d = read_deck(os.path.join(golden, "ref_run", "631.h2o", "patin.dat"))
    np.testing.assert_allclose(unomol_amd.moment_ints(d), O.moment_ints(d), atol=1e-13)


def test_deck_generator_deterministic(tmp_path):
    a, b = tmp_path / "a.dat", tmp_path / "b.dat"
    unomol_amd.generate_deck(a, "water_cluster", 16, 16001, "6-31G*", 0)
    unomol_amd.generate_deck(b, "water_cluster", 16, 16001, "6-31G*", 0)
    assert a.read_text() == b.read_text()
    d = read_deck(a)
    assert d["norb"] == 304 and d["nshell"] == 160 and d["nelec"] == 160
    xyz = d["xyz"].reshape(-1, 3)
    O = xyz[d["charge"] == 8.0] * 0.529177210903
    dist = np.linalg.norm(O[:, None] - O[None], axis=-1) + np.eye(len(O)) * 99
    assert dist.min() >= 2.5 - 1e-9
    L = (16 * 18.015 / 6.02214076e23) ** (1 / 3) * 1e8
    assert O.min() >= -1e-9 and O.max() <= L + 1e-9
    # rigid monomers: O-H 1.809 bohr, angle 104.52 deg
    for w in range(16):
        o, h1, h2 = xyz[3 * w], xyz[3 * w + 1], xyz[3 * w + 2]
        assert abs(np.linalg.norm(h1 - o) - 1.809) < 1e-12
        c = np.dot(h1 - o, h2 - o) / 1.809 ** 2
        assert abs(np.degrees(np.arccos(c)) - 104.52) < 1e-9


def test_config_sizes(tmp_path):
    for n, norb in ((1, 19), (64, 1216)):
        p = tmp_path / ("w%d.dat" % n)
        unomol_amd.generate_deck(p, "water_cluster", n, 64001, "6-31G*", 0)
        assert read_deck(p)["norb"] == norb
    p = tmp_path / "a.dat"
    unomol_amd.generate_deck(p, "water", 1, 0, "6-31G", 0)
    d = read_deck(p)
    assert d["norb"] == 13 and d["nshell"] == 9
    np.testing.assert_allclose(d["xyz"].reshape(-1, 3)[1], [0, 1.4305507125, 1.1072513982], atol=1e-9)


def test_deck_errors(tmp_path):
    p = tmp_path / "bad.dat"
    p.write_text("3 13 1 5\n10 20\n1e-8\n0 0\n2 0 0\n0 0 0\n")
    with pytest.raises(N.UnoError):
        unomol_amd.run_deck(tmp_path)   # no patin.dat there
    with pytest.raises(N.UnoError):
        unomol_amd.generate_deck(tmp_path / "x.dat", "molecule", 0, 1, "cc-pVDZ", 0)


def _read_inc_array(text, name, dtype):
    body = re.search(r"%s\[\d+\] = \{(.*?)\};" % name, text, re.S).group(1)
    vals = []
    for line in body.split("\n"):
        line = line.strip()
        if line and not line.startswith("//"):
            vals += [float(v.rstrip("f")) for v in line.split(",") if v.strip()]
    return np.array(vals, dtype=dtype)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_rys_power_tables_vs_mpmath(golden, prec):
    """The Rys tables the kernels evaluate (rys_mono.inc: the Chebyshev series of each unit
    interval re-expanded in powers of s = 2(x-k)-1), evaluated with Horner in the kernel's
    precision (Gauss-Hermite limit beyond X_n), against 40-digit mpmath roots and weights.
    Double: within 3e-15 relative (numpy rounds the product and the sum apart; the kernels
    fuse them).  Float: a few ulp on top of the 1.7e-8 truncation of the degree-6 series."""
    src = os.path.join(ROOT, "unomol_amd", "csrc")
    t64 = open(os.path.join(src, "rys_table.inc")).read()
    t32 = open(os.path.join(src, "rys_table_f32.inc")).read()
    mono = open(os.path.join(src, "rys_mono.inc")).read()
    xasym = [int(v) for v in re.search(r"uno_rys_xasym\[\d+\] = \{(.*?)\}", t64).group(1).split(",")]
    if prec == "f64":
        ft = np.float64
        off = [int(v) for v in re.search(r"uno_rys_off\[\d+\] = \{(.*?)\}", t64).group(1).split(",")]
        nc = int(re.search(r"#define UNO_RYS_NCOEF (\d+)", t64).group(1))
        coef = _read_inc_array(mono, "uno_rys_mono", ft)
        au = _read_inc_array(t64, "uno_rys_asym_u", ft)
        aw = _read_inc_array(t64, "uno_rys_asym_w", ft)
    else:
        ft = np.float32
        off = [int(v) for v in re.search(r"uno_rys_off_f\[\d+\] = \{(.*?)\}", t32).group(1).split(",")]
        nc = int(re.search(r"#define UNO_RYS_NCOEF_F (\d+)", t32).group(1))
        coef = _read_inc_array(mono, "uno_rys_mono_f", ft)
        au = _read_inc_array(t32, "uno_rys_asym_u_f", ft)
        aw = _read_inc_array(t32, "uno_rys_asym_w_f", ft)
    z = np.load(os.path.join(golden, "rys_mp.npz"))
    for n in range(1, 6):
        for i, x64 in enumerate(z["x"]):
            x = ft(x64)
            if x >= xasym[n]:
                o = n * (n - 1) // 2
                u = au[o:o + n] / x
                w = aw[o:o + n] / np.sqrt(x)
            else:
                k = int(x)
                s = ft(2) * (x - ft(k)) - ft(1)
                vals = []
                for fi in range(2 * n):
                    c = coef[off[n] + (k * 2 * n + fi) * nc: off[n] + (k * 2 * n + fi + 1) * nc]
                    v = c[nc - 1]
                    for m in range(nc - 2, -1, -1):
                        v = v * s + c[m]
                    vals.append(v)
                u, w = np.array(vals[:n]), np.array(vals[n:])
            if prec == "f64":
                rtol = 3e-15
            else:
                # the mpmath reference is taken at x rounded to float like the argument
                rtol = 2e-6 if x64 != float(x) else 1e-6
            np.testing.assert_allclose(u, z["u"][n, i, :n], rtol=rtol)
            np.testing.assert_allclose(w, z["w"][n, i, :n], rtol=rtol)


def test_rys_power_tables_are_chebyshev_re_expansion():
    """rys_mono.inc holds the Chebyshev series of rys_table.inc / rys_table_f32.inc in the
    power basis: at s = -1, 0, 1 the power sums are sum a_k (-1)^k, a_0 and sum a_k, and the
    Chebyshev sums sum c_m (-1)^m, sum c_m T_m(0) and sum c_m; they agree to rounding for
    every interval and function (a dropped or misplaced coefficient fails this)."""
    src = os.path.join(ROOT, "unomol_amd", "csrc")
    mono = open(os.path.join(src, "rys_mono.inc")).read()
    for tab, name, mname, ncn in [("rys_table.inc", "uno_rys_coef", "uno_rys_mono", "UNO_RYS_NCOEF"),
                                  ("rys_table_f32.inc", "uno_rys_coef_f", "uno_rys_mono_f", "UNO_RYS_NCOEF_F")]:
        t = open(os.path.join(src, tab)).read()
        nc = int(re.search(r"#define %s (\d+)" % ncn, t).group(1))
        c = _read_inc_array(t, name, np.float64).reshape(-1, nc)
        a = _read_inc_array(mono, mname, np.float64).reshape(-1, nc)
        assert c.shape == a.shape
        m = np.arange(nc)
        t0 = np.cos(m * np.pi / 2)   # T_m(0)
        t0[np.abs(t0) < 0.5] = 0.0
        scale = np.abs(c).sum(axis=1)
        for cheb, power in [(c @ ((-1.0) ** m), a @ ((-1.0) ** m)), (c @ t0, a[:, 0]), (c.sum(axis=1), a.sum(axis=1))]:
            tol = 4e-15 if name == "uno_rys_coef" else 4e-7
            assert np.all(np.abs(cheb - power) <= tol * scale)


def _rot(seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    if np.linalg.det(Q) < 0:
        Q[:, 0] = -Q[:, 0]
    return Q


def _rotated_deck(d, Q, shift=(0.0, 0.0, 0.0)):
    r = dict(d)
    r["xyz"] = (d["xyz"].reshape(-1, 3) @ Q.T + np.asarray(shift)).copy()
    return r


def _block_rotation(d, Q):
    """Block-diagonal map of all basis functions (shell order) under rotation Q."""
    n = d["norb"]
    M = np.zeros((n, n))
    o = 0
    for l in d["sh_l"]:
        nc = (l + 1) * (l + 2) // 2
        M[o:o + nc, o:o + nc] = unomol_amd.shell_rotation(int(l), Q)
        o += nc
    return M


def _dense(p, n):
    A = np.zeros((n, n))
    A[np.tril_indices(n)] = p
    return A + np.tril(A, -1).T


def _diis_ring(E, K, newest):
    """Error vectors E (rows, oldest first) placed in a ring of K slots ending at `newest`:
    the ring's B matrix."""
    nd = len(E)
    slots = [(newest - nd + 1 + i) % K for i in range(nd)]
    B = np.zeros((K, K))
    for a, sa in enumerate(slots):
        for b, sb in enumerate(slots):
            B[sa, sb] = E[a] @ E[b]
    return B, slots


def test_diis_coefficients_minimise_the_error():
    """Well-conditioned histories: the coefficients sum to one and satisfy the Lagrange
    conditions of min |sum c_i e_i| (E^T E c equal in every component), i.e. the bordered
    system solved by numpy; in a wrapped ring too."""
    rng = np.random.default_rng(3)
    for K, nd, newest in ((8, 2, 1), (8, 5, 4), (8, 8, 2), (6, 6, 5)):
        E = rng.standard_normal((nd, 30)) * np.logspace(0, -3, nd)[:, None]
        B, slots = _diis_ring(E, K, newest)
        c, used = unomol_amd.diis_coefficients(B, nd, newest)
        assert used == nd
        A = np.zeros((nd + 1, nd + 1))
        A[:nd, :nd] = E @ E.T
        A[:nd, nd] = A[nd, :nd] = -1.0
        rhs = np.zeros(nd + 1)
        rhs[nd] = -1.0
        ref = np.linalg.solve(A, rhs)[:nd]
        np.testing.assert_allclose(c[slots], ref, rtol=1e-9, atol=1e-12)
        assert abs(c.sum() - 1.0) < 1e-13
        assert np.count_nonzero(c) == nd
        g = E @ (E.T @ c[slots])
        assert np.ptp(g) < 1e-9 * np.abs(g).max()


def test_diis_coefficients_drop_dependent_history():
    """Error vectors that are all parallel (H2: one occupied-virtual rotation): with three
    or more of them the minimiser is not unique, so the two newest are used and their
    combination vanishes; identical vectors leave the newest Fock matrix alone."""
    u = np.random.default_rng(4).standard_normal(12)
    a = np.array([3e-2, -4e-3, 5e-4, 7e-5])
    B, slots = _diis_ring(a[:, None] * u[None, :], 8, 3)
    c, used = unomol_amd.diis_coefficients(B, 4, 3)
    assert used == 2
    assert np.count_nonzero(c) == 2 and c[3] != 0 and c[2] != 0
    assert abs(c.sum() - 1.0) < 1e-13
    assert abs(c[2] * a[2] + c[3] * a[3]) < 1e-14
    B1, _ = _diis_ring(np.tile(u, (3, 1)), 8, 2)
    c1, used1 = unomol_amd.diis_coefficients(B1, 3, 2)
    assert used1 == 1 and c1[2] == 1.0 and np.count_nonzero(c1) == 1


def test_shell_rotation_is_a_representation():
    Q1, Q2 = _rot(1), _rot(2)
    for l in range(3):
        M1, M2 = unomol_amd.shell_rotation(l, Q1), unomol_amd.shell_rotation(l, Q2)
        nc = M1.shape[0]
        np.testing.assert_allclose(M1 @ unomol_amd.shell_rotation(l, Q1.T), np.eye(nc), atol=1e-13)
        # one of the two composition orders holds for a representation of SO(3)
        M12 = unomol_amd.shell_rotation(l, Q1 @ Q2)
        assert min(np.abs(M12 - M1 @ M2).max(), np.abs(M12 - M2 @ M1).max()) < 1e-13
    np.testing.assert_allclose(unomol_amd.shell_rotation(1, Q1), Q1.T, atol=1e-15)


@pytest.mark.parametrize("deck_tag", ["water_631gs", "631.nh3"])
def test_shell_rotation_maps_overlap(golden, deck_tag):
    """Overlap of a rotated molecule: S = M S' M^T with M from shell_rotation (pins the
    Cartesian d normalisation and component order against the one-electron code)."""
    path = os.path.join(golden, "decks", deck_tag + ".dat")
    if not os.path.exists(path):
        path = os.path.join(golden, "ref_run", deck_tag, "patin.dat")
    d = read_deck(path)
    Q = _rot(3)
    n = d["norb"]
    S = _dense(unomol_amd.oneint(d)[0], n)
    Sr = _dense(unomol_amd.oneint(_rotated_deck(d, Q, (0.3, -1.1, 2.0)))[0], n)
    M = _block_rotation(d, Q)
    np.testing.assert_allclose(S, M @ Sr @ M.T, atol=1e-12)


def _w4(tmp_path):
    p = tmp_path / "w4.dat"
    unomol_amd.generate_deck(p, "water_cluster", 4, 77, "6-31G*", 0)
    return read_deck(p)


def test_guess_samples_assemble_round_trip(tmp_path):
    """Blocks taken from a density in the pair frames and assembled back give the density
    on every atom pair within the cutoff (and zero beyond it)."""
    d = _w4(tmp_path)
    n = d["norb"]
    rng = np.random.default_rng(4)
    A = rng.standard_normal((n, n))
    P = (A + A.T)[np.tril_indices(n)]
    X, Y, M = unomol_amd.guess_samples(d, P)
    Xf, pairs = unomol_amd.guess_features(d)
    np.testing.assert_array_equal(X, Xf)
    assert X.shape == (len(pairs), 128) and Y.shape == (len(pairs), 256)
    Pb = unomol_amd.guess_assemble(d, pairs, Y)
    Dp, Db = _dense(P, n), _dense(Pb, n)
    # function -> atom
    atom = np.concatenate([[c] * ((l + 1) * (l + 2) // 2) for l, c in zip(d["sh_l"], d["sh_cen"])])
    near = np.zeros((n, n), dtype=bool)
    for a, b in pairs:
        near[np.ix_(atom == a, atom == b)] = True
        near[np.ix_(atom == b, atom == a)] = True
    assert np.abs(Db - Dp)[near].max() < 1e-5 * np.abs(Dp).max()
    assert np.all(Db[~near] == 0.0)
    # every mask row covers exactly n_A x n_B entries
    nf = np.bincount(atom, minlength=d["ncen"])
    np.testing.assert_array_equal(M.sum(axis=1), nf[pairs[:, 0]] * nf[pairs[:, 1]])


def test_guess_descriptors_and_blocks_rotation_invariant(tmp_path):
    """Rotating and translating the cluster leaves the descriptors and the frame blocks of
    the correspondingly rotated density unchanged: the learned guess is covariant."""
    d = _w4(tmp_path)
    n = d["norb"]
    rng = np.random.default_rng(5)
    A = rng.standard_normal((n, n))
    D = A + A.T
    Q = _rot(6)
    dr = _rotated_deck(d, Q, (1.5, 0.2, -0.7))
    M = _block_rotation(d, Q)
    Dr = M.T @ D @ M
    X, Y, _ = unomol_amd.guess_samples(d, D[np.tril_indices(n)])
    Xr, Yr, _ = unomol_amd.guess_samples(dr, Dr[np.tril_indices(n)])
    np.testing.assert_allclose(Xr, X, atol=2e-5)
    np.testing.assert_allclose(Yr, Y, atol=2e-5 * np.abs(Y).max())


@pytest.mark.skipif(not os.path.isdir("/root/reference"), reason="needs the reference sources")
def test_reference_drivers_build_on_shim(tmp_path):
    """The reference's unchanged drivers (Unomol.cc, RHF.hpp, UHF.hpp and every other
    reference source but TwoElectronInts and SymmPack.cpp) compile and link against the
    drop-in shim (unomol_amd/shim) and libunomol_hip.so; the program's undefined symbols
    that the library provides are exactly C-ABI entry points declared in include/unomol.h."""
    import subprocess
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "shim", "OUT=%s" % tmp_path], check=True,
                   stdout=subprocess.DEVNULL)
    exe = tmp_path / "Unomol_shim"
    assert exe.exists()
    und = subprocess.run(["nm", "-D", "--undefined-only", str(exe)], check=True, capture_output=True,
                         text=True).stdout
    used = set(re.findall(r"\b(uno_[a-z_0-9]+)\b", und))
    hdr = open(os.path.join(ROOT, "include", "unomol.h")).read()
    declared = set(re.findall(r"\b(uno_[a-z_0-9]+)\s*\(", hdr))
    assert {"uno_ctx_create", "uno_fock_rhf", "uno_fock_uhf", "uno_eigh_packed", "uno_sp_trans"} <= used
    assert used <= declared
