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
