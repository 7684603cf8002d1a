#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ (run in the build container, where the
reference sources are mounted; the GPU host never runs this).

Sources, all of them the reference's own data or its own code:
  * ref_test/      the reference's committed regression decks and goldens
                   (test/patin.dat.*, scfout.dat.*, short.dat.*, moments.dat.*), copied as data;
  * ref_run/<tag>/ outputs of the reference rebuilt from its sources (oracle/_ref/Unomol)
                   on those decks and on this project's generated decks (config A water
                   6-31G, water 6-31G*, its +1 cation, a 2-water cluster, the 16-water
                   cluster and its cation, the DPM scans);
  * ref_run_x2/    the same runs with only the reference's 2-root fit (Rys.cpp:413-631)
                   replaced by exact roots (oracle/_ref/Unomol_x2);
  * eri_<tag>.npz  every canonical ERI of small decks from the reference's ERI code
                   (oracle/_ref/libref_harness.so) and from the same code with its 2-root
                   fit replaced by exact roots (libref_harness_x2.so), plus its S, T, H;
  * rys_mp.npz     Rys roots/weights and Boys F_m from 40-digit mpmath at probe points.

usage: make -C oracle && make -C oracle ref && python tests/golden/make_golden.py
       python tests/golden/make_golden.py --config-b    (the 16-water runs only, a few minutes)
       python tests/golden/make_golden.py --extra       (finite-field decks and the N atom)
       python tests/golden/make_golden.py --dpm         (DPM scans of water and its cation)
       python tests/golden/make_golden.py --outputs     (every output file of a few decks, with
                                                         printing switched on, and a restart)
"""
import ctypes as C
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "unomol_amd", "csrc"))

REF_TAGS = [b + "." + m for b in ("3g", "431", "631") for m in ("h2", "hf", "h2o", "nh3", "ch4", "n2", "co")]
REF_TAGS += ["b.dhdz", "f.dhdz", "c.dhdz", "o.dhdz"]
GEN_DECKS = {   # tag -> (kind, n, seed, basis, charge)
    "A_water_631g": ("water", 1, 0, "6-31G", 0),
    "water_631gs": ("water", 1, 0, "6-31G*", 0),
    "water_631gs_cation": ("water", 1, 0, "6-31G*", 1),
    "w2_631gs": ("water_cluster", 2, 2001, "6-31G*", 0),
}
ERI_TAGS = ["A_water_631g", "water_631gs"]
# the 16-water 6-31G* cluster (304 functions), RHF and its +1 UHF doublet, run for three
# SCF iterations from the core-Hamiltonian guess: the reference's damped SCF does not
# converge on this cluster within 300 iterations, so parity is taken on the trajectory
B_DECKS = {
    "B_w16_631gs": ("water_cluster", 16, 16001, "6-31G*", 0),
    "B_w16_631gs_cation": ("water_cluster", 16, 16001, "6-31G*", 1),
}
B_MAXITS = 3
# further reference decks: the finite-field analysis ones (int_flag[1] = 1: polarizabilities
# in finitefield.out) and the nitrogen atom (UHF quartet)
EXTRA_TAGS = ["n.dhdz", "d6s3p.h2", "dh95.h2", "dh95.c2h2", "dh95.co2", "tzp.ch4", "tz2p.sf6"]
# DPM polarization-potential scans (int_flag[0] = 1) of the 6-31G water and its cation: a
# small positron basis on the moving centre (posin.bas) and three grid points
# (pos.grid.dat); both input files are this project's synthetic choices
DPM_DECKS = {"dpm_water": ("water", 1, 0, "6-31G", 0), "dpm_water_cation": ("water", 1, 0, "6-31G", 1)}
POSIN = "4 6 1\n1 0\n2.0 1.0\n1 0\n0.5 1.0\n1 0\n0.125 1.0\n1 1\n0.3 1.0\n"
GRID = "3\n0.0 0.0 -4.0\n0.0 3.0 3.5\n4.5 0.5 0.5\n"


def run_ref(deck_text, outdir, exe="Unomol", extra_files=None):
    os.makedirs(outdir, exist_ok=True)
    with tempfile.TemporaryDirectory() as tmp:
        with open(os.path.join(tmp, "patin.dat"), "w") as f:
            f.write(deck_text)
        for name, text in (extra_files or {}).items():
            with open(os.path.join(tmp, name), "w") as f:
                f.write(text)
        subprocess.run([os.path.join(ROOT, "oracle", "_ref", exe)], cwd=tmp, check=True,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        for fn in ("short.gs.out", "scfout.gs.out", "moments.out", "finitefield.out", "vpol.out", "spol.out"):
            if fn in ("finitefield.out", "vpol.out", "spol.out") and not os.path.exists(os.path.join(tmp, fn)):
                continue
            shutil.copy(os.path.join(tmp, fn), os.path.join(outdir, fn))
        for name, text in (extra_files or {}).items():
            with open(os.path.join(outdir, name), "w") as f:
                f.write(text)
        with open(os.path.join(outdir, "patin.dat"), "w") as f:
            f.write(deck_text)


def harness_eris(lib, path):
    cnt = C.c_long()
    lib.ref_eri_all(path.encode(), C.c_long(0), None, None, C.byref(cnt))
    n = cnt.value
    v = np.zeros(n)
    f = np.zeros((n, 4), dtype=np.int32)
    lib.ref_eri_all(path.encode(), C.c_long(n), v.ctypes.data_as(C.c_void_p), f.ctypes.data_as(C.c_void_p),
                    C.byref(cnt))
    return v, f


def make_config_b():
    import unomol_amd
    gdir = os.path.join(HERE, "decks")
    os.makedirs(gdir, exist_ok=True)
    for tag, (kind, n, seed, basis, charge) in B_DECKS.items():
        path = os.path.join(gdir, tag + ".dat")
        unomol_amd.generate_deck(path, kind, n, seed, basis, charge, 1e-10, B_MAXITS)
        run_ref(open(path).read(), os.path.join(HERE, "ref_run", tag))
        run_ref(open(path).read(), os.path.join(HERE, "ref_run_x2", tag), exe="Unomol_x2")
        print("config B reference run done:", tag, flush=True)


def make_extra():
    dst = os.path.join(HERE, "ref_test")
    for tag in EXTRA_TAGS:
        src = os.path.join(REF, "test", "patin.dat." + tag)
        shutil.copy(src, os.path.join(dst, "patin.dat." + tag))
        run_ref(open(src).read(), os.path.join(HERE, "ref_run", tag))
        run_ref(open(src).read(), os.path.join(HERE, "ref_run_x2", tag), exe="Unomol_x2")
        print("extra reference run done:", tag, flush=True)


def make_dpm():
    import unomol_amd
    gdir = os.path.join(HERE, "decks")
    for tag, (kind, n, seed, basis, charge) in DPM_DECKS.items():
        path = os.path.join(gdir, tag + ".dat")
        unomol_amd.generate_deck(path, kind, n, seed, basis, charge, 1e-10, 300)
        lines = open(path).read().split("\n")
        # int_flag line (the fourth line): DPM on, finite field off
        assert lines[3].split() == ["0", "0"], lines[3]
        lines[3] = "1 0"
        text = "\n".join(lines)
        open(path, "w").write(text)
        for sub, exe in (("ref_run", "Unomol"), ("ref_run_x2", "Unomol_x2")):
            run_ref(text, os.path.join(HERE, sub, tag), exe=exe, extra_files={"posin.bas": POSIN, "pos.grid.dat": GRID})
        print("dpm reference run done:", tag, flush=True)


# full output sets of the reference (every file it writes, binary ones included) for the
# byte/format checks: the deck with eigenvector and Mulliken-overlap printing switched on
# (prt_flag[0] = prt_flag[2] = 1) and a restart from the reference's own PMATRIX.DAT
# (scf_flag[2] = 1, RHF.hpp:120-123)
# (the open-shell atoms b.dhdz / f.dhdz are left out: their unpaired electron sits along
# an arbitrary axis of a degenerate p level, so P, the vectors and the overlaps are fixed
# only up to a rotation of the atom)
OUT_TAGS = ["631.h2o", "431.co", "water_631gs_cation"]
OUT_FILES = ("short.gs.out", "scfout.gs.out", "moments.out", "intsout.out", "mol_dipmom.out", "overlaps.out",
             "PMATRIX.DAT", "MOMINTS.DAT", "RMOM.DAT")


def _set_line(text, idx, value):
    lines = text.split("\n")
    lines[idx] = value
    return "\n".join(lines)


def make_outputs():
    for tag in OUT_TAGS:
        src = os.path.join(HERE, "ref_run", tag, "patin.dat")
        text = open(src).read()
        assert len(open(src).read().split("\n")[5].split()) == 3
        text = _set_line(text, 5, " 1 0 1")
        for sub, deck, extra in (("print", text, None), ("restart", None, True)):
            outdir = os.path.join(HERE, "ref_out", tag, sub)
            os.makedirs(outdir, exist_ok=True)
            with tempfile.TemporaryDirectory() as tmp:
                d = deck
                if extra:
                    flags = text.split("\n")[4].split()
                    d = _set_line(text, 4, " %s %s 1" % (flags[0], flags[1]))
                    shutil.copy(os.path.join(HERE, "ref_out", tag, "print", "PMATRIX.DAT"),
                                os.path.join(tmp, "PMATRIX.DAT"))
                    shutil.copy(os.path.join(tmp, "PMATRIX.DAT"), os.path.join(outdir, "PMATRIX.in"))
                with open(os.path.join(tmp, "patin.dat"), "w") as f:
                    f.write(d)
                shutil.copy(os.path.join(tmp, "patin.dat"), os.path.join(outdir, "patin.dat"))
                subprocess.run([os.path.join(ROOT, "oracle", "_ref", "Unomol")], cwd=tmp, check=True,
                               stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
                for fn in OUT_FILES:
                    if os.path.exists(os.path.join(tmp, fn)):
                        shutil.copy(os.path.join(tmp, fn), os.path.join(outdir, fn))
        print("output set done:", tag, flush=True)


def main():
    if "--outputs" in sys.argv:
        make_outputs()
        return
    if "--dpm" in sys.argv:
        make_dpm()
        return
    if "--config-b" in sys.argv:
        make_config_b()
        return
    if "--extra" in sys.argv:
        make_extra()
        return
    import unomol_amd
    from unomol_amd.deck import read_deck
    # 1. reference test data
    dst = os.path.join(HERE, "ref_test")
    os.makedirs(dst, exist_ok=True)
    for tag in REF_TAGS:
        for pre in ("patin.dat.", "scfout.dat.", "short.dat.", "moments.dat."):
            src = os.path.join(REF, "test", pre + tag)
            if os.path.exists(src):
                shutil.copy(src, os.path.join(dst, pre + tag))
    # 2. generated decks
    gdir = os.path.join(HERE, "decks")
    os.makedirs(gdir, exist_ok=True)
    for tag, (kind, n, seed, basis, charge) in GEN_DECKS.items():
        unomol_amd.generate_deck(os.path.join(gdir, tag + ".dat"), kind, n, seed, basis, charge, 1e-10, 300)
    # 3. reference runs
    for tag in REF_TAGS:
        run_ref(open(os.path.join(dst, "patin.dat." + tag)).read(), os.path.join(HERE, "ref_run", tag))
    for tag in GEN_DECKS:
        run_ref(open(os.path.join(gdir, tag + ".dat")).read(), os.path.join(HERE, "ref_run", tag))
    # same runs with the reference's 2-root fit replaced by exact roots (Unomol_x2)
    for tag in REF_TAGS + list(GEN_DECKS):
        src = os.path.join(HERE, "ref_run", tag, "patin.dat")
        run_ref(open(src).read(), os.path.join(HERE, "ref_run_x2", tag), exe="Unomol_x2")
    # 4. ERIs and one-electron matrices from the reference's code
    h1 = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref_harness.so"))
    h2 = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref_harness_x2.so"))
    for tag in ERI_TAGS:
        path = os.path.join(gdir, tag + ".dat")
        v1, f1 = harness_eris(h1, path)
        v2, f2 = harness_eris(h2, path)
        assert (f1 == f2).all()
        d = read_deck(path)
        n2 = d["norb"] * (d["norb"] + 1) // 2
        S, T, H = np.zeros(n2), np.zeros(n2), np.zeros(n2)
        h1.ref_oneint(path.encode(), *[a.ctypes.data_as(C.c_void_p) for a in (S, T, H)])
        np.savez_compressed(os.path.join(HERE, "eri_%s.npz" % tag), ijkl=f1.astype(np.int16), ref=v1, ref_x2=v2,
                            S=S, T=T, H=H)
    # 5. mpmath Rys roots / Boys values
    import gen_rys_table as g
    import mpmath as mp
    xs = np.array([0.0, 1e-9, 3e-7, 0.37, 1.0, 2.5, 3.0, 5.0, 7.3, 10.0, 14.99, 15.0, 15.01, 18.0, 20.0, 23.0,
                   25.0, 29.0, 33.0, 35.0, 39.9, 40.0, 41.0, 47.5, 53.0, 59.0, 61.9, 62.0, 67.9, 68.0, 90.0,
                   300.0, 4.0e3])
    U = np.zeros((6, len(xs), 5))
    W = np.zeros((6, len(xs), 5))
    for nr in range(1, 6):
        for i, x in enumerate(xs):
            u, w = g.rys_exact(nr, x)
            U[nr, i, :nr] = [float(a) for a in u]
            W[nr, i, :nr] = [float(a) for a in w]
    F = np.array([[float(g.boys(m, x)) for m in range(12)] for x in xs])
    np.savez_compressed(os.path.join(HERE, "rys_mp.npz"), x=xs, u=U, w=W, boys=F)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
