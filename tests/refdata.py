"""Readers for the reference's output text files (formats of RHF.hpp:404-459,
UHF.hpp:399-485, Moments.cpp:252-272) and the golden fixture layout."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
RHF_TAGS = [b + "." + m for b in ("3g", "431", "631") for m in ("h2", "hf", "h2o", "nh3", "ch4", "n2", "co")]
UHF_TAGS = ["b.dhdz", "f.dhdz"]
# decks whose SCF solution depends on the eigenvector basis chosen inside a degenerate
# level at the occupation boundary (see test_oracle_scf_degenerate_start)
DEGENERATE_TAGS = ["3g.n2"]
GEN_TAGS = ["A_water_631g", "water_631gs", "water_631gs_cation", "w2_631gs"]


def deck_path(tag):
    p = os.path.join(GOLDEN, "ref_run", tag, "patin.dat")
    return p


def read_short(path):
    v = [float(x) for x in open(path).read().split()]
    return {"init": v[0], "energy": v[1], "ediff": v[2]}


def read_scfout(path):
    out = {"evals": [], "evals_b": [], "iterations": None, "converged": True}
    section = None
    for ln in open(path):
        if ln.startswith("WARNING CONVERGENCE"):
            out["converged"] = False
        elif ln.startswith("Final Iteration"):
            out["iterations"] = int(ln.split("=")[1])
        elif ln.startswith("Hartree Fock Energy"):
            out["energy"] = float(ln.split("=")[1].split()[0])
        elif ln.startswith("Kinetic Energy"):
            out["kinetic"] = float(ln.split("=")[1].split()[0])
        elif "Beta Orbital Energies" in ln:
            section = "b"
        elif "Orbital Energies" in ln:
            section = "a"
        elif ln.startswith("xxxx") or "Mulliken" in ln:
            if section in ("a", "b") and "Mulliken" in ln:
                section = None
        else:
            parts = ln.split()
            if section in ("a", "b") and len(parts) == 3 and parts[0].isdigit():
                (out["evals"] if section == "a" else out["evals_b"]).append(float(parts[1]))
    out["evals"] = np.array(out["evals"])
    out["evals_b"] = np.array(out["evals_b"])
    return out


def read_moments(path):
    vals = {}
    for ln in open(path):
        p = ln.split()
        if len(p) == 4 and p[0] in ("x", "y", "z", "xx", "xy", "xz", "yy", "yz", "zz"):
            vals[p[0]] = float(p[1])
        elif ln.strip().startswith("Dipole moment"):
            vals["dipole"] = float(ln.split("=")[1])
        elif ln.strip().startswith("Quadrupole moment"):
            vals["quadrupole"] = float(ln.split("=")[1])
    return vals


MOM_KEYS = ["x", "y", "z", "xx", "xy", "xz", "yy", "yz", "zz"]


def _quad(m):
    xx, xy, xz, yy, yz, zz = m
    return np.array([[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]])


def assert_moments(tag, mom, mref, tol=1e-7):
    """Moments to 1e-7 (the %15.7le precision of moments.out).  Open-shell atoms
    (UHF_TAGS) put their unpaired electron along an arbitrary axis inside a degenerate
    level, so only rotation invariants are compared there: |dipole| and the eigenvalues
    of the second-moment tensor."""
    ref = np.array([mref[k] for k in MOM_KEYS])
    mom = np.asarray(mom)
    if tag in UHF_TAGS:
        assert abs(np.linalg.norm(mom[:3]) - np.linalg.norm(ref[:3])) <= tol * 10
        ea = np.linalg.eigvalsh(_quad(mom[3:]))
        eb = np.linalg.eigvalsh(_quad(ref[3:]))
        assert np.all(np.abs(ea - eb) <= tol * np.maximum(1.0, np.abs(eb)) * 10)
        return
    for k in range(9):
        assert abs(mom[k] - ref[k]) <= tol * max(1.0, abs(ref[k])) + tol


def read_finitefield(path):
    """finitefield.out (RHF.hpp:259-270): field magnitude, per-axis (alpha, pol energy)."""
    out = {"alpha": [], "polnrg": []}
    for ln in open(path):
        t = ln.split()
        if "magnitude" in ln:
            out["field"] = float(t[-1])
        elif len(t) == 3 and t[0] in ("x", "y", "z"):
            out["alpha"].append(float(t[1]))
            out["polnrg"].append(float(t[2]))
    return out


# ---- the other output files (RHF.hpp:172-174,494-531,571-660; Moments.cpp:95-186,365-469) --
MOM_DTYPE = np.dtype([("v", "<f8", (9,)), ("ijr", "<u4"), ("pad", "<u4")])   # 80-byte MomInts record


def read_momints(path):
    return np.fromfile(path, dtype=MOM_DTYPE)


def read_intsout(path):
    """(text of the header and shell dump, [S, T, H] packed arrays) of intsout.out."""
    head, mats, cur = [], [], None
    for ln in open(path):
        if ln.strip().endswith("Matrix"):
            cur = []
            mats.append(cur)
        elif cur is not None:
            p = ln.split()
            if len(p) == 3:
                cur.append(float(p[2]))
        else:
            head.append(ln.rstrip())
    return "\n".join(head).strip(), [np.array(m) for m in mats]


def read_overlaps(path):
    rows = [ln.split() for ln in open(path)][2:]
    return np.array([[float(v) for v in r[2:]] for r in rows if r])


def read_mulliken(path):
    """(per-orbital populations [n x 1 or n x 2], per-centre net charges) from scfout."""
    pops, charges, mode = [], [], None
    for ln in open(path):
        if "Mulliken Populations" in ln:
            mode = "pop"
        elif "Mulliken Atomic Charges" in ln:
            mode = "chg"
        elif ln.startswith("xxxx") or "Eigenvectors" in ln:
            if mode == "chg":
                mode = None
        elif mode == "pop":
            p = ln.split()
            if p and p[0].isdigit():
                pops.append([float(v) for v in p[1:]])
        elif mode == "chg":
            p = ln.split()
            if len(p) == 3 and p[0].isdigit():
                charges.append(float(p[2]))
    return np.array(pops), np.array(charges)


def read_eigvecs(path, n):
    """Printed eigenvector blocks of scfout (prt_flag[0]): list per spin of (energies [n],
    coefficients [n x n], column i = vector i).  Only the first n values of a column are
    its coefficients: when n is not a multiple of 4 the reference's print loop
    (RHF.hpp:439-451) runs on past the column's last row and appends 4 - n % 4 + n % 4 values
    read beyond it."""
    out, cur, vals = [], None, []
    for ln in open(path):
        if "Eigenvectors" in ln:
            cur = {"e": [], "c": []}
            out.append(cur)
        elif cur is not None and ln.startswith("Column"):
            cur["e"].append(float(ln.split("EigenEnergy=")[1]))
            vals = []
            cur["c"].append(vals)
        elif cur is not None and ln.strip() and not ln.startswith("xxxx"):
            vals.extend(float(v) for v in ln.split())
    return [(np.array(b["e"]), np.array([c[:n] for c in b["c"]]).T) for b in out]


def read_mol_dipmom(path):
    """Blocks of MO transition dipoles: list per spin of {(i, j): (dx, dy, dz)}."""
    blocks, cur = [], None
    for ln in open(path):
        if "TRANSISITION DIPOLE MOMENTS" in ln:
            cur = {}
            blocks.append(cur)
        elif cur is not None:
            p = ln.split()
            if len(p) == 5 and p[0].isdigit():
                cur[(int(p[0]), int(p[1]))] = tuple(float(v) for v in p[2:])
    return blocks


def assert_output_files(got_dir, norb=None, uhf=False):
    """The drop-in run wrote every file the reference writes, with the reference's record
    layouts: PMATRIX.DAT packed doubles (alpha then beta), MOMINTS.DAT / RMOM.DAT one 80-byte
    record per packed function pair with its packed index, intsout.out / mol_dipmom.out."""
    got_dir = str(got_dir)
    for fn in ("PMATRIX.DAT", "MOMINTS.DAT", "RMOM.DAT", "intsout.out", "mol_dipmom.out"):
        assert os.path.exists(os.path.join(got_dir, fn)), fn
    if norb is not None:
        n2 = norb * (norb + 1) // 2
        assert os.path.getsize(os.path.join(got_dir, "PMATRIX.DAT")) == 8 * n2 * (2 if uhf else 1)
        for fn in ("MOMINTS.DAT", "RMOM.DAT"):
            m = read_momints(os.path.join(got_dir, fn))
            assert len(m) == n2
            assert sorted(m["ijr"].tolist()) == list(range(n2))
