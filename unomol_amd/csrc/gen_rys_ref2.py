#!/usr/bin/env python3
"""Reference-compatible two-root Rys table (rys_ref2.inc) and its golden samples.

The reference's two-root quadrature (Rys::root2, Rys.cpp:413-631) is accurate to about
1e-13 relative everywhere except 15 < x <= 33, where it evaluates the formula it fits for
33 < x <= 40 and is off the exact Rys quadrature by up to 3.5e-4 relative (1e-4 at x = 16,
1.5e-7 at x = 20).  ERIs of the two-root classes ((ps|ss), (pp|ss), (ps|ps), (ds|ss),
(pp|ps), (dp|ss), (ds|ps)) depend on it at the 1e-9 level, so a drop-in for the reference
must reproduce it there.

Method: the reference's own root2, compiled from its sources into
oracle/_ref/libref_harness.so (ref_rys_probe.cpp), is sampled in this container at the
Chebyshev points of the first kind of every unit interval [k, k+1), k = 15 .. 39; the
samples are converted to the kernels' convention (u = t^2 = r / (1 + r) from the
reference's r = t^2 / (1 - t^2); weights as they are) and fitted exactly like the exact
tables: a degree-12 Chebyshev series per interval (degree 6 for the float table),
re-expanded in powers of s = 2(x-k)-1 (gen_rys_monomial.py) for Horner evaluation.
Every other interval (x < 15 and 40 <= x < 48) is the exact table's, where the
reference's fit and the exact quadrature agree to 2e-13.  Each fitted interval is checked
against the reference at 16 interior points per interval; the worst relative differences
are written into the header.

Outputs:
  rys_ref2.inc                   uno_rys_ref2_mono (double) and uno_rys_ref2_mono_f (float):
                                 complete 2-root tables, same layout as the 2-root part of
                                 rys_mono.inc (48 intervals x 4 series x NCOEF)
  tests/golden/rys2_ref.npz      the reference's roots and weights on a dense grid of
                                 [14.5, 41] and at random points (the parity fixtures)

usage (needs the reference build, make -C oracle ref):
  python gen_rys_ref2.py > rys_ref2.inc
"""
import ctypes as C
import os
import re
import sys
from fractions import Fraction

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
import gen_rys_monomial as gm  # noqa: E402

K0, K1 = 15, 40          # fitted intervals [K0, K1)
DEG, DEGF = 12, 6
NR = 2
mp.mp.dps = 40


def ref_roots(lib, x):
    r = (C.c_double * 5)()
    w = (C.c_double * 5)()
    if lib.ref_rys_roots(C.c_int(NR), C.c_double(x), r, w) != 0:
        raise RuntimeError("ref_rys_roots failed at %r" % x)
    return [r[0], r[1]], [w[0], w[1]]


def kernel_form(r, w):
    """(u_1, u_2, w_1, w_2) with u = t^2 from the reference's t^2 / (1 - t^2)."""
    return [ri / (1.0 + ri) for ri in r] + list(w)


def cheb_fit(k, vals, deg):
    """Chebyshev coefficients (exact arithmetic on the samples) of one function on [k, k+1)."""
    npt = deg + 1
    c = []
    for m in range(npt):
        s = mp.fsum(mp.mpf(vals[j]) * mp.cos(mp.pi * m * (j + mp.mpf(1) / 2) / npt) for j in range(npt))
        c.append(s * 2 / npt)
    c[0] /= 2
    return [float(v) for v in c]


def nodes(k, deg):
    npt = deg + 1
    return [float(mp.mpf(k) + (1 + mp.cos(mp.pi * (j + mp.mpf(1) / 2) / npt)) / 2) for j in range(npt)]


def main():
    lib = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref_harness.so"))
    mono = open(os.path.join(HERE, "rys_mono.inc")).read()
    t64 = open(os.path.join(HERE, "rys_table.inc")).read()
    t32 = open(os.path.join(HERE, "rys_table_f32.inc")).read()
    nc64 = int(re.search(r"#define UNO_RYS_NCOEF (\d+)", t64).group(1))
    nc32 = int(re.search(r"#define UNO_RYS_NCOEF_F (\d+)", t32).group(1))
    off64 = gm.read_ints(t64, "uno_rys_off")[NR]
    off32 = gm.read_ints(t32, "uno_rys_off_f")[NR]
    xasym = gm.read_ints(t64, "uno_rys_xasym")[NR]
    assert nc64 == DEG + 1 and nc32 == DEGF + 1
    body = re.search(r"uno_rys_mono\[\d+\] = \{(.*?)\};", mono, re.S).group(1)
    all64 = [float(v) for v in body.replace("\n", "").split(",") if v.strip()]
    body = re.search(r"uno_rys_mono_f\[\d+\] = \{(.*?)\};", mono, re.S).group(1)
    all32 = [float(v.rstrip("f")) for v in body.replace("\n", "").split(",") if v.strip()]
    tab64 = all64[off64: off64 + xasym * 2 * NR * nc64]
    tab32 = all32[off32: off32 + xasym * 2 * NR * nc32]
    tp64, tp32 = gm.cheb_power(nc64), gm.cheb_power(nc32)
    r64 = lambda v: float(v)                              # noqa: E731
    r32 = lambda v: float(np.float32(float(v)))           # noqa: E731
    pts = [Fraction(2 * i + 1, 32) - 1 for i in range(16)]
    worst64 = worst32 = 0.0
    for k in range(K0, K1):
        samp = [kernel_form(*ref_roots(lib, x)) for x in nodes(k, DEG)]
        checks = []
        for s in pts:
            x = k + (1 + float(s)) / 2
            checks.append((s, kernel_form(*ref_roots(lib, x))))
        for f in range(2 * NR):
            c = cheb_fit(k, [v[f] for v in samp], DEG)
            a64 = [r64(v) for v in gm.to_power(c, tp64)]
            a32 = [r32(v) for v in gm.to_power(c[:nc32], tp32)]
            b64 = (k * 2 * NR + f) * nc64
            b32 = (k * 2 * NR + f) * nc32
            tab64[b64: b64 + nc64] = a64
            tab32[b32: b32 + nc32] = a32
            for s, ref in checks:
                worst64 = max(worst64, abs(float(gm.horner(a64, s, r64)) - ref[f]) / abs(ref[f]))
                worst32 = max(worst32, abs(float(gm.horner(a32, s, r32)) - ref[f]) / abs(ref[f]))
    o = sys.stdout
    o.write("// Generated by gen_rys_ref2.py (do not edit): the reference's two-root Rys quadrature\n")
    o.write("// (Rys::root2, sampled from the reference compiled from its sources) fitted per unit\n")
    o.write("// interval on [%d, %d) as degree-%d (double) / degree-%d (float) series in\n" % (K0, K1, DEG, DEGF))
    o.write("// s = 2(x-k)-1, power basis; the exact tables elsewhere.  Same layout as the 2-root part\n")
    o.write("// of rys_mono.inc.  Worst relative difference to the reference at 16 points per fitted\n")
    o.write("// interval: double %.2e, float %.2e\n" % (worst64, worst32))
    o.write("UNO_RYS_QUAL double uno_rys_ref2_mono[%d] = {\n" % len(tab64))
    for i in range(0, len(tab64), 4):
        o.write(",".join("%.17e" % v for v in tab64[i:i + 4]) + ",\n")
    o.write("};\n")
    o.write("UNO_RYS_QUAL float uno_rys_ref2_mono_f[%d] = {\n" % len(tab32))
    for i in range(0, len(tab32), 6):
        o.write(",".join("%.9ef" % v for v in tab32[i:i + 6]) + ",\n")
    o.write("};\n")
    # golden samples: a dense grid across the fitted range and its neighbours, plus random points
    rng = np.random.default_rng(20015)
    xs = np.concatenate([np.linspace(14.5, 41.0, 2651), rng.uniform(0.0, 60.0, 1000)])
    R = np.zeros((len(xs), NR))
    W = np.zeros((len(xs), NR))
    for i, x in enumerate(xs):
        r, w = ref_roots(lib, float(x))
        R[i], W[i] = r, w
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "rys2_ref.npz"), x=xs, r=R, w=W)
    print("double %.2e float %.2e" % (worst64, worst32), file=sys.stderr)


if __name__ == "__main__":
    main()
