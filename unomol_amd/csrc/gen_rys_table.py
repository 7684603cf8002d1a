#!/usr/bin/env python3
"""Generate the Rys quadrature root/weight tables used by the gfx950 ERI kernels.

Method (this project's own; no fitted constants are taken from anywhere):
  * For n roots and argument x, the Rys quadrature on u = t^2 in [0,1] with weight
    exp(-x u) u^(-1/2)/2 has moments mu_k = F_k(x) (Boys function).  F_k is evaluated
    in 40-digit arithmetic with mpmath (F_k(x) = gamma(k+1/2, x) / (2 x^(k+1/2))).
  * The recurrence coefficients come from Gautschi's Chebyshev algorithm on those
    moments, and the nodes/weights from the eigen-decomposition of the Jacobi matrix
    (Golub-Welsch), all in 40-digit arithmetic.
  * Each root u_i(x) = t_i^2 and weight w_i(x) is tabulated on unit intervals
    [k, k+1) as a degree-DEG Chebyshev series in s = 2(x-k)-1 (interpolation at the
    Chebyshev points of the first kind); the device evaluates their exact power-basis
    re-expansion (gen_rys_monomial.py) with Horner.
  * For x >= XASYM[n] the large-x limit is used: the integral over [0,1] becomes the
    half-line Gauss-Hermite rule with 2n points, u_i = h_i / x, w_i = omega_i / sqrt(x)
    (h_i = squared positive roots of H_2n, omega_i their Gauss-Hermite weights).
    XASYM[n] is the smallest integer where that limit is within 2 ulp of the exact
    values at the probed points (checked below and printed).

Every interval is validated against the exact values at 6 interior points after
rounding the coefficients to double and evaluating with double Clenshaw.

usage:  python gen_rys_table.py > rys_table.inc
"""
import sys
import multiprocessing as mproc

import mpmath as mp

mp.mp.dps = 40
NMAX = 5
DEG = 12                       # Chebyshev degree per unit interval
XASYM = {1: 40, 2: 48, 3: 56, 4: 62, 5: 68}


def boys(k, x):
    x = mp.mpf(x)
    if x == 0:
        return mp.mpf(1) / (2 * k + 1)
    half = mp.mpf(1) / 2
    return mp.gammainc(k + half, 0, x) / (2 * x ** (k + half))


def rys_exact(n, x):
    """(roots u_i ascending, weights w_i) in mp precision."""
    mu = [boys(k, x) for k in range(2 * n)]
    alpha = [mp.mpf(0)] * n
    beta = [mp.mpf(0)] * n
    sig_prev = [mp.mpf(0)] * (2 * n)
    sig = mu[:]
    alpha[0] = mu[1] / mu[0]
    beta[0] = mu[0]
    for k in range(1, n):
        new = [mp.mpf(0)] * (2 * n)
        for l in range(k, 2 * n - k):
            new[l] = sig[l + 1] - alpha[k - 1] * sig[l] - beta[k - 1] * sig_prev[l]
        alpha[k] = new[k + 1] / new[k] - sig[k] / sig[k - 1]
        beta[k] = new[k] / sig[k - 1]
        sig_prev, sig = sig, new
    jac = mp.zeros(n, n)
    for i in range(n):
        jac[i, i] = alpha[i]
        if i > 0:
            jac[i, i - 1] = jac[i - 1, i] = mp.sqrt(beta[i])
    ev, q = mp.eigsy(jac)
    pairs = sorted((ev[i], beta[0] * q[0, i] ** 2) for i in range(n))
    return [p[0] for p in pairs], [p[1] for p in pairs]


def hermite_half(n):
    """Positive nodes squared and weights of the 2n-point Gauss-Hermite rule."""
    # Golub-Welsch for Hermite (weight exp(-s^2)): alpha=0, beta_k = k/2
    m = 2 * n
    jac = mp.zeros(m, m)
    for k in range(1, m):
        jac[k, k - 1] = jac[k - 1, k] = mp.sqrt(mp.mpf(k) / 2)
    ev, q = mp.eigsy(jac)
    out = []
    for i in range(m):
        if ev[i] > 0:
            out.append((ev[i] ** 2, mp.sqrt(mp.pi) * q[0, i] ** 2))
    out.sort()
    return [o[0] for o in out], [o[1] for o in out]


def clenshaw(c, s):
    b1 = 0.0
    b2 = 0.0
    for ck in reversed(c[1:]):
        b1, b2 = ck + 2.0 * s * b1 - b2, b1
    return c[0] + s * b1 - b2


def fit_interval(args):
    n, k = args
    npt = DEG + 1
    nodes = [mp.mpf(k) + (1 + mp.cos(mp.pi * (j + mp.mpf(1) / 2) / npt)) / 2 for j in range(npt)]
    vals = [rys_exact(n, x) for x in nodes]
    coefs = []
    for fi in range(2 * n):
        f = [(v[0][fi] if fi < n else v[1][fi - n]) for v in vals]
        c = []
        for m in range(npt):
            s = sum(f[j] * mp.cos(mp.pi * m * (j + mp.mpf(1) / 2) / npt) for j in range(npt))
            c.append(s * 2 / npt)
        c[0] /= 2
        coefs.append([float(ci) for ci in c])
    # validate in double
    err = 0.0
    for s in (-0.97, -0.61, -0.2, 0.23, 0.66, 0.98):
        x = mp.mpf(k) + (1 + mp.mpf(s)) / 2
        u, w = rys_exact(n, x)
        for fi in range(2 * n):
            ex = float(u[fi] if fi < n else w[fi - n])
            ap = clenshaw(coefs[fi], s)
            err = max(err, abs(ap - ex) / abs(ex))
    return n, k, coefs, err


def asym_check(n):
    h, om = hermite_half(n)
    err = 0.0
    for x in (XASYM[n], XASYM[n] + 0.5, XASYM[n] + 3, XASYM[n] + 20):
        u, w = rys_exact(n, x)
        for i in range(n):
            eu = abs(float(h[i] / x) - float(u[i])) / float(u[i])
            ew = abs(float(om[i] / mp.sqrt(x)) - float(w[i])) / float(w[i])
            err = max(err, eu, ew)
    return h, om, err


def main():
    jobs = [(n, k) for n in range(1, NMAX + 1) for k in range(XASYM[n])]
    with mproc.Pool(8) as pool:
        res = pool.map(fit_interval, jobs, chunksize=4)
    res.sort(key=lambda r: (r[0], r[1]))
    worst = max(r[3] for r in res)
    out = sys.stdout
    out.write("// Rys quadrature tables for n = 1..%d roots (generated by gen_rys_table.py;\n" % NMAX)
    out.write("// do not edit).  u = t^2 roots and weights on unit intervals of x as Chebyshev\n")
    out.write("// series of degree %d in s = 2(x-k)-1; large-x Gauss-Hermite limit beyond.\n" % DEG)
    out.write("// worst relative error of the double evaluation at the validation points: %.3e\n" % worst)
    out.write("#define UNO_RYS_NMAX %d\n#define UNO_RYS_NCOEF %d\n" % (NMAX, DEG + 1))
    out.write("UNO_RYS_QUAL int uno_rys_xasym[%d] = {0" % (NMAX + 1))
    for n in range(1, NMAX + 1):
        out.write(", %d" % XASYM[n])
    out.write("};\n")
    offs = [0]
    acc = 0
    for n in range(1, NMAX + 1):
        offs.append(acc)
        acc += XASYM[n] * 2 * n * (DEG + 1)
    out.write("UNO_RYS_QUAL int uno_rys_off[%d] = {%s};\n" % (NMAX + 1, ", ".join(str(o) for o in offs)))
    out.write("UNO_RYS_QUAL double uno_rys_coef[%d] = {\n" % acc)
    for n, k, coefs, err in res:
        out.write("// n=%d x in [%d,%d)  maxrel %.2e\n" % (n, k, k + 1, err))
        for c in coefs:
            out.write(" " + ",".join("%.17e" % v for v in c) + ",\n")
    out.write("};\n")
    hu, hw = [], []
    aerr = 0.0
    for n in range(1, NMAX + 1):
        h, om, e = asym_check(n)
        aerr = max(aerr, e)
        hu += h
        hw += om
    out.write("// large-x limit: worst relative error at the probed points %.3e\n" % aerr)
    out.write("UNO_RYS_QUAL double uno_rys_asym_u[%d] = {%s};\n" % (len(hu), ",".join("%.17e" % float(v) for v in hu)))
    out.write("UNO_RYS_QUAL double uno_rys_asym_w[%d] = {%s};\n" % (len(hw), ",".join("%.17e" % float(v) for v in hw)))
    sys.stderr.write("worst interval error %.3e  asymptotic %.3e  ncoef %d\n" % (worst, aerr, acc))


if __name__ == "__main__":
    main()
