"""Algorithmic FLOP counts of the Fock-build ERI kernels, per class.

The count replays the arithmetic of one primitive quartet of the lane kernel
(``eri_block`` in csrc/hip/eri.hip with ``rys_roots``/``axis_table`` from eri_core.hpp)
and of one shell quartet's normalisation and RHF J/K digestion (``k_eri_bra``) on
symbolic values, the way the compiler sees the fully unrolled code:

* constants fold (products with 1 or 0, sums with 0, constant-only arithmetic);
* identical expressions are built once (common subexpressions);
* only operations whose results reach an output (the accumulated integrals, the
  primitive-cut test, the Coulomb/exchange sums) are counted (dead code is free);
* per-shell-pair quantities that do not change inside the primitive loops (the powers of
  AB and CD times the binomial coefficients, the ket primitive's Q - C) are hoisted and
  not counted per primitive quartet.

Counting: add, sub, mul = 1; fma = 2; a transcendental (the reciprocal-square-root
estimate) = 1.  The Rys roots are counted on the Chebyshev-table path (x below the
asymptotic switch).  The primitive-quartet count and the shell-quartet count that
multiply these figures are the kernel's own counters (quartets that survive the
screens and the primitive cut).

Classes up to ``LANE_MAX_COMP`` components run in the lane kernel; larger classes run
one quartet per wave with the same arithmetic spread over lanes, so the same count is
used for them (the wave kernel's extra index/LDS work is not arithmetic).
"""
import functools
import json

LANE_MAX_COMP = 162
NCOEF = {"f64": 13, "f32": 7}
RYS_ASYM = [0, 40, 48, 56, 62, 68]


def ncart(l):
    return (l + 1) * (l + 2) // 2


def cart(l):
    return [(x, y, l - x - y) for x in range(l, -1, -1) for y in range(l - x, -1, -1)]


def cart_nrm(l, k):
    return 0.57735026918962576451 if 2 in cart(l)[k] else 1.0


class _Graph:
    """Hash-consed expression DAG with constant folding."""

    def __init__(self):
        self.nodes = {}      # key -> id
        self.info = []       # id -> (flops, args)

    def var(self, name):
        key = ("var", name)
        if key not in self.nodes:
            self.nodes[key] = len(self.info)
            self.info.append((0, ()))
        return ("n", self.nodes[key])

    def _op(self, op, flops, *args):
        key = (op,) + tuple(args)
        if key not in self.nodes:
            self.nodes[key] = len(self.info)
            self.info.append((flops, tuple(a[1] for a in args if isinstance(a, tuple))))
        return ("n", self.nodes[key])

    @staticmethod
    def const(v):
        return float(v)

    def neg(self, a):
        if isinstance(a, float):
            return -a
        return self._op("neg", 0, a)   # a source modifier, no instruction

    def add(self, a, b):
        if isinstance(a, float) and isinstance(b, float):
            return a + b
        if a == 0.0:
            return b
        if b == 0.0:
            return a
        if isinstance(a, tuple) and isinstance(b, tuple) and a > b:
            a, b = b, a
        return self._op("add", 1, a, b)

    def sub(self, a, b):
        if isinstance(a, float) and isinstance(b, float):
            return a - b
        if b == 0.0:
            return a
        if a == 0.0:
            return self.neg(b)
        return self._op("sub", 1, a, b)

    def mul(self, a, b):
        if isinstance(a, float) and isinstance(b, float):
            return a * b
        if a == 0.0 or b == 0.0:
            return 0.0
        if a == 1.0:
            return b
        if b == 1.0:
            return a
        if a == -1.0:
            return self.neg(b)
        if b == -1.0:
            return self.neg(a)
        if isinstance(a, tuple) and isinstance(b, tuple) and a > b:
            a, b = b, a
        return self._op("mul", 1, a, b)

    def fma(self, a, b, c):
        if a == 0.0 or b == 0.0:
            return c
        if isinstance(a, float) and isinstance(b, float):
            return self.add(a * b, c)
        if c == 0.0:
            return self.mul(a, b)
        if a == 1.0:
            return self.add(b, c)
        if b == 1.0:
            return self.add(a, c)
        return self._op("fma", 2, a, b, c)

    def trans(self, name, a):
        return self._op(name, 1, a)

    def free(self, name, a):
        """A conversion or other non-arithmetic step: no FLOPs."""
        return self._op(name, 0, a)

    def count(self, outputs):
        seen, stack, total = set(), [o[1] for o in outputs if isinstance(o, tuple)], 0
        while stack:
            i = stack.pop()
            if i in seen:
                continue
            seen.add(i)
            fl, args = self.info[i]
            total += fl
            stack.extend(args)
        return total


def _axis_table(g, la, lb, lc, ld, C, Cp, B00, B1, B1p, abp, cdp):
    E, F = la + lb, lc + ld
    G = [[None] * (F + 1) for _ in range(E + 1)]
    G[0][0] = 1.0
    if E > 0:
        G[1][0] = C
    for e in range(1, E):
        G[e + 1][0] = g.add(g.mul(g.mul(float(e), B1), G[e - 1][0]), g.mul(C, G[e][0]))
    for f in range(F):
        for e in range(E + 1):
            v = g.mul(Cp, G[e][f])
            if f > 0:
                v = g.fma(g.mul(float(f), B1p), G[e][f - 1], v)
            if e > 0:
                v = g.fma(g.mul(float(e), B00), G[e - 1][f], v)
            G[e][f + 1] = v
    H = {}
    for e in range(E + 1):
        for c in range(lc + 1):
            for d in range(ld + 1):
                v = 0.0
                for j in range(d + 1):
                    v = g.fma(cdp[(d, j)], G[e][c + d - j], v)
                H[e, c, d] = v
    I = {}
    for a in range(la + 1):
        for b in range(lb + 1):
            for c in range(lc + 1):
                for d in range(ld + 1):
                    v = 0.0
                    for i in range(b + 1):
                        v = g.fma(abp[(b, i)], H[a + b - i, c, d], v)
                    I[a, b, c, d] = v
    return I


def _powers(g, name, l):
    """binom(n, j) * r^j as hoisted per-pair values (1 for j = 0)."""
    out = {}
    for n_ in range(l + 1):
        for j in range(n_ + 1):
            out[(n_, j)] = 1.0 if j == 0 else g.var("%s%d_%d" % (name, n_, j))
    return out


@functools.lru_cache(maxsize=None)
def prim_flops(la, lb, lc, ld, prec="f64"):
    """FLOPs of one primitive quartet in the compute precision (and the FP64 coordinate
    differences an FP32 kernel takes in double before rounding): (flops_T, flops_f64)."""
    g = _Graph()
    L = la + lb + lc + ld
    nr = L // 2 + 1
    v = g.var
    p, ipp, sp, Kp = v("p"), v("ip"), v("sp"), v("Kp")
    q, iqq, sq, Kq = v("q"), v("iq"), v("sq"), v("Kq")
    QC = [v("QC%d" % x) for x in range(3)]   # hoisted per ket primitive
    # rt = rsqrt_nr(p + q): estimate + Newton steps (f64: two; f32: the estimate only)
    s = g.add(p, q)
    y = g.trans("rsq", s)
    if prec == "f64":
        h = g.mul(0.5, s)
        for _ in range(2):
            r = g.fma(g.neg(g.mul(h, y)), y, 0.5)
            y = g.fma(y, r, y)
    rt = y
    it = g.mul(rt, rt)
    k25 = 34.986836655249725
    bare = g.mul(g.mul(g.mul(g.mul(g.mul(k25, sp), sq), ipp), iqq), rt)
    pref = g.mul(g.mul(g.mul(g.mul(g.mul(k25, Kp), Kq), ipp), iqq), rt)
    # PA, PQ: double subtractions, rounded to T (counted apart for f32)
    PA = [v("PA%d" % x) for x in range(3)]
    PQ = [v("PQ%d" % x) for x in range(3)]
    n64 = 6
    dot = g.fma(PQ[2], PQ[2], g.fma(PQ[1], PQ[1], g.mul(PQ[0], PQ[0])))
    x = g.mul(g.mul(g.mul(p, q), it), dot)
    # table path: k = (int)x, s = 2 (x - k) - 1, Horner over NC power-basis coefficients
    # per function
    xk = g.free("cvt", x)   # integer conversion and back
    ss = g.fma(2.0, g.sub(x, xk), -1.0)
    nc = NCOEF[prec]
    u, w = [], []
    for fidx in range(2 * nr):
        cf = [v("cf%d_%d" % (fidx, m)) for m in range(nc)]
        val = cf[nc - 1]
        for m in range(nc - 2, -1, -1):
            val = g.fma(val, ss, cf[m])
        (u if fidx < nr else w).append(val)
    abp = _powers(g, "ab", lb)
    cdp = _powers(g, "cd", ld)
    ca, cb, cc, cd_ = cart(la), cart(lb), cart(lc), cart(ld)
    e = {}
    for ia in range(len(ca)):
        for ib in range(len(cb)):
            for ic in range(len(cc)):
                for id_ in range(len(cd_)):
                    e[ia, ib, ic, id_] = v("e_%d_%d_%d_%d" % (ia, ib, ic, id_))
    for r in range(nr):
        fff = g.mul(u[r], it)
        B00 = g.mul(0.5, fff)
        B1 = g.mul(g.sub(0.5, g.mul(B00, q)), ipp)
        B1p = g.mul(g.sub(0.5, g.mul(B00, p)), iqq)
        I = []
        for ax in range(3):
            C = g.sub(PA[ax], g.mul(g.mul(q, PQ[ax]), fff))
            Cp = g.add(QC[ax], g.mul(g.mul(p, PQ[ax]), fff))
            I.append(_axis_table(g, la, lb, lc, ld, C, Cp, B00, B1, B1p, abp, cdp))
        wr = g.mul(w[r], pref)
        for (ia, ib, ic, id_), acc in list(e.items()):
            t = [(ca[ia][k], cb[ib][k], cc[ic][k], cd_[id_][k]) for k in range(3)]
            e[ia, ib, ic, id_] = g.fma(g.mul(wr, I[0][t[0]]), g.mul(I[1][t[1]], I[2][t[2]]), acc)
    flops = g.count(list(e.values()) + [bare])
    return (flops, n64) if prec == "f32" else (flops + n64, 0)


@functools.lru_cache(maxsize=None)
def shell_flops(la, lb, lc, ld, prec="f64"):
    """FLOPs of one shell quartet's normalisation and RHF J/K digestion in the lane
    kernel: (flops_T, flops_f64) (the wave-wide Coulomb reduction runs in double)."""
    g = _Graph()
    NA, NB, NC, ND = ncart(la), ncart(lb), ncart(lc), ncart(ld)
    f = g.mul(g.mul(g.var("h1"), g.var("h2")), g.var("h3"))
    e = {}
    for i in range(NA):
        for j in range(NB):
            for k in range(NC):
                for l in range(ND):
                    nrm = cart_nrm(la, i) * cart_nrm(lb, j) * cart_nrm(lc, k) * cart_nrm(ld, l)
                    e[i, j, k, l] = g.mul(g.var("e%d_%d_%d_%d" % (i, j, k, l)), g.mul(f, nrm))
    D = lambda a, b: g.var("D%s_%s" % (a, b))
    outs = []
    n64 = 0
    for i in range(NA):           # Coulomb, bra outputs (then a 6-step double wave sum)
        for j in range(NB):
            s = 0.0
            for k in range(NC):
                for l in range(ND):
                    s = g.fma(e[i, j, k, l], D("c%d" % k, "d%d" % l), s)
            outs.append(g.mul(4.0, s))
            n64 += 6
    for k in range(NC):           # Coulomb, ket outputs
        for l in range(ND):
            s = 0.0
            for i in range(NA):
                for j in range(NB):
                    s = g.fma(e[i, j, k, l], D("a%d" % i, "b%d" % j), s)
            outs.append(g.mul(4.0, s))
    groups = [((0, 2), (1, 3)), ((0, 3), (1, 2)), ((1, 2), (0, 3)), ((1, 3), (0, 2))]
    names = "abcd"
    dims = (NA, NB, NC, ND)
    for (o1, o2), (s1, s2) in groups:   # exchange: W_o1o2 -= sum e D_s1s2
        for a in range(dims[o1]):
            for b in range(dims[o2]):
                s = 0.0
                for c in range(dims[s1]):
                    for d in range(dims[s2]):
                        idx = [0] * 4
                        idx[o1], idx[o2], idx[s1], idx[s2] = a, b, c, d
                        s = g.fma(e[tuple(idx)], D("%s%d" % (names[s1], c), "%s%d" % (names[s2], d)), s)
                outs.append(s)
    return g.count(outs), n64


def class_flops(la, lb, lc, ld, prec, nprim, nshell):
    """Algorithmic FLOPs of a launch (or a sum of launches) that evaluated nprim primitive
    quartets and nshell shell quartets of the class."""
    pt, p64 = prim_flops(la, lb, lc, ld, prec)
    st, s64 = shell_flops(la, lb, lc, ld, prec)
    return nprim * (pt + p64) + nshell * (st + s64)


def table():
    """Per-class figures (f64 and f32) as a dict keyed by '(la lb|lc ld)'."""
    pairs = [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2)]
    out = {}
    for X in range(6):
        for Y in range(X + 1):
            la, lb = pairs[X]
            lc, ld = pairs[Y]
            key = "(%d%d|%d%d)" % (la, lb, lc, ld)
            out[key] = {prec: {"per_prim": prim_flops(la, lb, lc, ld, prec),
                               "per_shell": shell_flops(la, lb, lc, ld, prec)} for prec in ("f64", "f32")}
    return out


if __name__ == "__main__":
    print(json.dumps(table(), indent=1))
