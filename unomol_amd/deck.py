"""Deck (patin.dat) reader: the layout the reference parses in Basis.hpp:181-255
(header, centres, shells); raw, un-normalised coefficients."""
import numpy as np


def parse_deck(text):
    tok = text.split()
    it = iter(tok)
    nxt = lambda: next(it)
    d = {}
    d["nshell"] = int(nxt()); d["norb"] = int(nxt()); d["ncen"] = int(nxt()); d["maxl"] = int(nxt())
    d["nelec"] = int(nxt()); d["maxits"] = int(nxt())
    d["eps"] = float(nxt())
    d["int_flag"] = [int(nxt()), int(nxt())]
    d["scf_flag"] = [int(nxt()), int(nxt()), int(nxt())]
    d["prt_flag"] = [int(nxt()), int(nxt()), int(nxt())]
    q, xyz = [], []
    for _ in range(d["ncen"]):
        q.append(float(nxt())); xyz.append([float(nxt()), float(nxt()), float(nxt())])
    l, npr, cen, poff, alpha, coef = [], [], [], [], [], []
    for _ in range(d["nshell"]):
        n = int(nxt()); ll = int(nxt()); c = int(nxt())
        npr.append(n); l.append(ll); cen.append(c); poff.append(len(alpha))
        for _ in range(n):
            alpha.append(float(nxt())); coef.append(float(nxt()))
    d["charge"] = np.array(q, dtype=np.float64)
    d["xyz"] = np.array(xyz, dtype=np.float64).reshape(-1, 3)
    d["sh_l"] = np.array(l, dtype=np.int32)
    d["sh_npr"] = np.array(npr, dtype=np.int32)
    d["sh_cen"] = np.array(cen, dtype=np.int32)
    d["sh_poff"] = np.array(poff, dtype=np.int32)
    d["alpha"] = np.array(alpha, dtype=np.float64)
    d["coef"] = np.array(coef, dtype=np.float64)
    return d


def read_deck(path):
    with open(path) as f:
        return parse_deck(f.read())
