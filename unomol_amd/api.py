"""Mirror of the reference's C++ class surface over the C ABI.

Names, argument meaning and conventions follow the reference (packed lower-triangle
symmetric storage, half density for RHF, accumulate-into-G, rsp destroying its input);
errors raise ``UnoError`` where the reference would call exit().
"""
import ctypes as C
import os

import numpy as np

from . import _native as N
from .deck import parse_deck, read_deck


class Basis:
    """Deck-backed basis (reference Basis.hpp:168-352)."""

    def __init__(self, infile="patin.dat", text=None):
        self.deck = parse_deck(text) if text is not None else read_deck(infile)
        self._struct = N.basis_struct(self.deck)

    def number_of_orbitals(self):
        return self.deck["norb"]

    def number_of_electrons(self):
        return self.deck["nelec"]

    def number_of_shells(self):
        return self.deck["nshell"]

    def number_of_centers(self):
        return self.deck["ncen"]

    def maximum_iterations(self):
        return self.deck["maxits"]

    def scf_eps(self):
        return self.deck["eps"]

    def scf_flags(self, i):
        return self.deck["scf_flag"][i]

    @property
    def struct(self):
        return self._struct


class TwoElectronInts:
    """Direct-SCF replacement of TwoElectronInts (TwoElectronInts.hpp:79-110)."""

    def __init__(self, basis, start_shell=0, label="MINTS.DAT"):
        self.basis = basis
        self.ctx = C.c_void_p()
        N.check(N.lib().uno_ctx_create(C.byref(basis.struct), start_shell, C.byref(self.ctx)))
        self.norb = N.lib().uno_ctx_norb(self.ctx)

    def __del__(self):
        if getattr(self, "ctx", None) and self.ctx.value:
            N.lib().uno_ctx_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def recalculate(self, basis):
        self.basis = basis
        N.check(N.lib().uno_ctx_rebuild(self.ctx, C.byref(basis.struct)))

    def set(self, key, value):
        N.check(N.lib().uno_ctx_set(self.ctx, key.encode(), C.c_double(value)))

    def get(self, key):
        v = C.c_double()
        N.check(N.lib().uno_ctx_get(self.ctx, key.encode(), C.byref(v)))
        return v.value

    def autotune(self, P, PB=None, mixed=False):
        """Times the row-merge configurations of every quartet space on density P (packed;
        PB for UHF) and keeps the fastest; returns the tuned build's kernel milliseconds."""
        P = np.ascontiguousarray(P, dtype=np.float64)
        pb = None if PB is None else np.ascontiguousarray(PB, dtype=np.float64)
        ms = C.c_double()
        N.check(N.lib().uno_autotune(self.ctx, N.ptr(P), None if pb is None else N.ptr(pb), 1 if mixed else 0,
                                     C.byref(ms)))
        return ms.value

    def formGmatrix(self, *args):
        """formGmatrix(P, G) (RHF) or formGmatrix(PA, PB, GA, GB) (UHF); accumulates."""
        arrs = [a for a in args]
        for a in arrs:
            if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous):
                raise TypeError("packed float64 contiguous arrays expected")
        if len(arrs) == 2:
            N.check(N.lib().uno_fock_rhf(self.ctx, N.ptr(arrs[0]), N.ptr(arrs[1])))
        elif len(arrs) == 4:
            N.check(N.lib().uno_fock_uhf(self.ctx, *[N.ptr(a) for a in arrs]))
        else:
            raise TypeError("formGmatrix takes (P, G) or (PA, PB, GA, GB)")

    def eri_emit(self):
        """All canonical ERIs in the reference's order: (values, int32[n,4] indices)."""
        cnt = C.c_long()
        N.check(N.lib().uno_eri_emit(self.ctx, C.c_long(0), None, None, C.byref(cnt)))
        n = cnt.value
        val = np.zeros(n)
        f = np.zeros((n, 4), dtype=np.int32)
        N.check(N.lib().uno_eri_emit(self.ctx, C.c_long(n), N.ptr(val), N.ptr(f), C.byref(cnt)))
        return val, f

    def oneint(self):
        """S, T, H (packed) computed on the device."""
        n = self.norb
        S, T, H = [np.zeros(n * (n + 1) // 2) for _ in range(3)]
        N.check(N.lib().uno_oneint_device(self.ctx, N.ptr(S), N.ptr(T), N.ptr(H)))
        return S, T, H

    def schwarz(self):
        ns = self.basis.deck["nshell"]
        q = np.zeros(ns * (ns + 1) // 2)
        N.check(N.lib().uno_schwarz(self.ctx, N.ptr(q)))
        return q


class SymmPack:
    """SymmPack namespace (SymmPack.hpp:11-28) on the device."""

    @staticmethod
    def rsp(n, a, z, evals, tmp=None):
        N.check(N.lib().uno_eigh_packed(n, N.ptr(a), N.ptr(z), N.ptr(evals)))

    @staticmethod
    def sp_trans(n, a, z, tmp=None):
        N.check(N.lib().uno_sp_trans(n, N.ptr(a), N.ptr(np.ascontiguousarray(z, dtype=np.float64))))

    @staticmethod
    def TraceSymmPackProduct(a, b, n):
        idx = np.arange(n)
        diag = idx * (idx + 1) // 2 + idx
        return 2.0 * float(np.dot(a, b)) - float(np.dot(a[diag], b[diag]))

    @staticmethod
    def SymmPackDiffNorm(a, b, n):
        idx = np.arange(n)
        diag = idx * (idx + 1) // 2 + idx
        d = a - b
        return float(np.sqrt(2.0 * np.dot(d, d) - np.dot(d[diag], d[diag])) / n)


def scf_options(**kw):
    o = N.ScfOpts()
    N.lib().uno_scf_defaults(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class _HartreeFock:
    uhf = False

    def __init__(self, basis, tints):
        self.basis = basis
        self.tints = tints
        n = tints.norb
        ns = 2 if self.uhf else 1
        self.evals = np.zeros(ns * n)
        self.P = np.zeros(ns * n * (n + 1) // 2)
        self.C = np.zeros(ns * n * n)
        self.result = None

    def findEnergy(self, **opts):
        d = self.basis.deck
        o = scf_options(maxits=d["maxits"], eps=d["eps"], cflag=d["scf_flag"][0])
        for k, v in opts.items():
            setattr(o, k, v)
        r = N.ScfResult()
        N.check(N.lib().uno_scf_run(self.tints.ctx, d["nelec"], C.byref(o), C.byref(r),
                                    N.ptr(self.evals), N.ptr(self.P), N.ptr(self.C)),
                allow=(N.UNO_ENOCONV,))
        self.result = r
        return r.energy


class RestrictedHartreeFock(_HartreeFock):
    """RHF driver (RHF.hpp:22-685) running the device-resident loop."""

    def __init__(self, basis, tints):
        if basis.deck["nelec"] % 2:
            raise N.UnoError(N.UNO_EINVAL, "Odd number of electrons in RHF")
        super().__init__(basis, tints)


class UnRestrictedHartreeFock(_HartreeFock):
    """UHF driver (UHF.hpp:23-790); alpha gets the extra electron."""
    uhf = True


def run_deck(directory=".", **opts):
    """Drop-in run of patin.dat in ``directory`` (Unomol.cc:8-26); writes the reference outputs."""
    o = scf_options(**opts)
    r = N.ScfResult()
    N.check(N.lib().uno_run_deck(os.fspath(directory).encode(), C.byref(o), C.byref(r)),
            allow=(N.UNO_ENOCONV,))
    return r


def generate_deck(path, kind="water_cluster", n=1, seed=0, basis="6-31G*", charge=0, eps=1e-10, maxits=300):
    N.check(N.lib().uno_deck_generate(kind.encode(), n, C.c_ulonglong(seed), basis.encode(), charge,
                                      C.c_double(eps), maxits, os.fspath(path).encode()))


def rys_eval(n, x, fp32=False):
    x = np.ascontiguousarray(x, dtype=np.float64)
    u = np.zeros((len(x), n))
    w = np.zeros((len(x), n))
    f = N.lib().uno_rys_eval_f32 if fp32 else N.lib().uno_rys_eval
    N.check(f(n, len(x), N.ptr(x), N.ptr(u), N.ptr(w)))
    return u, w


def rys_eval_ref2(x, fp32=False):
    """Two roots from the reference-compatible table the ERI kernels use by default."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    u = np.zeros((len(x), 2))
    w = np.zeros((len(x), 2))
    N.check(N.lib().uno_rys_eval_ref2(1 if fp32 else 0, len(x), N.ptr(x), N.ptr(u), N.ptr(w)))
    return u, w


def diis_coefficients(B, filled, newest):
    """Pulay DIIS coefficients of the SCF's accel = 2 from the K x K error-overlap ring B:
    (coefficients[K], number of vectors used)."""
    B = np.ascontiguousarray(B, dtype=np.float64)
    K = B.shape[0]
    c = np.zeros(K)
    used = C.c_int()
    N.check(N.lib().uno_diis_coefficients(N.ptr(B), K, int(filled), int(newest), N.ptr(c), C.byref(used)))
    return c, used.value


def oneint(deck):
    b = N.basis_struct(deck)
    n = deck["norb"]
    n2 = n * (n + 1) // 2
    S, T, H = np.zeros(n2), np.zeros(n2), np.zeros(n2)
    N.check(N.lib().uno_oneint(C.byref(b), N.ptr(S), N.ptr(T), N.ptr(H)))
    return S, T, H


def moment_ints(deck):
    b = N.basis_struct(deck)
    n = deck["norb"]
    m = np.zeros(9 * n * (n + 1) // 2)
    N.check(N.lib().uno_moment_ints(C.byref(b), N.ptr(m)))
    return m.reshape(-1, 9)


# ---- learned density guess -------------------------------------------------------------
GUESS_IN, GUESS_OUT = 128, 256


def shell_rotation(l, Q):
    """(ncart x ncart) map of a shell's Cartesian functions under a rotation Q."""
    nc = (l + 1) * (l + 2) // 2
    M = np.zeros((nc, nc))
    N.check(N.lib().uno_shell_rotation(l, N.ptr(np.ascontiguousarray(Q, dtype=np.float64)), N.ptr(M)))
    return M


def guess_features(deck):
    """Descriptors [count x 128] and atom pairs [count x 2] of the pairs within 10 bohr."""
    b = N.basis_struct(deck)
    cnt = C.c_long()
    N.check(N.lib().uno_guess_features(C.byref(b), C.c_long(0), None, None, C.byref(cnt)))
    X = np.zeros((cnt.value, GUESS_IN), dtype=np.float32)
    pairs = np.zeros((cnt.value, 2), dtype=np.int32)
    N.check(N.lib().uno_guess_features(C.byref(b), C.c_long(cnt.value), N.ptr(X), N.ptr(pairs), C.byref(cnt)))
    return X, pairs


def guess_samples(deck, P):
    """Training samples from a packed (half) density: descriptors, blocks, masks."""
    b = N.basis_struct(deck)
    P = np.ascontiguousarray(P, dtype=np.float64)
    cnt = C.c_long()
    N.check(N.lib().uno_guess_samples(C.byref(b), N.ptr(P), C.c_long(0), None, None, None, C.byref(cnt)))
    n = cnt.value
    X = np.zeros((n, GUESS_IN), dtype=np.float32)
    Y = np.zeros((n, GUESS_OUT), dtype=np.float32)
    M = np.zeros((n, GUESS_OUT), dtype=np.float32)
    N.check(N.lib().uno_guess_samples(C.byref(b), N.ptr(P), C.c_long(n), N.ptr(X), N.ptr(Y), N.ptr(M),
                                      C.byref(cnt)))
    return X, Y, M


def guess_assemble(deck, pairs, Y):
    """Packed density from predicted blocks of the given atom pairs."""
    b = N.basis_struct(deck)
    n = deck["norb"]
    P = np.zeros(n * (n + 1) // 2)
    pairs = np.ascontiguousarray(pairs, dtype=np.int32)
    Y = np.ascontiguousarray(Y, dtype=np.float32)
    N.check(N.lib().uno_guess_assemble(C.byref(b), C.c_long(len(pairs)), N.ptr(pairs), N.ptr(Y), N.ptr(P)))
    return P


class MLP:
    """Density-guess network on the GPU (BF16 matrix-core training and inference)."""

    def __init__(self, in_dim=GUESS_IN, width=128, hidden_layers=3, out_dim=GUESS_OUT, seed=0, path=None):
        self.h = C.c_void_p()
        if path is not None:
            N.check(N.lib().uno_mlp_load(os.fspath(path).encode(), C.byref(self.h)))
        else:
            N.check(N.lib().uno_mlp_create(in_dim, width, hidden_layers, out_dim, C.c_ulonglong(seed),
                                           C.byref(self.h)))
        nd = C.c_int(16)
        dims = (C.c_int * 16)()
        N.check(N.lib().uno_mlp_dims(self.h, C.byref(nd), dims))
        self.dims = [dims[i] for i in range(nd.value)]

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            N.lib().uno_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def nparams(self):
        n = C.c_long()
        N.check(N.lib().uno_mlp_nparams(self.h, C.byref(n)))
        return n.value

    def get_params(self):
        p = np.zeros(self.nparams(), dtype=np.float32)
        N.check(N.lib().uno_mlp_get_params(self.h, N.ptr(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        N.check(N.lib().uno_mlp_set_params(self.h, N.ptr(p)))

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((len(X), self.dims[-1]), dtype=np.float32)
        N.check(N.lib().uno_mlp_forward(self.h, C.c_long(len(X)), N.ptr(X), N.ptr(Y)))
        return Y

    def loss_grad(self, X, T, M):
        X, T, M = [np.ascontiguousarray(a, dtype=np.float32) for a in (X, T, M)]
        loss = C.c_float()
        g = np.zeros(self.nparams(), dtype=np.float32)
        N.check(N.lib().uno_mlp_loss_grad(self.h, C.c_long(len(X)), N.ptr(X), N.ptr(T), N.ptr(M), C.byref(loss),
                                          N.ptr(g)))
        return loss.value, g

    def train(self, X, T, M, steps=2000, batch=1024, lr=2e-3, seed=0):
        X, T, M = [np.ascontiguousarray(a, dtype=np.float32) for a in (X, T, M)]
        loss = C.c_float()
        N.check(N.lib().uno_mlp_train(self.h, C.c_long(len(X)), N.ptr(X), N.ptr(T), N.ptr(M), steps, batch,
                                      C.c_float(lr), C.c_ulonglong(seed), C.byref(loss)))
        return loss.value

    def save(self, path):
        N.check(N.lib().uno_mlp_save(self.h, os.fspath(path).encode()))


def set_guess_model(tints, mlp):
    """Network used by findEnergy(guess=2) on this context (kept alive by the caller)."""
    N.check(N.lib().uno_ctx_set_guess(tints.ctx, mlp.h if mlp is not None else None))
    tints._guess_model = mlp


class Batch:
    """Batched SCF of many molecules (decks: list of deck dicts or paths) on the device."""

    def __init__(self, decks):
        self.decks = [read_deck(d) if isinstance(d, (str, os.PathLike)) else d for d in decks]
        self._structs = [N.basis_struct(d) for d in self.decks]
        arr = (N.Basis * len(self._structs))(*self._structs)
        self._arr = arr
        nelec = (C.c_int * len(self.decks))(*[d["nelec"] for d in self.decks])
        self.h = C.c_void_p()
        N.check(N.lib().uno_batch_create(len(self.decks), arr, nelec, C.byref(self.h)))

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            N.lib().uno_batch_destroy(self.h)
            self.h = C.c_void_p()

    def run(self, **opts):
        o = scf_options(**opts)
        res = (N.ScfResult * len(self.decks))()
        e = np.zeros(len(self.decks))
        N.check(N.lib().uno_batch_run(self.h, C.byref(o), res, N.ptr(e)), allow=(N.UNO_ENOCONV,))
        return e, list(res)

    def set(self, key, value):
        """Fock-build option: "graph" (capture and replay the build) or "queue" (work-queue
        launches), 1 or 0."""
        N.check(N.lib().uno_batch_set(self.h, key.encode(), C.c_double(value)))

    def get(self, key):
        """An option, or a counter of the last run: space_builds, graph_builds, graph_replays,
        launches_per_build."""
        v = C.c_double()
        N.check(N.lib().uno_batch_get(self.h, key.encode(), C.byref(v)))
        return v.value

    def density(self, m):
        """Final density of molecule m: dense (N, N) RHF half density, or (2, N, N) alpha/beta."""
        d = self.decks[m]
        n = d["norb"]
        ns = 2 if d["nelec"] % 2 else 1
        P = np.zeros(ns * n * n)
        N.check(N.lib().uno_batch_density(self.h, m, N.ptr(P)))
        return P.reshape(ns, n, n) if ns == 2 else P.reshape(n, n)
