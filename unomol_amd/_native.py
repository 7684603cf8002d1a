"""ctypes binding of libunomol_hip.so (C ABI declared in include/unomol.h).

The library is built in-tree by ``__graft_entry__.build()`` (``make -C unomol_amd/csrc``).
There is no CPU fallback: if the library is missing or cannot be loaded every entry
point raises.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libunomol_hip.so")

UNO_OK, UNO_EINVAL, UNO_EDEVICE, UNO_ENOMEM, UNO_ENOCONV, UNO_ESINGULAR, UNO_EIO, UNO_EEIGEN = range(8)


class UnoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("unomol_amd error %d: %s" % (code, msg))
        self.code = code


class Basis(C.Structure):
    _fields_ = [("nshell", C.c_int), ("ncen", C.c_int),
                ("sh_l", C.c_void_p), ("sh_npr", C.c_void_p), ("sh_cen", C.c_void_p),
                ("sh_poff", C.c_void_p), ("alpha", C.c_void_p), ("coef", C.c_void_p),
                ("xyz", C.c_void_p), ("charge", C.c_void_p), ("normalized", C.c_int)]


class ScfOpts(C.Structure):
    _fields_ = [("maxits", C.c_int), ("eps", C.c_double), ("cflag", C.c_int), ("accel", C.c_int),
                ("guess", C.c_int), ("precision", C.c_int), ("fp64_final", C.c_int),
                ("incremental", C.c_int), ("verbose", C.c_int), ("field", C.c_double * 3),
                ("single_start", C.c_int), ("guess_check", C.c_int)]


class ScfResult(C.Structure):
    _fields_ = [("energy", C.c_double), ("init_energy", C.c_double), ("ediff", C.c_double),
                ("pdiff", C.c_double), ("nucrep", C.c_double), ("kinetic", C.c_double),
                ("electronic", C.c_double), ("iterations", C.c_int), ("converged", C.c_int),
                ("uhf", C.c_int), ("nfock", C.c_int), ("wall_s", C.c_double), ("fock_s", C.c_double),
                ("diag_s", C.c_double), ("quartets", C.c_double), ("quartets32", C.c_double),
                ("nfull", C.c_int), ("guess_s", C.c_double), ("wall99_s", C.c_double),
                ("setup_s", C.c_double), ("diis_s", C.c_double), ("guess_fallback", C.c_int),
                ("guess_energy", C.c_double), ("guess_energy_coreh", C.c_double), ("purified", C.c_int), ("purify_steps", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None

# every symbol include/unomol.h declares
EXPORTS = ["uno_last_error", "uno_ctx_create", "uno_ctx_rebuild", "uno_ctx_destroy", "uno_ctx_set",
           "uno_ctx_get", "uno_ctx_norb", "uno_fock_rhf", "uno_fock_uhf", "uno_eigh_packed", "uno_eigh_packed_noalloc",
           "uno_sp_trans", "uno_scf_defaults", "uno_scf_run", "uno_eri_emit", "uno_rys_eval", "uno_rys_eval_f32", "uno_rys_eval_ref2",
           "uno_schwarz", "uno_profile", "uno_copy_probe", "uno_oneint", "uno_oneint_device", "uno_moment_ints", "uno_deck_generate", "uno_run_deck",
           "uno_mlp_create", "uno_mlp_destroy", "uno_mlp_nparams", "uno_mlp_dims", "uno_mlp_get_params",
           "uno_mlp_set_params", "uno_mlp_forward", "uno_mlp_loss_grad", "uno_mlp_train", "uno_mlp_save",
           "uno_mlp_load", "uno_guess_features", "uno_guess_samples", "uno_guess_assemble", "uno_ctx_set_guess",
           "uno_shell_rotation", "uno_autotune", "uno_batch_create", "uno_batch_destroy", "uno_batch_nmol", "uno_batch_run", "uno_batch_density",
           "uno_batch_set", "uno_batch_get", "uno_diis_coefficients"]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libunomol_hip.so is not built; run __graft_entry__.build() "
                               "(make -C unomol_amd/csrc)")
        _lib = C.CDLL(LIB_PATH)
        _lib.uno_last_error.restype = C.c_char_p
    return _lib


def check(rc, allow=()):
    if rc != UNO_OK and rc not in allow:
        raise UnoError(rc, lib().uno_last_error().decode())
    return rc


def ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def basis_struct(deck):
    """uno_basis view of a deck dict (keys as tests/deckio.py / unomol_amd.deck)."""
    keep = [np.ascontiguousarray(deck[k], dtype=t) for k, t in
            (("sh_l", np.int32), ("sh_npr", np.int32), ("sh_cen", np.int32), ("sh_poff", np.int32),
             ("alpha", np.float64), ("coef", np.float64), ("xyz", np.float64), ("charge", np.float64))]
    b = Basis(len(keep[0]), len(keep[7]), *[ptr(a) for a in keep])
    b._keep = keep
    return b
