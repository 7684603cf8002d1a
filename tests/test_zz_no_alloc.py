"""GPU test of uno_eigh_packed_noalloc, the eigensolver entry point that allocates no device
memory (its work space is a fixed array in the library), against uno_eigh_packed through
the C ABI: two calls in a row, the second on a smaller matrix, both equal to the
allocating entry point's results; and the size limit reported as an error."""
import ctypes as C

import numpy as np
import pytest

from unomol_amd import _native as N

pytestmark = pytest.mark.gpu


def _packed(A):
    return np.ascontiguousarray(A[np.tril_indices(A.shape[0])])


def _solve(fn, A):
    n = A.shape[0]
    a = _packed(A)
    z = np.zeros(n * n)
    e = np.zeros(n)
    rc = fn(C.c_int(n), N.ptr(a), N.ptr(z), N.ptr(e))
    return rc, e, z.reshape(n, n), a


def _matrix(n, seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = np.sort(rng.uniform(-20.0, 5.0, n))
    A = (Q * ev) @ Q.T
    return 0.5 * (A + A.T)


@pytest.mark.parametrize("sizes", [(200, 77), (256, 3)])
def test_eigh_noalloc_matches_allocating_entry_point(sizes):
    L = N.lib()
    for k, n in enumerate(sizes):
        print("[gpu-test] noalloc eigensolver n=%d (call %d)" % (n, k + 1), flush=True)
        A = _matrix(n, 100 * n + k)
        rc0, e0, Z0, _ = _solve(L.uno_eigh_packed, A)
        assert rc0 == N.UNO_OK
        rc1, e1, Z1, a1 = _solve(L.uno_eigh_packed_noalloc, A)
        assert rc1 == N.UNO_OK, L.uno_last_error()
        # same kernels on the same data: the same results
        np.testing.assert_allclose(e1, e0, rtol=0, atol=1e-13 * max(1.0, np.abs(e0).max()))
        np.testing.assert_allclose(Z1, Z0, rtol=0, atol=1e-12)
        # and a valid decomposition: residual, orthogonality, ascending, input destroyed
        scale = max(1.0, np.abs(e1).max())
        assert np.abs(A @ Z1 - Z1 * e1).max() < 1e-12 * scale * max(1, n / 10)
        assert np.abs(Z1.T @ Z1 - np.eye(n)).max() < 1e-12 * max(1, n / 10)
        assert np.all(np.diff(e1) >= 0)
        assert not a1.any()


def test_eigh_noalloc_rejects_sizes_beyond_its_work_space():
    L = N.lib()
    n = 257
    a = np.zeros(n * (n + 1) // 2)
    z = np.zeros(n * n)
    e = np.zeros(n)
    rc = L.uno_eigh_packed_noalloc(C.c_int(n), N.ptr(a), N.ptr(z), N.ptr(e))
    assert rc == N.UNO_EINVAL
    assert b"exceeds" in L.uno_last_error()
