"""The small dense symmetric eigensolve of the Rayleigh-Ritz steps (the
reference calls LAPACK DSYEV / DSTEV there: src/iterate/libiterate/pist/pist.f90,
src/iterate/libiterate/lanczos/lan.f90) against numpy's eigvalsh / eigh.

Eigenvalues only: Householder tridiagonalisation + Sturm-count multisection
(matrix in LDS up to n = 112, in device memory above), one problem or several
in one launch; with eigenvectors: cyclic Jacobi."""
import ctypes as C

import numpy as np
import pytest

from scalit_amd import _lib as L

pytestmark = pytest.mark.gpu

_dp = C.POINTER(C.c_double)


def _eig(a, vectors=False):
    n = a.shape[0]
    af = np.asfortranarray(a, dtype=np.float64)
    w = np.zeros(n)
    z = np.zeros((n, n), order="F") if vectors else None
    L.check(L.lib().scalit_debug_sym_eig(n, af.ctypes.data_as(_dp), n, w.ctypes.data_as(_dp),
                                         z.ctypes.data_as(_dp) if vectors else None, n), "sym_eig")
    return w, z


def _sym(n, seed):
    g = np.random.default_rng(seed).standard_normal((n, n))
    return (g + g.T) / 2


@pytest.mark.parametrize("n", [1, 2, 3, 17, 64, 111, 112, 113, 150, 200])
def test_eigvals_random(n):
    a = _sym(n, 100 + n)
    w, _ = _eig(a)
    ref = np.linalg.eigvalsh(a)
    scale = np.abs(ref).max()
    assert np.max(np.abs(w - ref)) <= 1e-12 * scale * max(1, n / 16)
    assert np.all(np.diff(w) >= 0)


def test_eigvals_structured():
    # only the upper triangle is read (DSYEV 'U'); the lower holds garbage
    n = 40
    a = _sym(n, 7)
    junk = a.copy()
    junk[np.tril_indices(n, -1)] = 1e30
    w, _ = _eig(junk)
    assert np.allclose(w, np.linalg.eigvalsh(a), rtol=0, atol=1e-12 * np.abs(a).max() * 4)
    # already tridiagonal (the Lanczos case), diagonal, and degenerate spectra
    d = np.linspace(-1, 2, 30)
    e = np.linspace(0.1, 0.5, 29)
    t = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    assert np.allclose(_eig(t)[0], np.linalg.eigvalsh(t), rtol=0, atol=1e-13 * 4)
    dd = np.diag(np.array([3.0, -1.0, 3.0, 0.5, -1.0]))
    assert np.array_equal(_eig(dd)[0], np.sort(np.diag(dd)))
    q, _ = np.linalg.qr(np.random.default_rng(3).standard_normal((24, 24)))
    lam = np.repeat([0.25, 1.5, 2.0], 8)
    m = q @ np.diag(lam) @ q.T
    assert np.max(np.abs(_eig(m)[0] - lam)) < 1e-13 * 16


@pytest.mark.parametrize("n", [5, 48])
def test_eigvectors(n):
    a = _sym(n, 11 + n)
    w, z = _eig(a, vectors=True)
    ref = np.linalg.eigvalsh(a)
    assert np.max(np.abs(w - ref)) <= 1e-12 * np.abs(ref).max() * 4
    assert np.max(np.abs(a @ z - z * w)) <= 1e-11 * np.abs(ref).max()
    assert np.max(np.abs(z.T @ z - np.eye(n))) <= 1e-12


def test_eig_arg_errors():
    w = np.zeros(2)
    a = np.eye(2)
    assert L.lib().scalit_debug_sym_eig(0, a.ctypes.data_as(_dp), 2, w.ctypes.data_as(_dp), None, 2) == L.SCALIT_EINVAL
    assert L.lib().scalit_debug_sym_eig(2, a.ctypes.data_as(_dp), 1, w.ctypes.data_as(_dp), None, 2) == L.SCALIT_EINVAL


@pytest.mark.parametrize("ns", [[12, 16, 11, 24, 1, 2], [96, 112, 40, 7], [100, 150, 200, 3]])
def test_eigvals_batch(ns):
    """Several problems in one launch (as the PIST windows' Rayleigh-Ritz
    steps run): each against eigvalsh, and bit for bit against the one-problem
    launch, which runs the same arithmetic."""
    lda = max(ns)
    a = np.zeros((len(ns), lda, lda))
    mats = []
    for j, n in enumerate(ns):
        m = _sym(n, 300 + 7 * j + n)
        mats.append(m)
        a[j, :n, :n] = m.T   # column-major blocks
        a[j, n:, :] = 9.9e30   # outside the leading block: never read
        a[j, :, n:] = 9.9e30
    w = np.zeros((len(ns), lda))
    nsa = (C.c_int * len(ns))(*ns)
    L.check(L.lib().scalit_debug_sym_eigvals_batch(len(ns), nsa, np.ascontiguousarray(a).ctypes.data_as(_dp), lda,
                                                   w.ctypes.data_as(_dp)), "sym_eigvals_batch")
    for j, n in enumerate(ns):
        ref = np.linalg.eigvalsh(mats[j])
        assert np.max(np.abs(w[j, :n] - ref)) <= 1e-12 * np.abs(ref).max() * max(1, n / 16)
        assert np.all(np.diff(w[j, :n]) >= 0)
        assert np.array_equal(w[j, :n], _eig(mats[j])[0])


def test_eigvals_batch_arg_errors():
    w = np.zeros(4)
    a = np.eye(2).reshape(-1)
    ns = (C.c_int * 2)(2, 3)
    assert L.lib().scalit_debug_sym_eigvals_batch(2, ns, a.ctypes.data_as(_dp), 2, w.ctypes.data_as(_dp)) \
        == L.SCALIT_EINVAL
    assert L.lib().scalit_debug_sym_eigvals_batch(0, ns, a.ctypes.data_as(_dp), 2, w.ctypes.data_as(_dp)) \
        == L.SCALIT_EINVAL
