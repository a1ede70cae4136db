"""The C-ABI library (CPU-side checks; no device calls): it loads, exports every
function include/scalit.h declares, and rejects bad arguments before touching
a GPU."""
import ctypes as C
import os
import re
import subprocess

import pytest

from scalit_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "scalit.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = "\n".join(l for l in text.splitlines() if not l.lstrip().startswith(("typedef", "#")))
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("defined",)))


def test_header_parses():
    names = declared_functions()
    assert "scalit_hx" in names and "qmr_" in names and "pist_conv_" in names and len(names) > 30


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True, text=True, check=True)
    exported = set(line.split()[-1] for line in out.stdout.splitlines() if line.strip())
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    for n in declared_functions():
        getattr(lib, n)
    # and the ctypes signature table covers the whole header
    assert set(declared_functions()) <= set(L.SIGNATURES)


def test_create_rejects_bad_arguments():
    lib = L.lib()
    h = C.c_void_p()
    sn = (C.c_int * 3)(4, 5, 6)
    assert lib.scalit_create(C.byref(h), 0, sn, (C.c_int * 3)(0, 0, 1), 0) == L.SCALIT_EINVAL
    assert lib.scalit_create(C.byref(h), 3, sn, (C.c_int * 3)(1, 0, 1), 0) == L.SCALIT_EINVAL   # sDEP(1) must be F
    assert lib.scalit_create(C.byref(h), 3, (C.c_int * 3)(4, 0, 6), (C.c_int * 3)(0, 0, 1), 0) == L.SCALIT_EINVAL
    assert lib.scalit_create(C.byref(h), 9, sn, (C.c_int * 3)(0, 0, 1), 0) == L.SCALIT_EINVAL


def test_null_context_is_rejected():
    lib = L.lib()
    assert lib.scalit_hx(None, 0, 1, None, 1, None, 1, None, None) == L.SCALIT_EINVAL
    assert lib.scalit_px(None, 0, 1, 1, None, 1, None, 1, None, None) == L.SCALIT_EINVAL
    assert lib.scalit_block_diag(None, 10, 1e-5) == L.SCALIT_EINVAL
    assert lib.scalit_destroy(None) == L.SCALIT_EINVAL
    n = C.c_long()
    assert lib.scalit_get_len(None, C.byref(n)) == L.SCALIT_EINVAL
    # complex level matrices (sCX): no context, no data
    z = (C.c_double * 4)()
    assert lib.scalit_set_h0_cx(None, 1, z, z) == L.SCALIT_EINVAL
    assert lib.scalit_set_dep_cx(None, 1, z, z) == L.SCALIT_EINVAL
    assert lib.scalit_set_outh_cx(None, z, z) == L.SCALIT_EINVAL
    assert lib.scalit_set_scx(None, 1) == L.SCALIT_EINVAL


def test_fortran_layer_without_bound_context():
    # no context bound: the drop-in entry points report failure in the reference's codes
    lib = L.lib()
    lib.scalit_bind(None)
    import numpy as np

    n = 5
    w = np.ones(n)
    vj = np.zeros((n, 2), order="F")
    ci = lambda v: C.byref(C.c_int(v))  # noqa: E731
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    assert lib.mgs_orth_(ci(n), dp(w), ci(1), dp(vj)) == 0
    eig = np.zeros(2)
    res = C.c_double()
    r = lib.pist_conv_(C.byref(C.c_double(0.0)), C.byref(C.c_double(1e-8)), ci(0), ci(n), dp(w), ci(2), ci(1), ci(1),
                       ci(2), ci(4), C.cast(lib.hx0_, L.FOP), C.cast(lib.osb_qmr_, L.FLINSOLV), dp(eig), C.byref(res))
    assert r == 0


def test_no_reference_or_oracle_in_product():
    # the product library never links the oracle nor the reference checker
    out = subprocess.run(["ldd", L.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle" not in out and "scalit_ref" not in out
    src = os.path.join(ROOT, "scalit_amd")
    for dp_, _, files in os.walk(src):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
                t = open(os.path.join(dp_, f)).read()
                assert "import oracle" not in t and "from oracle" not in t, f


def test_small_complex_eigenvalues_against_lapack():
    """scalit_small_eig_cx, the host replacement of ZGEEV('N', 'N') used by the
    complex PIST, against numpy's LAPACK on random general and complex
    symmetric matrices (each eigenvalue matched to its nearest, 1e-11 of the
    spectral radius), and a defective-looking Jordan-like block."""
    import numpy as np

    lib = L.lib()
    rng = np.random.default_rng(11)
    mats = []
    for n in (1, 2, 5, 17, 60):
        a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        mats += [a, a + a.T]
    mats.append(np.diag(np.full(6, 0.5 + 0.1j)) + np.diag(np.full(5, 1e-3), 1))
    for a in mats:
        n = a.shape[0]
        buf = np.zeros((n, n, 2))
        buf[..., 0], buf[..., 1] = a.real.T, a.imag.T     # column-major (i + n j) interleaved
        w = np.zeros(2 * n)
        rc = lib.scalit_small_eig_cx(n, buf.ctypes.data, n, w.ctypes.data)
        assert rc == 0
        got = w[0::2] + 1j * w[1::2]
        ref = np.linalg.eigvals(a)
        scale = max(1.0, np.abs(ref).max())
        tol = 1e-11 if n < 6 or not np.allclose(a, np.triu(a)) else 1e-5
        assert all(np.abs(ref - g).min() <= tol * scale for g in got), (n, got, ref)
        assert all(np.abs(got - r).min() <= tol * scale for r in ref), (n, got, ref)
