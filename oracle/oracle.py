"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this module, and only as the checker.  It wraps two
libraries built by ``oracle/Makefile``:

* ``liboracle.so`` -- the plain-C restatement of the reference path
  (scalit_oracle.c, every routine cites the Fortran it follows);
* ``_ref/libscalit_ref.so`` -- the reference's own LAPACK-free Fortran routines
  compiled where they lie (built only where the reference sources exist; tests
  that need it skip elsewhere).
"""
from __future__ import annotations

import ctypes as C
import types
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libscalit_ref.so")

FMAX = 8
TA0, TA1 = 0, 1
TOSB, TOSBD1, TOSBD2, TOSBW = 0, 1, 2, 3

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OrcOSB(C.Structure):
    _fields_ = [
        ("sF", C.c_int),
        ("sN", C.c_int * FMAX),
        ("sDEP", C.c_int * FMAX),
        ("myDim", C.c_int * FMAX),
        ("myBlk", C.c_int * FMAX),
        ("myLen", C.c_long),
        ("H0", _dp * FMAX),
        ("DEP", _dp * FMAX),
        ("RES", _dp),
        ("VOSB", _dp * FMAX),
        ("EIG0", _dp),
        ("E0", C.c_double),
        ("DE", C.c_double),
        ("beta", C.c_double),
        ("osb_kind", C.c_int),
        ("hc", C.c_int),
        ("pc", C.c_int),
        ("ndvr", C.c_int),
        ("OUTH", _dp),
    ]


class OrcSolverCtx(C.Structure):
    _fields_ = [
        ("o", C.POINTER(OrcOSB)),
        ("qmr_type", C.c_int),
        ("qmr_max", C.c_int),
        ("qmr_tol", C.c_double),
        ("hx_calls", C.c_long),
        ("px_calls", C.c_long),
        ("solves", C.c_long),
        ("qmr_iters", C.c_long),
        ("threaded", C.c_int),
        ("tmp", _dp),
    ]


OP_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_long, _dp, _dp)
LS_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_long, _dp, _dp, _dp)

_lib = None
_ref = None


def lib():
    """The oracle library; raises if it was not built (``make -C oracle``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        _lib = C.CDLL(LIB_PATH)
        _lib.orc_dot.restype = C.c_double
        _lib.orc_sum_off_diag.restype = C.c_double
        _lib.orc_mjacobi.restype = C.c_double
        _lib.orc_get_diff_max.restype = C.c_double
        _lib.orc_get_sep_max.restype = C.c_double
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    """The reference checker library (lazy binding: see oracle/Makefile)."""
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"reference checker missing: {REF_PATH} (run `make -C oracle ref`)")
        _ref = C.CDLL(REF_PATH, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
        for f in ("getdiffmax_", "getsepmax_", "mjacobi_", "sumoffdiag_"):
            getattr(_ref, f).restype = C.c_double
        for f in ("qmr_", "pisth0_", "lanh0_", "mgs_orth_", "gs_orth_", "lan_mgs_orth_", "refdrv_linsolv",
                  "refdrv_diag"):
            getattr(_ref, f).restype = C.c_int
    return _ref


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and (a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"])
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32
    return a.ctypes.data_as(_ip)


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def fmat(a) -> np.ndarray:
    """Column-major (Fortran order) float64 copy."""
    return np.asfortranarray(a, dtype=np.float64)


# --------------------------------------------------------------------------
# closed-form DVR pieces
# --------------------------------------------------------------------------
def sinc2_dvr(n: int, a0: float, b0: float, mass: float):
    x0 = np.zeros(n)
    h0 = np.zeros((n, n), order="F")
    lib().orc_sinc2_dvr(C.c_int(n), C.c_double(a0), C.c_double(b0), C.c_double(mass), dptr(x0), dptr(h0))
    return x0, h0


def gau_legendre(n: int, x1: float = -1.0, x2: float = 1.0):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().orc_gau_legendre(C.c_double(x1), C.c_double(x2), C.c_int(n), dptr(x), dptr(w))
    return x, w


def yj_polys(x: np.ndarray, l: int):
    x = f64(x)
    yj = np.zeros((len(x), l + 1), order="F")
    lib().orc_yj_polys(C.c_int(len(x)), C.c_int(l), dptr(x), dptr(yj))
    return yj


def legendre_l2(n: int):
    x = np.zeros(n)
    l2 = np.zeros((n, n), order="F")
    lib().orc_legendre_l2(C.c_int(n), dptr(x), dptr(l2))
    return x, l2


# --------------------------------------------------------------------------
# small dense helpers
# --------------------------------------------------------------------------
def sym_eigvals(a: np.ndarray) -> np.ndarray:
    a = fmat(a)
    n = a.shape[0]
    w = np.zeros(n)
    lib().orc_sym_eigvals(C.c_int(n), dptr(a), C.c_int(n), dptr(w))
    return w


def tridiag_eigvals(d: np.ndarray, e: np.ndarray) -> np.ndarray:
    d = f64(d)
    e = f64(np.concatenate([e, [0.0]]))
    w = np.zeros(len(d))
    lib().orc_tridiag_eigvals(C.c_int(len(d)), dptr(d), dptr(e), dptr(w))
    return w


def get_window(e0: float, vec: np.ndarray, m: int):
    vec = f64(vec)
    ind = np.zeros(m, dtype=np.int32)
    out = np.zeros(m)
    lib().orc_get_window0(C.c_double(e0), C.c_int(len(vec)), dptr(vec), C.c_int(m), iptr(ind), dptr(out))
    return ind, out


def get_diff_max(v1, v2) -> float:
    v1, v2 = f64(v1), f64(v2)
    return lib().orc_get_diff_max(C.c_int(len(v1)), dptr(v1), dptr(v2))


def get_sep_max(v1, v2) -> float:
    v1, v2 = f64(v1), f64(v2)
    return lib().orc_get_sep_max(C.c_int(len(v1)), dptr(v1), C.c_int(len(v2)), dptr(v2))


def mgs_orth(wj: np.ndarray, vj: np.ndarray, m: int) -> tuple[bool, np.ndarray]:
    """VJ(:, m+1) from WJ; vj is N x (>= m+1) column-major and is updated in place."""
    n = vj.shape[0]
    ok = lib().orc_mgs_orth(C.c_long(n), dptr(f64(wj)), C.c_int(m), dptr(vj))
    return bool(ok), vj


# --------------------------------------------------------------------------
# the OSB operator
# --------------------------------------------------------------------------
class Operator:
    """Direct-product DVR Hamiltonian with its OSB preconditioner.

    ``h0`` is a list of per-level n x n matrices, ``dep`` a list with either
    None or a vector of length myDim(level), ``res`` the diagonal potential on
    the full grid (level 1 fastest)."""

    def __init__(self, sn, h0, res, dep=None, e0: float = 0.0, outh=None):
        self.sF = len(sn)
        self.sN = [int(v) for v in sn]
        self.H0 = [fmat(h) for h in h0]
        dep = list(dep) if dep is not None else [None] * self.sF
        if outh is not None:
            dep[-1] = None   # sNDVR forces sDEP(sF) = F (src/iterate/module/osb.print.h:204)
        self.DEP = [None if d is None else f64(d) for d in dep]
        self.OUTH = None if outh is None else f64(outh).reshape(-1)
        self.RES = f64(res).reshape(-1)
        self.s = OrcOSB()
        self.s.sF = self.sF
        for i in range(self.sF):
            self.s.sN[i] = self.sN[i]
            self.s.sDEP[i] = 0 if self.DEP[i] is None else 1
            self.s.H0[i] = dptr(self.H0[i])
            if self.DEP[i] is not None:
                self.s.DEP[i] = dptr(self.DEP[i])
        lib().orc_cal_osb(C.byref(self.s))
        self.N = int(self.s.myLen)
        assert self.RES.size == self.N
        self.s.RES = dptr(self.RES)
        if self.OUTH is not None:
            assert self.OUTH.size == self.N * self.sN[-1]
            self.s.ndvr = 1
            self.s.OUTH = dptr(self.OUTH)
        self.s.E0 = e0
        self.s.DE = 1.0
        self.VOSB = None
        self.EIG0 = None

    @property
    def myDim(self):
        return [self.s.myDim[i] for i in range(self.sF)]

    @property
    def myBlk(self):
        return [self.s.myBlk[i] for i in range(self.sF)]

    def set_e0(self, e0: float):
        self.s.E0 = e0

    def set_shift(self, e0: float, de: float = 1.0, beta: float = 0.0):
        """sOSBW%mE0 / mDE / mBeta: the shift and the OSBD window."""
        self.s.E0 = e0
        self.s.DE = de
        self.s.beta = beta

    def hx(self, x: np.ndarray, ntype: int = TA0) -> np.ndarray:
        x = f64(x)
        y = np.zeros(self.N)
        lib().orc_myhx(C.byref(self.s), C.c_int(ntype), dptr(x), dptr(y))
        return y

    def hx_t(self, x: np.ndarray, ntype: int = TA0) -> np.ndarray:
        x = f64(x)
        y = np.zeros(self.N)
        lib().orc_myhx_t(C.byref(self.s), C.c_int(ntype), dptr(x), dptr(y))
        return y

    def hx_omp(self, x: np.ndarray, y: np.ndarray, tmp: np.ndarray, ntype: int = TA0):
        lib().orc_myhx_omp(C.byref(self.s), C.c_int(ntype), dptr(x), dptr(y), dptr(tmp))

    def diag(self, bjmax: int = 1000, bjtol: float = 1e-5):
        self.VOSB = [np.zeros(self.myBlk[l] * self.sN[l] ** 2) for l in range(self.sF)]
        self.EIG0 = np.zeros(self.N)
        for l in range(self.sF):
            self.s.VOSB[l] = dptr(self.VOSB[l])
        self.s.EIG0 = dptr(self.EIG0)
        ok = lib().orc_my_diag(C.byref(self.s), C.c_int(bjmax), C.c_double(bjtol))
        assert ok
        return self.VOSB, self.EIG0

    def set_precond(self, vosb, eig0):
        self.VOSB = [f64(v) for v in vosb]
        self.EIG0 = f64(eig0)
        for l in range(self.sF):
            self.s.VOSB[l] = dptr(self.VOSB[l])
        self.s.EIG0 = dptr(self.EIG0)

    def px(self, x: np.ndarray, npc: int = TA1, kind: int = TOSB) -> np.ndarray:
        assert self.VOSB is not None
        x = f64(x)
        y = np.zeros(self.N)
        lib().orc_mypx(C.byref(self.s), C.c_int(kind), C.c_int(npc), dptr(x), dptr(y))
        return y

    def px_omp(self, x: np.ndarray, npc: int = TA1, kind: int = TOSB) -> np.ndarray:
        assert self.VOSB is not None
        x = f64(x)
        y = np.zeros(self.N)
        lib().orc_mypx_omp(C.byref(self.s), C.c_int(kind), C.c_int(npc), dptr(x), dptr(y))
        return y

    def _ctx(self, qmr_type, qmr_max, qmr_tol, threaded=False):
        c = OrcSolverCtx()
        c.o = C.pointer(self.s)
        c.qmr_type, c.qmr_max, c.qmr_tol = qmr_type, qmr_max, qmr_tol
        if threaded:
            self._tmp = np.zeros(self.N)
            c.threaded = 1
            c.tmp = dptr(self._tmp)
        return c

    def qmr(self, b, qmr_type: int = -1, qmr_max: int = 1000, qmr_tol: float = 1e-3, x0=None):
        L = lib()
        b = f64(b)
        x = np.zeros(self.N) if x0 is None else f64(x0).copy()
        res = C.c_double(0.0)
        c = self._ctx(qmr_type, qmr_max, qmr_tol)
        hx = C.cast(L.orc_cb_hx_e, OP_FN)
        px = C.cast(L.orc_cb_px, OP_FN)
        r = L.orc_qmr(C.c_int(qmr_type), C.c_int(qmr_max), C.c_double(qmr_tol), C.c_long(self.N), dptr(b), hx, px,
                      C.byref(c), dptr(x), C.byref(res))
        return r, x, res.value, c

    def pist_conv(self, x, e0, etol, m, m_max, start, step, stepeig=1, ntype=0, qmr_type=-1, qmr_max=1000,
                  qmr_tol=1e-3, threaded=False):
        """PIST_CONV; `threaded` runs its H*X and P*X on the host threads
        (same results), as the reference's MPI path spreads them over ranks."""
        L = lib()
        self.set_e0(e0)
        x = f64(x)
        eig = np.zeros(m)
        res = C.c_double(0.0)
        c = self._ctx(qmr_type, qmr_max, qmr_tol, threaded)
        iters = C.c_int(0)
        r = L.orc_pist_conv(C.c_double(e0), C.c_double(etol), C.c_int(ntype), C.c_long(self.N), dptr(x),
                            C.c_int(start), C.c_int(step), C.c_int(stepeig), C.c_int(m), C.c_int(m_max),
                            C.cast(L.orc_cb_hx0, OP_FN), C.cast(L.orc_cb_linsolv, LS_FN), C.byref(c), dptr(eig),
                            C.byref(res), C.byref(iters))
        return r, eig, res.value, c

    def lan_conv(self, x, e0, etol, m, m_max, start, step, stepeig=1, ntype=0):
        L = lib()
        x = f64(x)
        eig = np.zeros(m)
        res = C.c_double(0.0)
        c = self._ctx(-1, 1, 1.0)
        r = L.orc_lan_conv(C.c_double(e0), C.c_double(etol), C.c_int(ntype), C.c_long(self.N), dptr(x),
                           C.c_int(start), C.c_int(step), C.c_int(stepeig), C.c_int(m), C.c_int(m_max),
                           C.cast(L.orc_cb_hx0, OP_FN), C.byref(c), dptr(eig), C.byref(res))
        return r, eig, res.value

    def dense(self) -> np.ndarray:
        """The full N x N matrix (tiny grids only): column k is H e_k."""
        n = self.N
        m = np.zeros((n, n))
        e = np.zeros(n)
        for k in range(n):
            e[:] = 0.0
            e[k] = 1.0
            m[:, k] = self.hx(e)
        return m


# --------------------------------------------------------------------------
# reference checker wrappers (Fortran by-reference calling convention)
# --------------------------------------------------------------------------
def _ci(v):
    return C.byref(C.c_int(int(v)))


def _cd(v):
    return C.byref(C.c_double(float(v)))


class RefOperator:
    """The same operator driven through the reference's compiled routines."""

    def __init__(self, op: Operator):
        self.op = op
        R = ref()
        sn = (C.c_int * FMAX)(*op.sN)
        sdep = (C.c_int * FMAX)(*[0 if d is None else 1 for d in op.DEP])
        h0 = (_dp * FMAX)(*[dptr(h) for h in op.H0])
        dep = (_dp * FMAX)(*[dptr(d) if d is not None else _dp() for d in op.DEP])
        outh = dptr(op.OUTH) if op.OUTH is not None else _dp()
        R.refdrv_setup(C.c_int(op.sF), sn, sdep, h0, dep, dptr(op.RES), C.c_double(op.s.E0), outh)
        self.vosb = None
        self.eig0 = None

    def set_e0(self, e0):
        op = self.op
        op.set_e0(e0)
        self.__init__(op)

    def hx(self, x, shifted=False):
        x = f64(x)
        y = np.zeros_like(x)
        f = ref().refdrv_hx if shifted else ref().refdrv_hx0
        f(_ci(len(x)), dptr(x), dptr(y))
        return y

    def diag(self, bjmax=1000, bjtol=1e-5):
        op = self.op
        self.vosb = [np.zeros(op.myBlk[l] * op.sN[l] ** 2) for l in range(op.sF)]
        self.eig0 = np.zeros(op.N)
        vp = (_dp * FMAX)(*[dptr(v) for v in self.vosb])
        ok = ref().refdrv_diag(C.c_int(bjmax), C.c_double(bjtol), vp, dptr(self.eig0))
        assert ok
        ref().refdrv_set_precond((_dp * FMAX)(*[dptr(v) for v in self.vosb]), dptr(self.eig0))
        return self.vosb, self.eig0

    def set_shift(self, e0, de=1.0, beta=0.0):
        """sOSBW%mE0 / mDE / mBeta of module osb (and the oracle's E0)."""
        self.op.set_e0(e0)
        ref().refdrv_set_shift(C.c_double(e0), C.c_double(de), C.c_double(beta))

    def px(self, x, kind=None, npc=TA1):
        """Module osb's MYPX: the TOSB / TA1 callback, or kernel `kind`
        (TOSB / TOSBD1 / TOSBD2) with preconditioner shift `npc`."""
        x = f64(x)
        y = np.zeros_like(x)
        if kind is None:
            ref().refdrv_px(_ci(len(x)), dptr(x), dptr(y))
        else:
            ref().refdrv_px_kind(C.c_int(kind), C.c_int(npc), dptr(x), dptr(y))
        return y

    def osbw(self, cnt, e0, de):
        """Module osb's window set-up (adjustWinSize, getOSBWIndex,
        getVOSBH0) after diag(): (0-based grid indices, width, H_W)."""
        d = C.c_double(de)
        n = C.c_int(0)
        ind = np.zeros(cnt, dtype=np.int32)
        hw = np.zeros((cnt, cnt), order="F")
        ref().refop_osbw(C.c_int(cnt), C.c_double(e0), C.byref(d), C.byref(n), iptr(ind), dptr(hw))
        assert n.value == cnt
        return ind - 1, d.value, hw

    def set_ap(self, which, a):
        ref().refop_set_ap(C.c_int(which), dptr(f64(a)))

    def hx_dx(self, x, ntype):
        """Module osb's MYHX_DX on a complex vector."""
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.zeros_like(x)
        ref().refop_hx_dx(C.c_int(ntype), C.c_void_p(x.ctypes.data), C.c_void_p(y.ctypes.data))
        return y

    def px_dx(self, x, kind, npc):
        """Module osb's MYPX_DX (kind TOSB / TOSBD1 / TOSBD2 as the oracle numbers them)."""
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.zeros_like(x)
        ref().refop_px_dx(C.c_int(kind + 1), C.c_int(npc), C.c_void_p(x.ctypes.data), C.c_void_p(y.ctypes.data))
        return y

    def qmr_dx(self, b, kind, hc, pc, ntype=-1, qmr_max=1000, qmr_tol=1e-3):
        """QMR_CX through OSB_QMRDX's operators: (code, x, res)."""
        b = np.ascontiguousarray(b, dtype=np.complex128)
        x = np.zeros_like(b)
        res = C.c_double(0.0)
        R = ref()
        R.refop_qmr_dx.restype = C.c_int
        r = R.refop_qmr_dx(C.c_int(kind + 1), C.c_int(hc), C.c_int(pc), C.c_int(ntype), C.c_int(qmr_max),
                           C.c_double(qmr_tol), C.c_void_p(b.ctypes.data), C.c_void_p(x.ctypes.data), C.byref(res))
        return r, x, res.value

    def osb_vector(self, k):
        """Module osb's getVi for the 0-based grid index k."""
        v = np.zeros(self.op.N)
        ref().refop_vi(C.c_int(k + 1), dptr(v))
        return v

    def qmr(self, b, qmr_type=-1, qmr_max=1000, qmr_tol=1e-3):
        R = ref()
        b = f64(b)
        x = np.zeros_like(b)
        res = C.c_double(0.0)
        FOP = C.CFUNCTYPE(None, _ip, _dp, _dp)
        r = R.qmr_(_ci(qmr_type), _ci(qmr_max), _cd(qmr_tol), _ci(len(b)), dptr(b), C.cast(R.refdrv_hx, FOP),
                   C.cast(R.refdrv_px, FOP), dptr(x), C.byref(res))
        return r, x, res.value

    def pist_h0(self, x, m_max, m_low, m_high, vj, hmat, qmr_type=-1, qmr_max=1000, qmr_tol=1e-3):
        R = ref()
        R.refdrv_set_qmr(C.c_int(qmr_type), C.c_int(qmr_max), C.c_double(qmr_tol))
        FOP = C.CFUNCTYPE(None, _ip, _dp, _dp)
        FLS = C.CFUNCTYPE(C.c_int, _ip, _dp, _dp, _dp)
        x = f64(x)
        return R.pisth0_(_ci(len(x)), dptr(x), _ci(m_max), _ci(m_low), _ci(m_high), C.cast(R.refdrv_hx0, FOP),
                         C.cast(R.refdrv_linsolv, FLS), dptr(vj), dptr(hmat))

    def lan_h0(self, x, vj, m_low, m_high, alpha, beta):
        R = ref()
        FOP = C.CFUNCTYPE(None, _ip, _dp, _dp)
        x = f64(x)
        return R.lanh0_(_ci(len(x)), dptr(x), dptr(vj), _ci(m_low), _ci(m_high), C.cast(R.refdrv_hx0, FOP),
                        dptr(alpha), dptr(beta))


# --------------------------------------------------------------------------
# sequential truncation (the checker for scalit_podvr / scalit_combine_2d)
# --------------------------------------------------------------------------
def column_signs(z: np.ndarray) -> np.ndarray:
    """Every eigenvector column with its largest-magnitude entry positive
    (DVR_Step3 leaves the signs to DSYEV; any choice gives an equivalent
    PO-DVR operator, and this one does not hang on rounding-level entries)."""
    z = np.array(z, dtype=np.float64, order="F")
    im = np.abs(z).argmax(axis=0)
    z[:, z[im, np.arange(z.shape[1])] < 0] *= -1.0
    return z


def podvr(x, t, v, m):
    """DVR_Step2 + DVR_Step3 (src/ham/libham/dvr/dvr.f90:84-160) with numpy's
    LAPACK eigensolver: h = T + diag(v), U its m lowest eigenvectors (each
    column's largest-magnitude entry positive: the signs of U decide the
    signs that normalise W),
    X(I,J) = sum_k U(k,I) x(k) U(k,J) = W diag(x0) W^T, H0 = W^T diag(E) W.
    Returns x0, H0, E (the m lowest eigenvalues of h)."""
    x = f64(x)
    h = np.array(t, dtype=np.float64) + np.diag(f64(v))
    e, u = np.linalg.eigh(0.5 * (h + h.T))
    um = column_signs(u[:, :m])
    xm = um.T @ (x[:, None] * um)
    x0, w = np.linalg.eigh(0.5 * (xm + xm.T))
    w = column_signs(w)
    h0 = w.T @ (e[:m, None] * w)
    return x0, 0.5 * (h0 + h0.T), e[:m]


def combine_2d(h1, v1, h2, v2, nkeep):
    """saveH0CBALL (src/ham/ja3/module/ja3.h0:154-199) with the cut placed so
    that exactly nkeep points stay: the product points in the reference's
    loop order (level 2 outer), the nkeep lowest v1[i] + v2[j] kept (stable in
    that order), H0c(a,b) = [j_a == j_b] h1(i_a,i_b) + [i_a == i_b] h2(j_a,j_b).
    Returns idx1, idx2 (0-based) and H0c."""
    v1, v2 = f64(v1), f64(v2)
    n1, n2 = len(v1), len(v2)
    e = (v1[None, :] + v2[:, None]).reshape(-1)       # loop order: a = i + n1 j
    keep = np.sort(np.argsort(e, kind="stable")[:nkeep])
    i, j = keep % n1, keep // n1
    h1, h2 = np.asarray(h1), np.asarray(h2)
    hc = (j[:, None] == j[None, :]) * h1[i[:, None], i[None, :]] + (i[:, None] == i[None, :]) * h2[j[:, None], j[None, :]]
    return i.astype(np.int32), j.astype(np.int32), np.asfortranarray(hc)


# --------------------------------------------------------------------------
# the windowed preconditioner TOSBW (the checker for scalit_set_osbw and the
# TOSBW kernel of scalit_px)
# --------------------------------------------------------------------------
def adjust_win_size(eig0, e0: float, cnt: int, de: float) -> float:
    """adjustWinSize (src/iterate/module/osb.osbw.h:117-145): double the
    width while fewer than cnt OSB states have |EIG0 - e0| <= width, halve
    towards the last smaller width while more do, until exactly cnt do."""
    d = np.abs(f64(eig0) - e0)
    de1 = 0.0
    k = int((d <= de).sum())
    it = 0
    while k != cnt:
        it += 1
        if it > 10000:
            raise RuntimeError("adjustWinSize does not settle (ties in EIG0)")
        if k < cnt:
            de1, de = de, 2.0 * de
        else:
            de = 0.5 * (de1 + de)
        k = int((d <= de).sum())
    return de


def osbw_index(eig0, e0: float, de: float, cnt: int) -> np.ndarray:
    """getOSBWIndex (src/iterate/module/osb.osbw.h:79-95), 0-based."""
    return np.nonzero(np.abs(f64(eig0) - e0) <= de)[0][:cnt]


def osb_vector(op: "Operator", k: int) -> np.ndarray:
    """getLevelIndex + getLevelVi (src/iterate/module/osb.vi.h:11-73): the
    OSB state k (0-based, level 1 fastest) is the tensor product of column
    c_l of level l's block b_l, c_l the state's level-l digit and b_l the
    number formed by its digits above level l."""
    v = np.ones(1)
    stride = 1
    for l in range(op.sF):
        n = op.sN[l]
        c = (k // stride) % n
        stride *= n
        b = k // stride
        blk = op.VOSB[l][b * n * n:(b + 1) * n * n].reshape(n, n, order="F")
        v = np.kron(blk[:, c], v)
    return v


def osbw_setup(op: "Operator", e0: float, de: float, cnt: int, npc: int = TA1):
    """initOSBW (src/iterate/module/osb.osbw.h:2-50) with getVOSBH0
    (src/iterate/module/osb.hij.h:133-160) and createWH
    (src/iterate/module/osb.pij.h:6-40): returns (indices, width, H_W, HWR,
    window vectors), HWR = H_W less e0 on the diagonal under TA1."""
    de = adjust_win_size(op.EIG0, e0, cnt, de)
    ind = osbw_index(op.EIG0, e0, de, cnt)
    vw = np.stack([osb_vector(op, int(k)) for k in ind], axis=1)
    hv = np.stack([op.hx(vw[:, j], TA0) for j in range(cnt)], axis=1)
    hw = vw.T @ hv
    hw = 0.5 * (hw + hw.T)
    hw[np.diag_indices(cnt)] = op.EIG0[ind]
    hwr = hw - (e0 if npc == TA1 else 0.0) * np.eye(cnt)
    return ind, de, hw, hwr, vw


def osbw_setup_cx(op, e0: float, de: float, cnt: int):
    """initOSBW for complex level matrices (an OperatorCX after diag()):
    adjustWinSize and getOSBWIndex as osbw_setup, and getVOSBH0_CX
    (src/iterate/module/osb.hijcx.h:135-164): H_W(i,j) = v_i . MYHX_CX(TA0)
    v_j for i < j, its conjugate below the diagonal, EIG0 on it.  Returns
    (indices, width, H_W (complex), None, window vectors)."""
    de = adjust_win_size(op.EIG0, e0, cnt, de)
    ind = osbw_index(op.EIG0, e0, de, cnt)
    vw = np.stack([osb_vector(op, int(k)) for k in ind], axis=1)
    hv = np.stack([op.hx(vw[:, j].astype(np.complex128), TA0) for j in range(cnt)], axis=1)
    up = np.triu(vw.T @ hv, 1)
    hw = up + np.conj(up).T + np.diag(op.EIG0[ind]).astype(np.complex128)
    return ind, de, hw, None, vw


def px_tosbw(op: "Operator", x, win, npc: int = TA1) -> np.ndarray:
    """MYPX with nOSB = TOSBW (src/iterate/module/osb.px.h:191-206): the
    rotated vector divided by EIG0 - E0 except on the window, whose
    components are replaced by HWR^-1 times them.  With z = V^T x this is
    P_TOSB x + V_W (HWR^-1 z_W - z_W / (EIG0_W - E0)), z_W = V_W^T x."""
    ind, _, _, hwr, vw = win
    x = f64(x)
    e0 = op.s.E0 if npc == TA1 else 0.0
    zw = vw.T @ x
    corr = np.linalg.solve(hwr, zw) - zw / (op.EIG0[ind] - e0)
    return op.px(x, npc=npc, kind=TOSB) + vw @ corr


# --------------------------------------------------------------------------
# complex (absorbing-potential) operators and solvers: the checker for
# scalit_hx_cx / scalit_px_cx / scalit_qmr_cx / scalit_pist_conv_sx
# --------------------------------------------------------------------------
TAAP0, TMAP0 = 2, 3                    # osb.data.h:18
TAAP, TAAPP, TAAPR = 100, 101, 102     # osb.data.h:20
TMAP, TMAPP, TMAPR = 200, 201, 202     # osb.data.h:22
AP, APP, APR = 0, 1, 2


def _imag_diag(ntype: int, aps):
    """The imaginary diagonal that nType / nPC adds: +-AP, +-APP, +-APR
    (osb.hxdx.h:196-218, osb.mypxdx.h)."""
    ap, app, apr = aps
    return {TAAP0: ap, TMAP0: -ap, TAAP: ap, TAAPP: app, TAAPR: apr,
            TMAP: -ap, TMAPP: -app, TMAPR: -apr}.get(ntype)


def _shift_of(op: "Operator", ntype: int) -> float:
    return 0.0 if ntype in (TA0, TAAP0, TMAP0) else op.s.E0


def hx_dx(op: "Operator", x, ntype: int, aps) -> np.ndarray:
    """MYHX_DX (src/iterate/module/osb.hxdx.h:177-240): Y = (RES - E0 +- i A) X
    plus the real level products applied to the real and imaginary parts."""
    x = np.asarray(x, dtype=np.complex128)
    y = op.hx(x.real, TA0) + 1j * op.hx(x.imag, TA0) - _shift_of(op, ntype) * x
    a = _imag_diag(ntype, aps)
    if a is not None:
        y = y + 1j * a * x
    return y


def osb_basis(op: "Operator") -> np.ndarray:
    """Every OSB state's vector as the columns of V (tiny grids)."""
    if getattr(op, "_vdense", None) is None:
        op._vdense = np.stack([osb_vector(op, k) for k in range(op.N)], axis=1)
    return op._vdense


def px_dx_denominator(op: "Operator", nosb: int, npc: int, aps) -> np.ndarray:
    """The complex divisor of MYPX_DX's kernel (osb.mypxdx.h): EIG0 - E0,
    clamped by TOSBD1 / TOSBD2 (DE, beta), plus +-i A by nPC."""
    e0 = _shift_of(op, npc)
    de = abs(op.s.DE)
    ade = op.EIG0 - e0
    if nosb == TOSBD1:
        ade = np.where(np.abs(ade) <= de, np.where(ade < 0, -de, de), ade)
    elif nosb == TOSBD2:
        beta = op.s.beta + 1.0
        alpha = -op.s.beta * de
        ade = np.where(np.abs(ade) <= de, alpha + beta * ade, ade)
    a = _imag_diag(npc, aps)
    return ade + (0j if a is None else 1j * a)


def px_dx(op: "Operator", x, nosb: int, npc: int, aps, win=None) -> np.ndarray:
    """MYPX_DX (src/iterate/module/osb.mypxdx.h): V (kernel)^-1 V^T X with the
    real OSB rotations applied to complex X.  nosb = TOSBW takes the window
    `win` of osbw_setup (px_tosbw_dx)."""
    if nosb == TOSBW:
        return px_tosbw_dx(op, x, win, npc, aps)
    v = osb_basis(op)
    z = v.T @ np.asarray(x, dtype=np.complex128)
    return v @ (z / px_dx_denominator(op, nosb, npc, aps))


def px_tosbw_dx(op: "Operator", x, win, npc: int, aps) -> np.ndarray:
    """MYPX_DX's TOSBW case (osb.mypxdx.h:127-170) with createWH_DXCX's window
    matrix (osb.pij.h:65-166) for a real H0: the rotated vector divided by
    EIG0 - E0 +- i A, except on the window, whose components are replaced by
    HWR^-1 times them, HWR = H_W + diag(-E0 +- i A(WIND)) (E0 = 0 under
    TA0; H_W with EIG0 on its diagonal, unshifted: win[2] of osbw_setup).
    With z = V^T x: P_TOSB x + V_W (HWR^-1 z_W - z_W / d_W)."""
    ind, _, hw, _, vw = win
    x = np.asarray(x, dtype=np.complex128)
    e0 = _shift_of(op, npc)
    a = _imag_diag(npc, aps)
    hwr = hw.astype(np.complex128) + np.diag(np.full(len(ind), -e0) + (0j if a is None else 1j * a[ind]))
    d = (op.EIG0 - e0 + (0j if a is None else 1j * a))[ind]
    zw = vw.T @ x
    corr = np.linalg.solve(hwr, zw) - zw / d
    return px_dx(op, x, TOSB, npc, aps) + vw @ corr


def dot_cx(x, y) -> complex:
    """DOT_CX (src/libutil/serial/base/vector_cx.f90:227-234): x^T y."""
    return complex(np.sum(x * y))


def qmr_cx(hx, px, b, ntype: int = -1, qmrmax: int = 1000, qmrtol: float = 1e-3, x0=None):
    """QMR_CX (src/iterate/libiterate/qmr/qmrcx.f90:4-179) for complex
    symmetric systems: the real QMR's recurrence with the bilinear x^T y for
    VZDOT / QSDOT and Hermitian norms.  Returns (code, x, res)."""
    b = np.asarray(b, dtype=np.complex128)
    n = b.size
    nrm = lambda v: np.sqrt(abs(np.vdot(v, v)))
    x = np.zeros(n, complex) if (ntype <= 0 or x0 is None) else np.array(x0, dtype=complex)
    zvrat = 1.0
    r = b.copy()
    v = b.copy()
    bnrm = nrm(b)
    vnrm = bnrm
    z = px(b)
    znrm = nrm(z)
    gamamo, eta = 1.0, -1.0 + 0j
    qsdot = 0j
    q = s = ux = ur = None
    thetamo = 0.0
    res = 0.0
    code = 0
    for i in range(1, qmrmax + 1):
        if vnrm == 0.0 or znrm == 0.0:
            break
        zvrat = zvrat * znrm / vnrm
        v = v / vnrm
        z = z / znrm
        vzdot = dot_cx(v, z)
        if abs(vzdot) == 0.0:
            break
        q = z.copy() if i == 1 else z - (vnrm * vzdot) / qsdot * q
        s = zvrat * hx(q)
        qsdot = dot_cx(q, s)
        beta = qsdot / vzdot
        if abs(qsdot) == 0.0 or abs(beta) == 0.0:
            break
        v = s - beta * v
        vnrmpo = nrm(v)
        z = px(v) / zvrat
        znrm = nrm(z)
        theta = vnrmpo / (gamamo * abs(beta))
        gama = 1.0 / np.sqrt(1.0 + theta * theta)
        eta = -eta * vnrm * gama * gama / (beta * gamamo * gamamo)
        if gama == 0.0:
            break
        if i == 1:
            ux = (eta * zvrat) * q
            ur = eta * s
        else:
            ux = (eta * zvrat) * q + ((thetamo * gama) ** 2) * ux
            ur = eta * s + ((thetamo * gama) ** 2) * ur
        x = x + ux
        r = r - ur
        t = abs(ntype)
        if t == 1:
            hx0 = hx(x)
            xnrm = nrm(hx0)
            res = nrm(hx0 - b) / xnrm
        elif t == 2:
            res = max(nrm(r) / bnrm, np.sqrt(abs(np.vdot(ux, ux) / np.vdot(x, x))),
                      np.sqrt(abs(np.vdot(ur, ur) / np.vdot(r, r))))
        else:
            res = nrm(r) / bnrm
        if abs(res) < qmrtol:
            code = i
            break
        code = -i
        vnrm, gamamo, thetamo = vnrmpo, gama, theta
    return code, x, float(res)


def mgs_orth_sx(w, vj: list) -> bool:
    """MGS_ORTH_SX (src/iterate/libiterate/orth/gsorth_sx.f90:47-74): one
    modified Gram-Schmidt step under x^T y, normalised by the complex
    sqrt(w^T w); appends to vj."""
    w = np.array(w, dtype=complex)
    for v in vj:
        w = w - dot_cx(v, w) * v
    nw = np.sqrt(complex(dot_cx(w, w)))
    if nw == 0:
        return False
    vj.append(w / nw)
    return True


def mgs_orth_cx(w, vj: list) -> bool:
    """MGS_ORTH_CX (src/iterate/libiterate/orth/gsorth_cx.f90:51-79): one
    modified Gram-Schmidt step under x^H y, normalised by the real norm;
    appends to vj."""
    w = np.array(w, dtype=complex)
    for v in vj:
        w = w - np.vdot(v, w) * v
    nw = np.sqrt(np.vdot(w, w).real)
    if nw == 0:
        return False
    vj.append(w / nw)
    return True


def _op_cx(vals: np.ndarray, ntype: int) -> np.ndarray:
    """getOpCX (src/libutil/serial/base/util_cx.f90:375-420) for the two
    measures PIST_CONV_SX uses: 1 = real part, 6 = |z|^2."""
    return vals.real.copy() if ntype == 1 else vals.real ** 2 + vals.imag ** 2


def pist_conv_sx(hx0, linsolv, x, e0, etol, ntype, m, m_max, startcnt, stepcnt, stepeig, eigvals=np.linalg.eigvals,
                 hermitian=False):
    """PIST_CONV_SX (src/iterate/libiterate/pist/pist_sx.f90:9-110) with
    pistH0_SX (pistHij_sx.f90:58-121): Krylov vectors from the shifted solves,
    MGS_ORTH_SX, HMAT(i,j) = v_i^T H0 v_j, complex eigenvalues (ZGEEV: here
    eigvals), windows by real part nearest e0 (getWindow_CX, nDiffType 1),
    convergence by getDiffMax_CX of |z|^2 (nConvType 6), ascending real parts
    (Reorder_CX).  ntype > 0 is refused (the reference passes getSepMax_CX
    five arguments for its four).  hermitian: PIST_CONV_CX (pist_cx.f90 with
    pistHij_cx.f90), the same loop with x^H y, real norms and MGS_ORTH_CX."""
    dot = (lambda a, b: complex(np.vdot(a, b))) if hermitian else dot_cx
    orth = mgs_orth_cx if hermitian else mgs_orth_sx
    if ntype > 0:
        raise ValueError("PIST_CONV_SX with NTYPE > 0 is not defined")
    if stepeig >= m_max or startcnt > m_max or m > m_max:
        return 0, np.zeros(m, complex), 0.0
    x = np.asarray(x, dtype=complex)
    cnt_low, cnt_high = 1, max(startcnt, m, stepeig + 1)
    vj, hmat = [], np.zeros((m_max, m_max), complex)
    code, res, eig = 0, 0.0, np.zeros(m, complex)
    while cnt_high <= m_max:
        # pistH0_SX
        if cnt_low == 1:
            t = np.sqrt(dot(x, x).real) if hermitian else np.sqrt(complex(dot_cx(x, x)))
            if t == 0:
                return -cnt_high, eig, res
            vj = [x / t]
            m_start = 1
        else:
            m_start = cnt_low - 1
        for j in range(m_start, cnt_high):
            num, w, _ = linsolv(vj[j - 1])
            if num <= 0:
                return -cnt_high, eig, res
            if not orth(w, vj):
                return -cnt_high, eig, res
        for j in range(cnt_low, cnt_high + 1):
            wj = hx0(vj[j - 1])
            for i in range(1, cnt_high + 1):
                hmat[i - 1, j - 1] = dot(vj[i - 1], wj)
        for j in range(1, cnt_low):
            wj = hx0(vj[j - 1])
            for i in range(cnt_low, cnt_high + 1):
                hmat[i - 1, j - 1] = dot(vj[i - 1], wj)
        num = cnt_high - stepeig
        tmp = eigvals(hmat[:num, :num])
        ind, _ = get_window(e0, _op_cx(tmp, 1), m)     # 1-based, as getWindow0's
        old = tmp[ind - 1]
        tmp = eigvals(hmat[:cnt_high, :cnt_high])
        ind, _ = get_window(e0, _op_cx(tmp, 1), m)
        eig = tmp[ind - 1]
        res = get_diff_max(_op_cx(eig, 6), _op_cx(old, 6))
        if res <= etol:
            code = cnt_high
            break
        code = -cnt_high
        cnt_low = cnt_high + 1
        if cnt_low >= m_max:
            code = -m_max
            break
        cnt_high = min(cnt_high + stepcnt, m_max)
    order = np.argsort(_op_cx(eig, 1), kind="stable")
    return code, eig[order], res


def lan_hij_dx(apply, x, m: int):
    """lanHij_dx (src/iterate/libiterate/lanczos/lanHij_dx.f90:59-97) with
    LAN_MGS_ORTH_CX (orth/gsorth_cx.f90:132-162): m Lanczos steps on a
    Hermitian operator with one full re-orthogonalisation pass per step.
    Returns (code, alpha, beta)."""
    x = np.asarray(x, dtype=complex)
    alpha, beta = np.zeros(m), np.zeros(m)
    t = np.sqrt(np.vdot(x, x).real)
    if t == 0.0:
        return 0, alpha, beta
    vj = [x / t]
    for j in range(1, m + 1):
        w = apply(vj[j - 1])
        if j != 1:
            w = w - beta[j - 2] * vj[j - 2]
        alpha[j - 1] = np.vdot(w, vj[j - 1]).real
        vn = w.copy()
        for v in vj:
            vn = vn - np.vdot(v, vn) * v
        nw = np.sqrt(np.vdot(vn, vn).real)
        beta[j - 1] = nw
        if nw == 0.0:
            return -j, alpha, beta
        vj.append(vn / nw)
    return m, alpha, beta


def lan_dx(apply, x, m: int):
    """LAN_DX (src/iterate/libiterate/lanczos/lan_dx.f90:299-325): lanHij_dx,
    then the tridiagonal's eigenvalues (DSTEV).  Returns (code, eigenvalues
    or None)."""
    code, alpha, beta = lan_hij_dx(apply, x, m)
    if code <= 0:
        return code, None
    return m, tridiag_eigvals(alpha, beta[:m - 1])


def crp_dx(op: "Operator", app, apr, energies, nm: int, x0s, qmr_type=-1, qmrmax=1000, qmrtol=1e-3, sap=True,
           kind: int = TOSB, hx=None, win=None):
    """osbCRP / calOSBCRP / CRP_HXDX (src/iterate/module/osbw.crp.h:7-131):
    N(E) = 4 sum of the nm Lanczos eigenvalues of APR^1/2 (H - i AP - E)^-1
    APP (H + i AP - E)^-1 APR^1/2, AP = APP + APR, each inverse by QMR_CX
    with the OSB preconditioner (including the absorbing term when sap).
    hx(op, u, ntype, aps): the complex H*X (MYHX_DX; an OperatorCX's hx for
    complex level matrices)."""
    hx = hx_dx if hx is None else hx
    app, apr = f64(app), f64(apr)
    ap, sq = app + apr, np.sqrt(apr)
    aps = (ap, app, apr)
    keep = op.s.E0
    out = []
    for e, x in zip(energies, x0s):
        op.s.E0 = e

        def apply(v):
            y = sq * v
            t = qmr_cx(lambda u: hx(op, u, TMAP, aps),
                       lambda u: px_dx(op, u, kind, TMAP if sap else TA1, aps, win), y, qmr_type, qmrmax, qmrtol)[1]
            y = app * t
            t = qmr_cx(lambda u: hx(op, u, TAAP, aps),
                       lambda u: px_dx(op, u, kind, TAAP if sap else TA1, aps, win), y, qmr_type, qmrmax, qmrtol)[1]
            return sq * t

        code, ev = lan_dx(apply, x, nm)
        out.append(4.0 * float(np.sum(ev)) if code > 0 else 0.0)
    op.s.E0 = keep
    return np.array(out)


# --------------------------------------------------------------------------
# complex level matrices (sCX = T): the checker for scalit_set_h0_cx,
# scalit_block_diag_cx and the complex operators on such an operator
# --------------------------------------------------------------------------
class OperatorCX:
    """The sCX = T operator of module osb (src/iterate/module/osb.data.h:63):
    complex level matrices H0CX (n x n, H[i, j] = H1(i, j)), complex DEPCX
    on the dependent levels, optional complex OUTHCX (sNDVR: the outermost
    level's n x n block per inner point, A(p, i, j) at p + nin (i + n j)),
    real RES and absorbing potentials.  Carries the attributes the complex
    preconditioner helpers read (sF, sN, VOSB, EIG0, N, s.E0 / DE / beta)."""

    def __init__(self, sn, h0, res, dep=None, outh=None, e0: float = 0.0):
        self.sF = len(sn)
        self.sN = [int(v) for v in sn]
        self.H0 = [np.asarray(h, dtype=np.complex128) for h in h0]
        dep = list(dep) if dep is not None else [None] * self.sF
        if outh is not None:
            dep[-1] = None
        self.DEP = [None if d is None else np.asarray(d, dtype=np.complex128).reshape(-1) for d in dep]
        self.OUTH = None if outh is None else np.asarray(outh, dtype=np.complex128).reshape(-1)
        self.RES = f64(res).reshape(-1)
        self.N = int(np.prod(self.sN))
        self.myDim = [int(np.prod(self.sN[:l])) for l in range(self.sF)]
        self.myBlk = [int(np.prod(self.sN[l + 1:])) for l in range(self.sF)]
        assert self.RES.size == self.N
        self.s = types.SimpleNamespace(E0=e0, DE=1.0, beta=0.0)
        self.VOSB = None
        self.EIG0 = None

    def set_shift(self, e0: float, de: float = 1.0, beta: float = 0.0):
        self.s.E0, self.s.DE, self.s.beta = e0, de, beta
        self._vdense = getattr(self, "_vdense", None)

    def level(self, l: int, x) -> np.ndarray:
        """Level l's product: H3X_CX for the sNDVR layer
        (src/libutil/serial/base/matrix_cx.f90:194-208), else H1X_XYZ_CX /
        H1X_DEP_CX (src/iterate/libiterate/hpx/hpx_cx.f90:6-39, with
        H1X_CX / H2X_CX of matrix_cx.f90:97-160): y(d, i, b) = [DEP(d)]
        sum_j H(i, j) x(d, j, b)."""
        ndk, n, blk = self.myDim[l], self.sN[l], self.myBlk[l]
        x = np.asarray(x, dtype=np.complex128)
        if self.OUTH is not None and l == self.sF - 1:
            a = self.OUTH.reshape(n, n, ndk)            # [j, i, p]
            return np.einsum("jip,jp->ip", a, x.reshape(n, ndk)).reshape(-1)
        y = np.einsum("ij,bjd->bid", self.H0[l], x.reshape(blk, n, ndk))
        if self.DEP[l] is not None:
            y = y * self.DEP[l][None, None, :]
        return y.reshape(-1)

    def hx(self, x, ntype: int, aps=None) -> np.ndarray:
        """MYHX_CX (src/iterate/module/osb.hxcx.h:191-245): (RES - E0 +- i A) X
        plus every level's product.  TMAPR takes -APR, as the parallel
        MYHX_CX does (src/iterate/module/mosb3.hxcx.h:92-94); the serial one
        (osb.hxcx.h:234-235) takes -APP there."""
        x = np.asarray(x, dtype=np.complex128)
        y = (self.RES - _shift_of(self, ntype)) * x
        a = _imag_diag(ntype, aps) if aps is not None else None
        if a is not None:
            y = y + 1j * a * x
        for l in range(self.sF):
            y = y + self.level(l, x)
        return y

    def dense(self, ntype: int = TA0, aps=None) -> np.ndarray:
        """The full N x N matrix from Kronecker products (tiny grids only)."""
        m = np.diag(self.RES - _shift_of(self, ntype)).astype(np.complex128)
        a = _imag_diag(ntype, aps) if aps is not None else None
        if a is not None:
            m = m + 1j * np.diag(a)
        for l in range(self.sF):
            ndk, n, blk = self.myDim[l], self.sN[l], self.myBlk[l]
            if self.OUTH is not None and l == self.sF - 1:
                a3 = self.OUTH.reshape(n, n, ndk)
                for p in range(ndk):
                    idx = p + ndk * np.arange(n)
                    m[np.ix_(idx, idx)] += a3[:, :, p].T
                continue
            d = np.ones(ndk) if self.DEP[l] is None else self.DEP[l]
            m = m + np.kron(np.eye(blk), np.kron(self.H0[l], np.diag(d)))
        return m

    def diag(self, bjmax: int = 1000, bjtol: float = 1e-5):
        """The sCX branch of myDiag (src/iterate/module/osb.diag.h:76-101):
        EIG0 = RES, then for levels sF..1 the level's HOSB (myDiagInitCX,
        osb.diag.h:36-49: OUTHCX, or HOSBINIT_DEP_CX / HOSBINIT_XYZ_CX of
        hinitdep.f90:38-63 / hinitxyz.f90:31-52) and HosbDiagDX
        (src/iterate/libiterate/diag/hosbdiag_dx.f90:6-73).  Returns
        (VOSB per level, EIG0)."""
        eig0 = self.RES.copy()
        vosb = [None] * self.sF
        for l in range(self.sF - 1, -1, -1):
            ndk, n, blk = self.myDim[l], self.sN[l], self.myBlk[l]
            if self.OUTH is not None and l == self.sF - 1:
                h = self.OUTH.reshape(1, n, n, ndk).copy()          # [b, j, k, d]
            else:
                d = np.ones(ndk) if self.DEP[l] is None else self.DEP[l]
                h = np.broadcast_to(self.H0[l].T[None, :, :, None] * d[None, None, None, :],
                                    (blk, n, n, ndk)).copy()
                e = eig0.reshape(blk, n, ndk)
                for k in range(n):
                    h[:, k, k, :] += e[:, k, :]
            v, ev = hosb_diag_dx(h, bjmax, bjtol)
            vosb[l] = v.reshape(-1)
            eig0 = ev.reshape(-1)
        self.VOSB, self.EIG0 = vosb, eig0
        self._vdense = None
        return vosb, eig0


def _sum_offdiag_cx(h: np.ndarray) -> np.ndarray:
    """sumOffDiagCX (src/iterate/libiterate/diag/sumdiag.f90:49-66) per block
    as HosbDiagDX calls it: its dummy HOUT is declared double precision
    (NIN, NOUT, NOUT) while the actual argument is the complex block, so it
    reads the first NIN NOUT^2 doubles of the block's interleaved (re, im)
    storage with real strides: 2 sum_{i>j} sum_d D(d + NIN (j + NOUT i))^2.
    h: [b, j, k, d] complex (element (d, k, j) of block b)."""
    blk, n, _, ndk = h.shape
    jj, ii = np.triu_indices(n, 1)                   # j < i
    dd = np.ascontiguousarray(h).reshape(blk, -1).view(np.float64)[:, :ndk * n * n].reshape(blk, n, n, ndk)
    return 2.0 * np.sum(dd[:, ii, jj, :] ** 2, axis=(1, 2))     # dd[b, i, j, d]


def _jacobi_angle(a, b, c):
    """MJACOBI's rotation angle (src/iterate/libiterate/diag/jacobi_cx.f90:39-56)."""
    pie = 3.14159265358979323846264338328
    x = np.zeros_like(a)
    eq = a == c
    x = np.where(eq & (b != 0.0), -np.copysign(pie / 8.0, b), x)
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.arctan(b / (a - c)) / 4.0
    t = np.where(a > c, np.where(t < 0, t + pie / 4.0, t - pie / 4.0), t)
    return np.where(eq, x, t)


def hosb_diag_dx(h: np.ndarray, bjmax: int, bjtol: float):
    """HosbDiagDX (src/iterate/libiterate/diag/hosbdiag_dx.f90:6-73) on every
    block at once: h is [b, j, k, d] complex (HOSB(d, k, j, b)), rotated in
    place.  Per block, sweeps of MJACOBI_CX (jacobi_cx.f90:5-96: one real
    angle per pair (p, q) from A = sum |H(p,q)|^2, B = 2 Re sum conj(H(p,q))
    (H(p,p) - H(q,q)) / 2, C = sum |(H(p,p) - H(q,q)) / 2|^2, a symmetric
    update of both triangles) over pairs i = 1..MAK-1, j = 1..MAK-i, pair
    (j, j+i), until |change of sumOffDiagCX| <= tol times its first value or
    BJMAX sweeps; VOSB accumulates the rotations (VUPDATE); EIG0 = real part
    of the diagonal.  A block that has stopped takes angle 0, which leaves
    it unchanged.  Returns (VOSB [b, q, r] = V(r, q), EIG0 [b, k, d])."""
    blk, n, _, ndk = h.shape
    v = np.zeros((blk, n, n))
    v[:, np.arange(n), np.arange(n)] = 1.0
    initye = _sum_offdiag_cx(h)
    ye = initye.copy()
    dye = initye.copy()
    k = np.ones(blk, dtype=int)
    active = (np.abs(dye) > initye * bjtol) & (k <= bjmax)
    pairs = [(j - 1, j + i - 1) for i in range(1, n) for j in range(1, n - i + 1)]
    while active.any():
        for p, q in pairs:
            odg = h[:, q, p, :].copy()
            ddg = 0.5 * h[:, p, p, :] - 0.5 * h[:, q, q, :]
            a = np.sum(np.conj(odg) * odg, axis=1).real
            b = 2.0 * np.sum(np.conj(odg) * ddg, axis=1).real
            c = np.sum(np.conj(ddg) * ddg, axis=1).real
            x = np.where(active, _jacobi_angle(a, b, c), 0.0)
            sn, cs = np.sin(x)[:, None], np.cos(x)[:, None]
            c2, s2, sc = cs ** 2, sn ** 2, sn * cs
            hpp, hqq = h[:, p, p, :].copy(), h[:, q, q, :].copy()
            newp = c2 * hpp + s2 * hqq - 2 * sc * odg
            newq = c2 * hqq + s2 * hpp + 2 * sc * odg
            newqp = (c2 - s2) * odg + 2 * sc * ddg
            h[:, p, p, :] = newp
            h[:, q, q, :] = newq
            h[:, q, p, :] = newqp
            h[:, p, q, :] = newqp
            r = np.array([t for t in range(n) if t != p and t != q], dtype=int)
            if r.size:
                hrp, hrq = h[:, p, r, :].copy(), h[:, q, r, :].copy()    # H(r, p), H(r, q)
                np_ = cs[:, None, :] * hrp - sn[:, None, :] * hrq
                nq_ = cs[:, None, :] * hrq + sn[:, None, :] * hrp
                h[:, p, r, :] = np_
                h[:, r, p, :] = np_
                h[:, q, r, :] = nq_
                h[:, r, q, :] = nq_
            vp, vq = v[:, p, :].copy(), v[:, q, :].copy()
            v[:, p, :] = cs * vp - sn * vq
            v[:, q, :] = cs * vq + sn * vp
        s = _sum_offdiag_cx(h)
        d_ = s - ye
        dye = np.where(active, d_, dye)
        ye = np.where(active, ye + d_, ye)
        k = np.where(active, k + 1, k)
        active = active & (np.abs(dye) > initye * bjtol) & (k <= bjmax)
    eig = np.stack([h[:, i, i, :].real for i in range(n)], axis=1)     # [b, k, d]
    return v, eig


_CXOP = C.CFUNCTYPE(None, _ip, C.c_void_p, C.c_void_p)


def ref_lanhij_dx(apply, x, m: int):
    """The reference's compiled lanHij_dx (lanHij_dx.f90:59-97) on a Python
    Hermitian operator: (code, alpha, beta)."""
    x = np.ascontiguousarray(x, dtype=np.complex128)
    n = x.size

    def cb(np_, xp, yp):
        k = np_[0]
        xv = np.ctypeslib.as_array(C.cast(xp, _dp), shape=(2 * k,)).view(np.complex128)
        yv = np.ctypeslib.as_array(C.cast(yp, _dp), shape=(2 * k,)).view(np.complex128)
        yv[:] = apply(xv.copy())

    f = _CXOP(cb)
    alpha, beta = np.zeros(m), np.zeros(m)
    R = ref()
    R.lanhij_dx_.restype = C.c_int
    code = R.lanhij_dx_(C.byref(C.c_int(n)), C.c_void_p(x.ctypes.data), C.byref(C.c_int(m)), f, dptr(alpha),
                        dptr(beta))
    return code, alpha, beta


def ref_mgs_orth_cx(w, vj: np.ndarray, m: int):
    """The reference's compiled MGS_ORTH_CX (gsorth_cx.f90:51-79)."""
    w = np.ascontiguousarray(w, dtype=np.complex128)
    R = ref()
    R.mgs_orth_cx_.restype = C.c_int
    ok = R.mgs_orth_cx_(C.byref(C.c_int(w.size)), C.c_void_p(w.ctypes.data), C.byref(C.c_int(m)),
                        C.c_void_p(vj.ctypes.data))
    return ok != 0


def ref_mgs_orth_sx(w, vj: np.ndarray, m: int):
    """The reference's compiled MGS_ORTH_SX (gsorth_sx.f90:47-74): vj is
    N x (m + 1) complex, column-major; column m + 1 is written."""
    w = np.ascontiguousarray(w, dtype=np.complex128)
    R = ref()
    R.mgs_orth_sx_.restype = C.c_int
    ok = R.mgs_orth_sx_(C.byref(C.c_int(w.size)), C.c_void_p(w.ctypes.data), C.byref(C.c_int(m)),
                        C.c_void_p(vj.ctypes.data))
    return ok != 0


class RefOperatorCX:
    """An OperatorCX driven through the reference's compiled module osb with
    sCX = T: MYHX_CX (osb.hxcx.h:191-245) and the sCX branch of myDiag
    (osb.diag.h:76-101, HosbDiagDX)."""

    def __init__(self, opc: OperatorCX):
        self.opc = opc
        real = Operator(opc.sN, [h.real for h in opc.H0], opc.RES,
                        [None if d is None else d.real for d in opc.DEP], e0=opc.s.E0,
                        outh=None if opc.OUTH is None else opc.OUTH.real)
        self.base = RefOperator(real)
        layer = opc.OUTH is not None
        h0 = [np.asarray(opc.H0[l], dtype=np.complex128).reshape(-1, order="F")
              for l in range(opc.sF) if not (layer and l == opc.sF - 1)]
        dep = [opc.DEP[l] for l in range(opc.sF) if opc.DEP[l] is not None and not (layer and l == opc.sF - 1)]
        self._h0 = np.ascontiguousarray(np.concatenate(h0) if h0 else np.zeros(1, complex))
        self._dep = np.ascontiguousarray(np.concatenate(dep) if dep else np.zeros(1, complex))
        self._outh = np.ascontiguousarray(opc.OUTH if layer else np.zeros(1, complex))
        ref().refop_set_cx(C.c_void_p(self._h0.ctypes.data), C.c_void_p(self._dep.ctypes.data),
                           C.c_void_p(self._outh.ctypes.data))
        ref().refop_set_shift(C.c_double(opc.s.E0), C.c_double(opc.s.DE), C.c_double(opc.s.beta))

    def set_ap(self, which, a):
        ref().refop_set_ap(C.c_int(which), dptr(f64(a)))

    def hx(self, x, ntype):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.zeros_like(x)
        ref().refop_hx_cx(C.c_int(ntype), C.c_void_p(x.ctypes.data), C.c_void_p(y.ctypes.data))
        return y

    def diag(self, bjmax=1000, bjtol=1e-5):
        opc = self.opc
        nv = [opc.myBlk[l] * opc.sN[l] ** 2 for l in range(opc.sF)]
        vosb = np.zeros(sum(nv))
        eig0 = np.zeros(opc.N)
        ref().refop_diag_cx(C.c_int(bjmax), C.c_double(bjtol), dptr(vosb), dptr(eig0))
        out, o = [], 0
        for k in nv:
            out.append(vosb[o:o + k].copy())
            o += k
        return out, eig0
