"""Python front end of the C ABI: a `Context` owns one operator on one GPU.

Vectors are torch tensors on the GPU of shape (nvec, N), contiguous, with the
grid's level 1 as the fastest index; torch is only the allocator and stream
provider here, every operation runs in libscalit_hip.so.
"""
from __future__ import annotations

import ctypes as C
import re

import numpy as np
import torch

from . import _lib as L
from ._lib import check

_dp = C.POINTER(C.c_double)


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _np_dp(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class Context:
    def __init__(self, sN, sDEP, device: int = 0):
        self.sN = [int(n) for n in sN]
        self.sF = len(self.sN)
        self.sDEP = [1 if d else 0 for d in sDEP]
        self.device = device
        self._h = C.c_void_p()
        sn = (C.c_int * len(self.sN))(*self.sN)
        sd = (C.c_int * len(self.sN))(*self.sDEP)
        check(L.lib().scalit_create(C.byref(self._h), self.sF, sn, sd, device), "scalit_create")
        n = C.c_long()
        check(L.lib().scalit_get_len(self._h, C.byref(n)))
        self.N = n.value
        self.dim = [int(np.prod(self.sN[:l])) for l in range(self.sF)]
        self.blk = [int(np.prod(self.sN[l + 1:])) for l in range(self.sF)]
        self.dev = torch.device("cuda", device)

    @classmethod
    def from_problem(cls, prob, device: int = 0, e0: float = 0.0, outh=None):
        """The operator of a Problem; with `outh` (the flat OUTH array of
        problems.outer_layer) the outermost level is the sNDVR = T dense layer."""
        dep = list(prob.dep)
        if outh is not None:
            dep[-1] = None
        ctx = cls(prob.sN, [d is not None for d in dep], device)
        for l, h in enumerate(prob.h0):
            if outh is None or l < len(prob.h0) - 1:
                ctx.set_h0(l + 1, h)
        for l, d in enumerate(dep):
            if d is not None:
                ctx.set_dep(l + 1, d)
        if outh is not None:
            ctx.set_outh(outh)
        ctx.set_res(prob.res)
        ctx.set_shift(e0)
        return ctx

    # ------------------------------------------------------------ files
    @staticmethod
    def layout(sN, sDEP, ndvr: bool = False):
        """calOSB's offsets and sizes (scalit_layout): where each level's piece
        starts in the concatenated H0 / DEP / VOSB files, and their lengths."""
        sF = len(sN)
        h0s, deps, vs, tot = (C.c_long * sF)(), (C.c_long * sF)(), (C.c_long * sF)(), (C.c_long * 5)()
        check(L.lib().scalit_layout(sF, (C.c_int * sF)(*sN), (C.c_int * sF)(*[1 if d else 0 for d in sDEP]),
                                    1 if ndvr else 0, h0s, deps, vs, tot), "scalit_layout")
        return {"h0_start": list(h0s), "dep_start": list(deps), "vosb_start": list(vs), "N": tot[0],
                "h0_len": tot[1], "dep_len": tot[2], "vosb_len": tot[3], "outh_len": tot[4]}

    @staticmethod
    def load_file(path: str, n: int, mode: str = "seq") -> np.ndarray:
        """n doubles from a file in one of the reference's formats: "seq"
        (sequential unformatted), "text" (list-directed) or "dir" (direct access)."""
        a = np.zeros(int(n))
        if mode == "dir":
            rc = L.lib().scalit_load_data_dir(path.encode(), a.size, _np_dp(a))
        else:
            rc = L.lib().scalit_load_data(path.encode(), a.size, _np_dp(a), 1 if mode == "seq" else 0)
        check(rc, f"load {path}")
        return a

    @staticmethod
    def load_file_cx(path: str, n: int, mode: str = "seq") -> np.ndarray:
        """n complex values from a file in one of the reference's formats
        (loadData_CX / loadDataDir_CX, src/libutil/serial/io/io_cx.f90:102-126,
        iodirect_cx.f90:68-90): a binary record holds the (re, im) pairs as
        2n doubles, so "seq" and "dir" read those; "text" reads list-directed
        complex constants "(re, im)" (with r* repeat counts)."""
        if mode == "text":
            with open(path) as f:
                text = f.read()
            num = r"[-+0-9.eEdD]+"
            vals = []
            for rep, re_, im_ in re.findall(rf"(?:(\d+)\*)?\(\s*({num})\s*,\s*({num})\s*\)", text):
                z = complex(float(re_.replace("d", "e").replace("D", "e")), float(im_.replace("d", "e").replace("D", "e")))
                vals.extend([z] * (int(rep) if rep else 1))
            if len(vals) < n:
                raise L.ScalitError(f"load {path}: {len(vals)} complex values, {n} expected")
            return np.array(vals[:n], dtype=np.complex128)
        return Context.load_file(path, 2 * int(n), mode).view(np.complex128)

    @staticmethod
    def load_grid_file(path: str, npts: int, ng: int) -> np.ndarray:
        """An OUTH / H0GM file of per-point direct records (P2SaveDataGrid,
        src/libutil/parallel/p_io/p2io_grid.f90:62-83) in the operator's
        layout: npts x ng x ng values, point index fastest."""
        a = np.zeros(int(npts) * ng * ng)
        check(L.lib().scalit_load_data_grid(path.encode(), int(npts), int(ng), _np_dp(a)), f"load {path}")
        return a

    @staticmethod
    def save_file(path: str, a: np.ndarray, mode: str = "seq"):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        if mode == "dir":
            rc = L.lib().scalit_save_data_dir(path.encode(), a.size, _np_dp(a))
        else:
            rc = L.lib().scalit_save_data(path.encode(), a.size, _np_dp(a), 1 if mode == "seq" else 0)
        check(rc, f"save {path}")

    @classmethod
    def from_files(cls, sN, sDEP, h0_file: str, res_file: str | None = None, dep_file: str | None = None,
                   outh_file: str | None = None, bin_mode: bool = True, device: int = 0, e0: float = 0.0,
                   outh_per_point: bool = False):
        """An operator from the files an iterate job reads (loadInitData,
        osb.io.h:2-62, 211-235): H0 of every level concatenated, sequential
        (binary when bin_mode, else text); RES, DEP (dependent levels
        concatenated) and, for the sNDVR = T layer, OUTH as one direct-access
        record (loadOutH) or, with outh_per_point, one record per inner grid
        point as the ja3 generator writes it.  With OUTH the job has no RES
        file and RES is zero."""
        ndvr = outh_file is not None
        lay = cls.layout(sN, sDEP, ndvr)
        h0 = cls.load_file(h0_file, lay["h0_len"], "seq" if bin_mode else "text")
        res = np.zeros(lay["N"]) if (ndvr and res_file is None) else cls.load_file(res_file, lay["N"], "dir")
        dep = cls.load_file(dep_file, lay["dep_len"], "dir") if lay["dep_len"] else None
        ctx = cls(sN, sDEP, device)
        for l, n in enumerate(ctx.sN):
            if ndvr and l == ctx.sF - 1:
                continue
            st = lay["h0_start"][l]
            ctx.set_h0(l + 1, h0[st:st + n * n].reshape(n, n, order="F"))
            if ctx.sDEP[l]:
                st = lay["dep_start"][l]
                ctx.set_dep(l + 1, dep[st:st + ctx.dim[l]])
        if ndvr:
            if outh_per_point:
                ctx.set_outh(cls.load_grid_file(outh_file, ctx.dim[-1], ctx.sN[-1]))
            else:
                ctx.set_outh(cls.load_file(outh_file, lay["outh_len"], "dir"))
        ctx.set_res(res)
        ctx.set_shift(e0)
        return ctx

    @classmethod
    def from_files_cx(cls, sN, sDEP, h0_file: str, res_file: str | None = None, dep_file: str | None = None,
                      outh_file: str | None = None, bin_mode: bool = True, device: int = 0, e0: float = 0.0):
        """An sCX = T operator from its files (loadH0 / loadDEP / loadOutH under
        sCX, osb.io.h:45-55,211-236): H0CX of every level concatenated,
        sequential (binary when bin_mode, else text); DEPCX and OUTHCX one
        direct-access record each; RES real (zero with OUTH and no RES file).
        The context holds only the complex level matrices, with sCX on: its
        complex operators and set-up run, its real H*X returns SCALIT_ESTATE."""
        ndvr = outh_file is not None
        lay = cls.layout(sN, sDEP, ndvr)
        h0 = cls.load_file_cx(h0_file, lay["h0_len"], "seq" if bin_mode else "text")
        res = np.zeros(lay["N"]) if (ndvr and res_file is None) else cls.load_file(res_file, lay["N"], "dir")
        dep = cls.load_file_cx(dep_file, lay["dep_len"], "dir") if lay["dep_len"] else None
        ctx = cls(sN, sDEP, device)
        if ndvr:
            ctx.set_outh_cx(cls.load_file_cx(outh_file, lay["outh_len"], "dir"))
        for l, n in enumerate(ctx.sN):
            if ndvr and l == ctx.sF - 1:
                continue
            st = lay["h0_start"][l]
            ctx.set_h0_cx(l + 1, h0[st:st + n * n].reshape(n, n, order="F"))
            if ctx.sDEP[l]:
                st = lay["dep_start"][l]
                ctx.set_dep_cx(l + 1, dep[st:st + ctx.dim[l]])
        ctx.set_res(res)
        ctx.set_shift(e0)
        ctx.set_scx(True)
        return ctx

    def save_precond(self, vosb_file: str, eig_file: str, mode: str = "dir"):
        """VOSB (levels concatenated) and EIG0, as saveVOSBEig writes them
        (osb.io.h:192-206): direct access, or sequential ("seq" / "text")."""
        vosb, eig0 = self.get_precond()
        self.save_file(vosb_file, np.concatenate(vosb), mode)
        self.save_file(eig_file, eig0, mode)

    def load_precond(self, vosb_file: str, eig_file: str, mode: str = "dir"):
        """The preconditioner from files (loadVOSBEig, osb.io.h:171-188)
        instead of a set-up."""
        tot = sum(self.blk[l] * self.sN[l] ** 2 for l in range(self.sF))
        flat = self.load_file(vosb_file, tot, mode)
        eig0 = self.load_file(eig_file, self.N, mode)
        parts, st = [], 0
        for l in range(self.sF):
            n = self.blk[l] * self.sN[l] ** 2
            parts.append(flat[st:st + n])
            st += n
        self.set_precond(parts, eig0)

    @staticmethod
    def qmr_multi(problems, ntype: int = -1, qmrmax: int = 1000, qmrtol: float = 1e-3):
        """Several operators' QMR solves in lock step, one graph launch per
        iteration for all of them (scalit_qmr_multi): `problems` is a list of
        (context, b tensor of shape (nvec, N), shifts or None).  Returns per
        problem (ret, x, res)."""
        jobs = (L.QmrJob * len(problems))()
        keep = []
        for j, (ctx, b, shifts) in enumerate(problems):
            b = b.contiguous()
            nv = b.shape[0]
            x = torch.zeros_like(b)
            sh = None if shifts is None else (C.c_double * nv)(*[float(v) for v in shifts])
            ret = (C.c_int * nv)()
            res = (C.c_double * nv)()
            keep.append((b, x, sh, ret, res))
            jobs[j].ctx, jobs[j].nvec, jobs[j].ld = ctx._h.value, nv, ctx.N
            jobs[j].b, jobs[j].x = b.data_ptr(), x.data_ptr()
            jobs[j].shifts = C.cast(sh, _dp) if sh is not None else None
            jobs[j].ret, jobs[j].res = C.cast(ret, C.POINTER(C.c_int)), C.cast(res, _dp)
        check(L.lib().scalit_qmr_multi(len(problems), C.cast(jobs, C.c_void_p), ntype, qmrmax, qmrtol),
              "scalit_qmr_multi")
        return [(np.array(k[3][:]), k[1], np.array(k[4][:])) for k in keep]

    @staticmethod
    def pist_conv_multi(problems, etol: float, m: int, m_max: int, startcnt: int, stepcnt: int, stepeig: int = 1,
                        ntype: int = 0, qmr_type: int = -1, qmrmax: int = 1000, qmrtol: float = 1e-3,
                        prec: int = L.FP64):
        """Several problems in one call (scalit_pist_conv_multi): `problems` is a
        list of (context, x0 tensor of shape (nwin, N), window centres); each
        runs its batched PIST on its own stream.  Returns per problem
        (ret, eig (nwin, m), res, stats)."""
        jobs = (L.PistJob * len(problems))()
        keep = []
        for j, (ctx, x0, e0s) in enumerate(problems):
            x0 = x0.contiguous()
            nw = x0.shape[0]
            e = (C.c_double * nw)(*[float(v) for v in e0s])
            eig = (C.c_double * (nw * m))()
            ret = (C.c_int * nw)()
            res = (C.c_double * nw)()
            keep += [x0, e, eig, ret, res]
            jobs[j].ctx, jobs[j].nwin, jobs[j].ld = ctx._h.value, nw, ctx.N
            jobs[j].e0 = C.cast(e, _dp)
            jobs[j].x = x0.data_ptr()
            jobs[j].eig, jobs[j].ret, jobs[j].res = C.cast(eig, _dp), C.cast(ret, C.POINTER(C.c_int)), C.cast(res, _dp)
        rc = L.lib().scalit_pist_conv_multi(len(problems), C.cast(jobs, C.c_void_p), etol, ntype, startcnt, stepcnt,
                                            stepeig, m, m_max, qmr_type, qmrmax, qmrtol, prec)
        check(rc, "scalit_pist_conv_multi")
        out = []
        for j, (ctx, x0, e0s) in enumerate(problems):
            nw = len(e0s)
            eig, ret, res = keep[5 * j + 2], keep[5 * j + 3], keep[5 * j + 4]
            out.append((np.array(ret[:]), np.array(eig[:]).reshape(nw, m), np.array(res[:]),
                        np.array(jobs[j].stats[:])))
        return out

    def close(self):
        if self._h:
            L.lib().scalit_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # ------------------------------------------------------------ operator data
    def set_h0(self, level: int, h0: np.ndarray):
        a = np.asfortranarray(h0, dtype=np.float64)
        check(L.lib().scalit_set_h0(self._h, level, _np_dp(a)), "scalit_set_h0")

    def set_dep(self, level: int, dep: np.ndarray):
        a = np.ascontiguousarray(dep, dtype=np.float64)
        check(L.lib().scalit_set_dep(self._h, level, _np_dp(a)), "scalit_set_dep")

    def set_outh(self, outh: np.ndarray):
        a = np.ascontiguousarray(outh, dtype=np.float64).reshape(-1)
        assert a.size == self.N * self.sN[-1]
        check(L.lib().scalit_set_outh(self._h, _np_dp(a)), "scalit_set_outh")

    def set_res(self, res: np.ndarray):
        a = np.ascontiguousarray(res, dtype=np.float64)
        check(L.lib().scalit_set_res(self._h, _np_dp(a)), "scalit_set_res")

    def set_shift(self, e0: float, de: float = 1.0, beta: float = 0.0):
        self.e0 = e0
        check(L.lib().scalit_set_shift(self._h, e0, de, beta))

    def block_diag(self, bjmax: int = 1000, bjtol: float = 1e-5):
        check(L.lib().scalit_block_diag(self._h, bjmax, bjtol), "scalit_block_diag")

    def get_precond(self):
        vosb = [np.zeros(self.blk[l] * self.sN[l] ** 2) for l in range(self.sF)]
        eig0 = np.zeros(self.N)
        arr = (_dp * self.sF)(*[_np_dp(v) for v in vosb])
        check(L.lib().scalit_get_precond(self._h, arr, _np_dp(eig0)), "scalit_get_precond")
        return vosb, eig0

    def set_precond(self, vosb, eig0):
        vs = [np.ascontiguousarray(v, dtype=np.float64) for v in vosb]
        e = np.ascontiguousarray(eig0, dtype=np.float64)
        arr = (_dp * self.sF)(*[_np_dp(v) for v in vs])
        check(L.lib().scalit_set_precond(self._h, arr, _np_dp(e)), "scalit_set_precond")

    # ------------------------------------------------------------ vectors
    def vectors(self, nvec: int, dtype=torch.float64):
        return torch.empty((nvec, self.N), dtype=dtype, device=self.dev)

    def to_device(self, a: np.ndarray, dtype=torch.float64):
        a = np.atleast_2d(a)
        return torch.from_numpy(np.ascontiguousarray(a)).to(device=self.dev, dtype=dtype)

    def _shifts(self, shifts, dtype):
        if shifts is None:
            return None
        return torch.as_tensor(np.asarray(shifts), dtype=dtype, device=self.dev).contiguous()

    def hx(self, x: torch.Tensor, ntype: int = L.TA0, shifts=None, out=None):
        x = x.contiguous()
        nv = x.shape[0]
        y = torch.empty_like(x) if out is None else out
        sh = self._shifts(shifts, torch.float64)
        check(L.lib().scalit_hx(self._h, ntype, nv, _ptr(x), self.N, _ptr(y), self.N, _ptr(sh), _stream()),
              "scalit_hx")
        return y

    def hx_lp(self, x: torch.Tensor, ntype: int = L.TA0, shifts=None, out=None):
        prec = {torch.float32: L.FP32, torch.bfloat16: L.BF16}[x.dtype]
        x = x.contiguous()
        y = torch.empty_like(x) if out is None else out
        sh = self._shifts(shifts, torch.float32)
        check(L.lib().scalit_hx_lp(self._h, prec, ntype, x.shape[0], _ptr(x), self.N, _ptr(y), self.N, _ptr(sh),
                                   _stream()), "scalit_hx_lp")
        return y

    def hx_variant(self, prec: int, nvec: int):
        """(pass-A, pass-B) variants of the fused H*X for this precision and
        batch size, as the autotuner (or SCALIT_HX_VARIANT_A/B) chose them."""
        va, vb = C.c_int(), C.c_int()
        check(L.lib().scalit_debug_hx_variant(self._h, prec, nvec, C.byref(va), C.byref(vb)),
              "scalit_debug_hx_variant")
        return va.value, vb.value

    def hx_adjoint(self, x: torch.Tensor, g: torch.Tensor, ntype: int = L.TA0, with_dv: bool = True):
        g = g.contiguous()
        x = x.contiguous()
        dx = torch.empty_like(g)
        dv = torch.zeros(self.N, dtype=torch.float64, device=self.dev) if with_dv else None
        check(L.lib().scalit_hx_adjoint(self._h, ntype, g.shape[0], _ptr(x), _ptr(g), self.N, _ptr(dx), _ptr(dv),
                                        _stream()), "scalit_hx_adjoint")
        return dx, dv

    def px(self, x: torch.Tensor, nosb: int = L.TOSB, npc: int = L.TA1, shifts=None, out=None):
        x = x.contiguous()
        y = torch.empty_like(x) if out is None else out
        sh = self._shifts(shifts, torch.float64)
        check(L.lib().scalit_px(self._h, nosb, npc, x.shape[0], _ptr(x), self.N, _ptr(y), self.N, _ptr(sh),
                                _stream()), "scalit_px")
        return y

    def px_lp(self, x: torch.Tensor, npc: int = L.TA1, shifts=None, out=None):
        prec = {torch.float32: L.FP32, torch.bfloat16: L.BF16}[x.dtype]
        x = x.contiguous()
        y = torch.empty_like(x) if out is None else out
        sh = self._shifts(shifts, torch.float32)
        check(L.lib().scalit_px_lp(self._h, prec, npc, x.shape[0], _ptr(x), self.N, _ptr(y), self.N, _ptr(sh),
                                   _stream()), "scalit_px_lp")
        return y

    # ------------------------------------------------------------ solvers
    def qmr(self, b: torch.Tensor, ntype: int = -1, qmrmax: int = 1000, qmrtol: float = 1e-3, shifts=None, x0=None):
        b = b.contiguous()
        nv = b.shape[0]
        x = torch.zeros_like(b) if x0 is None else x0.clone().contiguous()
        ret = (C.c_int * nv)()
        res = (C.c_double * nv)()
        sh = None if shifts is None else (C.c_double * nv)(*[float(s) for s in shifts])
        check(L.lib().scalit_qmr(self._h, ntype, qmrmax, qmrtol, nv, _ptr(b), _ptr(x), self.N, sh, ret, res,
                                 _stream()), "scalit_qmr")
        return np.array(ret[:]), x, np.array(res[:])

    def qmr_mixed(self, b: torch.Tensor, prec: int = L.FP32, qmrmax: int = 10, qmrtol: float = 1e-10,
                  inner_max: int = 1000, inner_tol: float = 1e-5, shifts=None):
        b = b.contiguous()
        nv = b.shape[0]
        x = torch.zeros_like(b)
        ret = (C.c_int * nv)()
        res = (C.c_double * nv)()
        sh = None if shifts is None else (C.c_double * nv)(*[float(s) for s in shifts])
        check(L.lib().scalit_qmr_mixed(self._h, prec, qmrmax, qmrtol, inner_max, inner_tol, nv, _ptr(b), _ptr(x),
                                       self.N, sh, ret, res, _stream()), "scalit_qmr_mixed")
        return np.array(ret[:]), x, np.array(res[:])

    def pist_conv(self, x0: torch.Tensor, e0: float, etol: float, m: int, m_max: int, startcnt: int, stepcnt: int,
                  stepeig: int = 1, ntype: int = 0, qmr_type: int = -1, qmrmax: int = 1000, qmrtol: float = 1e-3,
                  prec: int = L.FP64):
        x0 = x0.contiguous().reshape(-1)
        eig = (C.c_double * m)()
        res = C.c_double()
        stats = (C.c_int * 4)()
        r = L.lib().scalit_pist_conv(self._h, e0, etol, ntype, _ptr(x0), startcnt, stepcnt, stepeig, m, m_max,
                                     qmr_type, qmrmax, qmrtol, prec, eig, C.byref(res), stats)
        if r <= L.SCALIT_EINVAL and r >= L.SCALIT_ESTATE:
            check(r, "scalit_pist_conv")
        return r, np.array(eig[:]), res.value, np.array(stats[:])

    def pist_conv_batched(self, x0: torch.Tensor, e0s, etol: float, m: int, m_max: int, startcnt: int, stepcnt: int,
                          stepeig: int = 1, ntype: int = 0, qmr_type: int = -1, qmrmax: int = 1000,
                          qmrtol: float = 1e-3, prec: int = L.FP64):
        x0 = x0.contiguous()
        nw = x0.shape[0]
        e = (C.c_double * nw)(*[float(v) for v in e0s])
        eig = (C.c_double * (nw * m))()
        ret = (C.c_int * nw)()
        res = (C.c_double * nw)()
        stats = (C.c_int * 4)()
        check(L.lib().scalit_pist_conv_batched(self._h, nw, e, etol, ntype, _ptr(x0), self.N, startcnt, stepcnt,
                                               stepeig, m, m_max, qmr_type, qmrmax, qmrtol, prec, eig, ret, res,
                                               stats), "scalit_pist_conv_batched")
        return np.array(ret[:]), np.array(eig[:]).reshape(nw, m), np.array(res[:]), np.array(stats[:])

    def set_pist_orth(self, passes: int):
        """1: the reference's one MGS pass per Krylov vector; 2: twice."""
        check(L.lib().scalit_set_pist_orth(self._h, passes), "scalit_set_pist_orth")

    def set_qmr_persistent(self, on: int):
        """FP64 QMR on small fused 3-level operators: 1 one persistent launch
        per solve, 0 one kernel per step, -1 as SCALIT_QMR_PERSIST says."""
        check(L.lib().scalit_set_qmr_persistent(self._h, on), "scalit_set_qmr_persistent")

    def set_osb(self, nosb: int):
        """sOSB: the preconditioner kernel of QMR / PIST (TOSB, TOSBD1,
        TOSBD2, or TOSBW after set_osbw), with DE and beta from set_shift."""
        check(L.lib().scalit_set_osb(self._h, nosb), "scalit_set_osb")

    def set_osbw(self, e0: float, de: float, cnt: int, npc: int = L.TA1):
        """initOSBW (src/iterate/module/osb.osbw.h:2-50): the window of the
        cnt OSB states nearest e0 (adjustWinSize from width de) for the TOSBW
        kernel; returns (window size, width found)."""
        n, d = C.c_int(), C.c_double()
        check(L.lib().scalit_set_osbw(self._h, npc, e0, de, cnt, C.byref(n), C.byref(d)), "scalit_set_osbw")
        return n.value, d.value

    # ------------------------------------------------------------ complex jobs
    def to_device_cx(self, z) -> torch.Tensor:
        """Complex vectors (b, N) in the split layout: (b, 2, N) float64, the
        real plane then the imaginary plane of each vector."""
        z = np.atleast_2d(np.asarray(z, dtype=np.complex128))
        return self.to_device(np.stack([z.real, z.imag], axis=1).reshape(-1, self.N)).reshape(-1, 2, self.N)

    @staticmethod
    def from_device_cx(t: torch.Tensor) -> np.ndarray:
        a = t.detach().cpu().numpy().reshape(t.shape[0], 2, -1)
        return a[:, 0] + 1j * a[:, 1]

    def set_ap(self, which: int, a: np.ndarray):
        """AP / APP / APR, the absorbing potentials (osb.data.h:61)."""
        a = np.ascontiguousarray(a, dtype=np.float64)
        check(L.lib().scalit_set_ap(self._h, which, _np_dp(a)), "scalit_set_ap")

    def set_h0_cx(self, level: int, h0: np.ndarray):
        """H0CX of a level (1-based): an n x n complex matrix, H[i, j] = H1(i, j)
        (osb.data.h:63; complex symmetric, as the sCX set-up assumes)."""
        h = np.asarray(h0, dtype=np.complex128)
        re = np.ascontiguousarray(h.real.reshape(-1, order="F"))
        im = np.ascontiguousarray(h.imag.reshape(-1, order="F"))
        check(L.lib().scalit_set_h0_cx(self._h, level, _np_dp(re), _np_dp(im)), "scalit_set_h0_cx")

    def set_dep_cx(self, level: int, dep: np.ndarray):
        """DEPCX of a dependent level (1-based), myDim(level) complex values."""
        d = np.asarray(dep, dtype=np.complex128).reshape(-1)
        re, im = np.ascontiguousarray(d.real), np.ascontiguousarray(d.imag)
        check(L.lib().scalit_set_dep_cx(self._h, level, _np_dp(re), _np_dp(im)), "scalit_set_dep_cx")

    def set_outh_cx(self, outh: np.ndarray):
        """OUTHCX of an sNDVR operator: N n_sF complex values in OUTH's layout."""
        a = np.asarray(outh, dtype=np.complex128).reshape(-1)
        re, im = np.ascontiguousarray(a.real), np.ascontiguousarray(a.imag)
        check(L.lib().scalit_set_outh_cx(self._h, _np_dp(re), _np_dp(im)), "scalit_set_outh_cx")

    def set_scx(self, on: bool = True):
        """sCX: the complex level matrices replace the real ones in the complex
        operators and in block_diag's OSB set-up."""
        check(L.lib().scalit_set_scx(self._h, 1 if on else 0), "scalit_set_scx")

    def hx_cx(self, x: torch.Tensor, ntype: int = L.TMAP):
        """MYHX_DX (src/iterate/module/osb.hxdx.h:177-240) on split vectors, or
        MYHX_CX (osb.hxcx.h:191-245) when sCX is on."""
        x = x.contiguous()
        y = torch.empty_like(x)
        check(L.lib().scalit_hx_cx(self._h, ntype, x.shape[0], _ptr(x), _ptr(y), _stream()), "scalit_hx_cx")
        return y

    def px_cx(self, x: torch.Tensor, nosb: int = L.TOSB, npc: int = L.TMAP):
        """MYPX_DX (src/iterate/module/osb.mypxdx.h) on split vectors."""
        x = x.contiguous()
        y = torch.empty_like(x)
        check(L.lib().scalit_px_cx(self._h, nosb, npc, x.shape[0], _ptr(x), _ptr(y), _stream()), "scalit_px_cx")
        return y

    def qmr_cx(self, b: torch.Tensor, qmr_type: int = -1, hc: int = L.TMAP, pc: int = L.TMAP, qmrmax: int = 1000,
               qmrtol: float = 1e-3, x0=None):
        """OSB_QMRDX (src/iterate/module/osbw.qmrdx.h:16-27): QMR_CX for one
        split right-hand side (1, 2, N).  Returns (QMR_CX's value, x, res)."""
        b = b.contiguous()
        x = torch.zeros_like(b) if x0 is None else x0.clone().contiguous()
        ret, res = C.c_int(), C.c_double()
        check(L.lib().scalit_qmr_cx(self._h, qmr_type, hc, pc, qmrmax, qmrtol, _ptr(b), _ptr(x), C.byref(ret),
                                    C.byref(res)), "scalit_qmr_cx")
        return ret.value, x, res.value

    def pist_conv_sx(self, x0: torch.Tensor, e0: float, etol: float, m: int, m_max: int, startcnt: int, stepcnt: int,
                     stepeig: int = 1, ntype: int = 0, qmr_type: int = -1, qmrmax: int = 1000, qmrtol: float = 1e-3,
                     pc: int = L.TMAP, hermitian: bool = False):
        """OSB_PISTCONVDX1 (src/iterate/module/osbw.pistconv.h:98-140): complex
        resonance energies nearest e0 (by real part) of H - i AP; with
        hermitian, OSB_PISTCONVDX2's PIST_CONV_CX (x^H y products).  Returns
        (the PIST value, eigenvalues ascending by real part, res, stats)."""
        x0 = x0.contiguous()
        eig = (C.c_double * (2 * m))()
        ret, res = C.c_int(), C.c_double()
        stats = (C.c_int * 4)()
        fn = L.lib().scalit_pist_conv_cx if hermitian else L.lib().scalit_pist_conv_sx
        check(fn(self._h, e0, etol, ntype, _ptr(x0), startcnt, stepcnt, stepeig, m, m_max, qmr_type, qmrmax, qmrtol,
                 pc, eig, C.byref(ret), C.byref(res), stats), "scalit_pist_conv_sx/cx")
        e = np.array(eig[:])
        return ret.value, e[0::2] + 1j * e[1::2], res.value, np.array(stats[:])

    def crp(self, energies, nm: int, x0: torch.Tensor, qmr_type: int = -1, qmrmax: int = 1000, qmrtol: float = 1e-3,
            sap: bool = True):
        """osbCRP (src/iterate/module/osbw.crp.h:7-49): N(E) at each energy from
        nm Lanczos steps on the CRP operator; x0 (ncrp, 2, N) start vectors.
        Returns (N(E) per energy, [QMR iterations, QMR solves])."""
        e = np.ascontiguousarray(energies, dtype=np.float64)
        x0 = x0.contiguous()
        out = np.zeros(len(e))
        stats = (C.c_int * 2)()
        check(L.lib().scalit_crp(self._h, len(e), _np_dp(e), nm, _ptr(x0), qmr_type, qmrmax, qmrtol, 1 if sap else 0,
                                 _np_dp(out), stats), "scalit_crp")
        return out, np.array(stats[:])

    def pist_vectors(self, win: int, m: int):
        psi = torch.empty((m, self.N), dtype=torch.float64, device=self.dev)
        check(L.lib().scalit_pist_vectors(self._h, win, _ptr(psi), self.N), "scalit_pist_vectors")
        return psi

    def pist_vectors_all(self, nwin: int, m: int):
        """Ritz vectors of all nwin windows of the last batched solve, (nwin m, N)."""
        psi = torch.empty((nwin * m, self.N), dtype=torch.float64, device=self.dev)
        check(L.lib().scalit_pist_vectors_all(self._h, _ptr(psi), self.N, nwin * m), "scalit_pist_vectors_all")
        return psi

    def lan_conv(self, x0: torch.Tensor, e0: float, etol: float, m: int, m_max: int, startcnt: int, stepcnt: int,
                 stepeig: int = 1, ntype: int = 0):
        x0 = x0.contiguous().reshape(-1)
        eig = (C.c_double * m)()
        res = C.c_double()
        r = L.lib().scalit_lan_conv(self._h, e0, etol, ntype, _ptr(x0), startcnt, stepcnt, stepeig, m, m_max, eig,
                                    C.byref(res))
        return r, np.array(eig[:]), res.value

    def eig_gradients(self, psi: torch.Tensor, dvdp: torch.Tensor, noalloc: bool = False):
        """dE_k/dp_j for the nk rows of psi and the np rows of dvdp; with
        noalloc, through the entry point that allocates nothing."""
        psi = psi.contiguous()
        dvdp = dvdp.contiguous()
        nk, np_ = psi.shape[0], dvdp.shape[0]
        out = np.zeros((nk, np_))
        name = "scalit_eig_gradients_noalloc" if noalloc else "scalit_eig_gradients"
        check(getattr(L.lib(), name)(self._h, nk, _ptr(psi), self.N, np_, _ptr(dvdp), self.N, _np_dp(out)), name)
        return out

    def bind(self):
        check(L.lib().scalit_bind(self._h))
