"""The iterate job's parameter file and the PIST job it describes.

`read_iterate_input` restates the reference's reader myReadOSB
(src/iterate/module/osb.print.h:192-255; the MPI myReadMOSB,
mosb.print.h:236-290, reads the same file) with Fortran list-directed input:
items separated by blanks, commas or line ends, r*c repeat counts, T/F (or
.TRUE./.FALSE.) logicals, a D exponent accepted, '/' ending a list early; each
list-directed READ starts on a new line and drops the rest of its last line,
and each file name is one whole line.  The derived types read component by
component as declared in src/iterate/module/a_osbtype.f90:14-29.

`pist_params` restates setPISTConvParam (osbw.pistconv.h:303-333) and
`run_pist_job` the serial driver's bound-state path (src/iterate/prog/osb.drv.f90:
load, block Jacobi or load/save VOSB+EIG0, PIST) on this library.
"""
from __future__ import annotations

import dataclasses
import os
import re

import numpy as np

JOB_BOUND, JOB_RES1, JOB_RES2, JOB_CRP, JOB_CRP1, JOB_CRP2 = 1, 2, 3, 4, 5, 6   # osb.data.h:13-14
TOSB, TOSBD1, TOSBD2, TOSBW = 1, 2, 3, 4                                        # osb.data.h:10
SEQDIRNUM = 100                                                                  # osb.data.h:25


@dataclasses.dataclass
class IterateJob:
    sF: int
    sN: list
    sDEP: list
    sJOB: int
    sOSB: int
    sCX: bool
    sNDVR: bool
    sST: bool
    sAP: bool
    bj_max: int
    bj_tol: float
    qmr_max: int
    qmr_tol: float
    conv_e0: float
    conv_tol: float
    conv_start: int
    conv_step: int
    conv_max: int
    conv_num: int
    conv_gap: int
    osbw_e0: float
    osbw_de: float
    osbw_beta: float
    osbw_cnt: int
    sHOSB: int
    sVOSB: int
    sHW: int
    sVX: int
    sPT: int
    files: dict


class _Records:
    """Fortran formatted sequential input, one record (line) at a time."""

    def __init__(self, text: str):
        self.lines = text.splitlines()
        self.pos = 0

    def line(self) -> str:
        if self.pos >= len(self.lines):
            raise ValueError("iterate input: end of file")
        s = self.lines[self.pos]
        self.pos += 1
        return s

    def items(self, kinds):
        """One list-directed READ of len(kinds) items ('i', 'd' or 'l')."""
        out = [None] * len(kinds)
        k = 0
        while k < len(kinds):
            toks = re.findall(r"'[^']*'|\"[^\"]*\"|[^\s,]+|,", self.line())
            prev_sep = True
            for t in toks:
                if k >= len(kinds):
                    break
                if t == ",":
                    if prev_sep:   # null value: the item keeps its value
                        k += 1
                    prev_sep = True
                    continue
                prev_sep = False
                if t.startswith("/"):
                    return out
                rep, val = 1, t
                m = re.fullmatch(r"(\d+)\*(.*)", t)
                if m:
                    rep, val = int(m.group(1)), m.group(2)
                for _ in range(rep):
                    if k >= len(kinds):
                        break
                    if val != "":
                        out[k] = _convert(val, kinds[k])
                    k += 1
        return out


def _convert(tok: str, kind: str):
    if kind == "i":
        return int(tok)
    if kind == "d":
        return float(tok.replace("d", "e").replace("D", "E"))
    t = tok.upper().lstrip(".")
    if t.startswith("T"):
        return True
    if t.startswith("F"):
        return False
    raise ValueError(f"iterate input: not a logical: {tok!r}")


def read_iterate_input(path_or_text: str) -> IterateJob:
    text = path_or_text
    if "\n" not in path_or_text and os.path.exists(path_or_text):
        with open(path_or_text) as f:
            text = f.read()
    r = _Records(text)
    first = r.line()
    sF = int(re.findall(r"[^\s,]+", first)[0])
    r.pos -= 1
    head = r.items(["i"] * (sF + 1))
    sN = head[1:]
    sDEP = list(r.items(["l"] * sF))
    sDEP[0] = False
    sJOB, sOSB = r.items(["i", "i"])
    sCX, sNDVR, sST, sAP = r.items(["l"] * 4)
    if sNDVR:
        sDEP[sF - 1] = False
    bj_max, bj_tol, qmr_max, qmr_tol = r.items(["i", "d", "i", "d"])
    conv = r.items(["d", "d", "i", "i", "i", "i", "i"])
    osbw = r.items(["d", "d", "d", "i"])
    osbw[1] = abs(osbw[1])
    sHOSB, sVOSB, sHW, sVX, sPT = r.items(["i"] * 5)
    files = {"H0": r.line().rstrip()}
    if sNDVR:
        files["OUTH"] = r.line().rstrip()
    else:
        files["RES"] = r.line().rstrip()
    if any(sDEP):
        files["DEP"] = r.line().rstrip()
    if sJOB in (JOB_RES1, JOB_RES2):
        files["APP"] = r.line().rstrip()
    elif sJOB in (JOB_CRP, JOB_CRP1, JOB_CRP2):
        files["APP"] = r.line().rstrip()
        files["APR"] = r.line().rstrip()
    if sHOSB != 0:
        files["HOSB"] = r.line().rstrip()
    if sVOSB != 0:
        files["VOSB"] = r.line().rstrip()
        files["EIG"] = r.line().rstrip()
    if sHW != 0:
        files["HW"] = r.line().rstrip()
    if sVX != 0:
        files["VX"] = r.line().rstrip()
    if sPT != 0:
        files["PT"] = r.line().rstrip()
    return IterateJob(sF, sN, sDEP, sJOB, sOSB, sCX, sNDVR, sST, sAP, bj_max, bj_tol, qmr_max, qmr_tol, *conv,
                      *osbw, sHOSB, sVOSB, sHW, sVX, sPT, files)


def pist_params(job: IterateJob, m0: int):
    """setPISTConvParam: (E0, ETOL, start, step, stepeig, M_MAX) for M0 levels."""
    start = 2 * m0 if job.conv_start < m0 else job.conv_start
    step = 5 if job.conv_step < 1 else job.conv_step
    stepeig = 1 if job.conv_gap < 1 else job.conv_gap
    nmax = start + m0 if job.conv_max < start else job.conv_max
    return job.conv_e0, job.conv_tol, start, step, stepeig, nmax


def run_pist_job(job: IterateJob, base_dir: str = ".", device: int = 0, seed: int = 1):
    """The job of an iterate input on this library (progPistEig,
    src/iterate/module/osbw.progEig.h:6-80): operator from the job's files
    (complex level matrices when sCX, Context.from_files_cx), preconditioner
    from the set-up (or from the VOSB/EIG files when sVOSB < 0; saved there
    when sVOSB > 0), then

    * sJOB = 1, bound states: PIST for M0 = conv_num levels around conv_e0;
      returns (ret, eigenvalues, res).  With sCX, OSB_PISTCONVCX0
      (mosbp.pistconv.h:77-118): PIST_CONV_CX with H0X = MYHX_CX(TA0) and
      QMR_CX solves of MYHX_CX(TA1) with P = TA1 (here the TMAP0 / TMAP
      forms with AP = 0, the same products); returns complex Ritz values;
    * sJOB = 2 / 3, resonances (OSB_PISTCONVDX1 / OSB_PISTCONVDX2): AP = APP
      from its file, the complex-symmetric (2) or Hermitian-product (3) PIST
      for M0 energies of H - i AP whose real parts are nearest conv_e0, with
      the absorbing term in the preconditioner when sAP; returns (ret,
      complex energies, res);
    * sJOB = 4-6, CRP (osbCRP): APP and APR from their files, N(E) at
      conv_num energies conv_e0 + k conv_tol from conv_gap Lanczos steps
      each; returns (number of energies, N(E) per energy, 0.0).

    Start vectors are uniform on [0, 1) (real and imaginary parts for the
    complex jobs) like the reference's RANDOM_NUMBER / randMat_CX, from a
    seeded generator."""
    import torch

    from . import _lib as L
    from .context import Context

    if job.sJOB not in (JOB_BOUND, JOB_RES1, JOB_RES2, JOB_CRP, JOB_CRP1, JOB_CRP2):
        raise ValueError("sJOB must be bound states (1), resonances (2, 3) or CRP (4-6)")
    path = lambda k: os.path.join(base_dir, job.files[k]) if k in job.files else None
    nosb = {TOSB: L.TOSB, TOSBD1: L.TOSBD1, TOSBD2: L.TOSBD2, TOSBW: L.TOSBW}.get(job.sOSB)
    if nosb is None:
        raise ValueError(f"sOSB = {job.sOSB} is not a preconditioner kernel (1..4)")
    load = Context.from_files_cx if job.sCX else Context.from_files
    ctx = load(job.sN, job.sDEP, h0_file=path("H0"), res_file=path("RES"), dep_file=path("DEP"),
               outh_file=path("OUTH"), device=device)
    ctx.set_shift(job.osbw_e0, job.osbw_de, job.osbw_beta)
    mode = "dir" if abs(job.sVOSB) <= SEQDIRNUM else "seq"
    if job.sVOSB < 0:
        ctx.load_precond(path("VOSB"), path("EIG"), mode)
    else:
        ctx.block_diag(job.bj_max, job.bj_tol)
        if job.sVOSB > 0:
            ctx.save_precond(path("VOSB"), path("EIG"), mode)
    m0 = job.conv_num
    rng = np.random.default_rng(seed)

    def window():
        # initOSBW / initMOSBW with the job's OSBW line before the solver (every
        # OSB_PISTCONV* and osbCRP does); a window of no states fails the job
        if job.osbw_cnt < 1:
            return False
        ctx.set_osbw(job.osbw_e0, job.osbw_de, job.osbw_cnt, L.TA1)
        return True
    if job.sJOB in (JOB_RES1, JOB_RES2):
        # osb.io.h:20-23: APP from its file, APR = 0, AP = APP
        app = Context.load_file(path("APP"), ctx.N, "dir")
        ctx.set_ap(L.AP, app)
        ctx.set_ap(L.APP, app)
        ctx.set_ap(L.APR, np.zeros(ctx.N))
        if nosb == L.TOSBW and not window():
            ctx.close()
            return 0, np.zeros(m0, complex), 0.0
        ctx.set_osb(nosb)
        e0, etol, start, step, stepeig, nmax = pist_params(job, m0)
        x0 = rng.random(ctx.N) + 1j * rng.random(ctx.N)
        r, eig, res, _ = ctx.pist_conv_sx(ctx.to_device_cx(x0), e0, etol, m0, nmax, start, step, stepeig,
                                          qmrmax=job.qmr_max, qmrtol=job.qmr_tol, pc=L.TMAP if job.sAP else L.TA1,
                                          hermitian=job.sJOB == JOB_RES2)
        torch.cuda.synchronize()
        ctx.close()
        return r, eig, res
    if job.sJOB in (JOB_CRP, JOB_CRP1, JOB_CRP2):
        ctx.set_ap(L.APP, Context.load_file(path("APP"), ctx.N, "dir"))
        ctx.set_ap(L.APR, Context.load_file(path("APR"), ctx.N, "dir"))
        if nosb == L.TOSBW and not window():
            ctx.close()
            return 0, np.zeros(m0), 0.0
        ctx.set_osb(nosb)
        energies = [job.conv_e0 + k * job.conv_tol for k in range(m0)]   # osbw.progEig.h:66-68
        x0 = rng.random((m0, ctx.N)) + 1j * rng.random((m0, ctx.N))
        crp, _ = ctx.crp(energies, job.conv_gap, ctx.to_device_cx(x0), qmrmax=job.qmr_max, qmrtol=job.qmr_tol,
                         sap=bool(job.sAP))
        torch.cuda.synchronize()
        ctx.close()
        return m0, crp, 0.0
    if job.sCX:
        # OSB_PISTCONVCX0: sHC = sPC = TA1, no absorbing potential
        for k in (L.AP, L.APP, L.APR):
            ctx.set_ap(k, np.zeros(ctx.N))
        ctx.set_osb(nosb)
        e0, etol, start, step, stepeig, nmax = pist_params(job, m0)
        x0 = rng.random(ctx.N) + 1j * rng.random(ctx.N)
        r, eig, res, _ = ctx.pist_conv_sx(ctx.to_device_cx(x0), e0, etol, m0, nmax, start, step, stepeig,
                                          qmrmax=job.qmr_max, qmrtol=job.qmr_tol, pc=L.TA1, hermitian=True)
        torch.cuda.synchronize()
        ctx.close()
        return r, eig, res
    if nosb == L.TOSBW:
        # OSB_PISTCONV (osbw.pistconv.h:26-31): sPC = TA1, then initOSBW with
        # the job's window; a window of no states fails the job (returns 0)
        if not window():
            ctx.close()
            return 0, np.zeros(m0), 0.0
    ctx.set_osb(nosb)
    e0, etol, start, step, stepeig, nmax = pist_params(job, m0)
    x0 = rng.random((1, ctx.N))
    r, eig, res, _ = ctx.pist_conv(ctx.to_device(x0), e0, etol, m0, nmax, start, step, stepeig,
                                   qmrmax=job.qmr_max, qmrtol=job.qmr_tol, prec=L.FP64)
    torch.cuda.synchronize()
    ctx.close()
    return r, eig, res
