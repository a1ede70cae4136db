"""Writes the golden fixtures under tests/golden/ (run from the repo root).

    python tests/golden/make_golden.py [ref_small] [pist_config1]   (default: both)

* ref_small.npz -- input/output vectors of the reference's own compiled
  routines (oracle/_ref, built from the reference sources by
  `make -C oracle ref`) on a 6 x 7 x 5 operator: H*X (TA0, TA1) and P*X
  (TOSB, and TOSBD1 / TOSBD2 with a window DE around E0) from module osb's
  own MYHX / MYPX, the OSB set-up (VOSB, EIG0), and a QMR solve with its
  return code.
* pist_fp64_config1.json -- the lowest 20 levels of the 24 x 32 x 24
  jittered Morse+P2 triatomic (seed 1234): the oracle's PIST_CONV at tight
  tolerances and an independent scipy Lanczos (eigsh) of the same operator.
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from scalit_amd import problems as P  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def ref_small():
    prob = P.build_problem(6, 7, 5, P.gaussian_potential(3), name="small")
    e0 = -0.004
    op = O.Operator(prob.sN, prob.h0, prob.res, prob.dep, e0=e0)
    rop = O.RefOperator(op)
    x = P.random_vectors(op.N, 1, 11)[0]
    hx0 = rop.hx(x, shifted=False)
    hx1 = rop.hx(x, shifted=True)
    vosb, eig0 = rop.diag(1000, 1e-5)
    px = rop.px(x)
    b = P.random_vectors(op.N, 1, 12)[0]
    ret, xq, res = rop.qmr(b, qmr_type=-1, qmr_max=200, qmr_tol=1e-9)
    # the OSBD kernels: DE reaches the 31st eigenvalue of EIG0 above E0
    srt = np.sort(eig0)
    de = float(srt[30] - e0)
    beta = 0.3
    rop.set_shift(e0, de, beta)
    px_d1 = rop.px(x, kind=O.TOSBD1, npc=O.TA1)
    px_d2 = rop.px(x, kind=O.TOSBD2, npc=O.TA1)
    np.savez_compressed(
        os.path.join(HERE, "ref_small.npz"), grid=np.array(prob.sN), seed=3, e0=e0, x=x, hx0=hx0, hx1=hx1,
        vosb1=vosb[0], vosb2=vosb[1], vosb3=vosb[2], eig0=eig0, px=px, qmr_b=b, qmr_x=xq, qmr_ret=ret, qmr_res=res,
        de=de, beta=beta, px_d1=px_d1, px_d2=px_d2)


def pist_config1():
    import scipy.sparse.linalg as sla

    prob = P.config_problem("pist_fp64")
    e0 = float(prob.res.min())
    op = O.Operator(prob.sN, prob.h0, prob.res, prob.dep, e0=e0)
    op.diag(1000, 1e-5)
    x = P.random_vectors(op.N, 1, 77)[0]
    r, eig, res, ctx = op.pist_conv(x, e0, 1e-11, 20, m_max=200, start=40, step=10, stepeig=5, qmr_tol=1e-6)
    assert r > 0
    a = sla.LinearOperator((op.N, op.N), matvec=lambda v: op.hx(v), dtype=np.float64)
    w = np.sort(sla.eigsh(a, k=25, which="SA", tol=1e-13, ncv=80)[0])
    out = dict(grid=list(prob.sN), seed=prob.seed, start_seed=77, e0=e0, m=20, etol=1e-11, qmr_tol=1e-6, m_max=200,
               startcnt=40, stepcnt=10, stepeig=5, pist_ret=int(r), pist_eigs=[float(v) for v in eig],
               pist_qmr_iters=int(ctx.qmr_iters), pist_solves=int(ctx.solves), eigsh=[float(v) for v in w[:20]])
    with open(os.path.join(HERE, "pist_fp64_config1.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    which = sys.argv[1:] or ["ref_small", "pist_config1"]
    if "ref_small" in which:
        if not O.ref_available():
            raise SystemExit("ref_small needs the reference checker library (make -C oracle ref)")
        ref_small()
    if "pist_config1" in which:
        pist_config1()
    print("golden fixtures written")
