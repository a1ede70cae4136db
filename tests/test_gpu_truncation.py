"""Sequential truncation on the GPU through the C ABI: the PO-DVR and the
energy-cut combined layer against the oracle's restatement of DVR_Step2/3
and saveH0CBALL, the truncated operator (combined (r, R) layer of hundreds of
points as level 1, theta as level 2) through H*X, the OSB set-up, P*X and
PIST against the oracle and dense eigensolves, and the mixed-precision
workload: PIST for 100 levels near a seeded energy on the truncated
128 x 128 x 96 triatomic, every level's error against the FP64 solve.

Tolerances: PO-DVR points 1e-12 (relative to the grid), PO-DVR Hamiltonian
1e-11 (relative to its largest entry); H*X and P*X FP64 1e-12, P*X FP32
1e-4; FP64 eigenvalues 1e-10; mixed precision 1e-8.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from scalit_amd import _lib as L
from scalit_amd import problems as P
from scalit_amd.context import Context

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("n,m,coord", [(128, 40, "r"), (64, 20, "R"), (24, 24, "r")])
def test_podvr_matches_oracle(n, m, coord):
    print(f"podvr n={n} m={m} {coord}", flush=True)
    rng_, morse = (P.R_RANGE, P.MORSE_R) if coord == "r" else (P.BIGR_RANGE, P.MORSE_BIGR)
    x, t = P.sinc2_dvr(n, *rng_, P.MU)
    v = P.morse_1d(x, *morse)
    podvr, _ = P.library_truncation()
    x0, h0, e = podvr(x, t, v, m)
    xo, ho, eo = O.podvr(x, t, v, m)
    assert np.abs(x0 - xo).max() <= 1e-12 * np.abs(x).max()
    assert rel(h0, ho) <= 1e-11
    assert np.abs(e - eo).max() <= 1e-12 * np.abs(eo).max()


def test_combine_2d_matches_oracle():
    x1, t1 = P.sinc2_dvr(40, *P.R_RANGE, P.MU)
    x2, t2 = P.sinc2_dvr(48, *P.BIGR_RANGE, P.MU)
    a1, h1, _ = O.podvr(x1, t1, P.morse_1d(x1, *P.MORSE_R), 16)
    a2, h2, _ = O.podvr(x2, t2, P.morse_1d(x2, *P.MORSE_BIGR), 20)
    v1, v2 = P.morse_1d(a1, *P.MORSE_R), P.morse_1d(a2, *P.MORSE_BIGR)
    _, combine = P.library_truncation()
    for nk in (1, 100, 320):
        i1, i2, hc = combine(h1, v1, h2, v2, nk)
        j1, j2, hco = O.combine_2d(h1, v1, h2, v2, nk)
        assert np.array_equal(i1, j1) and np.array_equal(i2, j2) and np.array_equal(hc, hco)


def _truncated(grid, m12, nkeep, seed):
    pot = P.gaussian_potential(seed)
    tp = P.truncated_problem(*grid, pot, *m12, nkeep, *P.library_truncation(), seed=seed)
    to = P.truncated_problem(*grid, pot, *m12, nkeep, O.podvr, O.combine_2d, seed=seed)
    return tp, to


def test_truncated_operator_small():
    """40 x 40 x 8 contracted to 16 x 16 PO-DVR points, 160 kept: level-1
    blocks of 160 take the dense OSB set-up and the two-pass P*X."""
    tp, to = _truncated((40, 40, 8), (16, 16), 160, 3)
    assert rel(tp.h0[0], to.h0[0]) <= 1e-11 and rel(tp.res, to.res) <= 1e-11
    op = O.Operator(tp.sN, tp.h0, tp.res, tp.dep)
    ctx = Context.from_problem(tp)
    x = P.random_vectors(tp.N, 3, 5)
    y = ctx.hx(ctx.to_device(x)).cpu().numpy()
    for b in range(3):
        assert rel(y[b], op.hx(x[b])) <= 1e-12
    # OSB set-up: every level-1 block diagonalised, V orthogonal
    ctx.set_shift(-0.005, 0.002, 0.3)
    ctx.block_diag(1000, 1e-8)
    vosb, eig0 = ctx.get_precond()
    n1, n3 = tp.sN
    # level 2 (theta) was diagonalised first: EIG0 seen by level 1 is its result
    for b in (0, n3 // 2, n3 - 1):
        v = vosb[0][b * n1 * n1:(b + 1) * n1 * n1].reshape(n1, n1, order="F")
        assert np.abs(v.T @ v - np.eye(n1)).max() <= 1e-12
    op.set_precond(vosb, eig0)
    op.set_shift(-0.005, 0.002, 0.3)
    for kind in (O.TOSB, O.TOSBD2):
        yp = ctx.px(ctx.to_device(x), nosb=kind).cpu().numpy()
        for b in range(3):
            assert rel(yp[b], op.px(x[b], kind=kind)) <= 1e-12, kind
    y32 = ctx.px_lp(ctx.to_device(x, torch.float32)).double().cpu().numpy()
    assert rel(y32[0], op.px(x[0])) <= 1e-4
    # PIST: the lowest 8 levels against a dense eigensolve of the same operator
    ev = np.linalg.eigvalsh(op.dense())[:8]
    e0 = float(ev[0]) - 0.001
    ctx.set_shift(e0)
    ctx.block_diag(1000, 1e-8)
    x0 = ctx.to_device(P.random_vectors(tp.N, 1, 7))
    r, eig, res, stats = ctx.pist_conv(x0, e0, 1e-12, 8, 200, 30, 10, 4, qmrtol=1e-9)
    print("truncated PIST", r, stats, flush=True)
    assert r > 0 and (np.abs(eig - ev) / np.abs(ev)).max() <= 1e-10
    r, eig, res, stats = ctx.pist_conv(x0, e0, 1e-11, 8, 200, 30, 10, 4, qmrtol=1e-6, prec=L.FP32)
    print("truncated PIST fp32", r, stats, flush=True)
    assert r > 0 and (np.abs(eig - ev) / np.abs(ev)).max() <= 1e-8
    ctx.close()


def test_config3_mixed_precision_100_levels():
    """128 x 128 x 96 (1.57M primitive points) contracted to 40 x 40 PO-DVR
    points in r and R, 400 (r, R) points kept per theta: N = 38,400.  PIST for
    the 100 levels nearest a seeded energy in FP64 and with FP32 inner solves
    (FP64 refinement); every level of the FP32-inner solve within 1e-8 of the
    FP64 one.  The BF16-stored inner solve is run and its outcome printed."""
    tp = P.truncated_problem(128, 128, 96, P.gaussian_potential(1234), 40, 40, 400, *P.library_truncation(),
                             seed=1234)
    assert tp.N == 38_400
    ctx = Context.from_problem(tp)
    x0 = ctx.to_device(P.random_vectors(tp.N, 1, 11))
    # the lowest level, then a seeded energy above it
    ctx.set_shift(float(tp.res.min()) - 0.05)
    ctx.block_diag(1000, 1e-8)
    r1, e1, _, _ = ctx.pist_conv(x0, float(tp.res.min()) - 0.05, 1e-10, 1, 200, 40, 10, 5, qmrtol=1e-8)
    assert r1 > 0
    ec = float(e1[0]) + 0.004 + 0.002 * float(P.rng(1234).uniform())
    ctx.set_shift(ec)
    ctx.block_diag(1000, 1e-8)
    ctx.set_pist_orth(2)   # a Krylov space of hundreds of vectors
    args = (ec, 1e-10, 100, 800, 200, 50, 20)
    r64, ev64, res64, st64 = ctx.pist_conv(x0, *args, qmrtol=1e-6)
    print("config3 fp64", r64, res64, st64, ev64[[0, -1]], flush=True)
    assert r64 > 0
    psi = ctx.pist_vectors(0, 100)
    rr = ctx.hx(psi) - torch.as_tensor(ev64, device=psi.device)[:, None] * psi
    rn = (rr * rr).sum(dim=1).sqrt().cpu().numpy()
    print(f"config3 fp64 Ritz residuals max {rn.max():.1e}", flush=True)
    assert rn.max() <= 1e-8
    r32, ev32, res32, st32 = ctx.pist_conv(x0, *args, qmrtol=1e-6, prec=L.FP32)
    err = np.abs(ev32 - ev64) / np.abs(ev64)
    print("config3 fp32-inner", r32, res32, st32, flush=True)
    print("per-level relative error, FP32 inner vs FP64:", " ".join(f"{v:.1e}" for v in err), flush=True)
    assert r32 > 0 and err.max() <= 1e-8
    r16, ev16, res16, st16 = ctx.pist_conv(x0, *args, qmrtol=1e-6, prec=L.BF16)
    e16 = np.abs(ev16 - ev64) / np.abs(ev64) if r16 > 0 else None
    print("config3 bf16-stored inner", r16, res16, st16, None if e16 is None else f"max err {e16.max():.1e}",
          flush=True)
    ctx.close()
