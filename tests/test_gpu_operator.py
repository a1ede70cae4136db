"""GPU parity of the operator kernels against the CPU oracle, through the C ABI.

Tolerances: FP64 H*X and P*X within 1e-12 relative (max-norm, relative to the
oracle's max-norm); FP32 within 2e-6; BF16-stored within 2e-2."""
import dataclasses

import numpy as np
import pytest
import torch

from oracle import oracle as O
from scalit_amd import _lib as L
from scalit_amd import problems as P
from scalit_amd.context import Context

pytestmark = pytest.mark.gpu

GRIDS = [(24, 32, 24), (64, 96, 80), (48, 64, 40), (80, 96, 64), (128, 128, 96), (7, 9, 5), (17, 33, 19)]


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def setup(grid, seed=5, e0=0.0):
    prob = P.build_problem(*grid, P.gaussian_potential(seed), seed=seed)
    op = O.Operator(prob.sN, prob.h0, prob.res, prob.dep, e0=e0)
    ctx = Context.from_problem(prob, e0=e0)
    return prob, op, ctx


@pytest.mark.parametrize("grid", GRIDS)
def test_hx_fp64(grid):
    print(f"hx_fp64 grid={grid}", flush=True)
    prob, op, ctx = setup(grid, e0=-0.013)
    nv = 3 if prob.N < 600_000 else 1
    x = P.random_vectors(prob.N, nv, 7)
    xd = ctx.to_device(x)
    y = ctx.hx(xd, L.TA0).cpu().numpy()
    for b in range(nv):
        assert rel(y[b], op.hx(x[b], O.TA0)) <= 1e-12
    y1 = ctx.hx(xd, L.TA1).cpu().numpy()
    assert rel(y1[0], op.hx(x[0], O.TA1)) <= 1e-12
    ctx.close()


def test_hx_per_vector_shifts_and_generic_path():
    prob, op, ctx = setup((24, 32, 24))
    x = P.random_vectors(prob.N, 4, 3)
    xd = ctx.to_device(x)
    shifts = [0.0, 0.01, -0.02, 0.005]
    y = ctx.hx(xd, L.TA1, shifts=shifts).cpu().numpy()
    for b, s in enumerate(shifts):
        op.set_e0(s)
        assert rel(y[b], op.hx(x[b], O.TA1)) <= 1e-12
    L.lib().scalit_debug_force_generic(1)
    try:
        yg = ctx.hx(xd, L.TA0).cpu().numpy()
    finally:
        L.lib().scalit_debug_force_generic(0)
    op.set_e0(0.0)
    assert rel(yg[1], op.hx(x[1], O.TA0)) <= 1e-12
    ctx.close()


@pytest.mark.parametrize("levels", [(9,), (6, 11), (5, 4, 6, 3)])
def test_hx_other_level_counts(levels):
    # any number of levels: the generic path (dep on the last level when > 1 level)
    g = np.random.default_rng(1)
    mats = [(lambda a: a + a.T)(g.standard_normal((n, n))) for n in levels]
    n = int(np.prod(levels))
    dep = [None] * len(levels)
    if len(levels) > 1:
        dep[-1] = g.uniform(0.5, 1.5, int(np.prod(levels[:-1])))
    res = g.standard_normal(n)
    op = O.Operator(levels, mats, res, dep)
    ctx = Context(levels, [d is not None for d in dep])
    for l, m in enumerate(mats):
        ctx.set_h0(l + 1, m)
        if dep[l] is not None:
            ctx.set_dep(l + 1, dep[l])
    ctx.set_res(res)
    x = g.standard_normal((2, n))
    y = ctx.hx(ctx.to_device(x)).cpu().numpy()
    assert rel(y[0], op.hx(x[0])) <= 1e-12 and rel(y[1], op.hx(x[1])) <= 1e-12
    ctx.close()


@pytest.mark.parametrize("grid", [(24, 32, 24), (64, 96, 80)])
def test_hx_low_precision(grid):
    prob, op, ctx = setup(grid)
    x = P.random_vectors(prob.N, 2, 9)
    ref = op.hx(x[0])
    y32 = ctx.hx_lp(ctx.to_device(x, torch.float32)).double().cpu().numpy()
    assert rel(y32[0], ref) <= 2e-6
    xb = ctx.to_device(x, torch.bfloat16)
    y16 = ctx.hx_lp(xb).double().cpu().numpy()
    ref16 = op.hx(xb[0].double().cpu().numpy())
    assert rel(y16[0], ref16) <= 2e-2
    ctx.close()


@pytest.mark.parametrize("grid", [(24, 32, 24), (64, 96, 80)])
def test_hx_adjoint(grid):
    prob, op, ctx = setup(grid)
    x = P.random_vectors(prob.N, 2, 4)
    g = P.random_vectors(prob.N, 2, 5)
    dx, dv = ctx.hx_adjoint(ctx.to_device(x), ctx.to_device(g))
    dx = dx.cpu().numpy()
    for b in range(2):
        assert rel(dx[b], op.hx_t(g[b])) <= 1e-12
    assert rel(dv.cpu().numpy(), (x * g).sum(axis=0)) <= 1e-13
    # <g, H x> = <H^T g, x>
    hx = ctx.hx(ctx.to_device(x)).cpu().numpy()
    assert abs((g[0] * hx[0]).sum() - (dx[0] * x[0]).sum()) <= 1e-11 * abs((g[0] * hx[0]).sum())
    ctx.close()


@pytest.mark.parametrize("grid", [(24, 32, 24), (64, 96, 80), (17, 33, 19)])
def test_px_with_oracle_preconditioner(grid):
    print(f"px grid={grid}", flush=True)
    prob, op, ctx = setup(grid, e0=-0.02)
    if prob.N < 100_000:
        vosb, eig0 = op.diag(bjmax=200, bjtol=1e-5)
        ctx.set_precond(vosb, eig0)
    else:
        # large grid: the GPU's own OSB set-up, handed to the oracle
        ctx.block_diag(200, 1e-5)
        vosb, eig0 = ctx.get_precond()
        op.set_precond(vosb, eig0)
    nv = 2
    x = P.random_vectors(prob.N, nv, 8)
    xd = ctx.to_device(x)
    for kind in (O.TOSB, O.TOSBD1, O.TOSBD2):
        op.s.DE = 0.004
        op.s.beta = 0.3
        ctx.set_shift(-0.02, 0.004, 0.3)
        y = ctx.px(xd, nosb=kind).cpu().numpy()
        for b in range(nv):
            assert rel(y[b], op.px(x[b], kind=kind)) <= 1e-12, kind
    y32 = ctx.px_lp(ctx.to_device(x, torch.float32)).double().cpu().numpy()
    assert rel(y32[0], op.px(x[0])) <= 1e-4
    ctx.close()


@pytest.mark.parametrize("grid,nv", [((24, 32, 24), 19), ((17, 33, 19), 8), ((64, 96, 80), 16)])
def test_px_batch_per_vector_shifts(grid, nv):
    """Batches of 8 and more take the MFMA level-1 path (16 vectors per chunk;
    19 = one full chunk and a partial one); each vector has its own E0."""
    print(f"px batch grid={grid} nv={nv}", flush=True)
    prob, op, ctx = setup(grid, e0=-0.02)
    if prob.N < 100_000:
        vosb, eig0 = op.diag(bjmax=200, bjtol=1e-5)
        ctx.set_precond(vosb, eig0)
    else:
        ctx.block_diag(200, 1e-5)
        vosb, eig0 = ctx.get_precond()
        op.set_precond(vosb, eig0)
    x = P.random_vectors(prob.N, nv, 12)
    shifts = -0.02 + 0.003 * np.arange(nv)
    xd = ctx.to_device(x)
    for kind in (O.TOSB, O.TOSBD2):
        op.s.DE = 0.004
        op.s.beta = 0.3
        ctx.set_shift(-0.02, 0.004, 0.3)
        y = ctx.px(xd, nosb=kind, shifts=shifts).cpu().numpy()
        for b in range(nv):
            op.s.E0 = shifts[b]
            assert rel(y[b], op.px(x[b], kind=kind)) <= 1e-12, (kind, b)
    ctx.close()


@pytest.mark.parametrize("grid", [(24, 32, 24), (12, 10, 8), (32, 40, 24)])
def test_block_diag_matches_oracle(grid):
    print(f"block_diag grid={grid}", flush=True)
    prob, op, ctx = setup(grid, e0=0.0)
    vosb_o, eig0_o = op.diag(bjmax=1000, bjtol=1e-5)
    ctx.block_diag(1000, 1e-5)
    vosb_g, eig0_g = ctx.get_precond()
    # same rotation sequence, rounding-level differences only
    assert rel(eig0_g, eig0_o) <= 1e-9
    for l in range(3):
        assert np.abs(vosb_g[l] - vosb_o[l]).max() <= 1e-7
        n = prob.sN[l]
        v = vosb_g[l].reshape(-1, n, n)
        for b in range(0, v.shape[0], max(1, v.shape[0] // 7)):
            assert np.allclose(v[b] @ v[b].T, np.eye(n), atol=1e-12)
    ctx.close()


# ---------------------------------------------------------------- sNDVR = T
def setup_ndvr(grid, seed=5, e0=0.0):
    prob = P.build_problem(*grid, P.gaussian_potential(seed), seed=seed)
    outh = P.outer_layer(prob, seed + 100)
    op = O.Operator(prob.sN, prob.h0, prob.res, prob.dep, e0=e0, outh=outh)
    ctx = Context.from_problem(prob, e0=e0, outh=outh)
    return prob, op, ctx


@pytest.mark.parametrize("grid", [(24, 32, 24), (13, 11, 9), (64, 96, 80)])
def test_ndvr_hx(grid):
    """The outermost level as the dense per-point layer OUTH (H3X,
    matrix.f90:188-203 via osb.hx.h:93-94): fused and generic paths, shifts,
    the adjoint, and FP32."""
    print(f"ndvr hx grid={grid}", flush=True)
    prob, op, ctx = setup_ndvr(grid, e0=-0.011)
    nv = 5 if prob.N < 100_000 else 2
    x = P.random_vectors(prob.N, nv, 8)
    xd = ctx.to_device(x)
    y = ctx.hx(xd, L.TA0).cpu().numpy()
    y1 = ctx.hx(xd, L.TA1).cpu().numpy()
    for b in range(nv):
        assert rel(y[b], op.hx(x[b], O.TA0)) <= 1e-12
    assert rel(y1[nv - 1], op.hx(x[nv - 1], O.TA1)) <= 1e-12
    L.lib().scalit_debug_force_generic(1)
    try:
        yg = ctx.hx(xd, L.TA0).cpu().numpy()
    finally:
        L.lib().scalit_debug_force_generic(0)
    assert rel(yg[1], op.hx(x[1], O.TA0)) <= 1e-12
    g = P.random_vectors(prob.N, nv, 9)
    dx, dv = ctx.hx_adjoint(xd, ctx.to_device(g))
    dx = dx.cpu().numpy()
    assert rel(dx[0], op.hx_t(g[0], O.TA0)) <= 1e-12
    assert rel(dv.cpu().numpy(), (x * g).sum(axis=0)) <= 1e-12
    yf = ctx.hx_lp(ctx.to_device(x, torch.float32)).double().cpu().numpy()
    assert rel(yf[0], op.hx(x[0], O.TA0)) <= 2e-6
    ctx.close()


def test_ndvr_block_diag_px_qmr():
    """OSB set-up with HOSB(sF) = OUTH (osb.diag.h:22-23), then P*X and a QMR
    solve on the sNDVR operator against the oracle.  The potential sits on the
    diagonal of the OUTH blocks and RES = 0, the form in which the outer
    preconditioner block carries it."""
    prob = P.build_problem(12, 10, 8, P.gaussian_potential(5), seed=5)
    outh = P.outer_layer(prob, 105, potential=True)
    zero = np.zeros(prob.N)
    op = O.Operator(prob.sN, prob.h0, zero, prob.dep, outh=outh)
    ctx = Context.from_problem(dataclasses.replace(prob, res=zero), outh=outh)
    vosb_o, eig0_o = op.diag(bjmax=1000, bjtol=1e-5)
    ctx.block_diag(1000, 1e-5)
    vosb_g, eig0_g = ctx.get_precond()
    assert rel(eig0_g, eig0_o) <= 1e-9
    for l in range(3):
        assert np.abs(vosb_g[l] - vosb_o[l]).max() <= 1e-7
    e0 = float(np.sort(eig0_o)[3]) - 1e-3
    op.set_e0(e0)
    ctx.set_shift(e0)
    ctx.set_precond(vosb_o, eig0_o)   # the oracle's preconditioner: P*X parity at 1e-12
    x = P.random_vectors(prob.N, 2, 10)
    yp = ctx.px(ctx.to_device(x)).cpu().numpy()
    for b in range(2):
        assert rel(yp[b], op.px(x[b])) <= 1e-12
    b = P.random_vectors(prob.N, 1, 11)
    ret, xs, res = ctx.qmr(ctx.to_device(b), ntype=-1, qmrmax=300, qmrtol=1e-9)
    assert ret[0] > 0
    xs = xs.cpu().numpy()[0]
    r = op.hx(xs, O.TA1) - b[0]
    assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(op.hx(xs, O.TA1))
    ctx.close()


def _precond_from_gpu(ctx, op):
    ctx.block_diag(200, 1e-5)
    vosb, eig0 = ctx.get_precond()
    op.set_precond(vosb, eig0)


@pytest.mark.parametrize("grid,nv", [((80, 96, 64), 1), ((80, 96, 64), 16), ((128, 128, 96), 1)])
def test_px_large_level1_blocks(grid, nv):
    """Level-1 rotations larger than 64: the BIG forms of k_px_l1, on the MFMA
    path for n1 = 80 at batch 16 and on the VALU path for n1 = 80 at batch 1
    and n1 = 128, TOSB and TOSBD2, FP64 (1e-12) and FP32 (1e-4)."""
    print(f"px large blocks grid={grid} nv={nv}", flush=True)
    prob, op, ctx = setup(grid, e0=-0.02)
    _precond_from_gpu(ctx, op)
    x = P.random_vectors(prob.N, nv, 14)
    xd = ctx.to_device(x)
    check = sorted({0, nv // 2, nv - 1})
    for kind in (O.TOSB, O.TOSBD2):
        op.set_shift(-0.02, 0.004, 0.3)
        ctx.set_shift(-0.02, 0.004, 0.3)
        y = ctx.px(xd, nosb=kind).cpu().numpy()
        for b in check:
            assert rel(y[b], op.px(x[b], kind=kind)) <= 1e-12, (kind, b)
    ref = [op.px(x[b]) for b in check]
    y32 = ctx.px_lp(ctx.to_device(x, torch.float32)).double().cpu().numpy()
    for i, b in enumerate(check):
        assert rel(y32[b], ref[i]) <= 1e-4, b
    ctx.close()


def test_hx_bench_shape_every_variant(monkeypatch):
    """The bench workload's exact shape, 64 x 96 x 80 at batch 16, through
    every pass-A block shape and pass-B width the autotuner can pick (forced
    one pair at a time), FP64 within 1e-12 and FP32 within 2e-6 of the oracle
    on the first, a middle and the last vector of the batch."""
    grid, nv = (64, 96, 80), 16
    prob = P.build_problem(*grid, P.gaussian_potential(5), seed=5)
    op = O.Operator(prob.sN, prob.h0, prob.res, prob.dep)
    x = P.random_vectors(prob.N, nv, 21)
    check = (0, 7, 15)
    ref = {b: op.hx(x[b]) for b in check}
    for va in range(8):
        for vb in range(3):
            print(f"hx variant A={va} B={vb}", flush=True)
            monkeypatch.setenv("SCALIT_HX_VARIANT_A", str(va))
            monkeypatch.setenv("SCALIT_HX_VARIANT_B", str(vb))
            ctx = Context.from_problem(prob)
            y = ctx.hx(ctx.to_device(x)).cpu().numpy()
            got = ctx.hx_variant(L.FP64, nv)
            assert got == (va, vb), got
            for b in check:
                assert rel(y[b], ref[b]) <= 1e-12, (va, vb, b)
            y32 = ctx.hx_lp(ctx.to_device(x, torch.float32)).double().cpu().numpy()
            for b in check:
                assert rel(y32[b], ref[b]) <= 2e-6, (va, vb, b)
            ctx.close()
