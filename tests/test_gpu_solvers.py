"""GPU parity of QMR / PIST / Lanczos against the oracle and the golden
eigenvalues, through the C ABI.

Eigenvalue tolerance: 1e-10 relative (FP64 PIST and Lanczos against the
fixture), 1e-8 relative for mixed precision.  QMR: the same return code as
the oracle's restatement of qmr.f90 (+-1 where the last residual sits within
rounding of the tolerance) and solutions that agree to the solve tolerance."""
import ctypes as C
import json
import os

import numpy as np
import pytest
import torch

from oracle import oracle as O
from scalit_amd import _lib as L
from scalit_amd import problems as P
from scalit_amd.context import Context

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "pist_fp64_config1.json")


def gold():
    with open(GOLD) as f:
        return json.load(f)


def op_ctx(grid, e0, seed=1234, kind="jitter"):
    pot = P.jittered_potential(seed) if kind == "jitter" else P.gaussian_potential(seed)
    prob = P.build_problem(*grid, pot, seed=seed)
    op = O.Operator(prob.sN, prob.h0, prob.res, prob.dep, e0=e0)
    ctx = Context.from_problem(prob, e0=e0)
    return prob, op, ctx


@pytest.mark.parametrize("persist", [1, 0])
@pytest.mark.parametrize("qmr_type", [-1, 0, -2])
def test_qmr_matches_oracle(qmr_type, persist):
    """QMR at 24 x 32 x 24 through the persistent solver (one launch per
    solve) and through one kernel per step of the iteration."""
    print(f"qmr type={qmr_type} persistent={persist}", flush=True)
    prob, op, ctx = op_ctx((24, 32, 24), e0=-0.003)
    ctx.set_qmr_persistent(persist)
    vosb, eig0 = op.diag()
    ctx.set_precond(vosb, eig0)
    b = P.random_vectors(prob.N, 1, 31)
    r_o, x_o, res_o, _ = op.qmr(b[0], qmr_type=qmr_type, qmr_max=300, qmr_tol=1e-9)
    ret, x, res = ctx.qmr(ctx.to_device(b), ntype=qmr_type, qmrmax=300, qmrtol=1e-9)
    assert r_o > 0 and abs(int(ret[0]) - r_o) <= 1
    x = x.cpu().numpy()[0]
    assert np.abs(x - x_o).max() <= 1e-7 * np.abs(x_o).max()
    h = op.hx(x, O.TA1)
    assert np.linalg.norm(h - b[0]) <= 1e-8 * np.linalg.norm(h)
    ctx.close()


@pytest.mark.parametrize("persist", [1, 0])
def test_qmr_batched_shifts_and_breakdown(persist):
    print(f"qmr batched persistent={persist}", flush=True)
    prob, op, ctx = op_ctx((24, 32, 24), e0=-0.003)
    ctx.set_qmr_persistent(persist)
    vosb, eig0 = op.diag()
    ctx.set_precond(vosb, eig0)
    b = P.random_vectors(prob.N, 4, 32)
    b[2] = 0.0                       # zero right-hand side: QMR returns 0 at once
    shifts = [-0.003, -0.003, -0.003, -0.001]
    ret, x, res = ctx.qmr(ctx.to_device(b), ntype=-1, qmrmax=300, qmrtol=1e-8, shifts=shifts)
    assert ret[2] == 0
    x = x.cpu().numpy()
    for v in (0, 1, 3):
        op.set_e0(shifts[v])
        r_o, x_o, _, _ = op.qmr(b[v], qmr_type=-1, qmr_max=300, qmr_tol=1e-8)
        assert abs(int(ret[v]) - r_o) <= 1
        assert np.abs(x[v] - x_o).max() <= 1e-6 * np.abs(x_o).max()
    # iteration limit: -QMRMAX
    ret, _, _ = ctx.qmr(ctx.to_device(b[:1]), ntype=-1, qmrmax=2, qmrtol=1e-14)
    assert ret[0] == -2
    ctx.close()


@pytest.mark.parametrize("nosb", [O.TOSBD1, O.TOSBD2])
@pytest.mark.parametrize("persist", [1, 0])
def test_qmr_osbd_kernels(nosb, persist):
    """sOSB = TOSBD1 / TOSBD2: QMR's preconditioner divides by the OSBD
    kernel (window DE, beta) on both solver paths, against the oracle's QMR
    with the same kernel."""
    print(f"qmr osbd kind={nosb} persistent={persist}", flush=True)
    prob, op, ctx = op_ctx((24, 32, 24), e0=0.002)
    vosb, eig0 = op.diag()
    ctx.set_precond(vosb, eig0)
    ctx.set_qmr_persistent(persist)
    ctx.set_osb(nosb)
    ctx.set_shift(0.002, 0.0005, 0.3)
    op.set_shift(0.002, 0.0005, 0.3)
    op.s.osb_kind = nosb
    b = P.random_vectors(prob.N, 2, 33)
    ret, x, _ = ctx.qmr(ctx.to_device(b), ntype=-1, qmrmax=400, qmrtol=1e-9)
    x = x.cpu().numpy()
    for v in range(2):
        r_o, x_o, _, _ = op.qmr(b[v], qmr_type=-1, qmr_max=400, qmr_tol=1e-9)
        assert r_o > 0 and abs(int(ret[v]) - r_o) <= 1, (ret, r_o)
        assert np.abs(x[v] - x_o).max() <= 1e-7 * np.abs(x_o).max()
    ctx.close()


@pytest.mark.parametrize("npc", [L.TA1, L.TA0])
def test_tosbw_window_px_and_solvers(npc):
    """The windowed preconditioner TOSBW on 12 x 10 x 8: scalit_set_osbw
    against the oracle's restatement of initOSBW from the same EIG0 (width
    and window size exact), TOSBW P*X against the oracle's px_tosbw to 1e-10
    relative (the window inverse is formed from an eigendecomposition here and
    by a solve in the oracle), then QMR with sOSB = TOSBW (residual of the
    solution <= 1e-8 relative) and PIST for the 8 levels nearest E0 against a
    dense eigensolve (1e-10 relative).  The FP32 inner solves refuse TOSBW."""
    print(f"tosbw npc={npc}", flush=True)
    prob, op, ctx = op_ctx((12, 10, 8), e0=0.0)
    vosb, eig0 = op.diag()
    ctx.set_precond(vosb, eig0)
    srt = np.sort(eig0)
    e0 = float(0.5 * (srt[40] + srt[41]))
    ctx.set_shift(e0)
    op.set_shift(e0)
    x = P.random_vectors(prob.N, 3, 41)
    for cnt, de in ((16, 1e-6), (60, 2e-3)):
        n, d = ctx.set_osbw(e0, de, cnt, npc)
        win = O.osbw_setup(op, e0, de, cnt, npc)
        assert n == cnt and d == win[1], (n, d, win[1])
        print(f"window {cnt}: width {d:.3e}, cond(HWR) {np.linalg.cond(win[3]):.1e}", flush=True)
        y = ctx.px(ctx.to_device(x), nosb=L.TOSBW, npc=npc).cpu().numpy()
        for v in range(3):
            yo = O.px_tosbw(op, x[v], win, npc)
            assert np.abs(y[v] - yo).max() <= 1e-10 * np.abs(yo).max()
    if npc != L.TA1:
        ctx.close()
        return
    b = P.random_vectors(prob.N, 1, 43)
    ret0, _, _ = ctx.qmr(ctx.to_device(b), ntype=-1, qmrmax=400, qmrtol=1e-9)
    ctx.set_osb(L.TOSBW)
    ret, xs, _ = ctx.qmr(ctx.to_device(b), ntype=-1, qmrmax=400, qmrtol=1e-9)
    print(f"QMR iterations: TOSB {int(ret0[0])}, TOSBW {int(ret[0])}", flush=True)
    assert ret[0] > 0
    h = op.hx(xs.cpu().numpy()[0], O.TA1)
    assert np.linalg.norm(h - b[0]) <= 1e-8 * np.linalg.norm(b[0])
    ev = np.linalg.eigvalsh(op.dense())
    near = np.sort(ev[np.argsort(np.abs(ev - e0))[:8]])
    x0 = ctx.to_device(P.random_vectors(prob.N, 1, 7))
    r, eig, _, stats = ctx.pist_conv(x0, e0, 1e-12, 8, 200, 30, 10, 4, qmrtol=1e-9)
    print("PIST with TOSBW", r, stats, flush=True)
    assert r > 0 and (np.abs(np.sort(eig) - near) / np.abs(near)).max() <= 1e-10
    with pytest.raises(Exception):
        ctx.pist_conv(x0, e0, 1e-10, 8, 200, 30, 10, 4, qmrtol=1e-6, prec=L.FP32)
    ctx.close()


def test_tosbw_bench_grid():
    """TOSBW on the 64 x 96 x 80 operator (level-1 blocks of 64 on the fused
    MFMA kernel), a window of the 64 OSB states nearest E0: P*X of 4 vectors
    against the oracle's px_tosbw on the GPU's own VOSB / EIG0 (1e-10), and a
    batch-16 QMR with TOSBW to 1e-8."""
    print("tosbw 64x96x80", flush=True)
    prob = P.config_problem("hx_bench")
    e0 = float(np.sort(prob.res)[50])
    ctx = Context.from_problem(prob, e0=e0)
    ctx.block_diag(1000, 1e-5)
    vosb, eig0 = ctx.get_precond()
    op = O.Operator(prob.sN, prob.h0, prob.res, prob.dep, e0=e0)
    op.set_precond(vosb, eig0)
    n, d = ctx.set_osbw(e0, 1e-4, 64)
    win = O.osbw_setup(op, e0, 1e-4, 64)
    assert n == 64 and d == win[1]
    x = P.random_vectors(prob.N, 4, 51)
    y = ctx.px(ctx.to_device(x), nosb=L.TOSBW).cpu().numpy()
    for v in range(4):
        yo = O.px_tosbw(op, x[v], win)
        assert np.abs(y[v] - yo).max() <= 1e-10 * np.abs(yo).max(), v
    ctx.set_osb(L.TOSBW)
    b = ctx.to_device(P.random_vectors(prob.N, 16, 5))
    ret, _, res = ctx.qmr(b, ntype=-1, qmrmax=200, qmrtol=1e-8)
    print("TOSBW QMR", ret, max(res), flush=True)
    assert (np.array(ret) > 0).all()
    ctx.close()


@pytest.mark.parametrize("persist", [1, 0])
def test_pist_persistent_and_per_step_agree(persist):
    """The config-1 PIST (lowest 20 levels, 24 x 32 x 24) against the golden
    levels with either QMR path."""
    g = gold()
    prob = P.config_problem("pist_fp64")
    ctx = Context.from_problem(prob, e0=g["e0"])
    ctx.set_qmr_persistent(persist)
    ctx.block_diag(1000, 1e-5)
    x0 = ctx.to_device(P.random_vectors(prob.N, 1, g["start_seed"]))
    r, eig, res, stats = ctx.pist_conv(x0, g["e0"], g["etol"], g["m"], g["m_max"], g["startcnt"], g["stepcnt"],
                                       g["stepeig"], qmrtol=g["qmr_tol"])
    print("pist persistent" if persist else "pist per-step", r, stats, flush=True)
    ref = np.array(g["eigsh"])
    assert r > 0 and (np.abs(eig - ref) / np.abs(ref)).max() <= 1e-10
    ctx.close()


def test_pist_fp64_config1_against_golden():
    g = gold()
    prob = P.config_problem("pist_fp64")
    ctx = Context.from_problem(prob, e0=g["e0"])
    ctx.block_diag(1000, 1e-5)
    x0 = ctx.to_device(P.random_vectors(prob.N, 1, g["start_seed"]))
    r, eig, res, stats = ctx.pist_conv(x0, g["e0"], g["etol"], g["m"], g["m_max"], g["startcnt"], g["stepcnt"],
                                       g["stepeig"], qmrtol=g["qmr_tol"])
    print("pist", r, res, stats, flush=True)
    assert r > 0
    ref = np.array(g["eigsh"])
    assert (np.abs(eig - ref) / np.abs(ref)).max() <= 1e-10
    # Ritz vectors: unit norm, Rayleigh quotients reproduce the eigenvalues
    psi = ctx.pist_vectors(0, g["m"])
    hp = ctx.hx(psi)
    rq = (psi * hp).sum(dim=1).cpu().numpy()
    nrm = (psi * psi).sum(dim=1).cpu().numpy()
    assert np.allclose(nrm, 1.0, atol=1e-10)
    assert np.abs(rq - eig).max() <= 1e-9 * np.abs(eig).max()
    ctx.close()


def test_pist_parameter_errors():
    prob = P.config_problem("pist_fp64", grid=(8, 9, 6))
    ctx = Context.from_problem(prob, e0=0.0)
    ctx.block_diag()
    x0 = ctx.to_device(np.ones((1, prob.N)))
    assert ctx.pist_conv(x0, 0.0, 1e-8, 5, 5, 2, 1, stepeig=5)[0] == 0
    assert ctx.pist_conv(x0, 0.0, 1e-8, 6, 5, 2, 1)[0] == 0
    ctx.close()


def test_pist_mixed_precision_config1():
    g = gold()
    prob = P.config_problem("pist_fp64")
    ctx = Context.from_problem(prob, e0=g["e0"])
    ctx.block_diag(1000, 1e-5)
    x0 = ctx.to_device(P.random_vectors(prob.N, 1, g["start_seed"]))
    ref = np.array(g["eigsh"])
    for prec in (L.FP32, L.BF16):
        r, eig, res, stats = ctx.pist_conv(x0, g["e0"], 1e-10, g["m"], g["m_max"], g["startcnt"], g["stepcnt"],
                                           g["stepeig"], qmrtol=1e-6, prec=prec)
        print("mixed", prec, r, res, stats, flush=True)
        assert r > 0
        assert (np.abs(eig - ref) / np.abs(ref)).max() <= 1e-8
    ctx.close()


def test_pist_batched_windows():
    g = gold()
    prob = P.config_problem("pist_fp64")
    ctx = Context.from_problem(prob, e0=g["e0"])
    ctx.block_diag(1000, 1e-5)
    ref = np.array(g["eigsh"])
    e0s = [g["e0"], 0.0040, 0.0055]
    x0 = ctx.to_device(P.random_vectors(prob.N, 3, 5))
    ret, eig, res, stats = ctx.pist_conv_batched(x0, e0s, 1e-10, 4, 120, 20, 10, 3, qmrtol=1e-6)
    print("batched", ret, res, stats, flush=True)
    assert (ret > 0).all()
    for k, e0 in enumerate(e0s):
        # the 4 eigenvalues nearest E0 of the golden set
        want = np.sort(ref[np.argsort(np.abs(ref - e0), kind="stable")[:4]])
        assert (np.abs(eig[k] - want) / np.abs(want)).max() <= 1e-10
    ctx.close()


def test_pist_batched_one_level_windows():
    """20 windows of one level each (the batched layout the throughput figure
    uses): every golden level is the one its window finds, to 1e-10."""
    g = gold()
    prob = P.config_problem("pist_fp64")
    ctx = Context.from_problem(prob, e0=g["e0"])
    ctx.block_diag(1000, 1e-5)
    ref = np.array(g["eigsh"])
    gaps = np.diff(ref)
    e0s = [float(ref[k] + 0.25 * min(gaps[max(k - 1, 0)], gaps[k])) for k in range(19)]
    e0s.append(float(ref[19] - 0.25 * gaps[18]))
    x0 = ctx.to_device(P.random_vectors(prob.N, 20, 77))
    ret, eig, res, stats = ctx.pist_conv_batched(x0, e0s, 1e-11, 1, 200, 4, 4, 1, qmrtol=1e-6)
    print("one-level windows", ret, stats, flush=True)
    assert (ret > 0).all()
    assert (np.abs(eig.reshape(-1) - ref) / np.abs(ref)).max() <= 1e-10
    # the Ritz vectors of all windows in one call agree with one window at a
    # time, and their Rayleigh quotients give the eigenvalues
    psi = ctx.pist_vectors_all(20, 1)
    one = torch.cat([ctx.pist_vectors(k, 1) for k in range(20)])
    assert float((psi - one).abs().max()) <= 1e-12
    rq = (psi * ctx.hx(psi)).sum(dim=1).cpu().numpy()
    assert np.abs(rq - eig.reshape(-1)).max() <= 1e-9 * np.abs(ref).max()
    with pytest.raises(L.ScalitError):
        ctx.pist_vectors_all(10, 1)   # too few columns for 20 windows: refused
    ctx.close()


def _lan_abi(ctx, x, e0, etol, m, m_max, startcnt, stepcnt, stepeig, ntype=0):
    """Lan_Conv through the Fortran drop-in layer with the library's own H*X
    callback (hx0_), host arrays by reference."""
    lib = L.lib()
    ctx.bind()
    ci = lambda v: C.byref(C.c_int(v))  # noqa: E731
    cd = lambda v: C.byref(C.c_double(v))  # noqa: E731
    x = np.ascontiguousarray(x, dtype=np.float64)
    eig = np.zeros(m)
    res = C.c_double()
    r = lib.lan_conv_(cd(e0), cd(etol), ci(ntype), ci(x.size), O.dptr(x), ci(startcnt), ci(stepcnt), ci(stepeig),
                      ci(m), ci(m_max), C.cast(lib.hx0_, L.FOP), O.dptr(eig), C.byref(res))
    return r, eig, res.value


def test_lanczos_config1_against_golden():
    """GPU Lanczos on the 24 x 32 x 24 workload: the lowest 20 levels within
    1e-10 of the golden set, on the handle API and through lan_conv_."""
    g = gold()
    ref = np.array(g["eigsh"])
    prob = P.config_problem("pist_fp64")
    ctx = Context.from_problem(prob)
    x = P.random_vectors(prob.N, 1, 3)
    e0 = float(ref[0]) - 0.002
    # (the 20 lowest levels need about 1,300 Lanczos steps on this operator)
    r, eig, res = ctx.lan_conv(ctx.to_device(x), e0, 2e-12, 20, 1600, 400, 100, 10)
    print("lanczos config1", r, res, flush=True)
    assert r > 0, r
    assert (np.abs(eig - ref) / np.abs(ref)).max() <= 1e-10
    r2, eig2, _ = _lan_abi(ctx, x[0], e0, 2e-12, 20, 1600, 400, 100, 10)
    assert r2 == r and np.array_equal(eig2, eig)
    ctx.close()


def test_lanczos_per_step_path():
    """A grid above 20,480 points (24 x 32 x 28 = 21,504) takes the
    per-step re-orthogonalisation path (launch per projection, BETA through
    k_store_norm, a host check per step): converged, against scipy's eigsh
    of the oracle's H*X (1e-10) and through lan_conv_; and a short run that
    stops at M_MAX unconverged, step for step against the oracle's Lan_Conv
    (same return code, Ritz values within 1e-9)."""
    import scipy.sparse.linalg as sla

    prob = P.config_problem("pist_fp64", grid=(24, 32, 28))
    assert prob.N > 20480
    op = O.Operator(prob.sN, prob.h0, prob.res, prob.dep)
    a = sla.LinearOperator((op.N, op.N), matvec=lambda v: op.hx(v), dtype=np.float64)
    ev = np.sort(sla.eigsh(a, k=8, which="SA", tol=1e-13, ncv=60)[0])[:5]
    ctx = Context.from_problem(prob)
    x = P.random_vectors(prob.N, 1, 4)
    e0 = float(ev[0]) - 0.002
    # (the 5 lowest levels need about 650 steps here)
    r, eig, res = ctx.lan_conv(ctx.to_device(x), e0, 1e-12, 5, 800, 200, 100, 5)
    print("lanczos per-step", r, res, flush=True)
    assert r > 0, r
    assert (np.abs(eig - ev) / np.abs(ev)).max() <= 1e-10
    r2, eig2, _ = _lan_abi(ctx, x[0], e0, 1e-12, 5, 800, 200, 100, 5)
    assert r2 == r and np.array_equal(eig2, eig)
    r_o, eig_o, _ = op.lan_conv(x[0], e0, 1e-30, 3, 150, 60, 30, 5)
    r_s, eig_s, _ = ctx.lan_conv(ctx.to_device(x), e0, 1e-30, 3, 150, 60, 30, 5)
    print("lanczos per-step short", r_s, r_o, flush=True)
    assert r_s == r_o == -150
    assert np.abs(eig_s - eig_o).max() <= 1e-9 * np.abs(eig_o).max()
    ctx.close()


def test_lanczos_block_without_steps():
    """lanH0 returns 0 when a block would have no step (CNT_LOW >= CNT_HIGH,
    lanHij.f90:20) and Lan_Conv then returns -(CNT_LOW-1) (lan.f90:47-50):
    stepCnt 0 or 1 after an unconverged first block, and a first block of one
    step, give the oracle's return codes."""
    prob = P.config_problem("pist_fp64", grid=(8, 9, 6))
    op = O.Operator(prob.sN, prob.h0, prob.res, prob.dep)
    ctx = Context.from_problem(prob)
    x = P.random_vectors(prob.N, 1, 6)
    cases = [(1e-30, 3, 60, 20, 0, 2), (1e-30, 3, 60, 20, 1, 2), (1e-8, 1, 60, 1, 5, 0)]
    for etol, m, m_max, start, step, stepeig in cases:
        r_o, _, _ = op.lan_conv(x[0], 0.0, etol, m, m_max, start, step, stepeig)
        r, _, _ = ctx.lan_conv(ctx.to_device(x), 0.0, etol, m, m_max, start, step, stepeig)
        print("lanczos no-step block", (etol, m, m_max, start, step, stepeig), r, r_o, flush=True)
        assert r == r_o and r <= 0
    ctx.close()


def test_pist_many_levels_needs_second_orthogonalisation():
    """30 levels of the 24 x 32 x 24 Morse + P2 + Gaussians operator from
    E0 = min V grow a Krylov space of about 270 vectors.  With the reference's
    single MGS pass the basis loses orthogonality and some Ritz values are
    spurious (residuals far from zero); with the second pass every Ritz pair
    has a residual below 1e-9 and the eigenvalues match scipy's eigsh of the
    oracle's H*X within 1e-10."""
    import scipy.sparse.linalg as sla

    prob = P.build_problem(24, 32, 24, P.gaussian_potential(1234), seed=1234)
    op = O.Operator(prob.sN, prob.h0, prob.res, prob.dep)
    a = sla.LinearOperator((op.N, op.N), matvec=lambda v: op.hx(v), dtype=np.float64)
    ref = np.sort(sla.eigsh(a, k=36, which="SA", tol=1e-13, ncv=120)[0])[:30]
    e0 = float(prob.res.min())
    ctx = Context.from_problem(prob, e0=e0)
    ctx.block_diag(1000, 1e-5)
    x0 = ctx.to_device(P.random_vectors(prob.N, 1, 91))
    worst = {}
    for passes in (1, 2):
        ctx.set_pist_orth(passes)
        r, eig, res, stats = ctx.pist_conv(x0, e0, 1e-12, 30, 400, 90, 20, 10, qmrtol=1e-9)
        psi = ctx.pist_vectors(0, 30) if r > 0 else None
        if psi is not None:
            hp = ctx.hx(psi)
            rr = hp - torch.as_tensor(eig, device=psi.device)[:, None] * psi
            worst[passes] = float((rr * rr).sum(dim=1).sqrt().max())
        print("orth passes", passes, r, stats, worst.get(passes), flush=True)
        if passes == 2:
            assert r > 0 and worst[2] <= 1e-9
            assert (np.abs(eig - ref) / np.abs(ref)).max() <= 1e-10
    ctx.close()


def test_lanczos_small_grid():
    prob = P.config_problem("pist_fp64", grid=(12, 14, 10))
    op = O.Operator(prob.sN, prob.h0, prob.res, prob.dep)
    ev = np.linalg.eigvalsh(op.dense())
    ctx = Context.from_problem(prob)
    x0 = ctx.to_device(P.random_vectors(prob.N, 1, 3))
    r, eig, res = ctx.lan_conv(x0, float(ev[0]) - 0.01, 1e-12, 5, 600, 60, 30, 5)
    assert r > 0, r
    assert (np.abs(eig - ev[:5]) / np.abs(ev[:5])).max() <= 1e-10
    r_o, eig_o, _ = op.lan_conv(P.random_vectors(prob.N, 1, 3)[0], float(ev[0]) - 0.01, 1e-12, 5, 600, 60, 30, 5)
    assert np.abs(eig - eig_o).max() <= 1e-10 * np.abs(eig_o).max()
    ctx.close()


def test_hellmann_feynman_gradients():
    # dE_k/dp by HF on the GPU against central finite differences of the oracle's dense spectrum
    grid = (10, 12, 8)
    pot = P.gaussian_potential(9)
    prob = P.build_problem(*grid, pot, seed=9)
    op = O.Operator(prob.sN, prob.h0, prob.res, prob.dep)
    ev = np.linalg.eigvalsh(op.dense())
    e0 = float(ev[0]) - 0.002
    ctx = Context.from_problem(prob, e0=e0)
    ctx.block_diag(1000, 1e-8)
    x0 = ctx.to_device(P.random_vectors(prob.N, 1, 4))
    m = 4
    r, eig, _, _ = ctx.pist_conv(x0, e0, 1e-12, m, 150, 20, 10, 4, qmrtol=1e-9)
    assert r > 0
    psi = ctx.pist_vectors(0, m)
    dvdp = ctx.to_device(pot.gradient(prob.r, prob.R, prob.cos_theta))
    grad = ctx.eig_gradients(psi, dvdp)
    p = pot.params()
    h = 1e-5
    for j in (0, 2, 4, 6, 7):
        pp, pm = p.copy(), p.copy()
        pp[j] += h
        pm[j] -= h
        ep = np.linalg.eigvalsh(O.Operator(prob.sN, prob.h0, prob.with_potential(pot.with_params(pp)).res,
                                           prob.dep).dense())[:m]
        em = np.linalg.eigvalsh(O.Operator(prob.sN, prob.h0, prob.with_potential(pot.with_params(pm)).res,
                                           prob.dep).dense())[:m]
        fd = (ep - em) / (2 * h)
        assert np.abs(grad[:, j] - fd).max() <= 1e-6 * max(1.0, np.abs(fd).max()), (j, grad[:, j], fd)
    ctx.close()


@pytest.mark.parametrize("n", [700, 16385, 20480, 20481])
def test_mgs_orth_lengths(n):
    """MGS_ORTH through the Fortran-layer entry at lengths either side of the
    one-workgroup kernel's limits (8, 16 and 20 values per thread; 20481 takes
    the multi-launch path), against the oracle's restatement; a zero W fails."""
    prob, op, ctx = op_ctx((8, 9, 6), e0=0.0)
    ctx.bind()
    lib = L.lib()
    ci = lambda v: C.byref(C.c_int(v))  # noqa: E731
    m = 7
    q, _ = np.linalg.qr(np.random.default_rng(n).standard_normal((n, m)))
    vj = np.zeros((n, m + 1), order="F")
    vj[:, :m] = q
    w = q @ np.arange(1.0, m + 1) + 0.1 * np.random.default_rng(n + 1).standard_normal(n)
    vo = vj.copy(order="F")
    assert lib.mgs_orth_(ci(n), O.dptr(w), ci(m), O.dptr(vj)) == 1
    ok_o, _ = O.mgs_orth(w, vo, m)
    assert ok_o
    assert np.abs(vj - vo).max() <= 1e-11
    assert np.abs(vj.T @ vj - np.eye(m + 1)).max() <= 1e-12
    z = np.zeros(n)
    assert lib.mgs_orth_(ci(n), O.dptr(z), ci(m), O.dptr(vj)) == 0
    ctx.close()


def test_fortran_layer_drop_in():
    # the reference's call pattern: QMR(nType,QMRMAX,QMRTOL,N,B,HX,PX,X,RES) on host arrays
    prob, op, ctx = op_ctx((12, 14, 10), e0=-0.002)
    vosb, eig0 = op.diag()
    ctx.set_precond(vosb, eig0)
    ctx.bind()
    lib = L.lib()
    n = prob.N
    b = P.random_vectors(n, 1, 2)[0].copy()
    x = np.zeros(n)
    res = C.c_double(0)
    ci = lambda v: C.byref(C.c_int(v))  # noqa: E731
    ret = lib.qmr_(ci(-1), ci(300), C.byref(C.c_double(1e-9)), ci(n), O.dptr(b), C.cast(lib.hx_, L.FOP), C.cast(lib.px_, L.FOP),
                   O.dptr(x), C.byref(res))
    r_o, x_o, _, _ = op.qmr(b, qmr_type=-1, qmr_max=300, qmr_tol=1e-9)
    assert abs(ret - r_o) <= 1 and np.abs(x - x_o).max() <= 1e-7 * np.abs(x_o).max()
    # foreign host callbacks (the oracle's operator): the recurrence stays on the GPU
    x2 = np.zeros(n)

    def hx_cb(pn, px_in, py):
        xi = np.ctypeslib.as_array(px_in, shape=(pn[0],))
        np.ctypeslib.as_array(py, shape=(pn[0],))[:] = op.hx(xi, O.TA1)

    def px_cb(pn, px_in, py):
        xi = np.ctypeslib.as_array(px_in, shape=(pn[0],))
        np.ctypeslib.as_array(py, shape=(pn[0],))[:] = op.px(xi)

    f_hx, f_px = L.FOP(hx_cb), L.FOP(px_cb)
    ret2 = lib.qmr_(ci(-1), ci(300), C.byref(C.c_double(1e-9)), ci(n), O.dptr(b), f_hx, f_px, O.dptr(x2),
                    C.byref(res))
    assert abs(ret2 - r_o) <= 1 and np.abs(x2 - x_o).max() <= 1e-7 * np.abs(x_o).max()
    # MGS_ORTH
    vj = np.zeros((n, 4), order="F")
    q, _ = np.linalg.qr(np.random.default_rng(0).standard_normal((n, 3)))
    vj[:, :3] = q
    w = np.random.default_rng(1).standard_normal(n)
    vo = vj.copy(order="F")
    assert lib.mgs_orth_(ci(n), O.dptr(w), ci(3), O.dptr(vj)) == 1
    O.mgs_orth(w, vo, 3)
    assert np.abs(vj - vo).max() <= 1e-12
    # PIST_CONV through the Fortran layer with HX0 / OSB_QMR
    lib.scalit_set_qmr_params(-1, 300, 1e-8)
    eig = np.zeros(3)
    x0 = P.random_vectors(n, 1, 6)[0].copy()
    r = lib.pist_conv_(C.byref(C.c_double(-0.002)), C.byref(C.c_double(1e-12)), ci(0), ci(n), O.dptr(x0), ci(12),
                       ci(4), ci(2), ci(3), ci(100), C.cast(lib.hx0_, L.FOP), C.cast(lib.osb_qmr_, L.FLINSOLV), O.dptr(eig),
                       C.byref(res))
    ev = np.linalg.eigvalsh(op.dense())
    want = np.sort(ev[np.argsort(np.abs(ev + 0.002), kind="stable")[:3]])
    assert r > 0 and np.abs(eig - want).max() <= 1e-9 * np.abs(want).max()
    ctx.close()
