"""Several problems with uneven grids in one call (scalit_pist_conv_multi).

Each problem's result from the concurrent call must equal, bit for bit, its
own batched PIST run alone (the kernels are deterministic and every problem
keeps its own stream and state), and every window must converge to levels of
its problem.
"""
import numpy as np
import pytest

from scalit_amd import problems as P
from scalit_amd.context import Context

pytestmark = pytest.mark.gpu

GRIDS = [(24, 32, 24), (20, 28, 16), (16, 20, 12)]


def setup(grid, seed):
    prob = P.build_problem(*grid, P.jittered_potential(seed), seed=seed)
    e0 = float(np.sort(prob.res)[0]) + 0.003
    ctx = Context.from_problem(prob, e0=e0)
    ctx.block_diag(1000, 1e-5)
    x0 = ctx.to_device(P.random_vectors(prob.N, 1, seed))
    r, eig, _, _ = ctx.pist_conv(x0, e0, 1e-10, 6, 200, 20, 5, 2, qmrtol=1e-6)
    assert r > 0
    lev = np.sort(eig)
    centres = [float(lev[2 * k:2 * k + 2].mean()) for k in range(3)]
    return prob, ctx, lev, centres


def test_multi_equals_each_alone():
    runs = [setup(g, 11 + i) for i, g in enumerate(GRIDS)]
    args = dict(etol=1e-10, m=2, m_max=120, startcnt=12, stepcnt=4, stepeig=1, qmrtol=1e-6)
    alone, problems = [], []
    for prob, ctx, lev, centres in runs:
        x0 = ctx.to_device(P.random_vectors(prob.N, 3, 5))
        ret, eig, res, stats = ctx.pist_conv_batched(x0, centres, args["etol"], args["m"], args["m_max"],
                                                     args["startcnt"], args["stepcnt"], args["stepeig"],
                                                     qmrtol=args["qmrtol"])
        alone.append((np.array(ret), np.array(eig).reshape(3, 2)))
        problems.append((ctx, x0, centres))
    out = Context.pist_conv_multi(problems, **args)
    for (ret_a, eig_a), (ret_m, eig_m, res_m, stats_m), (prob, ctx, lev, centres) in zip(alone, out, runs):
        assert (ret_m > 0).all() and np.array_equal(ret_m, ret_a)
        assert np.array_equal(eig_m, eig_a)
        found = np.sort(eig_m.reshape(-1))
        assert np.max(np.abs(found - lev) / np.abs(lev)) <= 1e-9
    for _, ctx, _, _ in runs:
        ctx.close()


def test_multi_rejects_shared_context():
    prob, ctx, lev, centres = setup(GRIDS[2], 3)
    x0 = ctx.to_device(P.random_vectors(prob.N, 3, 5))
    with pytest.raises(Exception):
        Context.pist_conv_multi([(ctx, x0, centres), (ctx, x0, centres)], 1e-10, 2, 120, 12, 4)
    ctx.close()


def test_qmr_multi_one_graph_equals_each_alone():
    """scalit_qmr_multi: FP64 QMR on three operators of uneven grids, 4, 2
    and 3 right-hand sides at shifts inside each spectrum, every iteration of
    all three one graph launch.  Each solution and return code equals, bit
    for bit, the same solve alone on the one-kernel-per-step path, and each
    solves its system (residual against the oracle's H*X)."""
    from oracle import oracle as O

    nvs = [4, 2, 3]
    runs, problems = [], []
    for i, g in enumerate(GRIDS):
        prob = P.build_problem(*g, P.jittered_potential(21 + i), seed=21 + i)
        e0 = float(np.sort(prob.res)[20])
        ctx = Context.from_problem(prob, e0=e0)
        ctx.block_diag(1000, 1e-5)
        ctx.set_qmr_persistent(0)
        b = P.random_vectors(prob.N, nvs[i], 31 + i)
        shifts = [e0 + 1e-4 * k for k in range(nvs[i])]
        runs.append((prob, ctx, b, shifts))
        problems.append((ctx, ctx.to_device(b), shifts))
    multi = Context.qmr_multi(problems, ntype=-1, qmrmax=500, qmrtol=1e-9)
    for (prob, ctx, b, shifts), (ret_m, x_m, res_m) in zip(runs, multi):
        ret, x, res = ctx.qmr(ctx.to_device(b), ntype=-1, qmrmax=500, qmrtol=1e-9, shifts=shifts)
        print(prob.sN, "multi", ret_m, "alone", ret, flush=True)
        assert np.array_equal(np.array(ret), ret_m) and (ret_m > 0).all()
        assert bool((x == x_m).all())
        op = O.Operator(prob.sN, prob.h0, prob.res, prob.dep)
        xm = x_m.cpu().numpy()
        for v in range(len(shifts)):
            r = op.hx(xm[v]) - shifts[v] * xm[v] - b[v]
            assert np.linalg.norm(r) <= 1e-7 * np.linalg.norm(b[v])
    for _, ctx, _, _ in runs:
        ctx.close()
