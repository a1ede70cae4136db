"""The entry point that allocates nothing (scalit_eig_gradients_noalloc)
against scalit_eig_gradients, through the C ABI: two calls in a row, the
second on a smaller operator with fewer levels and parameters, each equal to
the allocating entry point's result (same kernels, same deterministic
reduction order) and to a host sum; oversized requests are refused."""
import ctypes as C

import numpy as np
import pytest
import torch

from scalit_amd import _lib as L
from scalit_amd import problems as P
from scalit_amd.context import Context

pytestmark = pytest.mark.gpu


def _case(grid, nk, npar, seed):
    prob = P.build_problem(*grid, P.gaussian_potential(seed), seed=seed)
    ctx = Context.from_problem(prob)
    rng = np.random.default_rng(seed)
    psi = rng.standard_normal((nk, prob.N))
    dvdp = rng.standard_normal((npar, prob.N))
    return ctx, psi, dvdp


def test_noalloc_gradients_match_allocating_entry():
    for grid, nk, npar, seed in (((24, 32, 24), 6, 12, 31), ((10, 12, 8), 3, 5, 32)):
        print(f"noalloc gradients grid={grid} nk={nk} np={npar}", flush=True)
        ctx, psi, dvdp = _case(grid, nk, npar, seed)
        dpsi, ddv = ctx.to_device(psi), ctx.to_device(dvdp)
        ref = ctx.eig_gradients(dpsi, ddv)
        got = ctx.eig_gradients(dpsi, ddv, noalloc=True)
        assert np.array_equal(got, ref), np.abs(got - ref).max()
        host = (psi ** 2) @ dvdp.T
        assert np.abs(got - host).max() <= 1e-12 * np.abs(host).max()
        ctx.close()


def test_noalloc_gradients_reject_oversize():
    ctx, psi, dvdp = _case((10, 12, 8), 1, 1, 33)
    # buffers large enough for the request, so that it is the size limit alone that refuses it
    dpsi = ctx.to_device(np.zeros((64, ctx.N)))
    out = np.zeros(4096)
    r = L.lib().scalit_eig_gradients_noalloc(ctx._h, 64, C.c_void_p(dpsi.data_ptr()), ctx.N, 64,
                                             C.c_void_p(dpsi.data_ptr()), ctx.N,
                                             out.ctypes.data_as(C.POINTER(C.c_double)))
    assert r == L.SCALIT_EINVAL
    torch.cuda.synchronize()
    ctx.close()
