"""The oracle against the reference-run fixture tests/golden/ref_small.npz.

The fixture holds the outputs of the reference's own compiled routines (module
osb's MYHX / MYPX, HosbDiag, QMR) on a 6 x 7 x 5 operator, written by
tests/golden/make_golden.py.  These tests need no reference build: they keep
the oracle pinned bit for bit wherever oracle/_ref is absent.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from scalit_amd import problems as P

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def fx():
    d = np.load(os.path.join(HERE, "ref_small.npz"))
    prob = P.build_problem(*[int(v) for v in d["grid"]], P.gaussian_potential(int(d["seed"])), name="small")
    op = O.Operator(prob.sN, prob.h0, prob.res, prob.dep, e0=float(d["e0"]))
    return d, op


def test_hx_matches_reference_run(fx):
    d, op = fx
    assert np.array_equal(op.hx(d["x"], O.TA0), d["hx0"])
    assert np.array_equal(op.hx(d["x"], O.TA1), d["hx1"])


def test_osb_setup_and_px_match_reference_run(fx):
    d, op = fx
    vosb, eig0 = op.diag(1000, 1e-5)
    for l in range(3):
        assert np.array_equal(vosb[l], d[f"vosb{l + 1}"])
    assert np.array_equal(eig0, d["eig0"])
    assert np.array_equal(op.px(d["x"]), d["px"])
    op.set_shift(float(d["e0"]), float(d["de"]), float(d["beta"]))
    # the window holds eigenvalues on both sides of E0 and leaves others out
    ade = eig0 - float(d["e0"])
    inside = np.abs(ade) <= float(d["de"])
    assert (inside & (ade < 0)).any() and (inside & (ade >= 0)).any() and (~inside).any()
    assert np.array_equal(op.px(d["x"], kind=O.TOSBD1), d["px_d1"])
    assert np.array_equal(op.px(d["x"], kind=O.TOSBD2), d["px_d2"])
    op.set_shift(float(d["e0"]))


def test_qmr_matches_reference_run(fx):
    d, op = fx
    op.set_precond([d["vosb1"], d["vosb2"], d["vosb3"]], d["eig0"])
    r, x, res, _ = op.qmr(d["qmr_b"], qmr_type=-1, qmr_max=200, qmr_tol=1e-9)
    assert r == int(d["qmr_ret"]) and r > 0
    assert np.array_equal(x, d["qmr_x"])
    assert res == float(d["qmr_res"])
