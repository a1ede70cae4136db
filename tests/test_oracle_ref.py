"""Pin the CPU oracle against the reference's own compiled Fortran routines.

The reference checker library (oracle/_ref) is built from the reference's
LAPACK-free hot-path sources; these tests skip where it is absent.  Where the
oracle restates a routine in the same evaluation order the comparison is
bit-exact; otherwise the tolerance is written in the test.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from scalit_amd import problems as P

pytestmark = pytest.mark.ref


def small_problem(n=(6, 7, 5), seed=3):
    return P.build_problem(*n, P.gaussian_potential(seed), name="small")


def make_ops(prob, e0=0.0):
    op = O.Operator(prob.sN, prob.h0, prob.res, prob.dep, e0=e0)
    return op, O.RefOperator(op)


@pytest.mark.parametrize("grid", [(6, 7, 5), (24, 32, 24)])
def test_hx_bit_exact(grid):
    prob = P.config_problem("pist_fp64", grid=grid)
    op, rop = make_ops(prob, e0=-0.01)
    x = P.random_vectors(op.N, 1, 11)[0]
    assert np.array_equal(op.hx(x, O.TA0), rop.hx(x, shifted=False))
    assert np.array_equal(op.hx(x, O.TA1), rop.hx(x, shifted=True))


def test_h1x_h2x_vx_vtx_direct():
    R = O.ref()
    g = np.random.default_rng(0)
    nin, nout, blk = 5, 7, 3
    h = O.fmat(g.standard_normal((nout, nout)))
    x = g.standard_normal(nin * nout * blk)
    corx = g.standard_normal(nin)
    y_ref = np.zeros_like(x)
    R.h1x_xyz_(O._ci(nin), O._ci(nout), O._ci(blk), O.dptr(h), O.dptr(x), O.dptr(y_ref))
    y = np.zeros_like(x)
    O.lib().orc_h1x_xyz(nin, nout, blk, O.dptr(h), O.dptr(x), O.dptr(y))
    assert np.array_equal(y, y_ref)
    R.h1x_dep_(O._ci(nin), O._ci(nout), O._ci(blk), O.dptr(h), O.dptr(corx), O.dptr(x), O.dptr(y_ref))
    O.lib().orc_h1x_dep(nin, nout, blk, O.dptr(h), O.dptr(corx), O.dptr(x), O.dptr(y))
    assert np.array_equal(y, y_ref)
    v = g.standard_normal(nout * nout * blk)
    for fr, fo in (("vx_", "orc_vx"), ("vtx_", "orc_vtx")):
        a = x.copy()
        b = x.copy()
        getattr(R, fr)(O._ci(nin), O._ci(nout), O._ci(blk), O.dptr(v), O.dptr(a))
        getattr(O.lib(), fo)(nin, nout, blk, O.dptr(v), O.dptr(b))
        assert np.array_equal(a, b)


def test_mjacobi_and_osb_diag_bit_exact():
    prob = P.config_problem("pist_fp64", grid=(12, 10, 8))
    op, rop = make_ops(prob, e0=0.0)
    vo, eo = op.diag(bjmax=1000, bjtol=1e-5)
    vr, er = rop.diag(bjmax=1000, bjtol=1e-5)
    for a, b in zip(vo, vr):
        assert np.array_equal(a, b)
    assert np.array_equal(eo, er)


def test_px_bit_exact():
    prob = P.config_problem("pist_fp64", grid=(12, 10, 8))
    op, rop = make_ops(prob, e0=-0.02)
    op.diag()
    rop.diag()
    x = P.random_vectors(op.N, 1, 5)[0]
    assert np.array_equal(op.px(x), rop.px(x))


@pytest.mark.parametrize("kind", [O.TOSB, O.TOSBD1, O.TOSBD2])
@pytest.mark.parametrize("npc", [O.TA0, O.TA1])
def test_px_kernels_bit_exact(kind, npc):
    """Every PX_KERNEL case of the reference's own MYPX (osb.px.h:156-188),
    called from module osb, against the oracle's orc_mypx: E0 sits inside
    the EIG0 spectrum and DE covers eigenvalues on both sides of it, so the
    clamped branches (ADE < 0 and >= 0) and the plain divide all run."""
    prob = P.config_problem("pist_fp64", grid=(12, 10, 8))
    op, rop = make_ops(prob, e0=0.0)
    _, eig0 = op.diag()
    rop.diag()
    srt = np.sort(eig0)
    e0 = float(0.5 * (srt[40] + srt[41]))
    # TA0 measures EIG0 from 0, TA1 from E0: DE reaches the 61st eigenvalue either way
    de = float(srt[60] - (e0 if npc == O.TA1 else 0.0))
    op.set_shift(e0, de, 0.3)
    rop.set_shift(e0, de, 0.3)
    ade = (eig0 - e0) if npc == O.TA1 else eig0
    assert ((np.abs(ade) <= de) & (ade < 0)).any() or npc == O.TA0
    assert ((np.abs(ade) <= de) & (ade >= 0)).any() and (np.abs(ade) > de).any()
    x = P.random_vectors(op.N, 1, 6)[0]
    assert np.array_equal(op.px(x, npc=npc, kind=kind), rop.px(x, kind=kind, npc=npc))


@pytest.mark.parametrize("qmr_type", [-1, 0, -2])
def test_qmr_bit_exact(qmr_type):
    prob = P.config_problem("pist_fp64", grid=(12, 10, 8))
    op, rop = make_ops(prob, e0=-0.005)
    op.diag()
    rop.diag()
    b = P.random_vectors(op.N, 1, 9)[0]
    r1, x1, res1, _ = op.qmr(b, qmr_type=qmr_type, qmr_max=200, qmr_tol=1e-8)
    r2, x2, res2 = rop.qmr(b, qmr_type=qmr_type, qmr_max=200, qmr_tol=1e-8)
    assert r1 == r2 and r1 > 0
    assert np.array_equal(x1, x2)
    assert res1 == res2


def test_pist_h0_bit_exact():
    prob = P.config_problem("pist_fp64", grid=(8, 9, 6))
    e0 = float(prob.res.min())
    op, rop = make_ops(prob, e0=e0)
    op.diag()
    rop.diag()
    x = P.random_vectors(op.N, 1, 21)[0]
    m_max = 12
    vj_o = np.zeros((op.N, m_max), order="F")
    hm_o = np.zeros((m_max, m_max), order="F")
    vj_r = np.zeros_like(vj_o)
    hm_r = np.zeros_like(hm_o)
    ctx = op._ctx(-1, 200, 1e-6)
    L = O.lib()
    # two calls, as PIST_CONV grows the Krylov space: 1..6 then 7..10
    for lo, hi in ((1, 6), (7, 10)):
        a = L.orc_pist_h0(C.c_long(op.N), O.dptr(x), m_max, lo, hi, C.cast(L.orc_cb_hx0, O.OP_FN),
                          C.cast(L.orc_cb_linsolv, O.LS_FN), C.byref(ctx), O.dptr(vj_o), O.dptr(hm_o), None)
        b = rop.pist_h0(x, m_max, lo, hi, vj_r, hm_r, qmr_type=-1, qmr_max=200, qmr_tol=1e-6)
        assert a == b == hi
    assert np.array_equal(vj_o[:, :10], vj_r[:, :10])
    assert np.array_equal(hm_o[:10, :10], hm_r[:10, :10])


def test_lan_h0_bit_exact():
    prob = P.config_problem("pist_fp64", grid=(8, 9, 6))
    op, rop = make_ops(prob)
    x = P.random_vectors(op.N, 1, 2)[0]
    m = 15
    vj_o = np.zeros((op.N, m + 1), order="F")
    vj_r = np.zeros_like(vj_o)
    al_o, be_o, al_r, be_r = (np.zeros(m) for _ in range(4))
    ctx = op._ctx(-1, 1, 1.0)
    L = O.lib()
    for lo, hi in ((1, 8), (9, 15)):
        a = L.orc_lan_h0(C.c_long(op.N), O.dptr(x), O.dptr(vj_o), lo, hi, C.cast(L.orc_cb_hx0, O.OP_FN),
                         C.byref(ctx), O.dptr(al_o), O.dptr(be_o))
        b = rop.lan_h0(x, vj_r, lo, hi, al_r, be_r)
        assert a == b == hi
    assert np.array_equal(al_o, al_r) and np.array_equal(be_o, be_r)
    assert np.array_equal(vj_o, vj_r)


def test_mgs_orth_bit_exact():
    g = np.random.default_rng(4)
    n, m = 50, 6
    vj = np.zeros((n, m + 1), order="F")
    q, _ = np.linalg.qr(g.standard_normal((n, m)))
    vj[:, :m] = q
    w = g.standard_normal(n)
    a = vj.copy(order="F")
    b = vj.copy(order="F")
    ok_o, _ = O.mgs_orth(w, a, m)
    ok_r = O.ref().mgs_orth_(O._ci(n), O.dptr(w), O._ci(m), O.dptr(b))
    assert ok_o and ok_r
    assert np.array_equal(a, b)
    # a zero vector fails in both
    z = np.zeros(n)
    ok_o, _ = O.mgs_orth(z, vj.copy(order="F"), m)
    ok_r = O.ref().mgs_orth_(O._ci(n), O.dptr(z), O._ci(m), O.dptr(vj.copy(order="F")))
    assert not ok_o and not ok_r


def test_window_helpers_bit_exact():
    R = O.ref()
    g = np.random.default_rng(7)
    v = g.uniform(-1, 1, 40)
    v[5] = v[9] = 0.25          # a tie: the first index wins
    for e0 in (0.0, 0.25, -2.0):
        for m in (1, 7, 40):
            ind, out = O.get_window(e0, v, m)
            ind_r = np.zeros(m, dtype=np.int32)
            out_r = np.zeros(m)
            R.getwindow0_(O._cd(e0), O._ci(len(v)), O.dptr(v), O._ci(m), O.iptr(ind_r), O.dptr(out_r))
            assert np.array_equal(ind, ind_r) and np.array_equal(out, out_r)
    a, b = g.uniform(size=9), g.uniform(size=9)
    assert O.get_diff_max(a, b) == R.getdiffmax_(O._ci(9), O.dptr(a), O.dptr(b))
    c = g.uniform(size=25)
    assert O.get_sep_max(a, c) == R.getsepmax_(O._ci(9), O.dptr(a), O._ci(25), O.dptr(c))


def test_dvr_builders_match_reference():
    R = O.ref()
    for n, a0, b0 in ((24, 1.5, 4.5), (32, 2.0, 8.0), (64, 1.5, 4.5)):
        x0, h0 = O.sinc2_dvr(n, a0, b0, P.MU)
        xr = np.zeros(n)
        hr = np.zeros((n, n), order="F")
        R.sinc2_dvr_(C.byref(C.c_int(0)), O._ci(n), O._cd(a0), O._cd(b0), O._cd(P.MU), O.dptr(xr), O.dptr(hr))
        assert np.array_equal(x0, xr) and np.array_equal(h0, hr)
        xp, hp = P.sinc2_dvr(n, a0, b0, P.MU)
        assert np.allclose(xp, xr, rtol=0, atol=1e-15) and np.allclose(hp, hr, rtol=1e-14, atol=0)
    for n in (24, 40, 80, 96):
        x, w = O.gau_legendre(n)
        xr, wr = np.zeros(n), np.zeros(n)
        R.gau_legendre_(O._cd(-1), O._cd(1), O._ci(n), O.dptr(xr), O.dptr(wr))
        assert np.array_equal(x, xr) and np.array_equal(w, wr)
        y = O.yj_polys(x, n - 1)
        yr = np.zeros((n, n), order="F")
        R.yjpolys_(O._ci(n), O._ci(n - 1), O.dptr(x), O.dptr(yr))
        assert np.array_equal(y, yr)


def _ndvr_ops(grid, e0=0.0, seed=21):
    prob = P.config_problem("pist_fp64", grid=grid)
    outh = P.outer_layer(prob, seed)
    op = O.Operator(prob.sN, prob.h0, prob.res, prob.dep, e0=e0, outh=outh)
    return prob, op, O.RefOperator(op)


def test_h3x_direct():
    R = O.ref()
    g = np.random.default_rng(1)
    nin, nout = 9, 6
    a = g.standard_normal(nin * nout * nout)
    x = g.standard_normal(nin * nout)
    y_ref = np.zeros_like(x)
    y = np.zeros_like(x)
    R.h3x_(O._ci(nin), O._ci(nout), O.dptr(a), O.dptr(x), O.dptr(y_ref))
    O.lib().orc_h3x(nin, nout, O.dptr(a), O.dptr(x), O.dptr(y))
    assert np.array_equal(y, y_ref)


@pytest.mark.parametrize("grid", [(6, 7, 5), (12, 10, 8)])
def test_ndvr_hx_diag_px_bit_exact(grid):
    """sNDVR = T: the outermost level is H3X with OUTH (osb.hx.h:93-94) and its
    OSB block is OUTH itself (osb.diag.h:22-23)."""
    _, op, rop = _ndvr_ops(grid, e0=-0.01)
    x = P.random_vectors(op.N, 1, 12)[0]
    assert np.array_equal(op.hx(x, O.TA0), rop.hx(x, shifted=False))
    assert np.array_equal(op.hx(x, O.TA1), rop.hx(x, shifted=True))
    vo, eo = op.diag(bjmax=1000, bjtol=1e-5)
    vr, er = rop.diag(bjmax=1000, bjtol=1e-5)
    for a, b in zip(vo, vr):
        assert np.array_equal(a, b)
    assert np.array_equal(eo, er)
    assert np.array_equal(op.px(x), rop.px(x))


def test_problem_legendre_builders_match_reference():
    """scalit_amd.problems' Gauss-Legendre nodes / weights and normalised
    Legendre polynomials (the inputs of every workload's L^2) against the
    reference's Gau_legendre (gauLege.f90:63-103) and YjPolys (yj.f90:16-40).
    Not bit-exact: problems.py iterates all nodes to 1e-15 together (the
    reference each node to 1e-14) and uses exact sqrt(1/2), sqrt(3/2) (the
    reference 14-digit constants)."""
    R = O.ref()
    for n in (24, 40, 80, 96, 128):
        x, w = P.gauss_legendre(n)
        xr, wr = np.zeros(n), np.zeros(n)
        R.gau_legendre_(O._cd(-1), O._cd(1), O._ci(n), O.dptr(xr), O.dptr(wr))
        assert np.abs(x - xr).max() <= 1e-15 and np.abs(w - wr).max() <= 1e-14
        y = P.normalized_legendre(xr, n - 1)
        yr = np.zeros((n, n), order="F")
        R.yjpolys_(O._ci(n), O._ci(n - 1), O.dptr(xr), O.dptr(yr))
        assert np.abs(y - yr).max() <= 1e-12 * np.abs(yr).max()


@pytest.mark.parametrize("cnt,de", [(1, 1e-6), (16, 1e-6), (16, 5.0), (60, 2e-3)])
def test_osbw_setup_matches_reference(cnt, de):
    """The window set-up of the reference's own module osb (adjustWinSize,
    getOSBWIndex, getVOSBH0 with HijX = MYHX(TA0)) against the oracle's
    restatement: width and indices bit-exact, H_W to 1e-13 of its largest
    entry (the oracle forms v_i . H v_j as one matrix product), every window
    state's vector (getVi) bit-exact."""
    prob = P.config_problem("pist_fp64", grid=(12, 10, 8))
    op, rop = make_ops(prob, e0=0.0)
    _, eig0 = op.diag()
    rop.diag()
    srt = np.sort(eig0)
    e0 = float(0.5 * (srt[40] + srt[41]))
    ind, de2, hw, _, vw = O.osbw_setup(op, e0, de, cnt)
    rind, rde, rhw = rop.osbw(cnt, e0, de)
    assert de2 == rde and np.array_equal(ind, rind)
    assert np.abs(hw - rhw).max() <= 1e-13 * np.abs(rhw).max()
    for j in range(0, cnt, max(1, cnt // 5)):
        assert np.array_equal(vw[:, j], rop.osb_vector(int(ind[j])))


def _cx_ops(grid=(8, 7, 6), e0=None):
    prob = P.config_problem("pist_fp64", grid=grid)
    op, rop = make_ops(prob, e0=0.0)
    _, eig0 = op.diag()
    rop.diag()
    srt = np.sort(eig0)
    e = float(0.5 * (srt[30] + srt[31])) if e0 is None else e0
    op.set_shift(e, float(srt[40] - srt[30]), 0.3)
    rop.set_shift(e, float(srt[40] - srt[30]), 0.3)
    aps = (P.absorbing_potential(prob, 0.02), P.absorbing_potential(prob, 0.015, 0.6),
           P.absorbing_potential(prob, 0.01, 0.5, coord="r"))
    for k, a in enumerate(aps):
        rop.set_ap(k, a)
    return prob, op, rop, aps


@pytest.mark.parametrize("ntype", [O.TA0, O.TA1, O.TAAP0, O.TMAP0, O.TAAP, O.TAAPP, O.TAAPR, O.TMAP, O.TMAPP,
                                   O.TMAPR])
def test_hx_dx_matches_reference(ntype):
    """MYHX_DX of the compiled module osb (osb.hxdx.h:177-240) against the
    oracle's restatement, every nType, to 1e-14 of the result (the oracle
    adds the diagonal after the level products)."""
    prob, op, rop, aps = _cx_ops()
    rng = P.rng(5)
    x = rng.standard_normal(op.N) + 1j * rng.standard_normal(op.N)
    y, yr = O.hx_dx(op, x, ntype, aps), rop.hx_dx(x, ntype)
    assert np.abs(y - yr).max() <= 1e-14 * np.abs(yr).max()


@pytest.mark.parametrize("kind", [O.TOSB, O.TOSBD1, O.TOSBD2])
@pytest.mark.parametrize("npc", [O.TA0, O.TA1, O.TAAP0, O.TMAP0, O.TAAP, O.TAAPP, O.TAAPR, O.TMAP, O.TMAPP,
                                 O.TMAPR])
def test_px_dx_matches_reference(kind, npc):
    """MYPX_DX of the compiled module osb (osb.mypxdx.h) against the oracle's
    restatement (dense OSB basis, complex divisor), every kernel and nPC,
    to 1e-12 of the result; E0 inside the EIG0 spectrum and DE wide enough
    that the OSBD branches run."""
    prob, op, rop, aps = _cx_ops()
    rng = P.rng(6)
    x = rng.standard_normal(op.N) + 1j * rng.standard_normal(op.N)
    y, yr = O.px_dx(op, x, kind, npc, aps), rop.px_dx(x, kind, npc)
    assert np.abs(y - yr).max() <= 1e-12 * np.abs(yr).max()


@pytest.mark.parametrize("ntype", [-1, 0, -2])
@pytest.mark.parametrize("pc", [O.TMAP, O.TA1])
def test_qmr_cx_matches_reference(ntype, pc):
    """QMR_CX (qmrcx.f90) as OSB_QMRDX runs it (HX_DX = MYHX_DX(TMAP), PX_DX =
    MYPX_DX(TOSB, pc)), compiled, against the oracle's restatement driven by
    the same compiled operators (so that only the recurrence differs) and by
    the oracle's own: the same return code for the criteria |R|/|B| and
    |HX - B|/|HX|; for the full criterion (-2), whose |UR|/|R| term is
    rounding noise once R is near the tolerance, within 5 iterations;
    solutions to 1e-8."""
    prob, op, rop, aps = _cx_ops()
    rng = P.rng(7)
    b = rng.standard_normal(op.N) + 1j * rng.standard_normal(op.N)
    r_ref, x_ref, _ = rop.qmr_dx(b, O.TOSB, O.TMAP, pc, ntype, 400, 1e-9)
    slack = 0 if ntype != -2 else 5
    for hx, px in ((lambda v: rop.hx_dx(v, O.TMAP), lambda v: rop.px_dx(v, O.TOSB, pc)),
                   (lambda v: O.hx_dx(op, v, O.TMAP, aps), lambda v: O.px_dx(op, v, O.TOSB, pc, aps))):
        r, x, _ = O.qmr_cx(hx, px, b, ntype, 400, 1e-9)
        assert r_ref > 0 and abs(r - r_ref) <= slack, (r, r_ref)
        assert np.abs(x - x_ref).max() <= 1e-8 * np.abs(x_ref).max()


def test_lan_hij_dx_and_mgs_orth_sx_match_reference():
    """lanHij_dx (lanHij_dx.f90:59-97) and MGS_ORTH_SX (gsorth_sx.f90:47-74),
    compiled, against the oracle's restatements: the Lanczos coefficients of
    a Hermitian operator of rank 12 for 8 steps to 1e-12, and one complex-
    symmetric Gram-Schmidt step to 1e-13."""
    rng = np.random.default_rng(3)
    n = 50
    b = rng.standard_normal((n, 12)) + 1j * rng.standard_normal((n, 12))
    h = b @ b.conj().T / 40.0
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    code_r, a_r, b_r = O.ref_lanhij_dx(lambda v: h @ v, x, 8)
    code, a, bb = O.lan_hij_dx(lambda v: h @ v, x, 8)
    assert code == code_r == 8
    assert np.abs(a - a_r).max() <= 1e-12 * np.abs(a_r).max() and np.abs(bb - b_r).max() <= 1e-12 * np.abs(b_r).max()
    vj = np.zeros((n, 4), dtype=np.complex128, order="F")
    vs = []
    for k in range(3):
        v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        assert O.mgs_orth_sx(v, vs)
        vj[:, k] = vs[-1]
    w = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert O.ref_mgs_orth_sx(w, vj, 3)
    assert O.mgs_orth_sx(w, vs)
    assert np.abs(vj[:, 3] - vs[3]).max() <= 1e-13
    # MGS_ORTH_CX (gsorth_cx.f90:51-79), the Hermitian twin
    vh = []
    vj = np.zeros((n, 4), dtype=np.complex128, order="F")
    for k in range(3):
        assert O.mgs_orth_cx(rng.standard_normal(n) + 1j * rng.standard_normal(n), vh)
        vj[:, k] = vh[-1]
    assert O.ref_mgs_orth_cx(w, vj, 3) and O.mgs_orth_cx(w, vh)
    assert np.abs(vj[:, 3] - vh[3]).max() <= 1e-13


# ---------------------------------------------------------------- sCX = T
def _scx_ops(grid=(6, 7, 5), ndvr=False):
    prob = P.config_problem("pist_fp64", grid=grid)
    outh = P.outer_layer(prob, 7) if ndvr else None
    parts = P.complex_levels(prob, 0.05, 1e-3, 11, outh)
    opc = O.OperatorCX(prob.sN, parts[0], prob.res, parts[1], outh=parts[2] if ndvr else None)
    srt = np.sort(prob.res)
    opc.set_shift(float(srt[len(srt) // 4]), 0.01, 0.3)
    ropc = O.RefOperatorCX(opc)
    aps = (P.absorbing_potential(prob, 0.02), P.absorbing_potential(prob, 0.015, 0.6),
           P.absorbing_potential(prob, 0.01, 0.5, coord="r"))
    for k, a in enumerate(aps):
        ropc.set_ap(k, a)
    return prob, opc, ropc, aps


@pytest.mark.parametrize("ndvr", [False, True])
@pytest.mark.parametrize("ntype", [O.TA0, O.TA1, O.TAAP0, O.TMAP0, O.TAAP, O.TAAPP, O.TAAPR, O.TMAP, O.TMAPP,
                                   O.TMAPR])
def test_hx_cx_matches_reference(ntype, ndvr):
    """MYHX_CX of the compiled module osb with sCX = T (osb.hxcx.h:191-245)
    against the oracle's restatement, every nType, with and without the
    sNDVR layer, to 1e-14 of the result.  The serial MYHX_CX adds -i APP
    under TMAPR (osb.hxcx.h:234-235) where the parallel one adds -i APR
    (mosb3.hxcx.h:92-94, which the oracle follows): there the compiled
    serial routine is compared with the oracle's TMAPP."""
    prob, opc, ropc, aps = _scx_ops(ndvr=ndvr)
    rng = P.rng(15)
    x = rng.standard_normal(opc.N) + 1j * rng.standard_normal(opc.N)
    y = opc.hx(x, O.TMAPP if ntype == O.TMAPR else ntype, aps)
    yr = ropc.hx(x, ntype)
    assert np.abs(y - yr).max() <= 1e-14 * np.abs(yr).max()
    if ntype == O.TMAPR:
        assert np.abs(opc.hx(x, ntype, aps) - yr).max() > 1e-6 * np.abs(yr).max()


@pytest.mark.parametrize("grid,ndvr", [((6, 7, 5), False), ((9, 8, 7), False), ((6, 7, 5), True)])
def test_block_diag_cx_matches_reference(grid, ndvr):
    """The sCX branch of myDiag (osb.diag.h:76-101: HOSBINIT_*_CX, HosbDiagDX
    with MJACOBI_CX and sumOffDiagCX), compiled, against the oracle's
    restatement: EIG0 to 1e-13 and VOSB to 1e-11 (same rotation sequence
    and sweep counts; the sums of A, B, C differ in order only).  The
    oracle's off-diagonal measure follows sumOffDiagCX reading the complex
    block through its double precision dummy; with the complex magnitudes
    instead the sweep counts change and VOSB moves by about 1e-5, which this
    bound would catch."""
    prob, opc, ropc, aps = _scx_ops(grid, ndvr)
    vr, er = ropc.diag(1000, 1e-5)
    v, e = opc.diag(1000, 1e-5)
    assert np.abs(e - er).max() <= 1e-13 * np.abs(er).max()
    for a, b in zip(v, vr):
        assert np.abs(a - b).max() <= 1e-11
