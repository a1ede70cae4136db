"""Synthetic triatomic operators for the benchmark workloads.

A J=0 triatomic in Jacobi coordinates (r, R, theta) on a direct-product DVR:

* level 1 = r, level 2 = R: Colbert-Miller radial sinc-DVR kinetic matrices,
  the [0, inf) form of src/ham/libham/dvr/sinc.f90:67-109 (H1_SINC2) with the
  grid of src/ham/libham/dvr/sinc_xh.f90:50-67 (Sinc2_DVR);
* level 3 = theta: L^2 in the Gauss-Legendre DVR (nodes by Newton iteration as
  src/ham/libham/poly/gauLege.f90:63-103, normalised Legendre recursion as
  src/ham/libham/poly/yj.f90:16-40), multiplied by the coordinate-dependent
  factor DEP(r,R) = 1/(2 mu_r r^2) + 1/(2 mu_R R^2);
* RES = V(r,R,theta) on the full grid (diagonal).

This is ScalIT's iterate-module operator with sF=3, sDEP=(F,F,T), sNDVR=F
(src/iterate/module/osb.hx.h:71-110).  The arrays are generated on the host:
they are the operator's input files, as ScalIT's step-2 generator writes them.

Every random draw uses numpy's Philox counter-based generator keyed by the
workload seed, so the operator and the input vectors are reproducible.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

MU = 7292.0
R_RANGE = (1.5, 4.5)
BIGR_RANGE = (2.0, 8.0)
MORSE_R = (0.2, 1.1, 2.3)      # De, a, re
MORSE_BIGR = (0.01, 0.8, 5.0)
P2_AMP = 0.005


def rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(key=seed))


# ----------------------------------------------------------------------------
# 1-D DVR pieces
# ----------------------------------------------------------------------------
def sinc2_dvr(n: int, a0: float, b0: float, mass: float):
    """Radial sinc-DVR grid and kinetic matrix in hartree (H1_SINC2 form).

    dx = (b0-a0)/n, i0 = int(a0/dx), x_i = dx*(i0+i), and
    T_ij = (-1)^(i-j) [2/(i-j)^2 - 2/(i+j+2 i0)^2] / (2 m dx^2),
    T_ii = [pi^2/3 - 1/(2 (i+i0)^2)] / (2 m dx^2)."""
    xmin, xmax = min(a0, b0), max(a0, b0)
    dx = (xmax - xmin) / n
    i0 = int(xmin / dx)
    i = np.arange(1, n + 1)
    x = dx * (i0 + i)
    ii, jj = np.meshgrid(i, i, indexing="ij")
    d = ii - jj
    s = ii + jj + 2 * i0
    with np.errstate(divide="ignore"):
        t = 2.0 / (d.astype(np.float64) ** 2) - 2.0 / (s.astype(np.float64) ** 2)
    t = np.where(d % 2 != 0, -t, t)
    diag = math.pi ** 2 / 3.0 - 0.5 / ((i + i0).astype(np.float64) ** 2)
    t[np.diag_indices(n)] = diag
    t *= 0.5 / (dx * dx * mass)
    return x, np.asfortranarray(t)


def gauss_legendre(n: int):
    """Gauss-Legendre nodes (ascending) and weights on [-1, 1] by Newton iteration."""
    m = (n + 1) // 2
    i = np.arange(1, m + 1)
    z = np.cos(math.pi * (i - 0.25) / (n + 0.5))
    for _ in range(100):
        p1 = np.ones_like(z)
        p2 = np.zeros_like(z)
        for j in range(1, n + 1):
            p3 = p2
            p2 = p1
            p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j
        pp = n * (z * p1 - p2) / (z * z - 1.0)
        z1 = z
        z = z1 - p1 / pp
        if np.all(np.abs(z - z1) <= 1e-15):
            break
    # one more evaluation of pp at the converged nodes
    p1 = np.ones_like(z)
    p2 = np.zeros_like(z)
    for j in range(1, n + 1):
        p3 = p2
        p2 = p1
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j
    pp = n * (z * p1 - p2) / (z * z - 1.0)
    x = np.zeros(n)
    w = np.zeros(n)
    x[:m] = -z
    x[n - m:] = z[::-1]
    wi = 2.0 / ((1.0 - z * z) * pp * pp)
    w[:m] = wi
    w[n - m:] = wi[::-1]
    return x, w


def normalized_legendre(x: np.ndarray, lmax: int) -> np.ndarray:
    """sqrt(l+1/2) P_l(x), l = 0..lmax, columns."""
    y = np.zeros((len(x), lmax + 1))
    y[:, 0] = math.sqrt(0.5)
    if lmax >= 1:
        y[:, 1] = math.sqrt(1.5) * x
    for i in range(3, lmax + 2):
        re0 = 1.0 / (i - 1.0)
        re1 = math.sqrt((2 * i - 1.0) * (2 * i - 3.0)) * re0
        re2 = math.sqrt((2 * i - 1.0) / (2 * i - 5.0)) * (1.0 - re0)
        y[:, i - 1] = re1 * x * y[:, i - 2] - re2 * y[:, i - 3]
    return y


def legendre_dvr_l2(n: int):
    """cos(theta) nodes and L^2 = T diag(j(j+1)) T^T, T_ij = sqrt(w_i) P~_j(x_i)."""
    x, w = gauss_legendre(n)
    t = np.sqrt(w)[:, None] * normalized_legendre(x, n - 1)
    j = np.arange(n, dtype=np.float64)
    l2 = (t * (j * (j + 1.0))[None, :]) @ t.T
    l2 = 0.5 * (l2 + l2.T)
    return x, np.asfortranarray(l2)


# ----------------------------------------------------------------------------
# potentials
# ----------------------------------------------------------------------------
@dataclasses.dataclass
class Potential:
    morse_r: tuple            # De, a, re
    morse_R: tuple
    p2: float
    gauss: np.ndarray          # (k, 4): amplitude, r0, R0, width

    def params(self) -> np.ndarray:
        """The 12 differentiable parameters: Morse De/a/re (r), (R), first 6 Gaussian amplitudes."""
        g = list(self.gauss[:6, 0]) + [0.0] * max(0, 6 - len(self.gauss))
        return np.array(list(self.morse_r) + list(self.morse_R) + g[:6], dtype=np.float64)

    def with_params(self, p: np.ndarray) -> "Potential":
        g = self.gauss.copy()
        k = min(6, len(g))
        g[:k, 0] = p[6:6 + k]
        return Potential(tuple(p[0:3]), tuple(p[3:6]), self.p2, g)

    def evaluate(self, r: np.ndarray, R: np.ndarray, c: np.ndarray) -> np.ndarray:
        """V on the grid, flattened with r fastest."""
        rr = r[:, None, None]
        RR = R[None, :, None]
        cc = c[None, None, :]
        De, a, re = self.morse_r
        v = De * (1.0 - np.exp(-a * (rr - re))) ** 2
        De, a, re = self.morse_R
        v = v + De * (1.0 - np.exp(-a * (RR - re))) ** 2
        v = v + self.p2 * 0.5 * (3.0 * cc * cc - 1.0)
        for amp, r0, R0, wd in self.gauss:
            v = v + amp * np.exp(-((rr - r0) ** 2 + (RR - R0) ** 2) / (2.0 * wd * wd))
        v = np.broadcast_to(v, (len(r), len(R), len(c)))
        return np.ascontiguousarray(v.transpose(2, 1, 0)).reshape(-1)

    def gradient(self, r: np.ndarray, R: np.ndarray, c: np.ndarray) -> np.ndarray:
        """dV/dp for params(), shape (12, N), r fastest."""
        rr = r[:, None, None]
        RR = R[None, :, None]
        shape = (len(r), len(R), len(c))
        out = []

        def morse_grad(q, De, a, re):
            e = np.exp(-a * (q - re))
            return [(1 - e) ** 2, 2 * De * (1 - e) * e * (q - re), -2 * De * (1 - e) * e * a]

        for g in morse_grad(rr, *self.morse_r):
            out.append(np.broadcast_to(g, shape))
        for g in morse_grad(RR, *self.morse_R):
            out.append(np.broadcast_to(g, shape))
        for k in range(6):
            if k < len(self.gauss):
                _, r0, R0, wd = self.gauss[k]
                out.append(np.broadcast_to(np.exp(-((rr - r0) ** 2 + (RR - R0) ** 2) / (2.0 * wd * wd)), shape))
            else:
                out.append(np.zeros(shape))
        return np.stack([np.ascontiguousarray(o.transpose(2, 1, 0)).reshape(-1) for o in out])


def base_potential() -> Potential:
    return Potential(MORSE_R, MORSE_BIGR, P2_AMP, np.zeros((0, 4)))


def jittered_potential(seed: int, frac: float = 0.05) -> Potential:
    """Morse + P2 with every parameter scaled by 1 + U(-frac, frac)."""
    g = rng(seed)
    j = 1.0 + g.uniform(-frac, frac, size=7)
    mr = tuple(np.array(MORSE_R) * j[0:3])
    mR = tuple(np.array(MORSE_BIGR) * j[3:6])
    return Potential(mr, mR, P2_AMP * j[6], np.zeros((0, 4)))


def gaussian_potential(seed: int, k: int = 8) -> Potential:
    """Morse base + P2 + k Gaussians in (r, R): amplitude U(-0.02, 0), width
    U(0.3, 1.0) bohr, centres uniform over the grid box."""
    g = rng(seed)
    amp = g.uniform(-0.02, 0.0, size=k)
    wd = g.uniform(0.3, 1.0, size=k)
    r0 = g.uniform(*R_RANGE, size=k)
    R0 = g.uniform(*BIGR_RANGE, size=k)
    return Potential(MORSE_R, MORSE_BIGR, P2_AMP, np.stack([amp, r0, R0, wd], axis=1))


# ----------------------------------------------------------------------------
# problems
# ----------------------------------------------------------------------------
@dataclasses.dataclass
class Problem:
    name: str
    sN: tuple
    h0: list                   # [T_r, T_R, L^2]
    dep: list                  # [None, None, DEP(r,R)]
    res: np.ndarray            # V on the grid
    r: np.ndarray
    R: np.ndarray
    cos_theta: np.ndarray
    potential: Potential
    seed: int

    @property
    def N(self) -> int:
        return int(np.prod(self.sN))

    def with_potential(self, pot: Potential) -> "Problem":
        return dataclasses.replace(self, potential=pot, res=pot.evaluate(self.r, self.R, self.cos_theta))


def build_problem(n1: int, n2: int, n3: int, pot: Potential, name: str = "custom", seed: int = 0) -> Problem:
    r, tr = sinc2_dvr(n1, *R_RANGE, MU)
    R, tR = sinc2_dvr(n2, *BIGR_RANGE, MU)
    c, l2 = legendre_dvr_l2(n3)
    dep = (1.0 / (2.0 * MU * r[:, None] ** 2) + 1.0 / (2.0 * MU * R[None, :] ** 2))
    dep = np.ascontiguousarray(dep.T).reshape(-1)   # index i1 + n1*i2
    res = pot.evaluate(r, R, c)
    return Problem(name, (n1, n2, n3), [tr, tR, l2], [None, None, dep], res, r, R, c, pot, seed)


CONFIG_GRIDS = {
    "pist_fp64": (24, 32, 24),
    "hx_bench": (64, 96, 80),
    "mixed": (128, 128, 96),
    "batched_a": (48, 64, 40),
    "batched_b": (80, 96, 64),
    "batched_c": (128, 128, 96),
    "adjoint": (64, 96, 80),
}


def config_problem(name: str, seed: int = 1234, grid=None) -> Problem:
    """The workloads: 'pist_fp64' (jittered Morse+P2), the others Morse+P2+8 Gaussians."""
    n1, n2, n3 = grid if grid is not None else CONFIG_GRIDS[name]
    if name == "pist_fp64":
        pot = jittered_potential(seed)
    else:
        pot = gaussian_potential(seed)
    return build_problem(n1, n2, n3, pot, name=name, seed=seed)


# ----------------------------------------------------------------------------
# sequential truncation
# ----------------------------------------------------------------------------
@dataclasses.dataclass
class TruncatedProblem:
    """The triatomic in a truncated basis: PO-DVR contraction of r and R
    (m1 and m2 points), then the energy-cut combined (r, R) layer of nkeep
    points as level 1 and theta as level 2 with DEP on it (sF = 2,
    sDEP = (F, T)): ScalIT's iterate operator for a ja3-type basis."""
    name: str
    sN: tuple                  # (nkeep, n3)
    h0: list                   # [H0 of the combined layer, L^2]
    dep: list                  # [None, DEP at the kept points]
    res: np.ndarray            # V - v_r - v_R on (kept point, theta), point fastest
    r: np.ndarray              # r and R of each kept point
    R: np.ndarray
    cos_theta: np.ndarray
    potential: Potential
    seed: int
    raw: tuple                 # the primitive grid (n1, n2, n3)
    contracted: tuple          # (m1, m2)

    @property
    def N(self) -> int:
        return int(np.prod(self.sN))

    def with_potential(self, pot: Potential) -> "TruncatedProblem":
        return dataclasses.replace(self, potential=pot, res=_residual(pot, self.r, self.R, self.cos_theta))


def morse_1d(q: np.ndarray, De: float, a: float, re: float) -> np.ndarray:
    return De * (1.0 - np.exp(-a * (q - re))) ** 2


def _residual(pot: Potential, r: np.ndarray, R: np.ndarray, c: np.ndarray) -> np.ndarray:
    """V(r_a, R_a, theta_k) - v_r(r_a) - v_R(R_a), kept point a fastest."""
    v = np.empty((len(c), len(r)))
    for k, ck in enumerate(c):
        v[k] = _point_potential(pot, r, R, ck)
    v -= (morse_1d(r, *pot.morse_r) + morse_1d(R, *pot.morse_R))[None, :]
    return np.ascontiguousarray(v).reshape(-1)


def _point_potential(pot: Potential, r: np.ndarray, R: np.ndarray, c: float) -> np.ndarray:
    v = morse_1d(r, *pot.morse_r) + morse_1d(R, *pot.morse_R) + pot.p2 * 0.5 * (3.0 * c * c - 1.0)
    for amp, r0, R0, wd in pot.gauss:
        v = v + amp * np.exp(-((r - r0) ** 2 + (R - R0) ** 2) / (2.0 * wd * wd))
    return v


def truncated_problem(n1: int, n2: int, n3: int, pot: Potential, m1: int, m2: int, nkeep: int, podvr, combine_2d,
                      name: str = "truncated", seed: int = 0) -> TruncatedProblem:
    """Sequential truncation of the n1 x n2 x n3 triatomic: the PO-DVR of r
    and R against the 1-D Morse reference potentials v_r, v_R of `pot`
    (DVR_Step2/3), then the nkeep (r, R) product points of lowest
    v_r + v_R as one combined layer (saveH0CBALL).  `podvr(x, t, v, m)` and
    `combine_2d(h1, v1, h2, v2, nkeep)` do the two steps (the library's
    scalit_podvr / scalit_combine_2d, see `library_truncation`)."""
    r, tr = sinc2_dvr(n1, *R_RANGE, MU)
    R, tR = sinc2_dvr(n2, *BIGR_RANGE, MU)
    c, l2 = legendre_dvr_l2(n3)
    xr, hr, _ = podvr(r, tr, morse_1d(r, *pot.morse_r), m1)
    xR, hR, _ = podvr(R, tR, morse_1d(R, *pot.morse_R), m2)
    i1, i2, hc = combine_2d(hr, morse_1d(xr, *pot.morse_r), hR, morse_1d(xR, *pot.morse_R), nkeep)
    rk, Rk = xr[i1], xR[i2]
    dep = 1.0 / (2.0 * MU * rk ** 2) + 1.0 / (2.0 * MU * Rk ** 2)
    res = _residual(pot, rk, Rk, c)
    return TruncatedProblem(name, (nkeep, n3), [np.asfortranarray(hc), l2], [None, dep], res, rk, Rk, c, pot, seed,
                            (n1, n2, n3), (m1, m2))


def library_truncation():
    """(podvr, combine_2d) through libscalit_hip.so (the eigensolves run on
    the current GPU)."""
    import ctypes as C

    from . import _lib as L

    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)

    def podvr(x, t, v, m):
        x = np.ascontiguousarray(x, dtype=np.float64)
        t = np.asfortranarray(t, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        x0, h0, e = np.zeros(m), np.zeros((m, m), order="F"), np.zeros(m)
        L.check(L.lib().scalit_podvr(len(x), x.ctypes.data_as(dp), t.ctypes.data_as(dp), v.ctypes.data_as(dp), m,
                                     x0.ctypes.data_as(dp), h0.ctypes.data_as(dp), e.ctypes.data_as(dp)),
                "scalit_podvr")
        return x0, h0, e

    def combine_2d(h1, v1, h2, v2, nkeep):
        h1 = np.asfortranarray(h1, dtype=np.float64)
        h2 = np.asfortranarray(h2, dtype=np.float64)
        v1 = np.ascontiguousarray(v1, dtype=np.float64)
        v2 = np.ascontiguousarray(v2, dtype=np.float64)
        i1, i2 = np.zeros(nkeep, dtype=np.int32), np.zeros(nkeep, dtype=np.int32)
        hc = np.zeros((nkeep, nkeep), order="F")
        L.check(L.lib().scalit_combine_2d(len(v1), h1.ctypes.data_as(dp), v1.ctypes.data_as(dp), len(v2),
                                          h2.ctypes.data_as(dp), v2.ctypes.data_as(dp), nkeep, i1.ctypes.data_as(ip),
                                          i2.ctypes.data_as(ip), hc.ctypes.data_as(dp)), "scalit_combine_2d")
        return i1, i2, hc

    return podvr, combine_2d


def outer_layer(prob: Problem, seed: int, eps: float = 2e-4, potential: bool = False) -> np.ndarray:
    """A dense per-point outermost layer (the sNDVR = T form ScalIT's ja3 writes,
    OUTH(NIN, n3, n3) of src/iterate/module/osb.data.h:62): for every inner grid
    point p = (r, R), A_p = DEP(p) L^2 + eps S_p with S_p a seeded symmetric
    N(0,1) matrix, so every block stays symmetric.  With `potential` the grid
    potential V(p, theta) is added on each block's diagonal (the operator is
    then used with RES = 0: the OSB block of this level is OUTH itself,
    src/iterate/module/osb.diag.h:22-23, so the potential must live in OUTH for
    the preconditioner to see it).  Returned flat, column-major (p fastest)."""
    n3 = prob.sN[2]
    nin = prob.sN[0] * prob.sN[1]
    g = rng(seed)
    s = g.standard_normal((nin, n3, n3))
    s = 0.5 * (s + s.transpose(0, 2, 1))
    a = prob.dep[2][:, None, None] * prob.h0[2][None, :, :] + eps * s
    if potential:
        v = prob.res.reshape(n3, nin).T          # V(p, theta)
        a[:, np.arange(n3), np.arange(n3)] += v
    return a.reshape(-1, order="F")   # a[p, i, j] at p + NIN (i + n3 j)


def absorbing_potential(prob: Problem, strength: float = 0.02, start: float = 0.7, coord: str = "R") -> np.ndarray:
    """A quadratic absorbing potential on the grid, strength ((x - x0) /
    (x_max - x0))^2 beyond x0 = x_min + start (x_max - x_min) along R (or r):
    the AP / APP / APR arrays of a resonance or CRP job (which the reference
    loads from files, src/iterate/module/osb.io.h:20-30)."""
    x = prob.R if coord == "R" else prob.r
    lo, hi = float(x.min()), float(x.max())
    x0 = lo + start * (hi - lo)
    a1 = np.where(x > x0, strength * ((x - x0) / (hi - x0)) ** 2, 0.0)
    n1, n2, n3 = prob.sN
    full = a1[None, :, None] if coord == "R" else a1[None, None, :]
    return np.ascontiguousarray(np.broadcast_to(full, (n3, n2, n1))).reshape(-1)


def complex_levels(prob: Problem, theta: float = 0.05, eps: float = 1e-3, seed: int = 11, outh=None):
    """Complex level matrices of an sCX = T job (H0CX, DEPCX and OUTHCX of
    src/iterate/module/osb.data.h:63, which the reference loads from its
    files): every level's kinetic matrix rotated by exp(-2 i theta), as
    complex scaling of the coordinates makes it, plus i eps S_l with S_l a
    seeded symmetric N(0,1) matrix scaled by max |H0_l|; DEP (and OUTH) times
    exp(-2 i theta).  Every matrix stays complex symmetric, the form the
    sCX set-up's real rotations assume.  Returns (h0 list, dep list[, outh])."""
    g = rng(seed)
    rot = np.exp(-2j * theta)
    h0 = []
    for h in prob.h0:
        h = np.asarray(h, dtype=np.float64)
        b = g.standard_normal(h.shape)
        h0.append(rot * h + 1j * eps * np.abs(h).max() * 0.5 * (b + b.T))
    dep = [None if d is None else rot * np.asarray(d, dtype=np.float64) for d in prob.dep]
    if outh is None:
        return h0, dep
    return h0, dep, rot * np.asarray(outh, dtype=np.float64)


def random_vectors(n: int, b: int, seed: int) -> np.ndarray:
    """B standard-normal vectors of length n, shape (b, n)."""
    return rng(seed).standard_normal((b, n))
