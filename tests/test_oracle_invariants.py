"""Closed-form pins of the oracle's explicit residuals, solvers and time step.

1. Strong-form identity: on an affine mesh with continuous polynomial states
   whose normal velocity vanishes on the walls, every face jump is zero and
   integration by parts turns the DG residuals of euler_operator.h into
   (value - div flux, phi) integrated exactly. The expectation is built with
   exact polynomial arithmetic (2D coefficient arrays) and 12-point Gauss
   quadrature, independently of the oracle; Ma and Fr differ from 1 so their
   placement is checked.
2. Global telescoping of the density residual on the hill mesh (any state).
3. Krylov solvers against dense direct solves.
4. IMEX tableau order conditions.
"""
import os
import sys

import numpy as np
import pytest
from numpy.polynomial import polynomial as P
from scipy.signal import convolve2d

from oracle.pyoracle import Oracle, lib, ptr, witch_of_agnesi

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import dense_ref as R  # noqa: E402

G = 2.0 - np.sqrt(2.0)
TAB = dict(a21=G, a31=0.5, a32=0.5, a21t=G / 2, a22t=G / 2, a31t=np.sqrt(2) / 4, a32t=np.sqrt(2) / 4,
           a33t=1 - np.sqrt(2) / 2, b1=np.sqrt(2) / 4, b2=np.sqrt(2) / 4, b3=1 - np.sqrt(2) / 2)
MA, FR, DT, GM = 0.7, 0.5, 0.3, 1.4
LX, HZ = 3.0, 2.0


def test_tableau_order_conditions():
    A = np.array([[0, 0, 0], [TAB["a21"], 0, 0], [TAB["a31"], TAB["a32"], 0]])
    At = np.array([[0, 0, 0], [TAB["a21t"], TAB["a22t"], 0], [TAB["a31t"], TAB["a32t"], TAB["a33t"]]])
    b = np.array([TAB["b1"], TAB["b2"], TAB["b3"]])
    c, ct = A.sum(1), At.sum(1)
    assert np.allclose(c, ct)  # stage times agree
    assert abs(b.sum() - 1) < 1e-15 and abs(b @ c - 0.5) < 1e-15
    assert np.allclose(b, At[2])  # stiffly accurate implicit part


# ---------------------------------------------------------- polynomial algebra
class Poly:
    def __init__(self, c):
        self.c = np.atleast_2d(np.asarray(c, dtype=float))

    def __add__(self, o):
        o = o if isinstance(o, Poly) else Poly([[o]])
        n = (max(self.c.shape[0], o.c.shape[0]), max(self.c.shape[1], o.c.shape[1]))
        a, b = np.zeros(n), np.zeros(n)
        a[:self.c.shape[0], :self.c.shape[1]] = self.c
        b[:o.c.shape[0], :o.c.shape[1]] = o.c
        return Poly(a + b)

    __radd__ = __add__

    def __mul__(self, o):
        if not isinstance(o, Poly):
            return Poly(self.c * o)
        return Poly(convolve2d(self.c, o.c))

    __rmul__ = __mul__

    def __sub__(self, o):
        return self + (-1.0) * o

    def __rsub__(self, o):
        return (-1.0) * self + o

    def dx(self):
        return Poly(P.polyder(self.c, axis=0)) if self.c.shape[0] > 1 else Poly([[0.0]])

    def dz(self):
        return Poly(P.polyder(self.c, axis=1)) if self.c.shape[1] > 1 else Poly([[0.0]])

    def __call__(self, x, z):
        return P.polyval2d(x, z, self.c)


def X():
    return Poly([[0.0], [1.0]])


def Z():
    return Poly([[0.0, 1.0]])


def make_state(seed):
    rng = np.random.default_rng(seed)
    x, z = X(), Z()
    rho = 1.0 + 0.1 * rng.uniform() * x + 0.1 * rng.uniform() * z
    U = (1 + rng.uniform()) * x * (LX - x) * (1.0 / LX ** 2)
    W = (0.5 + rng.uniform()) * z * (HZ - z) * (1.0 / HZ ** 2)
    p = 2.0 + 0.3 * rng.uniform() * x - 0.2 * rng.uniform() * z
    return rho, (U, W), p


def nodal(d, f):
    pts = R.node_coords(d)
    return f(pts[:, 0], pts[:, 1])


def project(d, f, nq=12):
    """(f, phi_i) integrated with nq^2 Gauss points per (affine) cell."""
    out = np.zeros(d.Ns)
    xg, wg = R.gauss(nq)
    for c in range(d.nc):
        v = d.m.verts(c)
        for b, eta in enumerate(xg):
            for a, xi in enumerate(xg):
                pt, J = R.bilinear(v, xi, eta)
                ph, _, _ = d.phi(xi, eta)
                out[c * d.nd2:(c + 1) * d.nd2] += f(pt[0], pt[1]) * ph * np.linalg.det(J) * wg[a] * wg[b]
    return out


def div(F):
    return F[0].dx() + F[1].dz()


def setup(k):
    nx, nz = 3, 2
    d = R.Disc(k, R.Mesh(nx, nz, LX, HZ, np.zeros(nx + 1), False))
    o = Oracle(k, nx, nz, Lx=LX, H=HZ, hx=np.zeros(nx + 1), periodic=False)
    o.set_params(Ma=MA, Fr=FR, dt=DT)
    return d, o


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("k", [2, 3])
@pytest.mark.parametrize("stage", [1, 2, 3])
def test_rhs_rho_strong_form(k, stage):
    d, o = setup(k)
    S = [make_state(10 + s) for s in range(stage)]
    w = {1: [TAB["a21"]], 2: [TAB["a31"], TAB["a32"]], 3: [TAB["b1"], TAB["b2"], TAB["b3"]]}[stage]
    strong = S[0][0]
    for s in range(stage):
        rho, (U, W), _ = S[s]
        strong = strong - (w[s] * DT) * div((rho * U, rho * W))
    srcs = []
    for s in range(stage):
        rho, (U, W), _ = S[s]
        srcs += [nodal(d, rho), np.concatenate([nodal(d, U), nodal(d, W)])]
    o.set_stage(stage, 1)
    assert rel(o.rhs_rho(srcs), project(d, strong)) < 1e-12


@pytest.mark.parametrize("k", [2, 3])
@pytest.mark.parametrize("stage", [1, 2, 3])
def test_rhs_pressure_strong_form(k, stage):
    d, o = setup(k)
    S = [make_state(20 + s) for s in range(stage)]
    rk, (Uk, Wk), _ = make_state(99)
    ma2, ifr2 = MA ** 2, 1.0 / FR ** 2
    if stage == 1:
        wa, wt, grav = [TAB["a21"]], [TAB["a21t"]], [(TAB["a21t"], 0), (TAB["a22t"], "k")]
    elif stage == 2:
        wa, wt = [TAB["a31"], TAB["a32"]], [TAB["a31t"], TAB["a32t"]]
        grav = [(TAB["a31t"], 0), (TAB["a32t"], 1), (TAB["a33t"], "k")]
    else:
        wa = wt = [TAB["b1"], TAB["b2"], TAB["b3"]]
        grav = [(TAB["b1"], 0), (TAB["b2"], 1), (TAB["b3"], 2)]
    rho0, (U0, W0), p0 = S[0]
    val = (1.0 / (GM - 1)) * p0 + (0.5 * ma2) * rho0 * (U0 * U0 + W0 * W0) - (0.5 * ma2) * rk * (Uk * Uk + Wk * Wk)
    for wgt, which in grav:
        r, (U, W) = (rk, (Uk, Wk)) if which == "k" else (S[which][0], S[which][1])
        val = val - (wgt * DT * ma2 * ifr2) * r * W  # e_r = (0, 1)
    Fx, Fz = Poly([[0.0]]), Poly([[0.0]])
    for s in range(stage):
        r, (U, W), p = S[s]
        kin = (0.5 * wa[s] * DT * ma2) * (U * U + W * W) * r
        ent = (wt[s] * DT * GM / (GM - 1)) * p
        Fx, Fz = Fx + (kin + ent) * U, Fz + (kin + ent) * W
    strong = val - div((Fx, Fz))
    srcs = []
    for s in range(stage):
        r, (U, W), p = S[s]
        srcs += [nodal(d, r), np.concatenate([nodal(d, U), nodal(d, W)]), nodal(d, p)]
    srcs += [nodal(d, rk), np.concatenate([nodal(d, Uk), nodal(d, Wk)])]
    o.set_stage(stage, 2)
    assert rel(o.rhs_pressure(srcs), project(d, strong)) < 1e-12


@pytest.mark.parametrize("k", [2, 3])
@pytest.mark.parametrize("stage", [1, 2, 3])
def test_rhs_velocity_strong_form(k, stage):
    d, o = setup(k)
    S = [make_state(30 + s) for s in range(stage)]
    rx, _, _ = make_state(77)
    ma2, ifr2 = MA ** 2, 1.0 / FR ** 2
    if stage == 1:
        wa, wt, grav = [TAB["a21"]], [TAB["a21t"]], [(TAB["a21t"], S[0][0]), (TAB["a22t"], rx)]
    elif stage == 2:
        wa, wt = [TAB["a31"], TAB["a32"]], [TAB["a31t"], TAB["a32t"]]
        grav = [(TAB["a31t"], S[0][0]), (TAB["a32t"], S[1][0]), (TAB["a33t"], rx)]
    else:
        wa = wt = [TAB["b1"], TAB["b2"], TAB["b3"]]
        grav = [(TAB["b1"], S[0][0]), (TAB["b2"], S[1][0]), (TAB["b3"], S[2][0])]
    rho0, (U0, W0), _ = S[0]
    expect = []
    for comp in range(2):
        val = rho0 * (U0 if comp == 0 else W0)
        if comp == 1:
            for wgt, r in grav:
                val = val - (wgt * DT * ifr2) * r
        Tx, Tz = Poly([[0.0]]), Poly([[0.0]])
        for s in range(stage):
            r, (U, W), p = S[s]
            uc = U if comp == 0 else W
            Tx = Tx + (wa[s] * DT) * r * uc * U + ((wt[s] * DT / ma2) * p if comp == 0 else 0.0)
            Tz = Tz + (wa[s] * DT) * r * uc * W + ((wt[s] * DT / ma2) * p if comp == 1 else 0.0)
        expect.append(project(d, val - div((Tx, Tz))))
    srcs = []
    for s in range(stage):
        r, (U, W), p = S[s]
        srcs += [nodal(d, r), np.concatenate([nodal(d, U), nodal(d, W)]), nodal(d, p)]
    if stage < 3:
        srcs += [nodal(d, rx)]
    o.set_stage(stage, 3)
    assert rel(o.rhs_velocity(srcs), np.concatenate(expect)) < 1e-12


@pytest.mark.parametrize("stage", [1, 2, 3])
def test_rhs_rho_telescopes_on_hill_mesh(stage):
    """sum_i rhs_i = integral of rho^n for any state (fluxes cancel pairwise,
    walls carry no flux)."""
    k, nx, nz = 3, 8, 4
    o = Oracle(k, nx, nz, hx=witch_of_agnesi(nx))
    o.set_params(dt=5.0)
    rng = np.random.default_rng(stage)
    srcs = []
    for s in range(stage):
        srcs += [1 + 0.2 * rng.uniform(-1, 1, o.Ns), rng.normal(size=2 * o.Ns) * 10]
    o.set_stage(stage, 1)
    rhs = o.rhs_rho(srcs)
    mass = o.vmult(srcs[0], 1).sum()
    assert abs(rhs.sum() - mass) < 1e-12 * abs(mass)


def test_dense_cg_and_gmres_match_direct_solves():
    rng = np.random.default_rng(0)
    n = 40
    Q = rng.normal(size=(n, n))
    A = Q @ Q.T + n * np.eye(n)
    b = rng.normal(size=n)
    x = np.zeros(n)
    its = lib().orc_dense_cg(n, ptr(A.ravel().copy()), ptr(b), ptr(x), 1e-13 * np.linalg.norm(b), 1000, 1)
    assert its > 0 and rel(x, np.linalg.solve(A, b)) < 1e-11
    N = A + 5 * rng.normal(size=(n, n))  # nonsymmetric
    x = np.zeros(n)
    its = lib().orc_dense_gmres(n, ptr(N.ravel().copy()), ptr(b), ptr(x), 1e-14 * np.linalg.norm(b), 2000, 10)
    assert its > 0 and rel(x, np.linalg.solve(N, b)) < 1e-10


def test_mg_prolongation_reproduces_polynomials():
    """DG embedding prolongation is exact for coarse-cell polynomials and the
    L2 projection down inverts it."""
    k, nx, nz = 3, 8, 4
    o = Oracle(k, nx, nz, hx=np.zeros(nx + 1), mg_min_cells=2)
    L = o.nlevels
    nd2 = (k + 1) ** 2
    nc = lib().orc_mg_level_cells(o._h, L - 2)
    coarse = np.random.default_rng(1).normal(size=nc * nd2)
    fine = np.zeros(lib().orc_mg_level_cells(o._h, L - 1) * nd2)
    lib().orc_prolongate(o._h, L - 1, ptr(fine), ptr(coarse))
    back = np.zeros_like(coarse)
    lib().orc_project_down(o._h, L - 1, ptr(back), ptr(fine))
    assert rel(back, coarse) < 1e-13


def test_mass_mg_solve_converges_and_matches_dense():
    k, nx, nz = 2, 8, 4
    o = Oracle(k, nx, nz, hx=witch_of_agnesi(nx), mg_min_cells=2)
    o.set_stage(1, 1)
    b = np.random.default_rng(2).normal(size=o.Ns)
    x, its = o.solve_mass_mg(1, b, tol=1e-13 * np.linalg.norm(b))
    assert 0 < its < 20
    assert rel(o.vmult(x, 1), b) < 1e-12


def test_imex_step_conserves_mass_and_rest_state():
    k, nx, nz = 2, 8, 4
    o = Oracle(k, nx, nz, hx=witch_of_agnesi(nx))
    o.set_params(dt=5.0, eps=1e-13, schur_inner_tol=1e-14, a_solve_tol=1e-14)
    rho, u, p = o.init_hill()
    r1, u1, p1, st = o.run(rho, u, p, 2)
    o.set_stage(1, 1)
    m0, m1 = o.vmult(rho, 1).sum(), o.vmult(r1, 1).sum()
    assert abs(m1 - m0) < 1e-11 * m0
    assert st["fp_its"] >= 4 and st["gmres_its"] > 0
    # a flat-bottom fluid at rest with constant density and pressure and no
    # gravity stays at rest
    o2 = Oracle(k, nx, nz, hx=np.zeros(nx + 1))
    o2.set_params(Fr=1e8, dt=5.0)
    r2, u2, p2, _ = o2.run(np.ones(o2.Ns), np.zeros(2 * o2.Ns), np.full(o2.Ns, 1e5), 1)
    assert np.abs(u2).max() < 1e-9 and np.abs(r2 - 1).max() < 1e-12 and np.abs(p2 / 1e5 - 1).max() < 1e-12
