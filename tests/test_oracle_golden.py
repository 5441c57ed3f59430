"""Pins the CPU oracle against the golden fixtures written by the independent
numpy brute-force assembly (tests/golden/make_golden.py) and against closed
forms of the basis and quadrature tables.
"""
import glob
import os

import numpy as np
import pytest

from oracle.pyoracle import Oracle, tables

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "dense_*.npz")))
A22T = 0.5 * (2.0 - np.sqrt(2.0))


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("k", [1, 2, 3])
def test_tables_closed_forms(k):
    gll, gq, gw, S, D = tables(k)
    x, w = np.polynomial.legendre.leggauss(2 * k + 1)
    assert np.allclose(np.sort((x + 1) / 2), gq, atol=1e-15, rtol=0)
    assert np.allclose(w[np.argsort(x)] / 2, gw, atol=1e-15, rtol=0)
    # Gauss-Lobatto: k=2 -> {0, 1/2, 1}; k=3 -> {0, (1 -/+ 1/sqrt5)/2, 1}
    ref = {1: [0, 1], 2: [0, 0.5, 1], 3: [0, (1 - 5 ** -0.5) / 2, (1 + 5 ** -0.5) / 2, 1]}[k]
    assert np.allclose(gll, ref, atol=1e-15, rtol=0)
    # partition of unity, exact derivative of the identity, exactness 2q-1
    assert np.allclose(S.sum(1), 1.0, atol=1e-14)
    assert np.allclose(D.sum(1), 0.0, atol=1e-12)
    assert np.allclose(D @ gll, 1.0, atol=1e-12)
    for m in range(2 * (2 * k + 1)):
        assert abs(gw @ gq ** m - 1.0 / (m + 1)) < 1e-14


def make(fx):
    k, nx, nz = int(fx["k"]), int(fx["nx"]), int(fx["nz"])
    o = Oracle(k, nx, nz, Lx=float(fx["Lx"]), H=float(fx["H"]), hx=fx["h"], periodic=bool(fx["periodic"]))
    o.set_params(Ma=float(fx["Ma"]), dt=float(fx["dt"]))
    return o


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_oracle_vs_dense_fixtures(path):
    fx = np.load(path, allow_pickle=False)
    o = make(fx)
    rng = np.random.default_rng(3)
    Ns = o.Ns
    M, A, B = fx["M"], fx["A"], fx["B"]
    gm = float(fx["gamma"])
    o.set_stage(1, 1)
    o.set_rho_for_fixed(fx["rho"])
    for _ in range(3):
        x = rng.normal(size=Ns)
        assert rel(o.vmult(x, 1), M @ x) < 1e-13
        assert rel(o.vmult(x, 6), M @ x / (gm - 1)) < 1e-13
        xu = rng.normal(size=2 * Ns)
        assert rel(o.vmult(xu, 3), A @ xu) < 1e-13
        assert rel(o.vmult_pressure(x), B @ x) < 1e-13
    o.set_pres_fixed(fx["pf"])
    assert rel(o.vmult_enthalpy(fx["u"]), fx["Cu"]) < 1e-13
    # inverse diagonals are the reciprocals of the brute-force diagonals
    assert rel(o.inverse_diagonal(1), 1 / np.diag(M)) < 1e-13
    assert rel(o.inverse_diagonal(3), 1 / np.diag(A)) < 1e-13


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_fixture_structure(path):
    """Properties of the brute-force matrices themselves (so a wrong fixture fails)."""
    fx = np.load(path, allow_pickle=False)
    M, A, B = fx["M"], fx["A"], fx["B"]
    Ns = M.shape[0]
    assert np.allclose(M, M.T, atol=1e-12 * np.abs(M).max())
    assert np.linalg.eigvalsh(M).min() > 0
    # DG mass matrix is block diagonal (cell-local)
    nd2 = (int(fx["k"]) + 1) ** 2
    off = M.copy()
    for c in range(Ns // nd2):
        off[c * nd2:(c + 1) * nd2, c * nd2:(c + 1) * nd2] = 0
    assert np.abs(off).max() == 0
    # A(rho = 1) = M (x) I
    # discrete divergence theorem with walls: B 1 = 0
    assert np.abs(B @ np.ones(Ns)).max() < 1e-11 * np.abs(B).max()
    # sum of all entries of M = domain area
    area = float(fx["Lx"]) * float(fx["H"]) - np.trapezoid(fx["h"], float(fx["Lx"]) * np.arange(len(fx["h"])) / (len(fx["h"]) - 1))
    assert abs(M.sum() - area) < 1e-9 * area


def test_C_equals_scaled_BT_for_constant_pfix():
    """C(u) = -(gamma P Ma^2/(gamma-1)) B^T u when p_fix = P is constant
    (integration by parts; the lambda term vanishes since [P] = 0)."""
    fx = np.load(FIXTURES[1], allow_pickle=False)
    o = make(fx)
    gm, P = float(fx["gamma"]), 7.0e4
    o.set_stage(1, 2)
    o.set_pres_fixed(np.full(o.Ns, P))
    u = np.random.default_rng(9).normal(size=2 * o.Ns)
    Bt = o.vmult_pressure  # noqa: F841
    B = fx["B"]
    expect = -(gm * P * float(fx["Ma"]) ** 2 / (gm - 1)) * (B.T @ u)
    assert rel(o.vmult_enthalpy(u), expect) < 1e-12


def test_c1_fixture_is_the_oracles_run():
    """tests/golden/c1_hill_40x20_Q2.npz (the 100-step 40x20 Q2 run in parity
    mode, written by tests/golden/make_c1_fixture.py): the first two steps are
    rerun here, and the mass integral (the sum of M rho over all dofs) of every
    stored state equals the initial one, the property the IMEX scheme
    conserves for any step count (the density residual telescopes)."""
    fx = np.load(os.path.join(HERE, "golden", "c1_hill_40x20_Q2.npz"))
    k, nx, nz = int(fx["k"]), int(fx["nx"]), int(fx["nz"])
    x = 40000.0 * np.arange(nx + 1) / nx
    o = Oracle(k, nx, nz, hx=400.0 / (1.0 + ((x - 20000.0) / 1000.0) ** 2), periodic=True)
    o.set_params(Ma=1.0, Fr=1.0 / np.sqrt(9.81), dt=float(fx["dt"]), eps=1e-13, schur_inner_tol=1e-14,
                 a_solve_tol=1e-14, fp_lock=3)
    r0, u0, p0 = o.init_hill(seed=int(fx["seed"]))
    for a, b in ((r0, fx["rho0"]), (u0, fx["u0"]), (p0, fx["p0"])):
        assert np.array_equal(a, b)
    r2, u2, p2, _ = o.run(r0, u0, p0, 2)
    for a, name in ((r2, "rho_2"), (u2, "u_2"), (p2, "p_2")):
        assert rel(a, fx[name]) < 1e-12, name
    m0 = o.vmult(fx["rho0"], 1).sum()
    for n in (2, 100):
        assert abs(o.vmult(fx[f"rho_{n}"], 1).sum() - m0) < 1e-12 * abs(m0), n
    assert np.all(np.isfinite(fx["p_100"])) and rel(fx["p_100"], fx["p0"]) > 1e-6  # the flow has evolved
