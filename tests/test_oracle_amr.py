"""Closed-form pins of the oracle on locally refined meshes (hanging faces,
multigrid across refinement levels).

The geometry of the cells is rebuilt here from the leaf list with numpy
(independently of both C++ mesh builders). On an affine base mesh with
walls, continuous polynomial states make every face jump vanish, also on
faces between cells of different size; the DG residuals then equal the
strong form integrated exactly, so a wrong split-face quadrature, a wrong
half-face trace or a missing subface contribution fails these tests.
"""
import numpy as np
import pytest

import test_oracle_invariants as T
from atmospheric_flows_amd.amr import Forest, hill_amr_leaves
from oracle.pyoracle import Oracle, tables, lib, ptr

LX, HZ = T.LX, T.HZ


def refined_leaves():
    """3 x 2 base mesh: base cell (1,0) refined twice in its top-right corner,
    the 2:1 balance then refines its right and upper neighbours once."""
    F = Forest(3, 2, periodic=False)
    F.refine((1, 0, 0, 0, 0))
    F.refine((1, 0, 1, 1, 1))
    F.balance()
    return F.ordered()


def cell_boxes(leaves, nx=3, nz=2):
    """(x0, x1, z0, z1) of each cell of the flat LX x HZ base mesh."""
    bi, bj, l, sx, sz = np.asarray(leaves).T
    n = 2.0 ** l
    dx, dz = LX / nx, HZ / nz
    x0 = bi * dx + sx / n * dx
    z0 = bj * dz + sz / n * dz
    return np.stack([x0, x0 + dx / n, z0, z0 + dz / n], axis=1)


def lagrange_1d(nodes, i, x):
    v = np.ones_like(np.asarray(x, dtype=float))
    for m, xm in enumerate(nodes):
        if m != i:
            v = v * (x - xm) / (nodes[i] - xm)
    return v


def nodal(boxes, k, f):
    gll = tables(k)[0]
    out = []
    for x0, x1, z0, z1 in boxes:
        for j in range(k + 1):
            for i in range(k + 1):
                out.append(f(x0 + gll[i] * (x1 - x0), z0 + gll[j] * (z1 - z0)))
    return np.array(out)


def project(boxes, k, f, nq=12):
    gll = tables(k)[0]
    xg, wg = np.polynomial.legendre.leggauss(nq)
    xg, wg = 0.5 * (xg + 1), 0.5 * wg
    out = []
    for x0, x1, z0, z1 in boxes:
        X, Z = np.meshgrid(x0 + xg * (x1 - x0), z0 + xg * (z1 - z0))
        W = np.outer(wg, wg) * (x1 - x0) * (z1 - z0)
        fv = f(X, Z) * W
        for j in range(k + 1):
            lz = lagrange_1d(gll, j, xg)
            for i in range(k + 1):
                lx = lagrange_1d(gll, i, xg)
                out.append(np.sum(fv * np.outer(lz, lx)))
    return np.array(out)


def setup(k):
    lv = refined_leaves()
    o = Oracle(k, 3, 2, Lx=LX, H=HZ, hx=np.zeros(4), periodic=False, leaves=lv)
    o.set_params(Ma=T.MA, Fr=T.FR, dt=T.DT)
    return cell_boxes(lv), o


def test_refined_mesh_has_hanging_faces_at_two_levels():
    lv = refined_leaves()
    assert sorted(set(lv[:, 2])) == [0, 1, 2]
    assert len(lv) == 3 + 7 + 4 + 4  # 3 untouched base cells, 3+4 in (1,0), 4 in each balanced neighbour


@pytest.mark.parametrize("k", [2, 3])
@pytest.mark.parametrize("stage", [1, 2, 3])
def test_rhs_rho_strong_form_on_refined_mesh(k, stage):
    boxes, o = setup(k)
    S = [T.make_state(10 + s) for s in range(stage)]
    w = {1: [T.TAB["a21"]], 2: [T.TAB["a31"], T.TAB["a32"]], 3: [T.TAB["b1"], T.TAB["b2"], T.TAB["b3"]]}[stage]
    strong = S[0][0]
    for s in range(stage):
        rho, (U, W), _ = S[s]
        strong = strong - (w[s] * T.DT) * T.div((rho * U, rho * W))
    srcs = []
    for s in range(stage):
        rho, (U, W), _ = S[s]
        srcs += [nodal(boxes, k, rho), np.concatenate([nodal(boxes, k, U), nodal(boxes, k, W)])]
    o.set_stage(stage, 1)
    assert T.rel(o.rhs_rho(srcs), project(boxes, k, strong)) < 1e-12


@pytest.mark.parametrize("k", [2, 3])
def test_rhs_pressure_and_velocity_strong_form_on_refined_mesh(k):
    boxes, o = setup(k)
    stage = 2
    S = [T.make_state(20 + s) for s in range(stage)]
    rk, (Uk, Wk), _ = T.make_state(99)
    ma2, ifr2, DT, GM, TAB = T.MA ** 2, 1.0 / T.FR ** 2, T.DT, T.GM, T.TAB
    wa, wt = [TAB["a31"], TAB["a32"]], [TAB["a31t"], TAB["a32t"]]
    grav = [(TAB["a31t"], 0), (TAB["a32t"], 1), (TAB["a33t"], "k")]
    rho0, (U0, W0), p0 = S[0]
    val = (1.0 / (GM - 1)) * p0 + (0.5 * ma2) * rho0 * (U0 * U0 + W0 * W0) - (0.5 * ma2) * rk * (Uk * Uk + Wk * Wk)
    for wgt, which in grav:
        r, (U, W) = (rk, (Uk, Wk)) if which == "k" else (S[which][0], S[which][1])
        val = val - (wgt * DT * ma2 * ifr2) * r * W
    Fx, Fz = T.Poly([[0.0]]), T.Poly([[0.0]])
    for s in range(stage):
        r, (U, W), p = S[s]
        kin = (0.5 * wa[s] * DT * ma2) * (U * U + W * W) * r
        ent = (wt[s] * DT * GM / (GM - 1)) * p
        Fx, Fz = Fx + (kin + ent) * U, Fz + (kin + ent) * W
    srcs = []
    for s in range(stage):
        r, (U, W), p = S[s]
        srcs += [nodal(boxes, k, r), np.concatenate([nodal(boxes, k, U), nodal(boxes, k, W)]), nodal(boxes, k, p)]
    srcs += [nodal(boxes, k, rk), np.concatenate([nodal(boxes, k, Uk), nodal(boxes, k, Wk)])]
    o.set_stage(stage, 2)
    assert T.rel(o.rhs_pressure(srcs), project(boxes, k, val - T.div((Fx, Fz)))) < 1e-12
    # momentum, stage 2
    rx, _, _ = T.make_state(77)
    gravu = [(TAB["a31t"], S[0][0]), (TAB["a32t"], S[1][0]), (TAB["a33t"], rx)]
    expect = []
    for comp in range(2):
        v = rho0 * (U0 if comp == 0 else W0)
        if comp == 1:
            for wgt, r in gravu:
                v = v - (wgt * DT * ifr2) * r
        Tx, Tz = T.Poly([[0.0]]), T.Poly([[0.0]])
        for s in range(stage):
            r, (U, W), p = S[s]
            uc = U if comp == 0 else W
            Tx = Tx + (wa[s] * DT) * r * uc * U + ((wt[s] * DT / ma2) * p if comp == 0 else 0.0)
            Tz = Tz + (wa[s] * DT) * r * uc * W + ((wt[s] * DT / ma2) * p if comp == 1 else 0.0)
        expect.append(project(boxes, k, v - T.div((Tx, Tz))))
    srcs_u = srcs[:6] + [nodal(boxes, k, rx)]
    o.set_stage(stage, 3)
    assert T.rel(o.rhs_velocity(srcs_u), np.concatenate(expect)) < 1e-12


@pytest.mark.parametrize("k", [2, 3])
def test_b_annihilates_constants_and_c_is_its_scaled_transpose(k):
    """B 1 = 0 on walls (discrete divergence theorem across hanging faces);
    for constant p_fix = P, C = -gamma P Ma^2/(gamma-1) B^T."""
    boxes, o = setup(k)
    o.set_stage(1, 2)
    ns = o.Ns
    assert np.abs(o.vmult_pressure(np.ones(ns))).max() < 1e-13
    P = 1.7
    o.set_pres_fixed(np.full(ns, P))
    Bm = np.stack([o.vmult_pressure(e) for e in np.eye(ns)], axis=1)
    Cm = np.stack([o.vmult_enthalpy(e) for e in np.eye(2 * ns)], axis=1)
    assert np.abs(Cm + T.GM * P * T.MA ** 2 / (T.GM - 1) * Bm.T).max() < 1e-12 * np.abs(Cm).max()


@pytest.mark.parametrize("stage", [1, 3])
def test_rhs_rho_telescopes_on_refined_hill_mesh(stage):
    lv, h = hill_amr_leaves(nx=8, nz=4, levels=2, Lx=8000.0, H=4000.0, xc=4000.0, band_half_width=1500.0,
                            band_top=1500.0)
    o = Oracle(3, 8, 4, Lx=8000.0, H=4000.0, hx=h, leaves=lv)
    o.set_params(dt=2.0)
    rng = np.random.default_rng(stage)
    srcs = []
    for s in range(stage):
        srcs += [1 + 0.2 * rng.uniform(-1, 1, o.Ns), rng.normal(size=2 * o.Ns) * 10]
    o.set_stage(stage, 1)
    rhs = o.rhs_rho(srcs)
    mass = o.vmult(srcs[0], 1).sum()
    assert abs(rhs.sum() - mass) < 1e-12 * abs(mass)


def test_mg_transfer_across_refinement_levels_reproduces_polynomials():
    """Prolongation from every coarser level of the hierarchy maps the nodal
    values of a degree-k polynomial to its nodal values on the finer level."""
    k = 2
    lv, h = hill_amr_leaves(nx=8, nz=4, levels=2, Lx=8000.0, H=4000.0, xc=4000.0, band_half_width=1500.0,
                            band_top=1500.0, h0=0.0)
    o = Oracle(k, 8, 4, Lx=8000.0, H=4000.0, hx=h, leaves=lv, mg_min_cells=2)
    L = o.nlevels
    assert L >= 3
    f = lambda x, z: 1.0 + 2e-4 * x - 3e-4 * z + 1e-8 * x * z - 2e-8 * x * x  # noqa: E731
    # nodal values on each level from the level's own cell boxes: recover them via the oracle's geometry
    for lfine in range(L - 1, 0, -1):
        nf, nc = lib().orc_mg_level_cells(o._h, lfine), lib().orc_mg_level_cells(o._h, lfine - 1)
        xf, zf = level_nodes(o, lfine, nf, k)
        xc_, zc_ = level_nodes(o, lfine - 1, nc, k)
        out = np.zeros(nf * (k + 1) ** 2)
        coarse = np.ascontiguousarray(f(xc_, zc_))
        lib().orc_prolongate(o._h, lfine, ptr(out), ptr(coarse))
        assert np.abs(out - f(xf, zf)).max() < 1e-12


def level_nodes(o, level, ncell, k):
    """Node coordinates of a level, from the oracle's level geometry."""
    xv, zv = np.zeros(4 * ncell), np.zeros(4 * ncell)
    lib().orc_level_vertices(o._h, level, ptr(xv), ptr(zv))
    gll = tables(k)[0]
    X, Z = [], []
    for c in range(ncell):
        x4, z4 = xv[4 * c:4 * c + 4], zv[4 * c:4 * c + 4]
        for j in range(k + 1):
            for i in range(k + 1):
                xi, eta = gll[i], gll[j]
                w = [(1 - xi) * (1 - eta), xi * (1 - eta), (1 - xi) * eta, xi * eta]
                X.append(np.dot(w, x4))
                Z.append(np.dot(w, z4))
    return np.array(X), np.array(Z)
