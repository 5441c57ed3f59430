"""FP32 multigrid sweeps on the matrix cores (k_cheb_sweep_mx, FP32 MFMA sum
factorization) against the same sweeps on the vector ALUs and against the
oracle, in the mixed form (precision 1).

Each MFMA outer-product step is one fused multiply-add per entry, and the
matrix-core sweep accumulates every contraction in the order the vector-ALU
sweep does (interpolation over i then j, integration over qz then qx), with
the same weight and update expressions: the two forms give bitwise equal
results, so CG + MG solves take the same iterations to the same solution and
IMEX steps agree bit for bit. Solutions at a tight tolerance are also compared
with the oracle's (1e-11), IMEX steps in parity mode (tight tolerances,
locked fixed-point count) with the oracle's (1e-11).
"""
import numpy as np
import pytest

import atmospheric_flows_amd as af
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

G = 9.81
FR = 1.0 / np.sqrt(G)


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def states(dev, seed):
    rng = np.random.default_rng(seed)
    rho0, u0, p0 = dev.init_hill()
    ns = dev.ns
    rho = rho0 * (1 + 0.1 * rng.uniform(-1, 1, ns))
    p = p0 * (1 + 0.1 * rng.uniform(-1, 1, ns))
    u = np.concatenate([10 + 3 * rng.uniform(-1, 1, ns), 3 * rng.uniform(-1, 1, ns)])
    return rho, u, p


@pytest.mark.parametrize("k,nx,nz", [(2, 32, 16), (3, 16, 16), (3, 64, 32)])
def test_cg_mg_solves_matrix_cores_vs_vector_alus(k, nx, nz):
    h = af.witch_of_agnesi(nx)
    orc = Oracle(k, nx, nz, hx=h, periodic=True, mg_min_cells=8)
    orc.set_params(Ma=1.0, Fr=FR, dt=5.0, eps=1e-10)
    its, sol = {}, {}
    for mx in (1, 0):
        dev = af.DeviceContext(k, nx, nz, h=h, Ma=1.0, Fr=FR, dt=5.0, precision=1, mg_matrix_cores=mx)
        assert dev.params.mg_matrix_cores == mx
        op = af.EULEROperator(dev)
        (rho, u, p), (rho2, _, _) = states(dev, 3), states(dev, 4)
        op.set_HYPERBOLIC_stage(1)
        orc.set_stage(1, 3)
        op.set_rho_for_fixed(dev.vector(af.BLOCK_DENSITY, rho2))
        orc.set_rho_for_fixed(rho2)
        for ns, block, b in ((1, af.BLOCK_DENSITY, rho), (3, af.BLOCK_VELOCITY, u), (6, af.BLOCK_PRESSURE, p)):
            x = dev.vector(block)
            n_it = op.solve_cg_mg(ns, x, dev.vector(block, b), 1e-13)
            its[(mx, ns)] = n_it
            sol[(mx, ns)] = x.download()
            if mx == 1:
                xr, its_r = orc.solve_mass_mg(ns, b, tol=1e-13 * np.linalg.norm(b))
                e = rel(x.download(), xr)
                print(f"k={k} {nx}x{nz} NS {ns}: matrix-core its {n_it}, oracle its {its_r}, rel L2 {e:.2e}")
                assert e < 1e-11, f"CG+MG NS {ns} against the oracle"
        dev.close()
    for ns in (1, 3, 6):
        print(f"NS {ns}: iterations matrix cores {its[(1, ns)]}, vector ALUs {its[(0, ns)]}")
        assert its[(1, ns)] == its[(0, ns)], f"NS {ns}: iteration counts differ"
        assert np.array_equal(sol[(1, ns)], sol[(0, ns)]), f"NS {ns}: solutions differ"


@pytest.mark.parametrize("k,nx,nz,dt", [(2, 16, 8, 5.0), (3, 16, 8, 4.0)])
def test_imex_steps_matrix_cores(k, nx, nz, dt):
    h = af.witch_of_agnesi(nx)
    out = {}
    for mx in (1, 0):
        dev = af.DeviceContext(k, nx, nz, h=h, Ma=1.0, Fr=FR, dt=dt, eps=1e-10, precision=1, mg_matrix_cores=mx)
        rho, u, p = dev.init_hill()
        dev.upload_state(rho, u, p)
        dev.imex_step(2)
        out[mx] = dev.download_state()
        print(mx, dev.stats())
        dev.close()
    for i, name in enumerate(("rho", "u", "p")):
        print(f"{name}: matrix cores vs vector ALUs rel L2 {rel(out[1][i], out[0][i]):.3e}")
        assert np.array_equal(out[1][i], out[0][i]), name
    # parity mode against the oracle
    tight = dict(eps=1e-13, schur_inner_tol=1e-14, a_solve_tol=1e-14, fp_lock=3)
    dev = af.DeviceContext(k, nx, nz, h=h, Ma=1.0, Fr=FR, dt=dt, precision=1, mg_matrix_cores=1, **tight)
    orc = Oracle(k, nx, nz, hx=h, periodic=True, mg_min_cells=8)
    orc.set_params(Ma=1.0, Fr=FR, dt=dt, **tight)
    rho, u, p = dev.init_hill()
    dev.upload_state(rho, u, p)
    dev.imex_step(2)
    got = dev.download_state()
    ref = orc.run(rho, u, p, 2)
    for i, name in enumerate(("rho", "u", "p")):
        e = rel(got[i], ref[i])
        print(f"{name}: matrix cores vs oracle rel L2 {e:.3e}")
        assert e < 1e-11, name
