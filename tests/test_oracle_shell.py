"""Pins of the spherical-shell oracle (oracle/atmo_oracle_shell.cpp), the
checker of the reference's own 3D case (Atmospheric_Flow.cc:279-289, 1277).

- geometry: volume and wall areas of the Q2-mapped shell converge to the
  sphere's at fourth order; neighbours pair up, walls only at the radii;
- B*1 = 0 (discrete divergence theorem: with the Q2 map and QGauss(3) the
  cofactor identity is integrated exactly) and C = -(gamma P Ma^2/(gamma-1)) B^T
  for a constant fixed pressure, on curved cells;
- density telescoping (sum of rhs_rho = mass) and mass conservation of the
  time step to round-off;
- the 3D forms reduce to the 2D oracle's (itself pinned by closed forms):
  on a box with walls and constant gravity along z, a state that does not
  depend on y and has no y velocity gives every residual and product of
  the x-z slice, scaled by the y extent of the test functions;
- the dense Schur complement: GMRES solves it;
- the analytic solid-body-rotation steady state (equation_data.h:47-176):
  the error after a fixed time falls under refinement.
"""
import numpy as np
import pytest

from oracle.pyoracle import Oracle
from oracle.pyoracle_shell import EARTH_RADIUS, ShellOracle

RNG = np.random.default_rng(2024)


def unit_shell(L, panels=6):
    return ShellOracle(L, panels=panels, r_inner=0.9, r_outer=1.0, planet_radius=1.0)


def random_state(o, rng, base=1.0, amp=0.2, uamp=0.5):
    rho = base + amp * rng.uniform(-1, 1, o.Ns)
    p = base + amp * rng.uniform(-1, 1, o.Ns)
    u = uamp * rng.uniform(-1, 1, 3 * o.Ns)
    return rho, u, p


@pytest.mark.parametrize("panels", [6, 12])
def test_geometry_converges_to_the_sphere(panels):
    a = EARTH_RADIUS
    errs = []
    for L in (1, 2, 3):
        o = ShellOracle(L, panels=panels)
        vol, a_in, a_out, a_all = o.measures()
        errs.append([abs(vol / (4 / 3 * np.pi * (a ** 3 - (0.9 * a) ** 3)) - 1),
                     abs(a_in / (4 * np.pi * (0.9 * a) ** 2) - 1), abs(a_out / (4 * np.pi * a ** 2) - 1)])
        assert abs(a_all - a_in - a_out) <= 1e-12 * a_all  # walls only at the two radii
    errs = np.array(errs)
    assert np.all(errs[1:] < errs[:-1] / 12), errs  # fourth order (factor 16 per halving)


@pytest.mark.parametrize("panels", [6, 12])
def test_neighbours_pair_up(panels):
    L = 2
    o = unit_shell(L, panels)
    n = 2 ** L
    nb, nbf = o.neighbours()
    for c in range(o.ncell):
        layer = (c // (n * n)) % n
        for f in range(6):
            if nb[c, f] < 0:
                assert (f == 4 and layer == 0) or (f == 5 and layer == n - 1), (c, f)
            else:
                assert nb[nb[c, f], nbf[c, f]] == c and nbf[nb[c, f], nbf[c, f]] == f
    assert (nb < 0).sum() == 2 * panels * n * n


@pytest.mark.parametrize("scale", [1.0, EARTH_RADIUS])
def test_divergence_theorem(scale):
    o = ShellOracle(2, r_inner=0.9 * scale, r_outer=scale, planet_radius=scale)
    o.set_params(Ma=0.3, Fr=1.0, dt=0.1)
    for hs in (1, 2):
        o.set_stage(hs, 2)
        b1 = o.vmult_pressure(np.ones(o.Ns))
        bp = o.vmult_pressure(RNG.uniform(0.5, 1.5, o.Ns))
        assert np.abs(b1).max() <= 1e-13 * np.abs(bp).max()


def dense(fn, n):
    cols = []
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        cols.append(fn(e))
    return np.array(cols).T


@pytest.mark.parametrize("panels", [6, 12])
def test_C_is_minus_kappa_B_transposed(panels):
    o = unit_shell(0, panels)
    Ma, gm, P = 0.3, 1.4, 1.7
    o.set_params(Ma=Ma, Fr=1.0, dt=0.2)
    for hs in (1, 2):
        o.set_stage(hs, 2)
        o.set_pres_fixed(np.full(o.Ns, P))
        B = dense(o.vmult_pressure, o.Ns)
        C = dense(o.vmult_enthalpy, 3 * o.Ns)
        kappa = gm * P * Ma ** 2 / (gm - 1)
        assert np.abs(C + kappa * B.T).max() <= 1e-13 * np.abs(C).max()


def test_mass_blocks():
    o = unit_shell(1)
    rho, u, p = random_state(o, RNG)
    o.set_rho_for_fixed(np.ones(o.Ns))
    x = RNG.uniform(-1, 1, o.Ns)
    Mx = o.vmult(1, x)
    Ax = o.vmult(3, np.concatenate([x, 2 * x, -x]))
    assert np.allclose(Ax, np.concatenate([Mx, 2 * Mx, -Mx]), rtol=0, atol=1e-15 * np.abs(Mx).max())
    assert np.abs(o.vmult(6, x) - Mx / 0.4).max() <= 1e-14 * np.abs(Mx).max()
    # symmetric: x . M y = y . M x; positive: x . M x > 0
    y = RNG.uniform(-1, 1, o.Ns)
    assert abs(x @ o.vmult(1, y) - y @ Mx) <= 1e-14 * abs(x @ Mx)
    o.set_rho_for_fixed(rho)
    for ns in (1, 3, 6):
        b = RNG.uniform(-1, 1, o.ncomp(ns) * o.Ns)
        assert np.allclose(o.vmult(ns, o.solve_block(ns, b)), b, rtol=0, atol=1e-12 * np.abs(b).max())
        d = o.inverse_diagonal(ns)
        e = np.zeros(o.ncomp(ns) * o.Ns)
        e[5] = 1.0
        assert np.isclose(o.vmult(ns, e)[5], 1.0 / d[5], rtol=1e-14)


def mass(o, rho):
    return o.vmult(1, rho).sum()


def test_density_telescoping_and_mass_conservation():
    o = unit_shell(1)
    o.set_params(Ma=1.0, Fr=1.0, dt=0.01)
    states = [random_state(o, RNG) for _ in range(3)]
    for hs in (1, 2, 3):
        o.set_stage(hs, 1)
        srcs = [a for st in states[:hs] for a in st[:2]]
        r = o.rhs_rho(*srcs)
        assert abs(r.sum() - mass(o, states[0][0])) <= 1e-13 * abs(r.sum())
    rho, u, p = o.sbr_state(Omega=0.5, a=1.0, g=1.0, RT=1.0)
    rho = rho * (1 + 0.01 * RNG.uniform(-1, 1, o.Ns))
    m0 = mass(o, rho)
    o.set_state(rho, u, p)
    o.step(2)
    r2, _, _ = o.get_state()
    assert abs(mass(o, r2) - m0) <= 1e-13 * m0


# ------------------------------------------------------- reduction to 2D
NX, NY, NZ, LX, LY, H = 4, 3, 3, 4.0, 2.0, 3.0


def lift(f2, nx=NX, ny=NY, nz=NZ):
    """y-independent 3D field of a 2D one (2D cell j*nx+i, dof jj*2+ii)."""
    f2 = np.asarray(f2).reshape(nz, nx, 2, 2)  # j, i, jj, ii
    f3 = np.zeros((nz, ny, nx, 2, 2, 2))       # j, jy, i, kk, yl, ii
    for yl in (0, 1):
        f3[:, :, :, :, yl, :] = f2[:, None, :, :, :]
    return f3.reshape(-1)


def lift_u(u2):
    n = len(u2) // 2
    return np.concatenate([lift(u2[:n]), np.zeros(n * 2 * NY), lift(u2[n:])])


def slices(f3):
    """(nz, ny, nx, kk, yl, ii) view of a 3D scalar field."""
    return np.asarray(f3).reshape(NZ, NY, NX, 2, 2, 2)


def scalar_dev(f3, f2):
    """largest deviation of every y-slice of f3 from f2 times the y extent
    of a test function, relative to the largest entry"""
    want = np.asarray(f2).reshape(NZ, NX, 2, 2) * (LY / NY) / 2
    got = slices(f3)
    return max(np.abs(got[:, jy, :, :, yl, :] - want).max() for jy in range(NY) for yl in (0, 1)) / \
        np.abs(want).max()


def vector_dev(f3, f2):
    n2, n3 = len(f2) // 2, len(f3) // 3
    return max(scalar_dev(f3[:n3], f2[:n2]), scalar_dev(f3[2 * n3:], f2[n2:]),
               np.abs(f3[n3:2 * n3]).max() / np.abs(f3).max())


def make_box_pair(library=None):
    o2 = Oracle(1, NX, NZ, Lx=LX, H=H, hx=np.zeros(NX + 1), periodic=False)
    o3 = ShellOracle(box=(NX, NY, NZ, LX, LY, H), library=library)
    kw = dict(Ma=0.4, Fr=0.8, dt=0.05)
    o2.set_params(**kw)
    o3.set_params(**kw)
    return o2, o3


@pytest.fixture(scope="module")
def box_pair():
    return make_box_pair()


def two_d_states(n, rng, ns):
    out = []
    for _ in range(n):
        rho = 1.0 + 0.2 * rng.uniform(-1, 1, ns)
        p = 1.0 + 0.2 * rng.uniform(-1, 1, ns)
        u = 0.5 * rng.uniform(-1, 1, 2 * ns)
        out.append((rho, u, p))
    return out


def box_residual_devs(o2, o3, hs):
    rng = np.random.default_rng(10 + hs)
    st2 = two_d_states(5, rng, o2.Ns)
    st3 = [(lift(r), lift_u(u), lift(p)) for r, u, p in st2]
    o2.set_stage(hs, 1)
    o3.set_stage(hs, 1)
    dev = {}
    # density: pairs (rho_j, u_j)
    s2 = [a for st in st2[:hs] for a in st[:2]]
    s3 = [a for st in st3[:hs] for a in st[:2]]
    dev["rhs_rho"] = scalar_dev(o3.rhs_rho(*s3), o2.rhs_rho(s2))
    # pressure: triples, then the kinetic pair (density, velocity)
    s2 = [a for st in st2[:hs] for a in st] + [st2[3][0], st2[4][1]]
    s3 = [a for st in st3[:hs] for a in st] + [st3[3][0], st3[4][1]]
    dev["rhs_pressure"] = scalar_dev(o3.rhs_pressure(*s3), o2.rhs_pressure(s2))
    # velocity: triples, then the gravity density of the implicit stage
    s2 = [a for st in st2[:hs] for a in st] + ([st2[3][0]] if hs < 3 else [])
    s3 = [a for st in st3[:hs] for a in st] + ([st3[3][0]] if hs < 3 else [])
    dev["rhs_velocity"] = vector_dev(o3.rhs_velocity(*s3), o2.rhs_velocity(s2))
    return dev


def box_product_devs(o2, o3, hs):
    rng = np.random.default_rng(20 + hs)
    (rho, u, p), = two_d_states(1, rng, o2.Ns)
    pf = 1.0 + 0.2 * rng.uniform(-1, 1, o2.Ns)
    for o, r, pp in ((o2, rho, pf), (o3, lift(rho), lift(pf))):
        o.set_stage(hs, 2)
        o.set_rho_for_fixed(r)
        o.set_pres_fixed(pp)
    return {"B": vector_dev(o3.vmult_pressure(lift(p)), o2.vmult_pressure(p)),
            "C": scalar_dev(o3.vmult_enthalpy(lift_u(u)), o2.vmult_enthalpy(u)),
            "M_rho": scalar_dev(o3.vmult(1, lift(p)), o2.vmult(p, 1)),
            "D": scalar_dev(o3.vmult(6, lift(p)), o2.vmult(p, 6)),
            "A": vector_dev(o3.vmult(3, lift_u(u)), o2.vmult(u, 3)),
            # the 2D oracle's inner CG of the Schur complement stops at 1e-12 relative
            "Schur": 1e-3 * scalar_dev(o3.vmult(2, lift(p)), o2.vmult(p, 2))}


def box_deviation(library=None):
    """largest deviation of the 3D forms on the box from the 2D oracle"""
    o2, o3 = make_box_pair(library)
    devs = [box_residual_devs(o2, o3, hs) for hs in (1, 2, 3)] + [box_product_devs(o2, o3, hs) for hs in (1, 2)]
    return max(v for d in devs for v in d.values())


@pytest.mark.parametrize("hs", [1, 2, 3])
def test_box_residuals_equal_the_2d_oracle(box_pair, hs):
    dev = box_residual_devs(*box_pair, hs)
    assert max(dev.values()) <= 1e-13, dev


@pytest.mark.parametrize("hs", [1, 2])
def test_box_products_equal_the_2d_oracle(box_pair, hs):
    dev = box_product_devs(*box_pair, hs)
    assert max(dev.values()) <= 1e-13, dev


def test_schur_gmres_against_a_dense_solve():
    """With a constant fixed pressure the Schur complement is linear (the
    lambda * [p_fix] term of C vanishes) and GMRES meets a dense solve.
    (With a varying one C depends on |u.n| through lambda, so the operator
    the reference hands to GMRES is not linear.)"""
    o = unit_shell(0)
    o.set_params(Ma=0.5, Fr=1.0, dt=0.05)
    rho, u, p = random_state(o, RNG)
    o.set_stage(1, 2)
    o.set_rho_for_fixed(rho)
    o.set_pres_fixed(np.full(o.Ns, 1.3))
    S = dense(lambda x: o.vmult(2, x), o.Ns)
    b = o.vmult(2, p)
    x, its = o.solve_schur(np.zeros(o.Ns), b, 1e-14)
    assert its > 0
    xd = np.linalg.solve(S, b)
    assert np.abs(xd - p).max() <= 1e-12
    assert np.abs(x - xd).max() <= 1e-11
    P = o.inverse_diagonal(2)
    assert np.linalg.norm(P * (b - o.vmult(2, x))) <= 1e-13 * np.linalg.norm(b)


def test_solid_body_rotation_converges():
    """Unit shell (a = 1, r in [0.9, 1], R T = 1 so Ma = 1, g = 1, Omega = 0.5):
    after t = 0.1 at acoustic CFL 0.5 the L2 error against the steady state
    falls by more than 2.5 per refinement in every field."""
    kw = dict(Omega=0.5, a=1.0, g=1.0, RT=1.0)
    errs = []
    for L in (1, 2):
        o = unit_shell(L)
        dt = 0.5 * (0.1 / 2 ** L) / np.sqrt(1.4)
        n = int(np.ceil(0.1 / dt))
        o.set_params(Ma=1.0, Fr=1.0, dt=0.1 / n)
        rho, u, p = o.sbr_state(**kw)
        o.set_state(rho, u, p)
        o.step(n)
        e = o.l2_errors(*o.get_state(), **kw)
        errs.append([e["u"][0], e["rho"][0], e["p"][0]])
    errs = np.array(errs)
    assert np.all(errs[1] < errs[0] / 2.5), errs


def tendency_ratios(library=None):
    """Velocity tendency |u^1 - u^0|_L2 / (dt g) of one IMEX step from the
    steady state (unit shell, g = 1, Omega = 0.5, dt = 1e-3) at 1, 2 and 3
    refinements; returns the two reduction factors. The steady state holds
    the pressure gradient, the centrifugal term and gravity g a^2/r^2 e_r in
    balance, so the tendency is discretization error and falls with h; a
    wrong gravity direction or factor leaves an O(1) imbalance."""
    kw = dict(Omega=0.5, a=1.0, g=1.0, RT=1.0)
    t = []
    for L in (1, 2, 3):
        o = ShellOracle(L, r_inner=0.9, r_outer=1.0, planet_radius=1.0, library=library)
        o.set_params(Ma=1.0, Fr=1.0, dt=1e-3)
        rho, u, p = o.sbr_state(**kw)
        o.set_state(rho, u, p)
        o.step(1)
        _, u1, _ = o.get_state()
        du2 = ((u1 - u) ** 2).reshape(3, -1).sum(0)
        t.append(np.sqrt(o.vmult(1, du2).sum()) / 1e-3)
    return [t[0] / t[1], t[1] / t[2]]


def test_steady_state_tendency_falls_with_h():
    r = tendency_ratios()
    assert all(x > 1.6 for x in r), r
