"""Pins the jump (Rusanov) terms of the oracle's three explicit residuals.

States that are constant on each cell, with one uniform velocity per source,
on a flat mesh (periodic in x, walls at bottom and top). Every face then
carries a jump while the velocity is continuous, so the local Lax-Friedrichs
flux of the reference, {f}.n + 1/2 lambda [w] with lambda = |u.n|
(euler_operator.h:519-526, 1154-1195, 1338-1355, 1938-1950), reduces to the
upwind flux of what the velocity carries:
  - density:  (u.n) rho_up;
  - energy:   Ma^2 (u.n) K_up + (u.n) (p_up/(gamma-1) + {p}), K = rho|u|^2/2
              (the internal energy is upwinded, the pressure work is central);
              the second source of stage 2 carries the reference's kinetic
              jump rho_2 (u_2.u_old)/2 instead of rho_2 |u_2|^2/2
              (euler_operator.h:1182-1183);
  - momentum: u_c (u.n) rho_up + {p}/Ma^2 n_c, and p/Ma^2 n_c at the walls
              (p_D = p, :2189-2305);
  and the density and energy residuals have no wall flux.
On a flat cell dx x dz the FE_DGQ(k) test functions phi_ab integrate exactly:
  int phi = dx dz w_a w_b, int dphi/dx = dz w_b (d_ak - d_a0), faces
  dz w_b d_ak (right) ... with w the Gauss-Lobatto weights on [0,1],
so each residual is V int phi + F . int grad phi - sum_faces Flux int_face phi
with cell constants V, F and the upwind fluxes above, evaluated here in numpy
with no quadrature and no face loop. The tableau is the reference's
(euler_operator.h:220-224).
"""
import numpy as np
import pytest

from oracle.pyoracle import Oracle

LX, H = 40000.0, 20000.0
GAM = 1.4
SQ2 = np.sqrt(2.0)
TG = 2.0 - SQ2
A21, A31, A32 = TG, 0.5, 0.5
A21T, A22T, A31T, A32T, A33T = 0.5 * TG, 0.5 * TG, SQ2 / 4.0, SQ2 / 4.0, 1.0 - SQ2 / 2.0
B = (A31T, A32T, A33T)


def gll_weights(k):
    """Gauss-Lobatto nodes and weights (integrals of the Lagrange basis) on [0, 1]."""
    inner = np.polynomial.legendre.Legendre.basis(k).deriv().roots() if k > 1 else np.array([])
    x = 0.5 * (np.concatenate([[-1.0], np.sort(inner.real), [1.0]]) + 1.0)
    xg, wg = np.polynomial.legendre.leggauss(2 * k + 2)
    xg, wg = 0.5 * (xg + 1.0), 0.5 * wg
    w = np.zeros(k + 1)
    for i in range(k + 1):
        li = np.ones_like(xg)
        for m in range(k + 1):
            if m != i:
                li *= (xg - x[m]) / (x[i] - x[m])
        w[i] = li @ wg
    return x, w


class Mesh:
    def __init__(self, k, nx, nz):
        self.k, self.nx, self.nz = k, nx, nz
        self.dx, self.dz = LX / nx, H / nz
        _, w = gll_weights(k)
        nd = k + 1
        d0, dk = np.zeros(nd), np.zeros(nd)
        d0[0], dk[k] = 1.0, 1.0
        # test-function integrals, [b, a] = (z index, x index), flattened x fastest
        self.I0 = (self.dx * self.dz * np.outer(w, w)).ravel()
        self.Ix = (self.dz * np.outer(w, dk - d0)).ravel()
        self.Iz = (self.dx * np.outer(dk - d0, w)).ravel()
        self.face = {  # (outward normal, integral of phi over the face)
            "R": ((1.0, 0.0), (self.dz * np.outer(w, dk)).ravel()),
            "L": ((-1.0, 0.0), (self.dz * np.outer(w, d0)).ravel()),
            "T": ((0.0, 1.0), (self.dx * np.outer(dk, w)).ravel()),
            "B": ((0.0, -1.0), (self.dx * np.outer(d0, w)).ravel()),
        }

    def neighbour(self, i, j, f):
        """neighbour cell index across face f, or None at a wall (periodic in x)"""
        if f == "R":
            return j * self.nx + (i + 1) % self.nx
        if f == "L":
            return j * self.nx + (i - 1) % self.nx
        if f == "T":
            return None if j == self.nz - 1 else (j + 1) * self.nx + i
        return None if j == 0 else (j - 1) * self.nx + i

    def assemble(self, V, F, flux, ncomp=1):
        """sum over cells of V int phi + F . int grad phi - sum_f Flux_f int_f phi.
        V[c][comp], F[c][comp] = (Fx, Fz), flux(c, nb, n)[comp] (nb None at walls)."""
        nd2 = (self.k + 1) ** 2
        nc = self.nx * self.nz
        out = np.zeros((ncomp, nc * nd2))
        for j in range(self.nz):
            for i in range(self.nx):
                c = j * self.nx + i
                for comp in range(ncomp):
                    r = V[c][comp] * self.I0 + F[c][comp][0] * self.Ix + F[c][comp][1] * self.Iz
                    for f, (n, If) in self.face.items():
                        r = r - flux(c, self.neighbour(i, j, f), n)[comp] * If
                    out[comp, c * nd2:(c + 1) * nd2] = r
        return out.ravel()


def cellwise(m, vals):
    """nodal vector of a field constant on each cell"""
    return np.repeat(vals, (m.k + 1) ** 2)


def velocity(m, U, W):
    n = m.nx * m.nz * (m.k + 1) ** 2
    return np.concatenate([np.full(n, U), np.full(n, W)])


def up(un, own, nb):
    return own if un > 0 else nb


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def make_case(k, library=None):
    """Flat 5 x 3 mesh of degree k, four stage states (cell-constant rho, p;
    one uniform velocity each, with signs that differ across sources)."""
    nx, nz = 5, 3
    m = Mesh(k, nx, nz)
    o = Oracle(k, nx, nz, library=library)  # flat: h = 0
    Ma, Fr, dt = 0.7, 0.4, 3.0
    o.set_params(Ma=Ma, Fr=Fr, dt=dt)
    rng = np.random.default_rng(100 + k)
    nc = nx * nz
    # four stage states: per-cell rho, p; one uniform velocity each (signs differ across sources)
    vel = [(7.0, 2.0), (-4.0, -1.5), (3.0, 0.5), (-2.5, 3.5)]
    st = []
    for s in range(4):
        rho = 1.0 + 0.3 * rng.uniform(-1, 1, nc)
        p = 1e5 * (1.0 + 0.3 * rng.uniform(-1, 1, nc))
        st.append(dict(rc=rho, pc=p, u=vel[s], rho=cellwise(m, rho), p=cellwise(m, p),
                       uv=velocity(m, *vel[s])))
    return m, o, st, Ma, Fr, dt


@pytest.fixture(params=[1, 2, 3])
def case(request):
    return make_case(request.param)


def dotn(u, n):
    return u[0] * n[0] + u[1] * n[1]


def density_errors(case):
    """relative L2 distance of the oracle's residual from the closed form, per stage"""
    errs = []
    m, o, st, Ma, Fr, dt = case
    nc = m.nx * m.nz
    for hs, w in ((1, (A21,)), (2, (A31, A32)), (3, B)):
        o.set_stage(hs, 1)
        srcs = [a for s in range(hs) for a in (st[s]["rho"], st[s]["uv"])]
        got = o.rhs_rho(srcs)
        V = [[st[0]["rc"][c]] for c in range(nc)]
        F = [[(sum(w[s] * dt * st[s]["rc"][c] * st[s]["u"][0] for s in range(hs)),
               sum(w[s] * dt * st[s]["rc"][c] * st[s]["u"][1] for s in range(hs)))] for c in range(nc)]

        def flux(c, nb, n):
            if nb is None:
                return [0.0]
            return [sum(w[s] * dt * dotn(st[s]["u"], n) * up(dotn(st[s]["u"], n), st[s]["rc"][c], st[s]["rc"][nb])
                        for s in range(hs))]
        want = m.assemble(V, F, flux)
        e = rel(got, want)
        print(f"k={m.k} stage {hs}: rhs_rho vs upwind closed form {e:.2e}")
        errs.append(e)
    return errs


def test_density_residual_is_upwind(case):
    assert max(density_errors(case)) < 1e-13


def energy_errors(case):
    """relative L2 distance of the oracle's residual from the closed form, per stage"""
    errs = []
    m, o, st, Ma, Fr, dt = case
    nc = m.nx * m.nz
    Ma2, iFr2 = Ma * Ma, 1.0 / (Fr * Fr)
    kin = lambda s, c: 0.5 * st[s]["rc"][c] * dotn(st[s]["u"], st[s]["u"])  # noqa: E731
    for hs in (1, 2, 3):
        o.set_stage(hs, 1)
        if hs == 1:
            wa, wt, ex = (A21,), (A21T,), [0]
            kin_s = 1                     # (rho_kin, u_kin) = source triple 1's rho and velocity
            grav = [(A21T, 0), (A22T, 1)]
        elif hs == 2:
            wa, wt, ex = (A31, A32), (A31T, A32T), [0, 1]
            kin_s = 2
            grav = [(A31T, 0), (A32T, 1), (A33T, 2)]
        else:
            wa, wt, ex = B, B, [0, 1, 2]
            kin_s = 3
            grav = [(B[0], 0), (B[1], 1), (B[2], 2)]
        srcs = [a for s in ex for a in (st[s]["rho"], st[s]["uv"], st[s]["p"])]
        srcs += [st[kin_s]["rho"], st[kin_s]["uv"]]
        got = o.rhs_pressure(srcs)
        V, F = [], []
        for c in range(nc):
            v = st[0]["pc"][c] / (GAM - 1.0) + Ma2 * kin(0, c) - Ma2 * kin(kin_s, c)
            for gw, s in grav:  # gravity e_r = (0, 1): rho (u . e_r)
                v -= gw * dt * Ma2 * iFr2 * st[s]["rc"][c] * st[s]["u"][1]
            V.append([v])
            fx = fz = 0.0
            for i, s in enumerate(ex):
                a = wa[i] * dt * Ma2 * kin(s, c) + wt[i] * dt * GAM / (GAM - 1.0) * st[s]["pc"][c]
                fx += a * st[s]["u"][0]
                fz += a * st[s]["u"][1]
            F.append([(fx, fz)])

        def flux(c, nb, n):
            if nb is None:
                return [0.0]
            tot = 0.0
            for i, s in enumerate(ex):
                un = dotn(st[s]["u"], n)
                if hs == 2 and i == 1:
                    # the reference's kinetic jump of the second source: rho_2 (u_2 . u_old)/2
                    u0 = st[0]["u"]
                    fk = un * 0.5 * (kin(s, c) + kin(s, nb)) + 0.5 * abs(un) * 0.5 * dotn(st[s]["u"], u0) * (
                        st[s]["rc"][c] - st[s]["rc"][nb])
                else:
                    fk = un * up(un, kin(s, c), kin(s, nb))
                fe = un * (up(un, st[s]["pc"][c], st[s]["pc"][nb]) / (GAM - 1.0) +
                           0.5 * (st[s]["pc"][c] + st[s]["pc"][nb]))
                tot += wa[i] * dt * Ma2 * fk + wt[i] * dt * fe
            return [tot]
        want = m.assemble(V, F, flux)
        e = rel(got, want)
        print(f"k={m.k} stage {hs}: rhs_pressure vs upwind closed form {e:.2e}")
        errs.append(e)
    return errs


def test_energy_residual_is_upwind(case):
    assert max(energy_errors(case)) < 1e-13


def momentum_errors(case):
    """relative L2 distance of the oracle's residual from the closed form, per stage"""
    errs = []
    m, o, st, Ma, Fr, dt = case
    nc = m.nx * m.nz
    Ma2, iFr2 = Ma * Ma, 1.0 / (Fr * Fr)
    for hs in (1, 2, 3):
        o.set_stage(hs, 1)
        if hs == 1:
            wa, wt, ex, grav = (A21,), (A21T,), [0], [(A21T, 0), (A22T, 1)]
        elif hs == 2:
            wa, wt, ex, grav = (A31, A32), (A31T, A32T), [0, 1], [(A31T, 0), (A32T, 1), (A33T, 2)]
        else:
            wa, wt, ex, grav = B, B, [0, 1, 2], [(B[0], 0), (B[1], 1), (B[2], 2)]
        srcs = [a for s in ex for a in (st[s]["rho"], st[s]["uv"], st[s]["p"])]
        if hs < 3:
            srcs.append(st[hs]["rho"])  # the implicit stage's density (rho_tmp_2 / rho_tmp_3)
        got = o.rhs_velocity(srcs)
        V, F = [], []
        for c in range(nc):
            vc, fc = [], []
            for comp in range(2):
                v = st[0]["rc"][c] * st[0]["u"][comp]
                for gw, s in grav:
                    v -= gw * dt * iFr2 * st[s]["rc"][c] * (1.0 if comp == 1 else 0.0)
                vc.append(v)
                fx = fz = 0.0
                for i, s in enumerate(ex):
                    ruc = st[s]["rc"][c] * st[s]["u"][comp]
                    fx += wa[i] * dt * ruc * st[s]["u"][0] + (wt[i] * dt / Ma2 * st[s]["pc"][c] if comp == 0 else 0.0)
                    fz += wa[i] * dt * ruc * st[s]["u"][1] + (wt[i] * dt / Ma2 * st[s]["pc"][c] if comp == 1 else 0.0)
                fc.append((fx, fz))
            V.append(vc)
            F.append(fc)

        def flux(c, nb, n):
            out = []
            for comp in range(2):
                tot = 0.0
                for i, s in enumerate(ex):
                    if nb is None:
                        tot += wt[i] * dt / Ma2 * st[s]["pc"][c] * n[comp]
                    else:
                        un = dotn(st[s]["u"], n)
                        tot += wa[i] * dt * st[s]["u"][comp] * un * up(un, st[s]["rc"][c], st[s]["rc"][nb])
                        tot += wt[i] * dt / Ma2 * 0.5 * (st[s]["pc"][c] + st[s]["pc"][nb]) * n[comp]
                out.append(tot)
            return out
        want = m.assemble(V, F, flux, ncomp=2)
        e = rel(got, want)
        print(f"k={m.k} stage {hs}: rhs_velocity vs upwind closed form {e:.2e}")
        errs.append(e)
    return errs


def test_momentum_residual_is_upwind(case):
    assert max(momentum_errors(case)) < 1e-13
