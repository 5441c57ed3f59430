"""Pins the stage wiring of the oracle's IMEX step against a second
restatement of EulerSolver::run's loop body (Atmospheric_Flow.cc:1115-1222)
written here in Python from the reference's stage functions and their
positional source lists (update_density :430-537 with sources :439-449,
pressure_fixed_point :544-673 with :555-568, update_velocity :680-781 with
:696-698, update_pressure :788-871 with :796-799), driving the oracle's
operators one by one (residuals, B, C, the Schur product, the mass-type
solves). Only the linear algebra between them is new: a GMRES with left
Jacobi preconditioning (diag D, :656-672) written here, with deal.II's
stopping rule: SolverControl(max_its, eps |rhs_pres|) against the norm of the
left-preconditioned residual, as SolverGMRES measures it (a tolerance in the
units of the unpreconditioned right-hand side).

With every tolerance tight and the fixed-point count locked to 3 on both
sides, one time step of the two restatements agrees to round-off; a wrong
source slot or a wrong stage hand-over in either gives an O(dt) difference.
The mesh has cells of unit size (8 x 4 m): the stopping rule above compares
|D^{-1} r| with eps |b|, and with diag D of order one that is a tight
tolerance, so both GMRES implementations stop at the same converged iterate
(on kilometre cells diag D^{-1} is about 1e-7, the effective tolerance is
loose and two GMRES codes stop at iterates that differ by about 1e-6).
The residuals themselves, the tableau and the order of the scheme are pinned
elsewhere (test_oracle_jumps.py, test_oracle_invariants.py,
test_oracle_imex_order.py)."""
import numpy as np

from oracle.pyoracle import Oracle, witch_of_agnesi

EPS = 1e-13


def gmres_left_jacobi(S, Pinv, b, x0, tol, restart=28, max_its=2000):
    """x with |P (b - S x)| <= tol |b| (modified Gram-Schmidt, restarted)."""
    x = x0.copy()
    bn = np.linalg.norm(b)
    its = 0
    while its < max_its:
        r = Pinv * (b - S(x))
        beta = np.linalg.norm(r)
        if beta <= tol * bn:
            return x
        V = [r / beta]
        Hm = np.zeros((restart + 1, restart))
        for j in range(restart):
            w = Pinv * S(V[j])
            for i in range(j + 1):
                Hm[i, j] = w @ V[i]
                w = w - Hm[i, j] * V[i]
            Hm[j + 1, j] = np.linalg.norm(w)
            its += 1
            e1 = np.zeros(j + 2)
            e1[0] = beta
            y = np.linalg.lstsq(Hm[:j + 2, :j + 1], e1, rcond=None)[0]
            res = np.linalg.norm(e1 - Hm[:j + 2, :j + 1] @ y)
            if res <= tol * bn or Hm[j + 1, j] == 0.0:
                break
            V.append(w / Hm[j + 1, j])
        x = x + sum(yi * vi for yi, vi in zip(y, V))
    raise RuntimeError("gmres: no convergence")


def python_step(o, rho_old, u_old, p_old, fp_its=3):
    def solve(ns, b, x0):
        x, _ = o.solve_mass_mg(ns, b, x0=x0, tol=EPS * np.linalg.norm(b))
        return x

    def update_density(hs, srcs, guess):
        o.set_stage(hs, 1 if hs < 3 else 4)
        return solve(1 if hs < 3 else 4, o.rhs_rho(srcs), guess)

    def fixed_point(hs, rho_s, p_start, u_start, srcs_p, srcs_u):
        o.set_rho_for_fixed(rho_s)
        p_fixed_old, u_fixed = p_start, u_start
        for _ in range(fp_its):
            # pressure_fixed_point: rhs_p - C A^{-1} rhs_u, then the Schur complement with GMRES
            o.set_stage(hs, 2)
            o.set_pres_fixed(p_fixed_old)
            rhs_p = o.rhs_pressure(srcs_p(u_fixed))
            rhs_u = o.rhs_velocity(srcs_u)
            o.set_stage(hs, 3)
            tmp1 = solve(3, rhs_u, np.zeros_like(rhs_u))
            o.set_stage(hs, 2)
            rhs_p = rhs_p - o.vmult_enthalpy(tmp1)
            Pinv = o.inverse_diagonal(6)  # diag D, the Schur complement's Jacobi preconditioner
            o.set_stage(hs, 2)
            p_fixed = gmres_left_jacobi(lambda v: o.vmult(v, 2), Pinv, rhs_p, p_fixed_old, EPS)
            # update_velocity: A u = rhs_u - B p_fixed
            o.set_stage(hs, 3)
            u_fixed = solve(3, rhs_u - o.vmult_pressure(p_fixed), u_fixed)
            p_fixed_old = p_fixed
        return p_fixed, u_fixed

    # stage 1
    rho2 = update_density(1, [rho_old, u_old], rho_old)
    p2, u2 = fixed_point(1, rho2, p_old, u_old,
                         lambda uf: [rho_old, u_old, p_old, rho2, uf], [rho_old, u_old, p_old, rho2])
    # stage 2
    rho3 = update_density(2, [rho_old, u_old, rho2, u2], rho2)
    p3, u3 = fixed_point(2, rho3, p2, u2,
                         lambda uf: [rho_old, u_old, p_old, rho2, u2, p2, rho3, uf],
                         [rho_old, u_old, p_old, rho2, u2, p2, rho3])
    # stage 3: density, velocity (NS 5), pressure (NS 6)
    rho_new = update_density(3, [rho_old, u_old, rho2, u2, rho3, u3], rho3)
    o.set_rho_for_fixed(rho_new)
    o.set_stage(3, 5)
    u_new = solve(5, o.rhs_velocity([rho_old, u_old, p_old, rho2, u2, p2, rho3, u3, p3]), u3)
    o.set_stage(3, 6)
    p_new = solve(6, o.rhs_pressure([rho_old, u_old, p_old, rho2, u2, p2, rho3, u3, p3, rho_new, u_new]), p3)
    return rho_new, u_new, p_new


def wiring_errors(library=None):
    k, nx, nz, dt, Lx, H = 2, 8, 4, 0.02, 8.0, 4.0
    h = witch_of_agnesi(nx, Lx=Lx, h0=0.4, a=0.8, xc=4.0)
    o = Oracle(k, nx, nz, Lx=Lx, H=H, hx=h, library=library)
    tight = dict(Ma=1.0, Fr=1.0 / np.sqrt(9.81), dt=dt, eps=EPS, schur_inner_tol=1e-14, a_solve_tol=1e-14, fp_lock=3)
    o.set_params(**tight)
    rng = np.random.default_rng(3)
    rho, u, p = o.init_hill(seed=3, U=12.0)
    rho = rho * (1.0 + 1e-3 * rng.uniform(-1, 1, rho.shape))
    u = u + 0.5 * rng.uniform(-1, 1, u.shape)
    ref = o.run(rho, u, p, 1)[:3]
    mine = python_step(o, rho, u, p)
    return [np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(mine, ref)]


def test_time_step_wiring_matches_second_restatement():
    errs = wiring_errors()
    print("rel L2 (rho, u, p):", ["%.2e" % e for e in errs])
    assert max(errs) < 1e-9
