"""Temporal self-convergence of the oracle's IMEX step (EulerSolver::run,
Atmospheric_Flow.cc:1115-1222, with the tableau of euler_operator.h:220-224).

The scheme is second order (sum b = 1, sum b c = 1/2), so with every solve
converged far below the time error the differences between runs to a fixed
final time with dt, dt/2, dt/4, dt/8 shrink by 4 per halving. A wrong stage
weight, a wrong source slot in a stage's residual or a wrong stage wiring
breaks an order condition (or consistency) and the ratio drops to about 2
or below. The state is the hill flow on a coarse steep-hill mesh with a
smooth density bubble (a random per-dof perturbation would excite grid-scale
modes that are not in the asymptotic range at these steps).
"""
import numpy as np

from oracle.pyoracle import Oracle, tables, witch_of_agnesi

LX, H = 40000.0, 20000.0


def bubble_state(o, k, nx, nz, h, amp=1e-2):
    rho, u, p = o.init_hill(pert=0.0, seed=5, U=10.0)
    gll = tables(k)[0]
    nd = k + 1
    dx = LX / nx
    X = np.zeros(o.Ns)
    Z = np.zeros(o.Ns)
    for j in range(nz):
        for i in range(nx):
            c = j * nx + i
            for jj in range(nd):
                for ii in range(nd):
                    hx = h[i] + (h[i + 1] - h[i]) * gll[ii]
                    X[c * nd * nd + jj * nd + ii] = (i + gll[ii]) * dx
                    Z[c * nd * nd + jj * nd + ii] = hx + (H - hx) * (j + gll[jj]) / nz
    bub = np.exp(-((X - 15000.0) / 6000.0) ** 2 - ((Z - 6000.0) / 4000.0) ** 2)
    return rho * (1.0 + amp * bub), u, p


def order_ratios(library=None, T=1.0, steps=(2, 4, 8, 16)):
    """Ratios of successive differences |x(dt) - x(dt/2)| per field (rho, u, p)."""
    k, nx, nz = 2, 8, 4
    h = witch_of_agnesi(nx, h0=2000.0, a=4000.0)
    o = Oracle(k, nx, nz, hx=h, library=library)
    s0 = bubble_state(o, k, nx, nz, h)
    res = {}
    for n in steps:
        o.set_params(Ma=1.0, Fr=1.0 / np.sqrt(9.81), dt=T / n, eps=1e-13, schur_inner_tol=1e-14,
                     a_solve_tol=1e-14, fp_tol=1e-14)
        res[n] = o.run(*s0, n)[:3]
    out = []
    for i in range(3):
        d = [np.linalg.norm(res[a][i] - res[b][i]) for a, b in zip(steps[:-1], steps[1:])]
        out.append([d[j] / d[j + 1] for j in range(len(d) - 1)])
    return out


def test_imex_step_is_second_order_in_time():
    ratios = order_ratios()
    for name, r in zip(("rho", "u", "p"), ratios):
        print(name, ["%.3f" % x for x in r])
        assert all(3.4 < x < 4.6 for x in r), (name, r)
