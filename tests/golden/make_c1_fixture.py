"""Writes c1_hill_40x20_Q2.npz: the oracle's state after 2 and after 100 IMEX
steps of the 40x20 Q2 hill case (dt = 5 s, U = 10 m/s, N = 0.01 1/s, seeded
density perturbation), solved in parity mode (GMRES to 1e-13, inner and A
solves to 1e-14, three fixed-point iterations per stage on both sides).

The 100-step run takes about two minutes on eight host threads, too long for
the test suite, so its result is kept as a fixture; the GPU test runs the same
100 steps and compares against it.

Run from the repo root:  python tests/golden/make_c1_fixture.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle.pyoracle import Oracle, witch_of_agnesi  # noqa: E402

K, NX, NZ, DT, SEED = 2, 40, 20, 5.0, 1234
PARITY = dict(eps=1e-13, schur_inner_tol=1e-14, a_solve_tol=1e-14, fp_lock=3)
CHECKPOINTS = (2, 100)


def main():
    o = Oracle(K, NX, NZ, hx=witch_of_agnesi(NX), periodic=True)
    o.set_params(Ma=1.0, Fr=1.0 / np.sqrt(9.81), dt=DT, **PARITY)
    rho, u, p = o.init_hill(seed=SEED)
    out = dict(k=K, nx=NX, nz=NZ, dt=DT, seed=SEED, rho0=rho, u0=u, p0=p)
    done = 0
    t0 = time.time()
    for n in CHECKPOINTS:
        rho, u, p, st = o.run(rho, u, p, n - done)
        done = n
        out[f"rho_{n}"], out[f"u_{n}"], out[f"p_{n}"] = rho, u, p
        out[f"gmres_its_{n}"] = st["gmres_its"]
        print(f"step {n}: {time.time() - t0:.1f} s, stats {st}", flush=True)
    np.savez_compressed(os.path.join(HERE, "c1_hill_40x20_Q2.npz"), **out)


if __name__ == "__main__":
    main()
