"""Writes the golden fixtures dense_*.npz from the independent numpy
brute-force assembly in dense_ref.py (no reference code is run: the
reference needs deal.II, which is not available; see the oracle header).

Run from the repo root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import dense_ref as R  # noqa: E402

GAMMA = 1.4
DT = 2.0
MA = 1.0
A22T = 0.5 * (2.0 - np.sqrt(2.0))

CASES = {
    # name: (k, nx, nz, periodic, Lx, H, terrain)
    "k1_hill_periodic": (1, 4, 2, True, 3000.0, 2000.0, "hill"),
    "k2_hill_periodic": (2, 3, 2, True, 3000.0, 2000.0, "hill"),
    "k3_rough_walls": (3, 2, 2, False, 3000.0, 2000.0, "rough"),
}


def terrain(kind, nx, Lx):
    x = Lx * np.arange(nx + 1) / nx
    if kind == "hill":
        h = 300.0 / (1.0 + ((x - Lx / 2) / 500.0) ** 2)
        h[-1] = h[0]
        return h
    rng = np.random.default_rng(5)
    return 150.0 * rng.uniform(0, 1, nx + 1)


def main():
    for name, (k, nx, nz, per, Lx, H, kind) in CASES.items():
        h = terrain(kind, nx, Lx)
        d = R.Disc(k, R.Mesh(nx, nz, Lx, H, h, per))
        rng = np.random.default_rng(sum(map(ord, name)))
        Ns = d.Ns
        rho = 1.0 + 0.2 * rng.uniform(-1, 1, Ns)
        pf = 1e5 * (1.0 + 0.1 * rng.uniform(-1, 1, Ns))
        u = np.concatenate([10 + 5 * rng.uniform(-1, 1, Ns), 4 * rng.uniform(-1, 1, Ns)])
        M = R.mass_matrix(d)
        A = R.A_matrix(d, rho)
        cf = A22T * DT / MA ** 2
        B = R.B_matrix(d, cf)
        Cu = R.C_apply(d, A22T * DT, GAMMA, pf, u)
        np.savez_compressed(os.path.join(HERE, f"dense_{name}.npz"), k=k, nx=nx, nz=nz, periodic=per, Lx=Lx, H=H,
                            h=h, rho=rho, pf=pf, u=u, M=M, A=A, B=B, Cu=Cu, dt=DT, Ma=MA, gamma=GAMMA)
        print(name, "Ns", Ns)


if __name__ == "__main__":
    main()
