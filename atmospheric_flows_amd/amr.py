"""Leaf lists of locally refined (quadtree) meshes over the terrain-following
base mesh, in the form atmo_ctx_create_amr takes: rows (bi, bj, level, sx,
sz), base cell by base cell in the order bj*nx + bi, depth-first inside a
base cell with children in the order (0,0), (1,0), (0,1), (1,1); 2:1
balanced across faces (the reference's triangulation is created with
limit_level_difference_at_vertices, Atmospheric_Flow.cc:225).
"""
import numpy as np


class Forest:
    """Quadtrees over an nx x nz base mesh (periodic in x when asked)."""

    def __init__(self, nx, nz, periodic=True):
        self.nx, self.nz, self.periodic = nx, nz, periodic
        # node (bi, bj, l, sx, sz) -> True for a leaf, False for a refined node
        self.nodes = {(i, j, 0, 0, 0): True for j in range(nz) for i in range(nx)}

    def leaves(self):
        return [n for n, leaf in self.nodes.items() if leaf]

    def refine(self, n):
        if not self.nodes.get(n, False):
            return
        bi, bj, l, sx, sz = n
        self.nodes[n] = False
        for b in range(2):
            for a in range(2):
                self.nodes[(bi, bj, l + 1, 2 * sx + a, 2 * sz + b)] = True

    def _neighbour(self, n, f):
        """Same-level position across face f (0 left, 1 right, 2 bottom, 3 top), or None."""
        bi, bj, l, sx, sz = n
        m = 1 << l
        if f == 0:
            sx -= 1
            if sx < 0:
                sx, bi = m - 1, bi - 1
        elif f == 1:
            sx += 1
            if sx == m:
                sx, bi = 0, bi + 1
        elif f == 2:
            sz -= 1
            if sz < 0:
                sz, bj = m - 1, bj - 1
        else:
            sz += 1
            if sz == m:
                sz, bj = 0, bj + 1
        if bj < 0 or bj >= self.nz:
            return None
        if bi < 0 or bi >= self.nx:
            if not self.periodic:
                return None
            bi %= self.nx
        return (bi, bj, l, sx, sz)

    def _too_fine_across(self, n, f):
        """True when a cell two levels finer than leaf n touches its face f."""
        nb = self._neighbour(n, f)
        if nb is None or self.nodes.get(nb, True):
            return False
        bi, bj, l, sx, sz = nb
        for a in range(2):
            cx = 2 * sx + (1 if f == 0 else 0) if f < 2 else 2 * sx + a
            cz = 2 * sz + a if f < 2 else 2 * sz + (1 if f == 2 else 0)
            if not self.nodes.get((bi, bj, l + 1, cx, cz), True):
                return True
        return False

    def balance(self):
        changed = True
        while changed:
            changed = False
            for n in self.leaves():
                if any(self._too_fine_across(n, f) for f in range(4)):
                    self.refine(n)
                    changed = True

    def ordered(self):
        out = []

        def walk(n):
            if self.nodes[n]:
                out.append(n)
                return
            bi, bj, l, sx, sz = n
            for b in range(2):
                for a in range(2):
                    walk((bi, bj, l + 1, 2 * sx + a, 2 * sz + b))

        for j in range(self.nz):
            for i in range(self.nx):
                walk((i, j, 0, 0, 0))
        return np.array(out, dtype=np.int32).reshape(-1, 5)


def cell_centres(leaves, nx, nz, Lx, H, h):
    """Physical centres of the cells (images of the reference sub-box centres)."""
    leaves = np.asarray(leaves)
    bi, bj, l, sx, sz = leaves.T
    n = 2.0 ** l
    xi, eta = (sx + 0.5) / n, (sz + 0.5) / n
    xl, dx = Lx * bi / nx, Lx / nx
    zb = [h[bi] + (H - h[bi]) * bj / nz, h[bi + 1] + (H - h[bi + 1]) * bj / nz,
          h[bi] + (H - h[bi]) * (bj + 1) / nz, h[bi + 1] + (H - h[bi + 1]) * (bj + 1) / nz]
    z = (1 - xi) * (1 - eta) * zb[0] + xi * (1 - eta) * zb[1] + (1 - xi) * eta * zb[2] + xi * eta * zb[3]
    return xl + xi * dx, z


def refine_where(nx, nz, levels, predicate, periodic=True, Lx=40000.0, H=20000.0, h=None):
    """Refine every cell whose centre satisfies predicate(x, z) (vectorized)
    up to `levels` times, then balance."""
    h = np.zeros(nx + 1) if h is None else np.asarray(h, dtype=float)
    F = Forest(nx, nz, periodic)
    for _ in range(levels):
        lv = np.array(F.leaves(), dtype=np.int64).reshape(-1, 5)
        x, z = cell_centres(lv, nx, nz, Lx, H, h)
        for row, flag in zip(lv, predicate(x, z)):
            if flag:
                F.refine(tuple(int(v) for v in row))
        F.balance()
    return F


def background_density(z, theta0=300.0, p0=1e5, N=0.01, g=9.81, R=287.0, gamma=1.4):
    """Density of the constant-N hill background at height z (the state the
    seeded perturbation of 1e-4 is applied to)."""
    cp = gamma * R / (gamma - 1.0)
    th = theta0 * np.exp(N * N * z / g)
    ex = 1.0 + g * g / (cp * theta0 * N * N) * (np.exp(-N * N * z / g) - 1.0)
    p = p0 * ex ** (cp / R)
    return p / (R * th * ex)


def hill_amr_leaves(nx=40, nz=20, levels=3, Lx=40000.0, H=20000.0, h0=400.0, a=1000.0, xc=20000.0,
                    band_half_width=3000.0, band_top=6000.0, grad_quantile=0.9):
    """The locally refined hill mesh: `levels` refinements of the cells whose
    centre lies in |x - xc| < band_half_width, z < band_top, plus one of the
    base cells whose density gradient |d rho/dz| at the centre exceeds the
    `grad_quantile` quantile over the base cells (the 1e-4 seeded
    perturbation changes that gradient by < 1 %, so the background decides),
    then 2:1 balance. Returns (leaves, h)."""
    x = Lx * np.arange(nx + 1) / nx
    h = h0 / (1.0 + ((x - xc) / a) ** 2)
    base = np.array([(i, j, 0, 0, 0) for j in range(nz) for i in range(nx)])
    _, zc = cell_centres(base, nx, nz, Lx, H, h)
    dz = 1e-3
    grad = np.abs(background_density(zc + dz) - background_density(zc - dz)) / (2 * dz)
    flagged = grad > np.quantile(grad, grad_quantile)
    F = Forest(nx, nz, True)
    for row, flag in zip(base, flagged):
        if flag:
            F.refine(tuple(int(v) for v in row))
    for _ in range(levels):
        lv = np.array(F.leaves(), dtype=np.int64).reshape(-1, 5)
        cx, cz = cell_centres(lv, nx, nz, Lx, H, h)
        inband = (np.abs(cx - xc) < band_half_width) & (cz < band_top) & (lv[:, 2] < levels)
        for row, flag in zip(lv, inband):
            if flag:
                F.refine(tuple(int(v) for v in row))
        F.balance()
    return F.ordered(), h
