"""Independent numpy brute-force assembly of the DG operators on tiny meshes.

Used to write the golden fixtures (make_golden.py) that pin the oracle. It
shares no code with the oracle: the Lagrange basis comes from a Vandermonde
inverse, quadrature from numpy's leggauss, basis functions are evaluated as
full 2D products at every quadrature point (no sum factorization), geometry
from bilinear shape-function derivatives, and face normals/lengths from the
physical edge vectors.

Weak forms (euler_operator.h):
  M   : (u, v)                                  :717-740
  A   : (rho u, v)                              :2328-2357
  B   : -cf (p, div v) + cf {p}[v.n] + cf p v.n :1510-1617, cf = c dt / Ma^2
  C(u): -cdt g/(g-1) (pf u, grad q)
        + cdt ({g/(g-1) pf u}.n + lam/2 [pf]/(g-1)) [q]   :1411-1507
"""
import numpy as np


def gll_nodes(k):
    if k == 1:
        r = np.array([-1.0, 1.0])
    else:
        inner = np.polynomial.legendre.Legendre.basis(k).deriv().roots()
        r = np.concatenate([[-1.0], np.sort(inner.real), [1.0]])
    return (r + 1.0) / 2.0


def lagrange_coeffs(nodes):
    V = np.vander(nodes, increasing=True)
    return np.linalg.inv(V)  # column j: monomial coefficients of l_j


def lag_eval(C, x, der=0):
    """values (der=0) or derivatives (der=1) of all basis functions at x."""
    n = C.shape[0]
    x = np.atleast_1d(x)
    P = np.zeros((x.size, n))
    for m in range(n):
        if der == 0:
            P[:, m] = x ** m
        elif m >= 1:
            P[:, m] = m * x ** (m - 1)
    return P @ C  # [point, basis j]


def gauss(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return (x + 1) / 2, w / 2


class Mesh:
    def __init__(self, nx, nz, Lx, H, h, periodic):
        self.nx, self.nz, self.periodic = nx, nz, periodic
        self.X = np.zeros((nz + 1, nx + 1))
        self.Z = np.zeros((nz + 1, nx + 1))
        for j in range(nz + 1):
            for i in range(nx + 1):
                self.X[j, i] = Lx * i / nx
                self.Z[j, i] = h[i] + (H - h[i]) * j / nz

    def verts(self, c):
        i, j = c % self.nx, c // self.nx
        return np.array([[self.X[j, i], self.Z[j, i]], [self.X[j, i + 1], self.Z[j, i + 1]],
                         [self.X[j + 1, i], self.Z[j + 1, i]], [self.X[j + 1, i + 1], self.Z[j + 1, i + 1]]])


def bilinear(v, xi, eta):
    N = np.array([(1 - xi) * (1 - eta), xi * (1 - eta), (1 - xi) * eta, xi * eta])
    dNx = np.array([-(1 - eta), (1 - eta), -eta, eta])
    dNe = np.array([-(1 - xi), -xi, (1 - xi), xi])
    pt = N @ v
    J = np.column_stack([dNx @ v, dNe @ v])  # J[a][b] = d x_a / d xi_b
    return pt, J


class Disc:
    def __init__(self, k, mesh):
        self.k, self.m = k, mesh
        self.nd = k + 1
        self.nd2 = self.nd ** 2
        self.nc = mesh.nx * mesh.nz
        self.Ns = self.nc * self.nd2
        self.C = lagrange_coeffs(gll_nodes(k))
        self.gq, self.gw = gauss(2 * k + 1)

    def phi(self, xi, eta):
        """values and reference gradients of the nd2 basis functions at one point."""
        lx, ly = lag_eval(self.C, xi)[0], lag_eval(self.C, eta)[0]
        dx, dy = lag_eval(self.C, xi, 1)[0], lag_eval(self.C, eta, 1)[0]
        v = np.outer(ly, lx).ravel()
        gx = np.outer(ly, dx).ravel()
        gy = np.outer(dy, lx).ravel()
        return v, gx, gy

    def cell_points(self, c):
        """yields (phi, grad_phi_physical[2,nd2], JxW, point) at the cell's Gauss points."""
        v = self.m.verts(c)
        for b, eta in enumerate(self.gq):
            for a, xi in enumerate(self.gq):
                pt, J = bilinear(v, xi, eta)
                det = np.linalg.det(J)
                ph, gx, gy = self.phi(xi, eta)
                grad = np.linalg.inv(J).T @ np.vstack([gx, gy])
                yield ph, grad, det * self.gw[a] * self.gw[b], pt

    def faces(self):
        """inner faces (cp, fp, cm, fm) and boundary faces (c, f); f: 0 left 1 right 2 bottom 3 top."""
        m = self.m
        inner, bdry = [], []
        for j in range(m.nz):
            for i in range(m.nx):
                c = j * m.nx + i
                if i + 1 < m.nx:
                    inner.append((c, 1, c + 1, 0))
                elif m.periodic:
                    inner.append((c, 1, j * m.nx, 0))
                else:
                    bdry.append((c, 1))
                if i == 0 and not m.periodic:
                    bdry.append((c, 0))
                if j + 1 < m.nz:
                    inner.append((c, 3, c + m.nx, 2))
                else:
                    bdry.append((c, 3))
                if j == 0:
                    bdry.append((c, 2))
        return inner, bdry

    def face_points(self, c, f):
        """(phi at the face points, outward normal, weights*length) with points
        ordered by the edge parameter (left->right or bottom->top)."""
        v = self.m.verts(c)
        ends = {0: (v[0], v[2]), 1: (v[1], v[3]), 2: (v[0], v[1]), 3: (v[2], v[3])}[f]
        t = ends[1] - ends[0]
        L = np.hypot(t[0], t[1])
        n = np.array([t[1], -t[0]]) / L  # right-hand normal of the edge direction
        if f in (0, 3):
            n = -n
        out = []
        for s, w in zip(self.gq, self.gw):
            xi, eta = {0: (0.0, s), 1: (1.0, s), 2: (s, 0.0), 3: (s, 1.0)}[f]
            ph, _, _ = self.phi(xi, eta)
            out.append((ph, w * L))
        return out, n

    def sl(self, c):
        return slice(c * self.nd2, (c + 1) * self.nd2)


def mass_matrix(d, coeff=1.0):
    M = np.zeros((d.Ns, d.Ns))
    for c in range(d.nc):
        s = d.sl(c)
        for ph, _, jw, _ in d.cell_points(c):
            M[s, s] += coeff * jw * np.outer(ph, ph)
    return M


def A_matrix(d, rho):
    A = np.zeros((2 * d.Ns, 2 * d.Ns))
    for c in range(d.nc):
        s = d.sl(c)
        for ph, _, jw, _ in d.cell_points(c):
            r = ph @ rho[s]
            blk = r * jw * np.outer(ph, ph)
            A[s, s] += blk
            A[d.Ns + s.start:d.Ns + s.stop, d.Ns + s.start:d.Ns + s.stop] += blk
    return A


def B_matrix(d, cf):
    Ns = d.Ns
    B = np.zeros((2 * Ns, Ns))
    for c in range(d.nc):
        s = d.sl(c)
        for ph, grad, jw, _ in d.cell_points(c):
            for comp in range(2):
                B[comp * Ns + s.start:comp * Ns + s.stop, s] += -cf * jw * np.outer(grad[comp], ph)
    inner, bdry = d.faces()
    for cp, fp, cm, fm in inner:
        P, n = d.face_points(cp, fp)
        Mq, _ = d.face_points(cm, fm)
        sp, sm = d.sl(cp), d.sl(cm)
        for (php, w), (phm, _) in zip(P, Mq):
            for comp in range(2):
                rp = slice(comp * Ns + sp.start, comp * Ns + sp.stop)
                rm = slice(comp * Ns + sm.start, comp * Ns + sm.stop)
                # {p}[v.n]: test v+ with +n, v- with -n
                B[rp, sp] += cf * w * 0.5 * n[comp] * np.outer(php, php)
                B[rp, sm] += cf * w * 0.5 * n[comp] * np.outer(php, phm)
                B[rm, sp] -= cf * w * 0.5 * n[comp] * np.outer(phm, php)
                B[rm, sm] -= cf * w * 0.5 * n[comp] * np.outer(phm, phm)
    for c, f in bdry:
        P, n = d.face_points(c, f)
        s = d.sl(c)
        for ph, w in P:
            for comp in range(2):
                B[comp * Ns + s.start:comp * Ns + s.stop, s] += cf * w * n[comp] * np.outer(ph, ph)
    return B


def C_apply(d, cdt, gm, pf, u):
    Ns = d.Ns
    out = np.zeros(Ns)
    for c in range(d.nc):
        s = d.sl(c)
        for ph, grad, jw, _ in d.cell_points(c):
            p = ph @ pf[s]
            uv = np.array([ph @ u[s], ph @ u[Ns + s.start:Ns + s.stop]])
            out[s] += -cdt * gm / (gm - 1) * p * jw * (uv @ grad)
    inner, _ = d.faces()
    for cp, fp, cm, fm in inner:
        P, n = d.face_points(cp, fp)
        Mq, _ = d.face_points(cm, fm)
        sp, sm = d.sl(cp), d.sl(cm)
        for (php, w), (phm, _) in zip(P, Mq):
            pp, pm = php @ pf[sp], phm @ pf[sm]
            up = np.array([php @ u[sp], php @ u[Ns + sp.start:Ns + sp.stop]])
            um = np.array([phm @ u[sm], phm @ u[Ns + sm.start:Ns + sm.stop]])
            lam = max(abs(up @ n), abs(um @ n))
            F = cdt * (0.5 * gm / (gm - 1) * (pp * up + pm * um) @ n + 0.5 * lam * (pp - pm) / (gm - 1))
            out[sp] += w * F * php
            out[sm] -= w * F * phm
    return out


def node_coords(d):
    """physical coordinates of every dof (nodal interpolation points)."""
    x = gll_nodes(d.k)
    pts = np.zeros((d.Ns, 2))
    for c in range(d.nc):
        v = d.m.verts(c)
        for jj in range(d.nd):
            for ii in range(d.nd):
                pt, _ = bilinear(v, x[ii], x[jj])
                pts[c * d.nd2 + jj * d.nd + ii] = pt
    return pts
