"""Pins the oracle's per-step diagnostics (Atmospheric_Flow.cc:991-1087).

The reference evaluates the fields at the QGaussLobatto(k+1) points
(quadrature_density/velocity/temperature, Atmospheric_Flow.cc:86-88,
233-235). Fields that are polynomials of degree <= k per direction are
represented exactly, so their extrema over those points follow from the
Gauss-Lobatto nodes alone, computed here from the roots of P_k' (numpy),
independently of the oracle's tables and interpolation. The fields are
chosen so that the Gauss points of the same count give other extrema, so a
diagnostic evaluated at the wrong point set fails."""
import numpy as np
import pytest

from oracle.pyoracle import Oracle, tables, witch_of_agnesi

LX, H = 40000.0, 20000.0


def gll_unit(k):
    """Gauss-Lobatto nodes of k+1 points on [0, 1], ascending: the end points
    and the roots of the derivative of the Legendre polynomial P_k."""
    inner = np.polynomial.legendre.Legendre.basis(k).deriv().roots() if k > 1 else np.array([])
    x = np.concatenate([[-1.0], np.sort(inner.real), [1.0]])
    return 0.5 * (x + 1.0)


def gauss_unit(k):
    return np.sort(0.5 * (np.polynomial.legendre.leggauss(k + 1)[0] + 1.0))


def nodal(o, k, nx, nz, f):
    """Nodal values of f(x, z) on a flat mesh, [cell][dof] layout."""
    gll = gll_unit(k)
    dx, dz = LX / nx, H / nz
    out = np.zeros(o.Ns)
    nd = k + 1
    for j in range(nz):
        for i in range(nx):
            c = j * nx + i
            for jj in range(nd):
                for ii in range(nd):
                    out[c * nd * nd + jj * nd + ii] = f((i + gll[ii]) * dx, (j + gll[jj]) * dz)
    return out


def extremum(k, nx, nz, f, pts, op):
    """op (min or max) of f over the points pts (on [0,1]) of every cell of a flat mesh."""
    dx, dz = LX / nx, H / nz
    xs = np.concatenate([(i + pts) * dx for i in range(nx)])
    zs = np.concatenate([(j + pts) * dz for j in range(nz)])
    X, Z = np.meshgrid(xs, zs)
    return op(f(X, Z))


@pytest.mark.parametrize("k", [1, 2, 3])
def test_gll_table_is_gauss_lobatto(k):
    np.testing.assert_allclose(tables(k)[0], gll_unit(k), rtol=0, atol=1e-15)


@pytest.mark.parametrize("k", [2, 3])
def test_extrema_at_gauss_lobatto_points(k):
    nx, nz = 4, 2
    o = Oracle(k, nx, nz)
    # interior extrema: a parabola in x with its vertex inside a cell (degree 2 <= k)
    x0, z0 = 0.2 * LX / nx, 0.15 * H / nz
    fr = lambda x, z: 1.0 + 1e-9 * (x - x0) ** 2 + 2e-9 * (z - z0) ** 2
    fp = lambda x, z: 1e5 * (1.0 - 1e-10 * (x - x0) ** 2)
    fu = lambda x, z: 10.0 - 1e-8 * (x - x0) ** 2
    fw = lambda x, z: 1e-4 * (z - z0)
    rho, p = nodal(o, k, nx, nz, fr), nodal(o, k, nx, nz, fp)
    u = np.concatenate([nodal(o, k, nx, nz, fu), nodal(o, k, nx, nz, fw)])
    g = 1.4
    mn, mx, cel, vel, bad = o.diagnostics(rho, u, p, g)
    gl, ga = gll_unit(k), gauss_unit(k)
    want_min = extremum(k, nx, nz, fr, gl, np.min)
    want_cel = extremum(k, nx, nz, lambda x, z: np.sqrt(g * fp(x, z) / fr(x, z)), gl, np.max)
    want_vel = extremum(k, nx, nz, lambda x, z: np.hypot(fu(x, z), fw(x, z)), gl, np.max)
    assert mn == pytest.approx(want_min, rel=1e-14)
    assert mx == pytest.approx(np.max(np.abs(rho)), rel=0)  # linfty_norm of the dofs
    assert cel == pytest.approx(want_cel, rel=1e-13)
    assert vel == pytest.approx(want_vel, rel=1e-14)
    assert bad == 0
    # the Gauss points of the same count give other extrema: the test sees the point set
    assert abs(extremum(k, nx, nz, fr, ga, np.min) - want_min) > 1e-12 * want_min
    assert abs(extremum(k, nx, nz, lambda x, z: np.hypot(fu(x, z), fw(x, z)), ga, np.max) - want_vel) > 1e-10


def test_nonfinite_values_are_counted_not_hidden():
    k, nx, nz = 3, 4, 2
    o = Oracle(k, nx, nz)
    rho = np.ones(o.Ns)
    p = np.full(o.Ns, 1e5)
    u = np.zeros(2 * o.Ns)
    rho[5] = np.nan
    mn, mx, cel, vel, bad = o.diagnostics(rho, u, p)
    # the NaN dof is its own point's value: counted by the density minimum, the celerity and linfty_norm
    assert bad == 3
    assert mn == pytest.approx(1.0, rel=1e-14) and mx == 1.0


def test_min_cell_diameter():
    nx, nz = 16, 8
    h = witch_of_agnesi(nx)
    o = Oracle(2, nx, nz, hx=h)
    dx = LX / nx
    best = np.inf
    for j in range(nz):
        for i in range(nx):
            z = [h[i] + (H - h[i]) * j / nz, h[i + 1] + (H - h[i + 1]) * j / nz,
                 h[i] + (H - h[i]) * (j + 1) / nz, h[i + 1] + (H - h[i + 1]) * (j + 1) / nz]
            best = min(best, max(np.hypot(dx, z[3] - z[0]), np.hypot(dx, z[2] - z[1])))
    assert o.min_cell_diameter() == pytest.approx(best, rel=1e-15)
    flat = Oracle(2, nx, nz)
    assert flat.min_cell_diameter() == pytest.approx(np.hypot(dx, H / nz), rel=1e-15)
