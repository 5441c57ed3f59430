"""Pins the oracle's per-step diagnostics (Atmospheric_Flow.cc:991-1087) with
closed forms: fields that are polynomials of degree <= k are represented
exactly, so their extrema over the QGauss(k+1) points follow from the Gauss
nodes alone (numpy leggauss), independently of the oracle's interpolation."""
import numpy as np
import pytest

from oracle.pyoracle import Oracle, tables, witch_of_agnesi

LX, H = 40000.0, 20000.0


def nodal(o, k, nx, nz, f):
    """Nodal values of f(x, z) on a flat mesh, [cell][dof] layout."""
    gll = tables(k)[0]
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


@pytest.mark.parametrize("k", [2, 3])
def test_extrema_of_polynomial_fields(k):
    nx, nz = 4, 2
    o = Oracle(k, nx, nz)
    a, b = 2e-6, 3e-6
    rho = nodal(o, k, nx, nz, lambda x, z: 1.0 + a * x + b * z)
    c = 500.0
    p = nodal(o, k, nx, nz, lambda x, z: c * (1.0 + a * x + b * z) ** 2)
    u = np.concatenate([nodal(o, k, nx, nz, lambda x, z: 10.0 + 1e-4 * z), nodal(o, k, nx, nz, lambda x, z: -2.0)])
    g = 1.4
    mn, mx, cel, vel, bad = o.diagnostics(rho, u, p, g)
    xg = 0.5 * (1.0 - np.polynomial.legendre.leggauss(k + 1)[0][::-1])  # ascending on [0,1]
    dx, dz = LX / nx, H / nz
    x_lo, z_lo = xg[0] * dx, xg[0] * dz
    x_hi, z_hi = LX - dx + xg[-1] * dx, H - dz + xg[-1] * dz
    assert mn == pytest.approx(1.0 + a * x_lo + b * z_lo, rel=1e-14)
    assert mx == pytest.approx(1.0 + a * LX + b * H, rel=1e-14)  # GLL nodes include the far corner
    assert cel == pytest.approx(np.sqrt(g * c * (1.0 + a * x_hi + b * z_hi)), rel=1e-13)
    assert vel == pytest.approx(np.hypot(10.0 + 1e-4 * z_hi, 2.0), rel=1e-14)
    assert bad == 0


def test_nonfinite_values_are_counted_not_hidden():
    k, nx, nz = 3, 4, 2
    o = Oracle(k, nx, nz)
    rho = np.ones(o.Ns)
    p = np.full(o.Ns, 1e5)
    u = np.zeros(2 * o.Ns)
    rho[5] = np.nan
    mn, mx, cel, vel, bad = o.diagnostics(rho, u, p)
    # the NaN dof enters all 16 points of its cell (density and celerity) and itself once
    assert bad == 16 + 16 + 1
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
