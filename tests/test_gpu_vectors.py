"""Device vector operations of the C ABI (atmo_vec_axpby / atmo_vec_dot /
atmo_vec_linf), as the reference's fixed-point test uses them
(Atmospheric_Flow.cc:1138-1150: pres_tmp.equ(1, pres_fixed);
pres_tmp.add(-1, pres_fixed_old); L-infinity norms at the QGaussLobatto points),
against numpy on the same data."""
import numpy as np
import pytest

import atmospheric_flows_amd as af

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return af.DeviceContext(3, 16, 8, h=af.witch_of_agnesi(16), Ma=1.0, Fr=1.0 / np.sqrt(9.81), dt=4.0)


@pytest.mark.parametrize("block", [af.BLOCK_VELOCITY, af.BLOCK_PRESSURE])
def test_axpby_dot_linf_against_numpy(ctx, block):
    rng = np.random.default_rng(5 + block)
    n = ctx.vector(block).size
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    dx, dy = ctx.vector(block, x), ctx.vector(block, y)
    # scalar product: fixed-order sum, reproducible bit for bit
    d1, d2 = dx * dy, dx * dy
    assert d1 == d2
    assert d1 == pytest.approx(float(x @ y), rel=1e-13, abs=1e-13 * np.sqrt(n))
    assert dx.linfty_norm() == np.abs(x).max()
    # y = 0.5 y - 2 x (sadd), then equ and add as the fixed-point test uses them
    dy.sadd(0.5, -2.0, dx)
    np.testing.assert_allclose(dy.download(), 0.5 * y - 2.0 * x, rtol=1e-15, atol=1e-15)
    t = ctx.vector(block)
    t.equ(1.0, dx)
    assert np.array_equal(t.download(), x)
    t.add(-1.0, dx)
    assert t.linfty_norm() == 0.0


def test_equ_overwrites_nonfinite_destination(ctx):
    """equ(a, x) sets the destination, whatever it held (b = 0 reads nothing)."""
    n = ctx.vector(af.BLOCK_PRESSURE).size
    x = np.linspace(-1.0, 1.0, n)
    t = ctx.vector(af.BLOCK_PRESSURE, np.full(n, np.inf))
    t.equ(3.0, ctx.vector(af.BLOCK_PRESSURE, x))
    assert np.array_equal(t.download(), 3.0 * x)


def test_fixed_point_error_on_device(ctx):
    """The relative L-infinity change of the fixed-point loop formed with the
    device vector operations (Atmospheric_Flow.cc:1140-1148)."""
    rng = np.random.default_rng(9)
    n = ctx.vector(af.BLOCK_PRESSURE).size
    p_new = 1e5 * (1 + 0.01 * rng.standard_normal(n))
    p_old = p_new * (1 + 1e-9 * rng.standard_normal(n))
    dn, do = ctx.vector(af.BLOCK_PRESSURE, p_new), ctx.vector(af.BLOCK_PRESSURE, p_old)
    tmp = ctx.vector(af.BLOCK_PRESSURE)
    den = dn.linfty_norm()
    tmp.equ(1.0, dn)
    tmp.add(-1.0, do)
    err = tmp.linfty_norm() / (den + 1e-10)
    assert err == pytest.approx(np.abs(p_new - p_old).max() / (np.abs(p_new).max() + 1e-10), rel=1e-15)


def test_mismatched_blocks_are_rejected(ctx):
    with pytest.raises(af.AtmoError):
        ctx.vector(af.BLOCK_VELOCITY) * ctx.vector(af.BLOCK_PRESSURE)
    with pytest.raises(af.AtmoError):
        ctx.vector(af.BLOCK_PRESSURE).add(1.0, ctx.vector(af.BLOCK_VELOCITY))
