"""VTU output (output_results, Atmospheric_Flow.cc:886-918) through the C ABI.

The host writer is checked without a GPU on states whose exact values are
known at every output point: the nodal interpolation of a function that is
linear in x and z is exact on the bilinear cells (x is linear and z bilinear
in the reference coordinates, so the function is bilinear there), so the
written values must equal the function at the written coordinates. The GPU
test writes a context's state after a step and compares the Lagrange-cell
nodes with the downloaded dofs.
"""
import base64
import ctypes
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import atmospheric_flows_amd as af
from atmospheric_flows_amd import _native as N

DT = {"Float64": np.float64, "Int32": np.int32, "UInt8": np.uint8}


def read_vtu(path):
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    out = {"npoints": int(piece.get("NumberOfPoints")), "ncells": int(piece.get("NumberOfCells"))}
    for da in root.iter("DataArray"):
        raw = base64.b64decode(da.text.strip())
        nbytes = int(np.frombuffer(raw[:4], dtype=np.uint32)[0])
        a = np.frombuffer(raw[4:4 + nbytes], dtype=DT[da.get("type")])
        nc = int(da.get("NumberOfComponents"))
        out[da.get("Name") or "points"] = a.reshape(-1, nc) if nc > 1 else a
    return out


def gll(k):
    return {1: [0.0, 1.0], 2: [0.0, 0.5, 1.0], 3: [0.0, (1 - 5 ** -0.5) / 2, (1 + 5 ** -0.5) / 2, 1.0]}[k]


def nodal_state(k, nx, nz, h, Lx, H, f):
    """f interpolated at the FE_DGQ(k) nodes of every cell (dof = jj*(k+1) + ii)."""
    xs = gll(k)
    nd = k + 1
    vals = np.zeros(nx * nz * nd * nd)
    for c in range(nx * nz):
        i, j = c % nx, c // nx
        dx = Lx / nx
        z = lambda ii, jj: h[ii] + (H - h[ii]) * jj / nz  # noqa: E731
        for jj in range(nd):
            for ii in range(nd):
                xi, eta = xs[ii], xs[jj]
                x = i * dx + xi * dx
                zz = ((1 - xi) * (1 - eta) * z(i, j) + xi * (1 - eta) * z(i + 1, j) + (1 - xi) * eta * z(i, j + 1)
                      + xi * eta * z(i + 1, j + 1))
                vals[c * nd * nd + jj * nd + ii] = f(x, zz)
    return vals


@pytest.mark.parametrize("k,high_order", [(2, False), (3, False), (3, True), (1, True)])
def test_host_vtu_values_and_cells(tmp_path, k, high_order):
    nx, nz, Lx, H = 3, 2, 3000.0, 2000.0
    x = Lx * np.arange(nx + 1) / nx
    h = 300.0 / (1.0 + ((x - Lx / 2) / 500.0) ** 2)
    h[-1] = h[0]
    f = lambda x, z: 1.0 + 1e-3 * x + 2e-3 * z  # noqa: E731
    rho = nodal_state(k, nx, nz, h, Lx, H, f)
    u = np.concatenate([2 * rho, -rho])
    p = 3 * rho
    mesh = N.MeshDesc(k, nx, nz, 1, Lx, H, h.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    path = tmp_path / "s.vtu"
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
    N.check(N.load().atmo_write_vtu_host(ctypes.byref(mesh), dp(rho), dp(u), dp(p), str(path).encode(),
                                         1 if high_order else 0))
    v = read_vtu(path)
    ncell, ppc = nx * nz, (k + 1) ** 2
    assert v["npoints"] == ncell * ppc
    assert v["ncells"] == (ncell if high_order else ncell * k * k)
    pts = v["points"]
    exact = f(pts[:, 0], pts[:, 1])
    assert np.abs(v["rho"] - exact).max() < 1e-12 * np.abs(exact).max()
    assert np.abs(v["u"][:, 0] - 2 * exact).max() < 1e-12 * np.abs(exact).max()
    assert np.abs(v["u"][:, 1] + exact).max() < 1e-12 * np.abs(exact).max()
    assert np.all(v["u"][:, 2] == 0) and np.all(pts[:, 2] == 0)
    assert np.abs(v["p"] - 3 * exact).max() < 1e-12 * np.abs(exact).max()
    conn, offs, types = v["connectivity"], v["offsets"], v["types"]
    assert conn.min() == 0 and conn.max() == ncell * ppc - 1
    if high_order:
        assert np.all(types == 70) and np.array_equal(offs, ppc * np.arange(1, ncell + 1))
        # VTK Lagrange order: the four corners first, counterclockwise from (xi, eta) = (0, 0)
        for c in range(ncell):
            i, j = c % nx, c // nx
            cp = pts[conn[c * ppc:c * ppc + 4]]
            xs = np.array([i, i + 1, i + 1, i]) * Lx / nx
            assert np.allclose(cp[:, 0], xs, rtol=0, atol=1e-9)
            zc = [h[i] + (H - h[i]) * j / nz, h[i + 1] + (H - h[i + 1]) * j / nz,
                  h[i + 1] + (H - h[i + 1]) * (j + 1) / nz, h[i] + (H - h[i]) * (j + 1) / nz]
            assert np.allclose(cp[:, 1], zc, rtol=0, atol=1e-9)
    else:
        assert np.all(types == 9) and np.array_equal(offs, 4 * np.arange(1, ncell * k * k + 1))
        q = pts[conn.reshape(-1, 4)]
        # every quadrilateral counterclockwise (positive shoelace area), together tiling the domain
        area = 0.5 * np.sum(q[:, :, 0] * np.roll(q[:, :, 1], -1, 1) - np.roll(q[:, :, 0], -1, 1) * q[:, :, 1], axis=1)
        assert np.all(area > 0)
        hmid = np.trapezoid(h, x)
        assert abs(area.sum() - (Lx * H - hmid)) < 1e-6 * Lx * H


def test_host_vtu_rejects_bad_arguments(tmp_path):
    h = np.zeros(3)
    mesh = N.MeshDesc(4, 2, 2, 1, 1.0, 1.0, h.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    a = np.zeros(2 * 4 * 16)  # two components of the largest mesh below (4 cells, 16 dofs each at degree 3)
    dp = a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert N.load().atmo_write_vtu_host(ctypes.byref(mesh), dp, dp, dp, str(tmp_path / "x.vtu").encode(), 0) \
        == N.ATMO_INVALID_ARG
    mesh.degree = 2
    assert N.load().atmo_write_vtu_host(ctypes.byref(mesh), dp, dp, dp, b"/nonexistent_dir/x.vtu", 0) \
        == N.ATMO_INVALID_ARG


@pytest.mark.gpu
@pytest.mark.parametrize("high_order", [True, False])
def test_context_vtu_after_a_step(tmp_path, high_order):
    """The context's writer (device state copied to the host, the context's
    mesh) gives the same file content as the host writer, pinned above, on the
    downloaded state."""
    k, nx, nz = 2, 8, 4
    h = af.witch_of_agnesi(nx)
    dev = af.DeviceContext(k, nx, nz, h=h, Ma=1.0, Fr=1.0 / np.sqrt(9.81), dt=5.0)
    rho, u, p = dev.init_hill()
    dev.upload_state(rho, u, p)
    dev.imex_step(1)
    r1, u1, p1 = dev.download_state()
    dev.write_vtu(tmp_path / "ctx.vtu", high_order=high_order)
    mesh = N.MeshDesc(k, nx, nz, 1, 40000.0, 20000.0, h.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
    N.check(N.load().atmo_write_vtu_host(ctypes.byref(mesh), dp(r1), dp(u1), dp(p1),
                                         str(tmp_path / "host.vtu").encode(), 1 if high_order else 0))
    a, b = read_vtu(tmp_path / "ctx.vtu"), read_vtu(tmp_path / "host.vtu")
    assert a.keys() == b.keys()
    for key in a:
        assert np.array_equal(a[key], b[key]), key
    assert np.all(np.isfinite(a["rho"])) and rel_change(r1, rho) > 0


def rel_change(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)
