"""ctypes front-end of the CPU FP64 oracle (oracle/atmo_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker. The product package never
imports this module.

Field layout everywhere: [component][cell][dof], cell = j*nx + i (x fastest),
dof = jj*(k+1) + ii (lexicographic, x fastest). Velocity has 2 components.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "libatmo_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)


def build(force=False):
    src = os.path.join(_HERE, "atmo_oracle.cpp")
    if force or not os.path.exists(_SO) or os.path.getmtime(_SO) < os.path.getmtime(src):
        subprocess.check_call(["make", "-s", "-C", _HERE, "-B" if force else "all"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        _lib = load_library(_SO)
    return _lib


def load_library(path):
    """The oracle's entry points from the shared library at `path` (the
    tests also load variants of the oracle built elsewhere)."""
    L = ctypes.CDLL(path)
    L.orc_create.restype = ctypes.c_void_p
    L.orc_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                             ctypes.c_double, _dp, ctypes.c_int]
    L.orc_destroy.argtypes = [ctypes.c_void_p]
    L.orc_create_amr.restype = ctypes.c_void_p
    L.orc_create_amr.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_double, _dp, ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int]
    for name in ("orc_ncells", "orc_nlevels"):
        getattr(L, name).argtypes = [ctypes.c_void_p]
        getattr(L, name).restype = ctypes.c_int
    L.orc_mg_level_cells.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.orc_level_vertices.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
    L.orc_mg_level_cells.restype = ctypes.c_int
    L.orc_set_threads.argtypes = [ctypes.c_int]
    L.orc_get_threads.restype = ctypes.c_int
    L.orc_set_params.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_int, ctypes.c_double, ctypes.c_int]
    L.orc_tables.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
    L.orc_set_stage.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    L.orc_set_rho_for_fixed.argtypes = [ctypes.c_void_p, _dp]
    L.orc_set_pres_fixed.argtypes = [ctypes.c_void_p, _dp]
    for name in ("orc_rhs_rho", "orc_rhs_pressure", "orc_rhs_velocity"):
        getattr(L, name).argtypes = [ctypes.c_void_p, _dp, ctypes.POINTER(_dp), ctypes.c_int]
        getattr(L, name).restype = ctypes.c_int
    L.orc_vmult_pressure.argtypes = [ctypes.c_void_p, _dp, _dp]
    L.orc_vmult_enthalpy.argtypes = [ctypes.c_void_p, _dp, _dp]
    L.orc_vmult.argtypes = [ctypes.c_void_p, _dp, _dp]
    L.orc_vmult.restype = ctypes.c_int
    L.orc_inverse_diagonal.argtypes = [ctypes.c_void_p, _dp]
    L.orc_dense_cg.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_int]
    L.orc_dense_cg.restype = ctypes.c_int
    L.orc_dense_gmres.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_int]
    L.orc_dense_gmres.restype = ctypes.c_int
    L.orc_solve_mass_mg.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_double]
    L.orc_solve_mass_mg.restype = ctypes.c_int
    L.orc_prolongate.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
    L.orc_project_down.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
    L.orc_init_hill.argtypes = [ctypes.c_void_p] + [ctypes.c_double] * 8 + [ctypes.c_uint64, _dp, _dp, _dp]
    L.orc_run.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_int, ctypes.POINTER(ctypes.c_long)]
    L.orc_run.restype = ctypes.c_int
    L.orc_diagnostics.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_double, _dp]
    L.orc_min_cell_diameter.argtypes = [ctypes.c_void_p]
    L.orc_min_cell_diameter.restype = ctypes.c_double
    return L


def ptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def tables(k):
    nd, nq = k + 1, 2 * k + 1
    gll, gq, gw = np.zeros(nd), np.zeros(nq), np.zeros(nq)
    S, D = np.zeros((nq, nd)), np.zeros((nq, nd))
    lib().orc_tables(k, ptr(gll), ptr(gq), ptr(gw), ptr(S), ptr(D))
    return gll, gq, gw, S, D


def witch_of_agnesi(nx, Lx=40000.0, h0=400.0, a=1000.0, xc=20000.0):
    x = Lx * np.arange(nx + 1) / nx
    return h0 / (1.0 + ((x - xc) / a) ** 2)


class Oracle:
    """One discretization (mesh + degree) of the oracle with its MG hierarchy."""

    def __init__(self, k, nx, nz, Lx=40000.0, H=20000.0, hx=None, periodic=True, mg_min_cells=8, leaves=None,
                 library=None):
        """leaves: optional (n, 5) int rows (bi, bj, level, sx, sz) of a
        locally refined mesh over the nx x nz base mesh. library: the loaded
        oracle to use (load_library), default the one built here."""
        self._L = library if library is not None else lib()
        self.k, self.nx, self.nz = k, nx, nz
        self.nd2 = (k + 1) ** 2
        self.ncell = nx * nz if leaves is None else len(leaves)
        self.Ns = self.ncell * self.nd2
        hx = np.zeros(nx + 1) if hx is None else np.ascontiguousarray(hx, dtype=np.float64)
        if leaves is None:
            self._h = self._L.orc_create(k, nx, nz, 1 if periodic else 0, Lx, H, ptr(hx), mg_min_cells)
        else:
            lv = np.ascontiguousarray(leaves, dtype=np.int32).reshape(-1, 5)
            self._h = self._L.orc_create_amr(k, nx, nz, 1 if periodic else 0, Lx, H, ptr(hx),
                                           lv.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), lv.shape[0],
                                           mg_min_cells)
        if not self._h:
            raise RuntimeError("orc_create failed")
        self.set_params()

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.orc_destroy(self._h)
            self._h = None

    @property
    def nlevels(self):
        return self._L.orc_nlevels(self._h)

    def set_params(self, Ma=1.0, Fr=1.0 / np.sqrt(9.81), dt=5.0, eps=1e-10, max_its=30000, schur_inner_tol=1e-12,
                   a_solve_tol=1e-12, fp_max_iter=100, fp_tol=1e-10, fp_lock=0):
        self._L.orc_set_params(self._h, Ma, Fr, dt, eps, max_its, schur_inner_tol, a_solve_tol, fp_max_iter, fp_tol,
                             fp_lock)

    def set_stage(self, hyperbolic_stage, ns_stage=1):
        self._hs = hyperbolic_stage
        self._L.orc_set_stage(self._h, hyperbolic_stage, ns_stage)

    def set_rho_for_fixed(self, rho):
        self._L.orc_set_rho_for_fixed(self._h, ptr(np.ascontiguousarray(rho)))

    def set_pres_fixed(self, p):
        self._L.orc_set_pres_fixed(self._h, ptr(np.ascontiguousarray(p)))

    def _srcs(self, srcs):
        srcs = [np.ascontiguousarray(s, dtype=np.float64) for s in srcs]
        arr = (_dp * len(srcs))(*[ptr(s) for s in srcs])
        return srcs, arr

    def rhs_rho(self, srcs):
        keep, arr = self._srcs(srcs)
        dst = np.zeros(self.Ns)
        if self._L.orc_rhs_rho(self._h, ptr(dst), arr, len(keep)):
            raise ValueError("rhs_rho: bad sources")
        return dst

    def rhs_pressure(self, srcs):
        keep, arr = self._srcs(srcs)
        dst = np.zeros(self.Ns)
        if self._L.orc_rhs_pressure(self._h, ptr(dst), arr, len(keep)):
            raise ValueError("rhs_pressure: bad sources")
        return dst

    def rhs_velocity(self, srcs):
        keep, arr = self._srcs(srcs)
        dst = np.zeros(2 * self.Ns)
        if self._L.orc_rhs_velocity(self._h, ptr(dst), arr, len(keep)):
            raise ValueError("rhs_velocity: bad sources")
        return dst

    def vmult_pressure(self, p):
        dst = np.zeros(2 * self.Ns)
        self._L.orc_vmult_pressure(self._h, ptr(dst), ptr(np.ascontiguousarray(p)))
        return dst

    def vmult_enthalpy(self, u):
        dst = np.zeros(self.Ns)
        self._L.orc_vmult_enthalpy(self._h, ptr(dst), ptr(np.ascontiguousarray(u)))
        return dst

    def vmult(self, src, ns):
        self_ns_size = 2 * self.Ns if ns in (3, 5) else self.Ns
        self._L.orc_set_stage(self._h, self._hs, ns)
        dst = np.zeros(self_ns_size)
        if self._L.orc_vmult(self._h, ptr(dst), ptr(np.ascontiguousarray(src))):
            raise RuntimeError("vmult failed")
        return dst

    _hs = 1

    def inverse_diagonal(self, ns):
        n = 2 * self.Ns if ns in (3, 5) else self.Ns
        self._L.orc_set_stage(self._h, self._hs, ns)
        dst = np.zeros(n)
        self._L.orc_inverse_diagonal(self._h, ptr(dst))
        return dst

    def solve_mass_mg(self, ns, b, x0=None, tol=1e-12):
        x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64)
        its = self._L.orc_solve_mass_mg(self._h, ns, ptr(x), ptr(np.ascontiguousarray(b)), tol)
        if its < 0:
            raise RuntimeError("MG solve did not converge")
        return x, its

    def init_hill(self, theta0=300.0, p0=1e5, N=0.01, U=10.0, g=9.81, R=287.0, gamma=1.4, pert=1e-4, seed=1234):
        rho, u, p = np.zeros(self.Ns), np.zeros(2 * self.Ns), np.zeros(self.Ns)
        self._L.orc_init_hill(self._h, theta0, p0, N, U, g, R, gamma, pert, seed, ptr(rho), ptr(u), ptr(p))
        return rho, u, p

    def run(self, rho, u, p, nsteps):
        rho, u, p = (np.array(a, dtype=np.float64) for a in (rho, u, p))
        st = (ctypes.c_long * 7)()
        rc = self._L.orc_run(self._h, ptr(rho), ptr(u), ptr(p), nsteps, st)
        if rc:
            raise RuntimeError("oracle IMEX run failed (no convergence)")
        keys = ("cg_its", "gmres_its", "inner_cg_its", "fp_its", "mg_cycles", "coarse_its", "schur_matvecs")
        return rho, u, p, dict(zip(keys, list(st)))

    def diagnostics(self, rho, u, p, gamma=1.4):
        """(min rho at QGaussLobatto(k+1) points, max |rho| over dofs, max celerity,
        max |u|, number of non-finite values) of one state."""
        out = np.zeros(5)
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (rho, u, p)]
        self._L.orc_diagnostics(self._h, ptr(a[0]), ptr(a[1]), ptr(a[2]), gamma, ptr(out))
        return out

    def min_cell_diameter(self):
        return self._L.orc_min_cell_diameter(self._h)
