"""ctypes front-end of the CPU FP64 oracle of the spherical-shell case
(oracle/atmo_oracle_shell.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's cpu_baseline
leg as the checker. The product package never imports this module.

Field layout: [component][cell][dof], dof = i + 2j + 4k at the cell's
vertex (xi, eta, zeta) = (i, j, k); velocity has 3 components. Shell cells
are numbered ((panel*n + layer)*n + row)*n + column with n = 2^refinements.
"""
import ctypes
import os

import numpy as np

from oracle import pyoracle

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "libatmo_oracle_shell.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)

# EquationData (equation_data.h:23-39) and the 3D run's parameter file (parameter-file.prm:5-24)
EARTH_RADIUS = 6.37122e6
GRAVITY = 9.80616
R_GAS = 287.0
T_BAR = 273.0
OMEGA = 2.0 * np.pi / 86400.0
MACH = 0.00357254516
FROUDE = 0.31933793491
DT = 8.64


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            pyoracle.build()
        _lib = load_library(_SO)
    return _lib


def load_library(path):
    L = ctypes.CDLL(path)
    L.orc3_last_error.restype = ctypes.c_char_p
    L.orc3_create_shell.restype = ctypes.c_void_p
    L.orc3_create_shell.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double]
    L.orc3_create_box.restype = ctypes.c_void_p
    L.orc3_create_box.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 3
    L.orc3_destroy.argtypes = [ctypes.c_void_p]
    L.orc3_ncells.argtypes = [ctypes.c_void_p]
    L.orc3_ncells.restype = ctypes.c_int
    L.orc3_set_threads.argtypes = [ctypes.c_int]
    L.orc3_set_params.argtypes = [ctypes.c_void_p] + [ctypes.c_double] * 5 + [ctypes.c_int, ctypes.c_double] + \
        [ctypes.c_int] * 3
    L.orc3_set_stage.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    for n in ("orc3_set_rho_for_fixed", "orc3_set_pres_fixed"):
        getattr(L, n).argtypes = [ctypes.c_void_p, _dp]
    for n in ("orc3_rhs_rho", "orc3_rhs_pressure", "orc3_rhs_velocity"):
        getattr(L, n).argtypes = [ctypes.c_void_p, _dp, ctypes.POINTER(_dp), ctypes.c_int]
    for n in ("orc3_vmult_pressure", "orc3_vmult_enthalpy"):
        getattr(L, n).argtypes = [ctypes.c_void_p, _dp, _dp]
    L.orc3_vmult.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
    L.orc3_inverse_diagonal.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
    L.orc3_solve_block.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
    L.orc3_solve_schur.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_double, _ip]
    L.orc3_set_state.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
    L.orc3_get_state.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
    L.orc3_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.orc3_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_long)]
    L.orc3_dof_points.argtypes = [ctypes.c_void_p, _dp]
    L.orc3_neighbours.argtypes = [ctypes.c_void_p, _ip, _ip]
    L.orc3_measures.argtypes = [ctypes.c_void_p, _dp]
    L.orc3_sbr_state.argtypes = [ctypes.c_void_p] + [ctypes.c_double] * 4 + [_dp] * 3
    L.orc3_l2_errors.argtypes = [ctypes.c_void_p] + [ctypes.c_double] * 4 + [_dp] * 4
    for n in ("orc3_rhs_rho", "orc3_rhs_pressure", "orc3_rhs_velocity", "orc3_vmult_pressure",
              "orc3_vmult_enthalpy", "orc3_vmult", "orc3_inverse_diagonal", "orc3_solve_block",
              "orc3_solve_schur", "orc3_step"):
        getattr(L, n).restype = ctypes.c_int
    return L


def _p(a):
    return a.ctypes.data_as(_dp)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class ShellOracle:
    """EULEROperator + IMEX time step in 3D on the spherical shell (or, for
    pinning, on a box with walls and constant gravity along z)."""

    def __init__(self, refinements=None, panels=6, r_inner=0.9 * EARTH_RADIUS, r_outer=EARTH_RADIUS,
                 planet_radius=EARTH_RADIUS, box=None, library=None):
        self._L = library or lib()
        if box is not None:
            nx, ny, nz, Lx, Ly, H = box
            self.h = self._L.orc3_create_box(nx, ny, nz, Lx, Ly, H)
        else:
            self.h = self._L.orc3_create_shell(refinements, panels, r_inner, r_outer, planet_radius)
        if not self.h:
            raise RuntimeError(self._L.orc3_last_error().decode())
        self.ncell = self._L.orc3_ncells(self.h)
        self.Ns = 8 * self.ncell
        self.set_params()

    def __del__(self):
        if getattr(self, "h", None):
            self._L.orc3_destroy(self.h)
            self.h = None

    def _ok(self, rc):
        if rc != 0:
            raise RuntimeError(self._L.orc3_last_error().decode())

    def set_params(self, Ma=MACH, Fr=FROUDE, dt=DT, gamma_gas=1.4, eps=1e-10, max_its=30000, fp_tol=1e-10,
                   fp_max_iter=100, fp_lock=0, gmres_restart=28):
        self._L.orc3_set_params(self.h, Ma, Fr, dt, gamma_gas, eps, max_its, fp_tol, fp_max_iter, fp_lock,
                                gmres_restart)

    def set_stage(self, hs, ns=1):
        self._L.orc3_set_stage(self.h, hs, ns)

    def set_rho_for_fixed(self, rho):
        self._L.orc3_set_rho_for_fixed(self.h, _p(_c(rho)))

    def set_pres_fixed(self, p):
        self._L.orc3_set_pres_fixed(self.h, _p(_c(p)))

    def _rhs(self, fn, ncomp, srcs):
        srcs = [_c(s) for s in srcs]
        arr = (_dp * len(srcs))(*[_p(s) for s in srcs])
        out = np.zeros(ncomp * self.Ns)
        self._ok(fn(self.h, _p(out), arr, len(srcs)))
        return out

    def rhs_rho(self, *srcs):
        return self._rhs(self._L.orc3_rhs_rho, 1, srcs)

    def rhs_pressure(self, *srcs):
        return self._rhs(self._L.orc3_rhs_pressure, 1, srcs)

    def rhs_velocity(self, *srcs):
        return self._rhs(self._L.orc3_rhs_velocity, 3, srcs)

    def vmult_pressure(self, p):
        out = np.zeros(3 * self.Ns)
        self._ok(self._L.orc3_vmult_pressure(self.h, _p(out), _p(_c(p))))
        return out

    def vmult_enthalpy(self, u):
        out = np.zeros(self.Ns)
        self._ok(self._L.orc3_vmult_enthalpy(self.h, _p(out), _p(_c(u))))
        return out

    def ncomp(self, ns):
        return 3 if ns in (3, 5) else 1

    def vmult(self, ns, src):
        out = np.zeros(self.ncomp(ns) * self.Ns)
        self._ok(self._L.orc3_vmult(self.h, ns, _p(out), _p(_c(src))))
        return out

    def inverse_diagonal(self, ns):
        out = np.zeros(self.ncomp(ns) * self.Ns)
        self._ok(self._L.orc3_inverse_diagonal(self.h, ns, _p(out)))
        return out

    def solve_block(self, ns, b):
        out = np.zeros(self.ncomp(ns) * self.Ns)
        self._ok(self._L.orc3_solve_block(self.h, ns, _p(out), _p(_c(b))))
        return out

    def solve_schur(self, x0, b, tol_factor):
        x = _c(x0).copy()
        its = ctypes.c_int(0)
        self._ok(self._L.orc3_solve_schur(self.h, _p(x), _p(_c(b)), tol_factor, ctypes.byref(its)))
        return x, its.value

    def set_state(self, rho, u, p):
        self._L.orc3_set_state(self.h, _p(_c(rho)), _p(_c(u)), _p(_c(p)))

    def get_state(self):
        rho, u, p = np.zeros(self.Ns), np.zeros(3 * self.Ns), np.zeros(self.Ns)
        self._L.orc3_get_state(self.h, _p(rho), _p(u), _p(p))
        return rho, u, p

    def step(self, n=1):
        self._ok(self._L.orc3_step(self.h, n))

    def stats(self):
        out = (ctypes.c_long * 2)()
        self._L.orc3_stats(self.h, out)
        return {"gmres_its": out[0], "fp_its": out[1]}

    def dof_points(self):
        out = np.zeros(self.Ns * 3)
        self._L.orc3_dof_points(self.h, _p(out))
        return out.reshape(self.ncell, 8, 3)

    def neighbours(self):
        nb = np.zeros(self.ncell * 6, dtype=np.int32)
        nbf = np.zeros(self.ncell * 6, dtype=np.int32)
        self._L.orc3_neighbours(self.h, nb.ctypes.data_as(_ip), nbf.ctypes.data_as(_ip))
        return nb.reshape(-1, 6), nbf.reshape(-1, 6)

    def measures(self):
        """(volume, inner wall area, outer wall area, all wall area)"""
        out = np.zeros(4)
        self._L.orc3_measures(self.h, _p(out))
        return out

    def sbr_state(self, Omega=OMEGA, a=EARTH_RADIUS, g=GRAVITY, RT=R_GAS * T_BAR):
        """The solid-body-rotation steady state at the dofs (non-dimensional)."""
        rho, u, p = np.zeros(self.Ns), np.zeros(3 * self.Ns), np.zeros(self.Ns)
        self._L.orc3_sbr_state(self.h, Omega, a, g, RT, _p(rho), _p(u), _p(p))
        return rho, u, p

    def l2_errors(self, rho, u, p, Omega=OMEGA, a=EARTH_RADIUS, g=GRAVITY, RT=R_GAS * T_BAR):
        """analyze_results: {u, rho, p} -> (absolute, relative) L2 errors."""
        out = np.zeros(6)
        self._L.orc3_l2_errors(self.h, Omega, a, g, RT, _p(_c(rho)), _p(_c(u)), _p(_c(p)), _p(out))
        return {"u": (out[0], out[1]), "rho": (out[2], out[3]), "p": (out[4], out[5])}
