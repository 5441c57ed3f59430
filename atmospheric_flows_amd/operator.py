"""Host-side mirror of the reference's operator/solver interface over the C ABI.

`DeviceContext` plays the role of EulerSolver's setup (mesh, MatrixFree
storage, MG hierarchy, Atmospheric_Flow.cc:217-409); `EULEROperator` keeps the
member names of euler_operator.h:24-67 (setters, vmult_rhs_*, vmult_pressure,
vmult_enthalpy, vmult per NS_stage, compute_diagonal); `DeviceVector` is
MatrixFree::initialize_dof_vector(v, block). All arithmetic runs in
libatmo_hip.so on the GPU.
"""
import ctypes

import numpy as np

from . import _native as N

BLOCK_VELOCITY, BLOCK_PRESSURE, BLOCK_DENSITY = 0, 1, 2

PARAM_NAMES = [f[0] for f in N.Params._fields_]


def witch_of_agnesi(nx, Lx=40000.0, h0=400.0, a=1000.0, xc=20000.0):
    """Terrain heights h(x_i) = h0/(1+((x_i-xc)/a)^2) at the nx+1 vertex columns."""
    x = Lx * np.arange(nx + 1) / nx
    return h0 / (1.0 + ((x - xc) / a) ** 2)


def default_params(**overrides):
    p = N.Params()
    N.load().atmo_params_default(ctypes.byref(p))
    for k, v in overrides.items():
        if k not in PARAM_NAMES:
            raise KeyError(f"unknown parameter {k}")
        setattr(p, k, v)
    return p


def tuning_dict(t):
    return dict(arnoldi_us=list(t.arnoldi_us), sweep_us_vector_alus=t.sweep_us[0], sweep_us_matrix_cores=t.sweep_us[1],
                arnoldi_waves_per_simd=t.arnoldi_waves_per_simd, mg_matrix_cores=t.mg_matrix_cores)


def _ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class DeviceVector:
    def __init__(self, ctx, block):
        self.ctx, self.block = ctx, block
        h = ctypes.c_void_p()
        N.check(N.load().atmo_vec_alloc(ctx.handle, block, ctypes.byref(h)), ctx.handle)
        self.handle = h
        n = ctypes.c_size_t()
        N.load().atmo_vec_size(h, ctypes.byref(n))
        self.size = n.value

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        N.check(N.load().atmo_vec_upload(self.ctx.handle, self.handle, _ptr(a), a.size), self.ctx.handle)
        return self

    def download(self):
        out = np.empty(self.size)
        N.check(N.load().atmo_vec_download(self.ctx.handle, self.handle, _ptr(out), out.size), self.ctx.handle)
        return out

    # deal.II LinearAlgebra::distributed::Vector operations, on the device
    def sadd(self, s, a, x):
        """self = s * self + a * x"""
        N.check(N.load().atmo_vec_axpby(self.ctx.handle, float(a), x.handle, float(s), self.handle), self.ctx.handle)
        return self

    def equ(self, a, x):
        """self = a * x"""
        return self.sadd(0.0, a, x)

    def add(self, a, x):
        """self += a * x"""
        return self.sadd(1.0, a, x)

    def __mul__(self, other):
        """scalar product self . other"""
        r = ctypes.c_double()
        N.check(N.load().atmo_vec_dot(self.ctx.handle, self.handle, other.handle, ctypes.byref(r)), self.ctx.handle)
        return r.value

    def linfty_norm(self):
        r = ctypes.c_double()
        N.check(N.load().atmo_vec_linf(self.ctx.handle, self.handle, ctypes.byref(r)), self.ctx.handle)
        return r.value

    def __del__(self):
        if getattr(self, "handle", None) and getattr(self.ctx, "handle", None):
            N.load().atmo_vec_free(self.ctx.handle, self.handle)
            self.handle = None


class DeviceContext:
    """One simulation on the GPU. `leaves` (optional, int array (n, 5) of
    (bi, bj, level, sx, sz)) makes the mesh a locally refined one over the
    nx x nz base mesh (atmo_ctx_create_amr)."""

    def __init__(self, degree, nx, nz, Lx=40000.0, H=20000.0, h=None, periodic=True, leaves=None, **params):
        L = N.load()
        self.degree, self.nx, self.nz = degree, nx, nz
        self._h = np.zeros(nx + 1) if h is None else np.ascontiguousarray(h, dtype=np.float64)
        self.mesh = N.MeshDesc(degree, nx, nz, 1 if periodic else 0, Lx, H, _ptr(self._h))
        self.params = default_params(**params)
        h_ = ctypes.c_void_p()
        if leaves is None:
            N.check(L.atmo_ctx_create(ctypes.byref(self.mesh), ctypes.byref(self.params), ctypes.byref(h_)))
        else:
            lv = np.ascontiguousarray(leaves, dtype=np.int32).reshape(-1, 5)
            N.check(L.atmo_ctx_create_amr(ctypes.byref(self.mesh), lv.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                          lv.shape[0], ctypes.byref(self.params), ctypes.byref(h_)))
        self.handle = h_
        self._info()

    def _info(self):
        L = N.load()
        nc, ns, nl, d = ctypes.c_int(), ctypes.c_long(), ctypes.c_int(), ctypes.c_int()
        L.atmo_ctx_info(self.handle, ctypes.byref(nc), ctypes.byref(ns), ctypes.byref(nl))
        L.atmo_ctx_dim(self.handle, ctypes.byref(d))
        self.ncells, self.ns, self.nlevels, self.dim = nc.value, ns.value, nl.value, d.value

    def close(self):
        if getattr(self, "handle", None):
            N.load().atmo_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()

    def set_params(self, **kw):
        for k, v in kw.items():
            if k not in PARAM_NAMES:
                raise KeyError(f"unknown parameter {k}")
            setattr(self.params, k, v)
        N.check(N.load().atmo_set_params(self.handle, ctypes.byref(self.params)), self.handle)

    def vector(self, block, data=None):
        v = DeviceVector(self, block)
        if data is not None:
            v.upload(data)
        return v

    def sync(self):
        N.check(N.load().atmo_sync(self.handle), self.handle)

    def init_hill(self, theta0=300.0, p0=1e5, N_bv=0.01, U=10.0, g=9.81, R=287.0, gamma=1.4, pert=1e-4, seed=1234):
        rho, u, p = np.zeros(self.ns), np.zeros(2 * self.ns), np.zeros(self.ns)
        N.check(N.load().atmo_ctx_init_hill_state(self.handle, theta0, p0, N_bv, U, g, R, gamma, pert, seed,
                                                  _ptr(rho), _ptr(u), _ptr(p)), self.handle)
        return rho, u, p

    def upload_state(self, rho, u, p):
        rho, u, p = (np.ascontiguousarray(a, dtype=np.float64) for a in (rho, u, p))
        N.check(N.load().atmo_state_upload(self.handle, _ptr(rho), _ptr(u), _ptr(p)), self.handle)

    def download_state(self):
        rho, u, p = np.zeros(self.ns), np.zeros(self.dim * self.ns), np.zeros(self.ns)
        N.check(N.load().atmo_state_download(self.handle, _ptr(rho), _ptr(u), _ptr(p)), self.handle)
        return rho, u, p

    def write_vtu(self, path, high_order=False):
        """output_results (Atmospheric_Flow.cc:886-918): the current state as VTU."""
        N.check(N.load().atmo_write_vtu(self.handle, str(path).encode(), 1 if high_order else 0), self.handle)

    def imex_step(self, nsteps=1):
        N.check(N.load().atmo_imex_step(self.handle, nsteps), self.handle)

    def step_diagnostics(self, of_state=False):
        """Diagnostics of the last step (min/max density per stage, max
        celerity and velocity, Courant numbers, non-finite count); with
        of_state, recomputed on the current state first."""
        d = N.StepDiag()
        N.check(N.load().atmo_step_diagnostics(self.handle, 1 if of_state else 0, ctypes.byref(d)), self.handle)
        out = {n: getattr(d, n) for n, _ in N.StepDiag._fields_}
        out["min_density"] = list(d.min_density)
        out["max_density"] = list(d.max_density)
        return out

    def stats(self):
        s = N.Stats()
        N.load().atmo_get_stats(self.handle, ctypes.byref(s))
        return {n: getattr(s, n) for n, _ in N.Stats._fields_}

    def device_bytes(self):
        b = ctypes.c_size_t()
        N.check(N.load().atmo_ctx_device_bytes(self.handle, ctypes.byref(b)), self.handle)
        return b.value

    def reset_stats(self):
        N.load().atmo_reset_stats(self.handle)

    def time_kernel(self, which, nrep=20):
        ms, b = ctypes.c_double(), ctypes.c_double()
        N.check(N.load().atmo_time_kernel(self.handle, which, nrep, ctypes.byref(ms), ctypes.byref(b)), self.handle)
        return ms.value, b.value

    def tune_kernels(self, nrep=20):
        """Kernel-shape autotuner (atmo_tune_kernels): times the Arnoldi kernel
        at 3 and 4 waves per SIMD and the FP32 multigrid sweep on the vector
        ALUs and the matrix cores on this context's mesh, keeps the fastest in
        the context's parameters and returns the timings and choices."""
        t = N.KernelTuning()
        N.check(N.load().atmo_tune_kernels(self.handle, nrep, ctypes.byref(t)), self.handle)
        self.params.arnoldi_waves_per_simd = t.arnoldi_waves_per_simd or self.params.arnoldi_waves_per_simd
        if t.mg_matrix_cores >= 0:
            self.params.mg_matrix_cores = t.mg_matrix_cores
        return tuning_dict(t)


# EquationData (equation_data.h:23-33) of the reference's 3D test
EARTH_RADIUS = 6.37122e6
GRAVITY = 9.80616
R_GAS = 287.0
T_BAR = 273.0
OMEGA_SBR = 2.0 * np.pi / 86400.0


class ShellContext(DeviceContext):
    """The reference's own 3D case on the GPU (atmo_ctx_create_shell):
    FE_DGQ(1) on the shell between r_inner and r_outer (0.9 a and a,
    Atmospheric_Flow.cc:283-284), `refinements` global refinements of
    `panels` coarse cells, radial gravity g a^2/r^2 with a = `planet`.
    Velocity fields have 3 components; every operator, solver and IMEX
    method of DeviceContext applies."""

    def __init__(self, refinements=3, panels=6, r_inner=0.9 * EARTH_RADIUS, r_outer=EARTH_RADIUS,
                 planet=EARTH_RADIUS, **params):
        L = N.load()
        self.degree = 1
        self.desc = N.ShellDesc(refinements, panels, r_inner, r_outer, planet)
        self.params = default_params(**params)
        h_ = ctypes.c_void_p()
        N.check(L.atmo_ctx_create_shell(ctypes.byref(self.desc), ctypes.byref(self.params), ctypes.byref(h_)))
        self.handle = h_
        self._info()

    def dof_points(self):
        out = np.zeros(self.ns * 3)
        N.check(N.load().atmo_shell_dof_points(self.handle, _ptr(out)), self.handle)
        return out.reshape(self.ncells, 8, 3)

    def sbr_state(self, Omega=OMEGA_SBR, a=EARTH_RADIUS, g=GRAVITY, RT=R_GAS * T_BAR):
        """The solid-body-rotation steady state at the dofs (non-dimensional)."""
        rho, u, p = np.zeros(self.ns), np.zeros(3 * self.ns), np.zeros(self.ns)
        N.check(N.load().atmo_shell_sbr_state(self.handle, Omega, a, g, RT, _ptr(rho), _ptr(u), _ptr(p)),
                self.handle)
        return rho, u, p

    def l2_errors(self, Omega=OMEGA_SBR, a=EARTH_RADIUS, g=GRAVITY, RT=R_GAS * T_BAR):
        """analyze_results on the current state: {u, rho, p} -> (absolute, relative) L2 errors."""
        out = np.zeros(6)
        N.check(N.load().atmo_shell_l2_errors(self.handle, Omega, a, g, RT, _ptr(out)), self.handle)
        return {"u": (out[0], out[1]), "rho": (out[2], out[3]), "p": (out[4], out[5])}


def shell_device_bytes(refinements=3, panels=6, r_inner=0.9 * EARTH_RADIUS, r_outer=EARTH_RADIUS,
                       planet=EARTH_RADIUS, **params):
    d = N.ShellDesc(refinements, panels, r_inner, r_outer, planet)
    p = default_params(**params)
    b = ctypes.c_size_t()
    N.check(N.load().atmo_shell_device_bytes(ctypes.byref(d), ctypes.byref(p), ctypes.byref(b)))
    return b.value


class EULEROperator:
    """euler_operator.h:24-67 over the device context."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.HYPERBOLIC_stage, self.NS_stage = 1, 1
        self._inv_diag = None

    def _L(self):
        return N.load()

    def set_dt(self, dt):
        self.ctx.set_params(dt=dt)

    def set_Mach(self, Ma):
        self.ctx.set_params(Ma=Ma)

    def set_Froude(self, Fr):
        self.ctx.set_params(Fr=Fr)

    def _stages(self):
        N.check(self._L().atmo_set_stages(self.ctx.handle, self.HYPERBOLIC_stage, self.NS_stage), self.ctx.handle)

    def set_HYPERBOLIC_stage(self, stage):
        if stage not in (1, 2, 3):
            raise ValueError("HYPERBOLIC_stage must be 1..3")
        self.HYPERBOLIC_stage = stage
        self._stages()

    def set_NS_stage(self, stage):
        if not 1 <= stage <= 6:
            raise ValueError("NS_stage must be 1..6")
        self.NS_stage = stage
        self._stages()

    def set_rho_for_fixed(self, v):
        N.check(self._L().atmo_set_rho_for_fixed(self.ctx.handle, v.handle), self.ctx.handle)

    def set_pres_fixed(self, v):
        N.check(self._L().atmo_set_pres_fixed(self.ctx.handle, v.handle), self.ctx.handle)

    def set_u_fixed(self, v):
        """Kept for interface parity: the reference stores u_fixed but its
        operators read the fixed-point velocity from the rhs source list."""
        self._u_fixed = v

    def _src(self, src):
        return (ctypes.c_void_p * len(src))(*[s.handle.value for s in src])

    def vmult_rhs_rho_update(self, dst, src):
        self._stages()
        N.check(self._L().atmo_rhs_rho(self.ctx.handle, dst.handle, self._src(src), len(src)), self.ctx.handle)

    def vmult_rhs_pressure(self, dst, src):
        self._stages()
        N.check(self._L().atmo_rhs_pressure(self.ctx.handle, dst.handle, self._src(src), len(src)), self.ctx.handle)

    def vmult_rhs_velocity_fixed(self, dst, src):
        self._stages()
        N.check(self._L().atmo_rhs_velocity(self.ctx.handle, dst.handle, self._src(src), len(src)), self.ctx.handle)

    def vmult_pressure(self, dst, src):
        self._stages()
        N.check(self._L().atmo_vmult_pressure(self.ctx.handle, dst.handle, src.handle), self.ctx.handle)

    def vmult_enthalpy(self, dst, src):
        self._stages()
        N.check(self._L().atmo_vmult_enthalpy(self.ctx.handle, dst.handle, src.handle), self.ctx.handle)

    def vmult(self, dst, src):
        self._stages()
        N.check(self._L().atmo_vmult(self.ctx.handle, dst.handle, src.handle), self.ctx.handle)

    def compute_diagonal(self):
        self._stages()
        blk = BLOCK_VELOCITY if self.NS_stage in (3, 5) else (BLOCK_DENSITY if self.NS_stage in (1, 4)
                                                               else BLOCK_PRESSURE)
        self._inv_diag = self.ctx.vector(blk)
        N.check(self._L().atmo_inverse_diagonal(self.ctx.handle, self._inv_diag.handle), self.ctx.handle)

    def get_matrix_diagonal_inverse(self):
        return self._inv_diag

    # solver entry points of EulerSolver's stage functions
    def solve_cg_mg(self, ns_stage, x, b, tol_factor):
        its = ctypes.c_int()
        N.check(self._L().atmo_solve_cg_mg(self.ctx.handle, ns_stage, x.handle, b.handle, tol_factor,
                                           ctypes.byref(its)), self.ctx.handle)
        return its.value

    def solve_schur_gmres(self, x, b, tol_factor):
        self._stages()
        its = ctypes.c_int()
        N.check(self._L().atmo_solve_schur_gmres(self.ctx.handle, x.handle, b.handle, tol_factor, ctypes.byref(its)),
                self.ctx.handle)
        return its.value
