"""ctypes binding of libatmo_hip.so (declarations in include/atmo_hip.h).

Loading fails loudly when the shared library is missing: there is no CPU
fallback for any operator of this package.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libatmo_hip.so")

ATMO_OK = 0
ATMO_INVALID_ARG = 1
ATMO_NO_CONVERGENCE = 2
ATMO_NONFINITE = 3
ATMO_OOM = 4
ATMO_DEVICE = 5

_dp = ctypes.POINTER(ctypes.c_double)


class MeshDesc(ctypes.Structure):
    _fields_ = [("degree", ctypes.c_int), ("nx", ctypes.c_int), ("nz", ctypes.c_int), ("periodic", ctypes.c_int),
                ("Lx", ctypes.c_double), ("H", ctypes.c_double), ("h", _dp)]


class Params(ctypes.Structure):
    _fields_ = [("Ma", ctypes.c_double), ("Fr", ctypes.c_double), ("dt", ctypes.c_double),
                ("gamma_gas", ctypes.c_double), ("grav_x", ctypes.c_double), ("grav_z", ctypes.c_double),
                ("grav_factor", ctypes.c_double), ("eps", ctypes.c_double), ("max_iterations", ctypes.c_int),
                ("schur_inner_tol", ctypes.c_double), ("schur_inner_max", ctypes.c_int),
                ("a_solve_tol", ctypes.c_double), ("fp_max_iter", ctypes.c_int), ("fp_tol", ctypes.c_double),
                ("fp_lock", ctypes.c_int), ("mg_min_cells", ctypes.c_int), ("cheb_degree", ctypes.c_int),
                ("cheb_range", ctypes.c_double), ("cheb_eig_its", ctypes.c_int), ("gmres_restart", ctypes.c_int),
                ("precision", ctypes.c_int), ("mg_matrix_cores", ctypes.c_int),
                ("arnoldi_waves_per_simd", ctypes.c_int)]


class KernelTuning(ctypes.Structure):
    _fields_ = [("arnoldi_us", ctypes.c_double * 2), ("sweep_us", ctypes.c_double * 2),
                ("arnoldi_waves_per_simd", ctypes.c_int), ("mg_matrix_cores", ctypes.c_int)]


class ShellDesc(ctypes.Structure):
    _fields_ = [("refinements", ctypes.c_int), ("panels", ctypes.c_int), ("r_inner", ctypes.c_double),
                ("r_outer", ctypes.c_double), ("planet", ctypes.c_double)]


class Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_long) for n in ("cg_its", "gmres_its", "inner_cg_its", "fp_its", "mg_cycles",
                                               "coarse_its", "schur_matvecs", "host_syncs")]


class StepDiag(ctypes.Structure):
    _fields_ = [("min_density", ctypes.c_double * 3), ("max_density", ctypes.c_double * 3),
                ("max_celerity", ctypes.c_double), ("max_velocity", ctypes.c_double),
                ("min_cell_diameter", ctypes.c_double), ("cfl_c", ctypes.c_double), ("cfl_u", ctypes.c_double),
                ("nonfinite", ctypes.c_long)]


class MemberDesc(ctypes.Structure):
    _fields_ = [("mesh", MeshDesc), ("params", Params), ("theta0", ctypes.c_double), ("p0", ctypes.c_double),
                ("N", ctypes.c_double), ("U", ctypes.c_double), ("g", ctypes.c_double), ("R", ctypes.c_double),
                ("gamma_gas", ctypes.c_double), ("perturbation", ctypes.c_double), ("seed", ctypes.c_uint64),
                ("nsteps", ctypes.c_int), ("arrival_s", ctypes.c_double)]


class MemberResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("arrival_s", ctypes.c_double), ("admit_s", ctypes.c_double),
                ("start_s", ctypes.c_double), ("finish_s", ctypes.c_double), ("steps_done", ctypes.c_int),
                ("device_bytes", ctypes.c_size_t), ("stats", Stats), ("error", ctypes.c_char * 256)]


# (name, restype, argtypes) of every exported symbol
SIGNATURES = [
    ("atmo_params_default", None, [ctypes.POINTER(Params)]),
    ("atmo_version", ctypes.c_char_p, []),
    ("atmo_last_error", ctypes.c_char_p, [ctypes.c_void_p]),
    ("atmo_ctx_create", ctypes.c_int, [ctypes.POINTER(MeshDesc), ctypes.POINTER(Params),
                                       ctypes.POINTER(ctypes.c_void_p)]),
    ("atmo_ctx_create_amr", ctypes.c_int, [ctypes.POINTER(MeshDesc), ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                           ctypes.POINTER(Params), ctypes.POINTER(ctypes.c_void_p)]),
    ("atmo_ctx_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("atmo_ctx_init_hill_state", ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_double] * 8 +
     [ctypes.c_uint64, _dp, _dp, _dp]),
    ("atmo_set_params", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Params)]),
    ("atmo_ctx_info", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_long),
                                     ctypes.POINTER(ctypes.c_int)]),
    ("atmo_sync", ctypes.c_int, [ctypes.c_void_p]),
    ("atmo_get_stats", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Stats)]),
    ("atmo_reset_stats", ctypes.c_int, [ctypes.c_void_p]),
    ("atmo_vec_alloc", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    ("atmo_vec_free", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("atmo_vec_size", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t)]),
    ("atmo_vec_upload", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, _dp, ctypes.c_size_t]),
    ("atmo_vec_download", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, _dp, ctypes.c_size_t]),
    ("atmo_vec_device_ptr", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    ("atmo_vec_axpby", ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p, ctypes.c_double,
                                      ctypes.c_void_p]),
    ("atmo_vec_dot", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, _dp]),
    ("atmo_vec_linf", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, _dp]),
    ("atmo_dot_noalloc", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                                        ctypes.c_void_p]),
    ("atmo_set_stages", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    ("atmo_set_rho_for_fixed", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("atmo_set_pres_fixed", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("atmo_rhs_rho", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.c_int]),
    ("atmo_rhs_pressure", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p),
                                         ctypes.c_int]),
    ("atmo_rhs_velocity", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p),
                                         ctypes.c_int]),
    ("atmo_vmult_pressure", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("atmo_vmult_enthalpy", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("atmo_vmult", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("atmo_inverse_diagonal", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("atmo_solve_cg_mg", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_double, ctypes.POINTER(ctypes.c_int)]),
    ("atmo_solve_schur_gmres", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                              ctypes.POINTER(ctypes.c_int)]),
    ("atmo_state_upload", ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp]),
    ("atmo_state_download", ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp]),
    ("atmo_imex_step", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("atmo_step_diagnostics", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(StepDiag)]),
    ("atmo_init_hill_state", ctypes.c_int, [ctypes.POINTER(MeshDesc)] + [ctypes.c_double] * 8 +
     [ctypes.c_uint64, _dp, _dp, _dp]),
    ("atmo_write_vtu", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]),
    ("atmo_write_vtu_host", ctypes.c_int, [ctypes.POINTER(MeshDesc), _dp, _dp, _dp, ctypes.c_char_p, ctypes.c_int]),
    ("atmo_time_kernel", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _dp]),
    ("atmo_tune_kernels", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(KernelTuning)]),
    ("atmo_ensemble_tune_kernels", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(KernelTuning),
                                                  ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    ("atmo_device_bytes", ctypes.c_int, [ctypes.POINTER(MeshDesc), ctypes.POINTER(Params),
                                         ctypes.POINTER(ctypes.c_size_t)]),
    ("atmo_ctx_device_bytes", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t)]),
    ("atmo_ctx_create_shell", ctypes.c_int, [ctypes.POINTER(ShellDesc), ctypes.POINTER(Params),
                                             ctypes.POINTER(ctypes.c_void_p)]),
    ("atmo_ctx_dim", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    ("atmo_shell_dof_points", ctypes.c_int, [ctypes.c_void_p, _dp]),
    ("atmo_shell_sbr_state", ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_double] * 4 + [_dp, _dp, _dp]),
    ("atmo_shell_l2_errors", ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_double] * 4 + [_dp]),
    ("atmo_shell_device_bytes", ctypes.c_int, [ctypes.POINTER(ShellDesc), ctypes.POINTER(Params),
                                               ctypes.POINTER(ctypes.c_size_t)]),
    ("atmo_ensemble_create", ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_void_p)]),
    ("atmo_ensemble_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("atmo_ensemble_submit", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(MemberDesc), ctypes.POINTER(ctypes.c_int)]),
    ("atmo_ensemble_autotune", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                              ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                              ctypes.POINTER(ctypes.c_int), _dp]),
    ("atmo_ensemble_run", ctypes.c_int, [ctypes.c_void_p, _dp]),
    ("atmo_ensemble_result", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(MemberResult), _dp,
                                            ctypes.c_int]),
    ("atmo_ensemble_state", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, ctypes.c_size_t]),
]

_lib = None


class AtmoError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"atmo status {status}: {msg}")
        self.status = status


class NoConvergence(AtmoError):
    pass


def load():
    """Loads libatmo_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m atmospheric_flows_amd.build` "
                          "(the package has no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(status, ctx=None):
    if status == ATMO_OK:
        return
    msg = load().atmo_last_error(ctx).decode(errors="replace")
    if status == ATMO_NO_CONVERGENCE:
        raise NoConvergence(status, msg)
    raise AtmoError(status, msg)
