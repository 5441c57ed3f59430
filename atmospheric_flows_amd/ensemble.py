"""Ensemble / serving form over the C ABI (atmo_ensemble_*): many independent
hill-flow members, each with its own mesh, parameters, initial state, step
count and arrival time, run concurrently on one GPU under an HBM budget.

The reference runs one simulation per process; this is a build-defined
batching layer around its IMEX step (EulerSolver::run's loop body,
Atmospheric_Flow.cc:1115-1222), so there is no reference interface to mirror.
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from .operator import _ptr, default_params, tuning_dict, witch_of_agnesi


def _mesh(degree, nx, nz, h, periodic=True, Lx=40000.0, H=20000.0):
    h = np.ascontiguousarray(h, dtype=np.float64)
    return N.MeshDesc(degree, nx, nz, 1 if periodic else 0, Lx, H, _ptr(h)), h


def device_bytes(degree, nx, nz, h=None, periodic=True, Lx=40000.0, H=20000.0, **params):
    """Device bytes a context of this mesh and parameters allocates (no GPU needed)."""
    h = np.zeros(nx + 1) if h is None else h
    m, _keep = _mesh(degree, nx, nz, h, periodic, Lx, H)
    p = default_params(**params)
    out = ctypes.c_size_t()
    N.check(N.load().atmo_device_bytes(ctypes.byref(m), ctypes.byref(p), ctypes.byref(out)))
    return out.value


@dataclass
class Member:
    degree: int
    nx: int
    nz: int
    nsteps: int
    h0: float = 400.0       # hill height (Witch of Agnesi, half width 1 km, centre 20 km)
    U: float = 10.0
    N_bv: float = 0.01
    seed: int = 1234
    arrival_s: float = 0.0
    params: dict = field(default_factory=dict)  # atmo_params overrides (dt, Ma, Fr, eps, precision, ...)


class Ensemble:
    def __init__(self, slots=8, hbm_budget_bytes=256e9, keep_states=False):
        h = ctypes.c_void_p()
        N.check(N.load().atmo_ensemble_create(slots, float(hbm_budget_bytes), 1 if keep_states else 0,
                                              ctypes.byref(h)))
        self.handle = h
        self.members = []
        self._keep = []

    def submit(self, m: Member):
        h = witch_of_agnesi(m.nx, h0=m.h0)
        mesh, harr = _mesh(m.degree, m.nx, m.nz, h)
        self._keep.append(harr)
        prm = dict(Ma=1.0, Fr=1.0 / np.sqrt(9.81))
        prm.update(m.params)
        d = N.MemberDesc(mesh, default_params(**prm), 300.0, 1e5, m.N_bv, m.U, 9.81, 287.0, 1.4, 1e-4, m.seed,
                         m.nsteps, m.arrival_s)
        i = ctypes.c_int()
        N.check(N.load().atmo_ensemble_submit(self.handle, ctypes.byref(d), ctypes.byref(i)))
        self.members.append(m)
        return i.value

    def autotune(self, candidates=(4, 8, 12, 16), probe_steps=8, sample=32, rate_tolerance=0.0):
        """Picks the number of concurrent members (slots) by measurement:
        the first `sample` submitted members run `probe_steps` steps with each
        candidate; the fewest slots whose aggregate DoF-steps/s is within
        `rate_tolerance` (a fraction; 0 = the fastest) of the best rate becomes
        this ensemble's slot count. Returns (best, {candidate: DoF-steps/s})."""
        c = (ctypes.c_int * len(candidates))(*candidates)
        rates = np.zeros(len(candidates))
        best = ctypes.c_int()
        N.check(N.load().atmo_ensemble_autotune(self.handle, c, len(candidates), probe_steps, sample,
                                                float(rate_tolerance), ctypes.byref(best), _ptr(rates)))
        return best.value, dict(zip(candidates, rates.tolist()))

    def tune_kernels(self, nrep=20):
        """Kernel-shape autotuner per class of members (same degree, mesh size
        and precision): times the tunable kernel shapes on one engine of the
        class and applies the choices to every member of it. Returns one dict
        per class, in order of first appearance."""
        cap = 64
        out = (N.KernelTuning * cap)()
        n = ctypes.c_int()
        N.check(N.load().atmo_ensemble_tune_kernels(self.handle, nrep, out, cap, ctypes.byref(n)))
        return [tuning_dict(out[c]) for c in range(min(n.value, cap))]

    def run(self):
        w = ctypes.c_double()
        N.check(N.load().atmo_ensemble_run(self.handle, ctypes.byref(w)))
        return w.value

    def result(self, i):
        r = N.MemberResult()
        n = self.members[i].nsteps
        ms = np.zeros(max(n, 1))
        N.check(N.load().atmo_ensemble_result(self.handle, i, ctypes.byref(r), _ptr(ms), n))
        out = {k: getattr(r, k) for k, _ in N.MemberResult._fields_ if k not in ("stats", "error")}
        out["stats"] = {k: getattr(r.stats, k) for k, _ in N.Stats._fields_}
        out["error"] = r.error.decode(errors="replace")
        out["step_ms"] = ms[:r.steps_done].copy()
        return out

    def state(self, i):
        m = self.members[i]
        ns = m.nx * m.nz * (m.degree + 1) ** 2
        rho, u, p = np.zeros(ns), np.zeros(2 * ns), np.zeros(ns)
        N.check(N.load().atmo_ensemble_state(self.handle, i, _ptr(rho), _ptr(u), _ptr(p), ns))
        return rho, u, p

    def close(self):
        if getattr(self, "handle", None):
            N.load().atmo_ensemble_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


def latency_summary(results):
    """p50/p99 of per-step and per-request latency over finished members."""
    steps = np.concatenate([r["step_ms"] for r in results if r["status"] == 0 and len(r["step_ms"])])
    req = np.array([1e3 * (r["finish_s"] - r["arrival_s"]) for r in results if r["status"] == 0])
    return dict(step_ms_p50=float(np.percentile(steps, 50)), step_ms_p99=float(np.percentile(steps, 99)),
                request_ms_p50=float(np.percentile(req, 50)), request_ms_p99=float(np.percentile(req, 99)),
                steps=int(steps.size), requests=int(req.size))
