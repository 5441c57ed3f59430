"""The C++ driver (reference-shaped main + EulerSolver::run) on the GPU: the
40x20 Q2 hill example runs to final_time and its last state's Courant
numbers match the Python path's diagnostics of the same steps."""
import os
import re
import subprocess

import numpy as np
import pytest

import atmospheric_flows_amd as af

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "atmospheric_flows_amd", "atmo_run")
PRM = os.path.join(ROOT, "atmospheric_flows_amd", "driver", "hill_40x20_Q2.prm")


def test_driver_runs_example_and_matches_python_path():
    if not os.path.exists(DRIVER):
        pytest.fail("atmo_run is not built (python -m atmospheric_flows_amd.build)")
    r = subprocess.run([DRIVER, PRM], capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    assert "Apparently everything went fine!" in r.stdout
    steps = re.findall(r"Step = (\d+) Time = ([0-9.eE+-]+)", r.stdout)
    assert len(steps) == 10 and float(steps[-1][1]) == pytest.approx(50.0)
    cfl_c = [float(x) for x in re.findall(r"CFL_c = ([0-9.eE+-]+)", r.stdout)]
    # same run through the Python mirror
    ctx = af.DeviceContext(2, 40, 20, h=af.witch_of_agnesi(40), Ma=1.0, Fr=0.31927542840705515, dt=5.0, eps=1e-10)
    rho, u, p = ctx.init_hill(N_bv=0.01, U=10.0, g=9.81, seed=20240601)
    ctx.upload_state(rho, u, p)
    ctx.imex_step(10)
    d = ctx.step_diagnostics()
    assert cfl_c[-1] == pytest.approx(d["cfl_c"], rel=1e-9)
    assert d["nonfinite"] == 0 and 0 < d["cfl_c"] < 100
