"""The pins of the spherical-shell oracle (test_oracle_shell.py) catch the
plausible mistakes they are meant to pin. Each case builds a copy of
oracle/atmo_oracle_shell.cpp with one planted mistake (a text substitution,
compiled into a temporary directory) and asserts that the reduction to the
2D oracle on the box or the steady-state tendency check rejects it: a jump
term, the reference's kinetic-jump branch, a wall term, a gravity weight,
the radial gravity's direction or a^2/r^2 factor, the neighbour's trace
coordinates. The oracle in the tree is not modified.
"""
import concurrent.futures
import os
import shutil
import subprocess
import tempfile

import pytest

from oracle.pyoracle_shell import load_library

import test_oracle_shell as pins

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "atmo_oracle_shell.cpp")

# name -> (exact text in the oracle, planted replacement)
MUTANTS = {
    "density jump sign": ("acc += w[s] * dt * (fn + 0.5 * lam * (rp[s][0][q] - rm[s][0][q]));",
                          "acc += w[s] * dt * (fn - 0.5 * lam * (rp[s][0][q] - rm[s][0][q]));"),
    "kinetic-jump branch of stage 2 dropped": ("if (st == 2 && s == 1) {", "if (false) {"),
    "momentum wall term factor": ("acc += -wt[s] * dt / Ma2 * 0.5 * (pP + pD) * fg.n[q][comp];",
                                  "acc += -wt[s] * dt / Ma2 * 0.25 * (pP + pD) * fg.n[q][comp];"),
    "B wall trace": ("const double pP = fpo[f][0][q], pM = wall ? pP : fpn[f][0][q];",
                     "const double pP = fpo[f][0][q], pM = wall ? 0.0 : fpn[f][0][q];"),
    "enthalpy jump factor": ("val[q] = coeff * P.dt * (an + 0.5 * lam * jump);",
                             "val[q] = coeff * P.dt * (an + lam * jump);"),
    "energy gravity weight (stage 1)": ("gwt[1] = tb.a22t; gsl[1] = -1; ngr = 2; }",
                                        "gwt[1] = tb.a21t; gsl[1] = 0; ngr = 2; }"),
    "gravity factor a/r": ("G = m.planet * m.planet / (r * r);", "G = m.planet / r;"),
    "gravity direction": ("er[0] = x[0] / r;\n  er[1] = x[1] / r;\n  er[2] = x[2] / r;",
                          "er[0] = -x[0] / r;\n  er[1] = -x[1] / r;\n  er[2] = -x[2] / r;"),
    "neighbour trace at own coordinates": (
        "if (cn >= 0) nbv[comp][s] = eval(f + (size_t)comp * Ns + (size_t)cn * 8, fg.xinb[s]);",
        "if (cn >= 0) nbv[comp][s] = eval(f + (size_t)comp * Ns + (size_t)cn * 8, fg.xi[s]);"),
}


@pytest.fixture(scope="module")
def mutants():
    src = open(SRC).read()
    tmp = tempfile.mkdtemp(prefix="orc3_mut_")

    def build(item):
        i, (name, (old, new)) = item
        assert src.count(old) == 1, f"mutation site of '{name}' not found once in the oracle"
        path = os.path.join(tmp, f"m{i}.cpp")
        with open(path, "w") as f:
            f.write(src.replace(old, new))
        so = os.path.join(tmp, f"libm{i}.so")
        subprocess.check_call(["g++", "-O2", "-fopenmp", "-fPIC", "-std=c++17", "-shared", "-o", so, path])
        return name, so

    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        built = dict(ex.map(build, enumerate(MUTANTS.items())))
    yield built
    shutil.rmtree(tmp, ignore_errors=True)


def detected(so):
    L = load_library(so)
    found = []
    dev = pins.box_deviation(library=L)
    if not dev <= 1e-13:
        found.append(f"box vs 2D {dev:.1e}")
        return found
    try:
        r = pins.tendency_ratios(library=L)
        if not all(x > 1.6 for x in r):
            found.append(f"steady-state tendency {r}")
    except RuntimeError as e:
        found.append(f"time step fails: {e}")
    return found


@pytest.mark.parametrize("name", list(MUTANTS))
def test_planted_mistake_is_caught(mutants, name):
    found = detected(mutants[name])
    print(f"{name}: caught by {found}")
    assert found, f"no check catches: {name}"
