"""The closed-form checks of the jump terms (test_oracle_jumps.py) and the
temporal order check of the IMEX step (test_oracle_imex_order.py) catch the
plausible mistakes they are meant to pin, together with the second
restatement of the time step (test_oracle_imex_wiring.py). Each case below builds a copy of
the oracle with one planted mistake (a text substitution in
oracle/atmo_oracle.cpp, compiled into a temporary directory) and asserts that
at least one of those checks rejects it: a jump's sign or factor, the
reference's kinetic-jump branch, a tableau weight, a stage source slot, or
the stage wiring of the time step. The oracle in the tree is not modified.
"""
import concurrent.futures
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from oracle.pyoracle import load_library

import test_oracle_imex_order as order
import test_oracle_imex_wiring as wiring
import test_oracle_jumps as jumps

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "atmo_oracle.cpp")

# name -> (exact text in the oracle, planted replacement)
MUTANTS = {
    "density jump sign": ("acc += w[s] * dt * (ax * n0 + az * n1 + 0.5 * lam * jump);",
                          "acc += w[s] * dt * (ax * n0 + az * n1 - 0.5 * lam * jump);"),
    "kinetic jump factor": ("acc += -wa[s] * dt * Ma2 * (akx * n0 + akz * n1 + 0.5 * lam * jkin) -",
                            "acc += -wa[s] * dt * Ma2 * (akx * n0 + akz * n1 + 1.0 * lam * jkin) -"),
    "internal-energy jump sign": ("wt[s] * dt * (aex * n0 + aez * n1 + 0.5 * lam * je);",
                                  "wt[s] * dt * (aex * n0 + aez * n1 - 0.5 * lam * je);"),
    "kinetic-jump branch of stage 2 dropped": ("if (st == 2 && s == 1) {", "if (false) {"),
    "stage-3 energy jump factor": ("wa[s] * dt * 0.5 * lam * jrE;", "wa[s] * dt * 0.25 * lam * jrE;"),
    "momentum jump factor": ("- wa[s] * dt * 0.5 * lam * jr;", "- wa[s] * dt * lam * jr;"),
    "tableau weight a32": ("a32 = 0.5;", "a32 = 0.45;"),
    "tableau weight b1": ("b1 = a31t;", "b1 = 0.9 * a31t;"),
    "momentum gravity source slot (stage 2)": ("gw_[2] = tb.a33t; gr[2] = src[6]; ng = 3; }",
                                               "gw_[2] = tb.a33t; gr[2] = src[3]; ng = 3; }"),
    "energy kinetic source slot": ("const double* rho_kin = src[3 * nex];", "const double* rho_kin = src[0];"),
    "wiring: stage-2 density sources": (
        "const double* s[4] = {rho_old.data(), u_old.data(), rho2.data(), u2.data()};",
        "const double* s[4] = {rho_old.data(), u_old.data(), rho2.data(), u_old.data()};"),
    "wiring: stage-2 momentum sources": (
        "const double* su[7] = {rho_old.data(), u_old.data(), p_old.data(), rho2.data(), u2.data(), p2.data(), "
        "rho3.data()};",
        "const double* su[7] = {rho_old.data(), u_old.data(), p_old.data(), rho2.data(), u2.data(), p2.data(), "
        "rho2.data()};"),
    "wiring: stage-3 energy sources": ("rho3.data(), u3.data(), p3.data(), rho_curr.data(), u_curr.data()};",
                                       "rho3.data(), u3.data(), p3.data(), rho_curr.data(), u3.data()};"),
}


@pytest.fixture(scope="module")
def mutants():
    src = open(SRC).read()
    tmp = tempfile.mkdtemp(prefix="orc_mut_")

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
    """Which checks reject the oracle built at `so`."""
    L = load_library(so)
    found = []
    for k in (2, 3):
        case = jumps.make_case(k, library=L)
        for name, fn in (("density", jumps.density_errors), ("energy", jumps.energy_errors),
                         ("momentum", jumps.momentum_errors)):
            if max(fn(case)) >= 1e-13:
                found.append(f"{name} jumps k={k}")
    try:
        ratios = order.order_ratios(library=L)
        if not all(3.4 < x < 4.6 for r in ratios for x in r):
            found.append("temporal order " + str(np.round(ratios, 2).tolist()))
    except RuntimeError as e:  # a run that no longer converges is rejected too
        found.append(f"time step fails: {e}")
    if found:  # the step restatement (about 5 s) only for what the cheaper checks let through
        return found
    try:
        e = max(wiring.wiring_errors(library=L))
        if not e < 1e-9:
            found.append(f"wiring {e:.1e}")
    except RuntimeError as e:
        found.append(f"wiring step fails: {e}")
    return found


@pytest.mark.parametrize("name", list(MUTANTS))
def test_planted_mistake_is_caught(mutants, name):
    found = detected(mutants[name])
    print(f"{name}: caught by {found}")
    assert found, f"no check catches: {name}"
