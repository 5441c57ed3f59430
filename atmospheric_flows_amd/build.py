"""Builds libatmo_hip.so (gfx950) in-tree with hipcc.

Every translation unit is compiled for gfx950 only, with packed-FP32
instructions disabled, and linked into atmospheric_flows_amd/libatmo_hip.so.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libatmo_hip.so")
DRIVER = os.path.join(HERE, "atmo_run")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
SOURCES = ["ops.hip", "krylov.hip", "mg.hip", "diag.hip", "schur.hip", "amr.cpp", "engine.hip", "shell.cpp", "shell_ops.hip", "shell_engine.hip", "ensemble.cpp", "vtu.cpp", "capi.cpp", "noalloc.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-std=c++17", "-O3", "-fPIC", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _compile(src):
    obj = os.path.join(OBJ, src + ".o")
    cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{msgs}")
    return obj, msgs


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    jobs = min(len(SOURCES), int(os.environ.get("MAX_JOBS", "8")))
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_compile, SOURCES))
    for obj, msgs in results:
        if verbose and msgs.strip():
            print(msgs, file=sys.stderr)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    build_driver()
    return LIB


def build_driver():
    """The C++ command-line driver (driver/atmo_run.cpp) linked against the library."""
    cxx = os.environ.get("CXX", "g++")
    tmp = DRIVER + ".tmp"
    cmd = [cxx, "-std=c++17", "-O2", "-Wall", "-I" + INCLUDE, os.path.join(HERE, "driver", "atmo_run.cpp"),
           "-L" + HERE, "-latmo_hip", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("driver build failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, DRIVER)
    return DRIVER


if __name__ == "__main__":
    print(build(verbose=True))
