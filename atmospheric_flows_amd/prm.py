"""Reader of the reference's parameter file format (runtime_parameters.h:16-178).

Same subsections, keys, defaults and pattern checks as Data_Storage
(deal.II ParameterHandler syntax: `subsection`, `set key = value`, `end`,
`#` comments); unknown keys and out-of-pattern values raise. An extra
`Device` subsection carries the settings this build adds (degree, mesh,
precision).
"""
import os
from dataclasses import dataclass

# (subsection, key) -> (attribute, type, default, check)
_ENTRIES = {
    ("Physical data", "initial_time"): ("initial_time", float, 0.0, lambda v: v >= 0.0),
    ("Physical data", "final_time"): ("final_time", float, 1.0, lambda v: v >= 0.0),
    ("Physical data", "Mach"): ("Mach", float, 1.0, lambda v: v >= 0.0),
    ("Physical data", "Froude"): ("Froude", float, 1.0, lambda v: v >= 0.0),
    ("Time step data", "dt"): ("dt", float, 5e-4, lambda v: v >= 0.0),
    ("Space discretization", "n_of_refines"): ("n_global_refines", int, 3, lambda v: 0 <= v <= 15),
    ("Data solve", "max_iterations"): ("max_iterations", int, 1000, lambda v: 1 <= v <= 30000),
    ("Data solve", "eps"): ("eps", float, 1e-12, lambda v: v >= 0.0),
    ("", "verbose"): ("verbose", bool, True, lambda v: True),
    ("", "output_interval"): ("output_interval", int, 1, lambda v: v >= 1),
    ("", "saving directory"): ("dir", str, "SimTest", lambda v: True),
    # build-defined
    ("Device", "degree"): ("degree", int, 3, lambda v: 1 <= v <= 3),
    ("Device", "nx"): ("nx", int, 512, lambda v: v >= 1),
    ("Device", "nz"): ("nz", int, 256, lambda v: v >= 1),
    ("Device", "Lx"): ("Lx", float, 40000.0, lambda v: v > 0),
    ("Device", "H"): ("H", float, 20000.0, lambda v: v > 0),
    ("Device", "hill_height"): ("hill_height", float, 400.0, lambda v: v >= 0),
    ("Device", "hill_half_width"): ("hill_half_width", float, 1000.0, lambda v: v > 0),
    ("Device", "hill_center"): ("hill_center", float, 20000.0, lambda v: True),
    ("Device", "precision"): ("precision", str, "fp64", lambda v: v in ("fp64", "mixed", "fp64_direct")),
    ("Device", "gravity"): ("gravity", float, 9.81, lambda v: v > 0),
    ("Device", "wind"): ("wind", float, 10.0, lambda v: True),
    ("Device", "brunt_vaisala"): ("brunt_vaisala", float, 0.01, lambda v: v > 0),
    ("Device", "seed"): ("seed", int, 20240601, lambda v: v >= 0),
}


@dataclass
class DataStorage:
    initial_time: float = 0.0
    final_time: float = 1.0
    Mach: float = 1.0
    Froude: float = 1.0
    dt: float = 5e-4
    n_global_refines: int = 3
    max_iterations: int = 1000
    eps: float = 1e-12
    verbose: bool = True
    output_interval: int = 1
    dir: str = "SimTest"
    degree: int = 3
    nx: int = 512
    nz: int = 256
    Lx: float = 40000.0
    H: float = 20000.0
    hill_height: float = 400.0
    hill_half_width: float = 1000.0
    hill_center: float = 20000.0
    precision: str = "fp64"
    gravity: float = 9.81
    wind: float = 10.0
    brunt_vaisala: float = 0.01
    seed: int = 20240601


def _convert(typ, text):
    if typ is bool:
        t = text.strip().lower()
        if t in ("true", "1", "yes"):
            return True
        if t in ("false", "0", "no"):
            return False
        raise ValueError(f"not a Bool: {text}")
    if typ is int:
        return int(text.strip())
    if typ is float:
        return float(text.strip())
    return text.strip()


def read_data(filename):
    """Data_Storage::read_data (runtime_parameters.h:137-176)."""
    if not os.path.exists(filename):
        raise FileNotFoundError(filename)  # AssertThrow(file, ExcFileNotOpen) at :138-139
    d = DataStorage()
    stack = []
    with open(filename) as f:
        for ln, raw in enumerate(f, 1):
            line = raw.split("#", 1)[0].strip()
            if not line:
                continue
            if line.startswith("subsection "):
                stack.append(line[len("subsection "):].strip())
            elif line == "end":
                if not stack:
                    raise ValueError(f"{filename}:{ln}: 'end' without subsection")
                stack.pop()
            elif line.startswith("set "):
                key, sep, val = line[4:].partition("=")
                if not sep:
                    raise ValueError(f"{filename}:{ln}: expected 'set key = value'")
                sec = stack[-1] if stack else ""
                ent = _ENTRIES.get((sec, key.strip()))
                if ent is None:
                    raise ValueError(f"{filename}:{ln}: unknown entry '{key.strip()}' in subsection '{sec}'")
                attr, typ, _, ok = ent
                v = _convert(typ, val)
                if not ok(v):
                    raise ValueError(f"{filename}:{ln}: value {v!r} of '{key.strip()}' does not match its pattern")
                setattr(d, attr, v)
            else:
                raise ValueError(f"{filename}:{ln}: cannot parse '{line}'")
    if stack:
        raise ValueError(f"{filename}: unterminated subsection '{stack[-1]}'")
    if d.dt <= 0.0 or d.dt > 0.5 * d.final_time:
        # EulerSolver constructor check (Atmospheric_Flow.cc:256, T = final_time at :220)
        raise ValueError(f"The time step {d.dt} is out of range (0, {0.5 * d.final_time}]")
    return d
