"""Shared helpers of the GSIS tests: package/oracle loaders, case builders for
the reference's shipped inputs, field comparisons."""
from __future__ import annotations

import importlib.util
import pathlib
import sys

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[1]
PKG_DIR = ROOT / "hybrid2d-2--sis---shakov---coette--9-6_amd"
GOLDEN = ROOT / "tests" / "golden"


def load_pkg():
    if "gsis_amd" in sys.modules:
        return sys.modules["gsis_amd"]
    spec = importlib.util.spec_from_file_location("gsis_amd", PKG_DIR / "__init__.py",
                                                  submodule_search_locations=[str(PKG_DIR)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["gsis_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def load_oracle():
    if "gsis_oracle_py" in sys.modules:
        return sys.modules["gsis_oracle_py"]
    spec = importlib.util.spec_from_file_location("gsis_oracle_py", ROOT / "oracle" / "oracle.py")
    mod = importlib.util.module_from_spec(spec)
    sys.modules["gsis_oracle_py"] = mod
    spec.loader.exec_module(mod)
    return mod


def solver_mod():
    load_pkg()
    import importlib
    return importlib.import_module("gsis_amd.solver")


def cylinder_case():
    g = load_pkg()
    return g.cases.shipped_case(GOLDEN / "cylinder", "cylinder40")


def channel_case():
    g = load_pkg()
    return g.cases.shipped_case(GOLDEN / "channel", "channel50x6")


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    den = np.linalg.norm(b)
    if den == 0.0:
        return float(np.linalg.norm(a - b))
    return float(np.linalg.norm(a - b) / den)


def cells(case, lo=0, hi=0):
    """Fortran-array slices of the cells lo..n-? for S (lower bound -3)."""
    return slice(3 + lo, case.nx + 3 + hi + 1), slice(3 + lo, case.ny + 3 + hi + 1)
