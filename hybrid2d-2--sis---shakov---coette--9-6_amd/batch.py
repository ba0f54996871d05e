"""Batches of GSIS instances stepped together (gsis_batch_*): one launch of the
column-wavefront kinetic kernel over every instance's strips per step, the
whole step of all instances captured once as a hipGraph, and the project's
autotuner choosing strip width and job order.  No counterpart in the
reference, which runs one case per process.

``instances`` draws the heterogeneous set of the batched configuration:
Couette or cavity with equal odds, mesh side uniform in [64, 256], Kn
log-uniform in [1e-3, 10], wall or lid speed uniform in [0.05, 0.3], velocity
grid side uniform in [24, 64].
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import capi, cases


def instances(count: int, seed: int = 20261202, side=(64, 256), nv=(24, 64), kn=(1e-3, 10.0),
              speed=(0.05, 0.3), precision_scheme=None):
    """`count` Case objects drawn from `seed` (ranges inclusive)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        couette = bool(rng.integers(0, 2))
        n = int(rng.integers(side[0], side[1] + 1))
        v = int(rng.integers(nv[0], nv[1] + 1))
        knv = float(math.exp(rng.uniform(math.log(kn[0]), math.log(kn[1]))))
        sp = float(rng.uniform(speed[0], speed[1]))
        if couette:
            c = cases.couette_case(ny_cells=n, nx_cells=4, nv=v, kn=knv, uwall=sp, dt=1.0,
                                   name=f"b{k}_couette{n}_v{v}_kn{knv:.3g}")
        else:
            c = cases.cavity_case(n, v, kn=knv, lid=sp, dt=1.0, name=f"b{k}_cavity{n}_v{v}_kn{knv:.3g}")
        if precision_scheme is not None:
            c.const["scheme"] = precision_scheme
        out.append(c)
    return out


class Batch:
    """A gsis_batch over GsisSolver handles (one precision and flux scheme;
    the synthetic solves of M1 instances run inside the step's graph, one
    workgroup each, or after it, one cooperative launch each).  The solvers
    stay usable on their own between batch steps."""

    def __init__(self, solvers):
        self.lib = capi.load_library()
        self.solvers = list(solvers)
        arr = (ctypes.c_void_p * len(self.solvers))(*[s.h.value for s in self.solvers])
        b = ctypes.c_void_p()
        capi.check(self.lib, self.lib.gsis_batch_create(arr, len(self.solvers), ctypes.byref(b)))
        self.b = b

    def step(self, n: int = 1):
        capi.check(self.lib, self.lib.gsis_batch_step(self.b, int(n)))

    def tune(self, steps_per_trial: int = 2):
        """Autotuner: every strip width x job order runs steps_per_trial real
        steps; the fastest is kept.  Returns (nw, order, ms per step);
        4 * steps_per_trial steps are taken."""
        nw, order = ctypes.c_int32(), ctypes.c_int32()
        ms = ctypes.c_double()
        capi.check(self.lib, self.lib.gsis_batch_tune(self.b, int(steps_per_trial), ctypes.byref(nw),
                                                      ctypes.byref(order), ctypes.byref(ms)))
        return nw.value, order.value, ms.value

    def set_ns_graph(self, mode: int):
        """Synthetic solves inside the graph: -1 automatic, 0 after the graph, 1 inside (gsis_batch_set_ns_graph)."""
        capi.check(self.lib, self.lib.gsis_batch_set_ns_graph(self.b, int(mode)))

    def ns_graph(self) -> int:
        """The number of instances whose synthetic solve runs inside the graph."""
        n = ctypes.c_int32()
        capi.check(self.lib, self.lib.gsis_batch_ns_graph(self.b, ctypes.byref(n)))
        return n.value

    def close(self):
        if getattr(self, "b", None) is not None and self.b.value:
            self.lib.gsis_batch_destroy(self.b)
            self.b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
