"""The learned HoT closure of the synthetic NS solve (no counterpart in the
reference): the network handle over the C-ABI, training-data harvest from
the solver's own runs, training on the GPU, and the trained network shipped
with the package.

Convert_NS builds the high-order terms (HoT) that the synthetic equations add
to the NS stress and heat flux from the DVM moments
(src/Hybrid_Sub.f90:317-349 for i-faces, :419-448 for j-faces).  The network
maps the 8 local inputs of an NS face (strain rates du/dx, dv/dy,
du/dy + dv/dx, temperature gradient, face density and temperature,
log10 Kn) to the five HoT tau_11, tau_22, tau_12, q_x, q_y of that face.
"""
from __future__ import annotations

import ctypes
import json
import pathlib

import numpy as np

from . import capi

NIN, NOUT = capi.MLP_NIN, capi.MLP_NOUT
WEIGHTS = pathlib.Path(__file__).resolve().parent / "closure_hot.npz"


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _f32(a, cols=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if cols is not None and (a.ndim != 2 or a.shape[1] != cols):
        raise ValueError(f"expected an (n, {cols}) array, got {a.shape}")
    return a


class Mlp:
    """One network on the GPU (gsis_mlp_*)."""

    def __init__(self, params=None, scaling=None):
        self.lib = capi.load_library()
        h = ctypes.c_void_p()
        capi.check(self.lib, self.lib.gsis_mlp_create(ctypes.byref(h)))
        self.h = h
        if params is not None:
            self.set_params(params)
        if scaling is not None:
            self.set_scaling(scaling)

    @staticmethod
    def nparams() -> int:
        return int(capi.load_library().gsis_mlp_nparams())

    def set_params(self, p):
        p = _f32(p)
        if p.size != self.nparams():
            raise ValueError(f"expected {self.nparams()} parameters, got {p.size}")
        capi.check(self.lib, self.lib.gsis_mlp_set_params(self.h, _fp(p)))

    def params(self) -> np.ndarray:
        p = np.zeros(self.nparams(), np.float32)
        capi.check(self.lib, self.lib.gsis_mlp_get_params(self.h, _fp(p)))
        return p

    def set_scaling(self, s):
        s = _f32(s)
        if s.size != 2 * NIN + 2 * NOUT:
            raise ValueError("scaling holds 26 values")
        self.scaling = s.copy()
        capi.check(self.lib, self.lib.gsis_mlp_set_scaling(self.h, _fp(s)))

    def forward(self, x) -> np.ndarray:
        x = _f32(x, NIN)
        y = np.zeros((x.shape[0], NOUT), np.float32)
        capi.check(self.lib, self.lib.gsis_mlp_forward(self.h, _fp(x), x.shape[0], _fp(y)))
        return y

    def loss_grad(self, x, y):
        x, y = _f32(x, NIN), _f32(y, NOUT)
        loss = ctypes.c_double()
        g = np.zeros(self.nparams(), np.float32)
        capi.check(self.lib, self.lib.gsis_mlp_loss_grad(self.h, _fp(x), _fp(y), x.shape[0], ctypes.byref(loss),
                                                         _fp(g)))
        return loss.value, g

    def train(self, x, y, epochs: int, batch: int = 4096, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        x, y = _f32(x, NIN), _f32(y, NOUT)
        hist = np.zeros(max(epochs, 1))
        capi.check(self.lib, self.lib.gsis_mlp_train(self.h, _fp(x), _fp(y), x.shape[0], int(epochs), int(batch),
                                                     float(lr), int(seed),
                                                     hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return hist[:epochs]

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.gsis_mlp_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def init_params(seed: int = 0) -> np.ndarray:
    """Uniform(+-1/sqrt(fan_in)) weights and zero biases, in the parameter order of gsis_mlp_nparams."""
    rng = np.random.default_rng(seed)
    rows = [capi.MLP_WIDTH] * capi.MLP_HIDDEN + [NOUT]
    cols = [NIN] + [capi.MLP_WIDTH] * capi.MLP_HIDDEN
    parts = []
    for r, c in zip(rows, cols):
        lim = 1.0 / np.sqrt(c)
        parts += [rng.uniform(-lim, lim, r * c), np.zeros(r)]
    return np.concatenate(parts).astype(np.float32)


def scaling_of(x, y) -> np.ndarray:
    """Per-feature mean and standard deviation of inputs and targets (floors
    keep a constant feature, such as log10 Kn of a single-Kn set, finite)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    sx = np.maximum(x.std(axis=0), 1e-6)
    sy = np.maximum(y.std(axis=0), 1e-12)
    return np.concatenate([x.mean(axis=0), sx, y.mean(axis=0), sy]).astype(np.float32)


def samples(solver) -> tuple[np.ndarray, np.ndarray]:
    """Network inputs and DVM HoT of the closure faces of a GsisSolver's last
    Convert_NS (gsis_closure_samples)."""
    lib = solver.lib
    n = ctypes.c_int64()
    capi.check(lib, lib.gsis_closure_samples(solver.h, None, None, 0, ctypes.byref(n)))
    x = np.zeros((n.value, NIN), np.float32)
    y = np.zeros((n.value, NOUT), np.float32)
    if n.value:
        capi.check(lib, lib.gsis_closure_samples(solver.h, _fp(x), _fp(y), n.value, ctypes.byref(n)))
    return x, y


def set_closure(solver, net: Mlp | None, gate: float = 1.0, every: int = 50) -> None:
    """Use ``net`` in the synthetic solve of ``solver`` (gsis_set_closure)."""
    lib = solver.lib
    capi.check(lib, lib.gsis_set_closure(solver.h, None if net is None else net.h, float(gate), int(every)))
    solver._closure_net = net  # the network must outlive its use


def set_warmup(solver, net: Mlp | None, nsteps: int, every: int = 50) -> None:
    """The network's HoT in place of the DVM's for the next nsteps synthetic
    solves of ``solver`` (gsis_set_closure_warmup)."""
    lib = solver.lib
    capi.check(lib, lib.gsis_set_closure_warmup(solver.h, None if net is None else net.h, int(nsteps), int(every)))
    solver._closure_warm_net = net


def stats(solver) -> dict:
    lib = solver.lib
    f, a, r, x = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32()
    capi.check(lib, lib.gsis_closure_stats(solver.h, ctypes.byref(f), ctypes.byref(a), ctypes.byref(r),
                                           ctypes.byref(x)))
    return {"faces": f.value, "accepted": a.value, "refreshes": r.value, "rejected": x.value}


def save(path, params, scaling, meta: dict) -> None:
    np.savez(path, params=np.asarray(params, np.float32), scaling=np.asarray(scaling, np.float32),
             meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8))


def load(path=WEIGHTS):
    """(params, scaling, meta) of a saved network (no pickle)."""
    with np.load(path, allow_pickle=False) as z:
        return z["params"].copy(), z["scaling"].copy(), json.loads(bytes(z["meta"]).decode())


def trained_network(path=WEIGHTS) -> Mlp:
    p, s, _ = load(path)
    return Mlp(p, s)
