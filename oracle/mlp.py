"""CPU FP64 restatement of the learned HoT closure's network.  TEST
INFRASTRUCTURE ONLY: what the GPU training and inference kernels are checked
against; the product never imports it.

The network (no counterpart in the reference) maps 8 standardised inputs
through 4 hidden layers of 128 SiLU units to 5 standardised outputs; the
parameter vector is [W1 b1 W2 b2 W3 b3 W4 b4 W5 b5], each W row-major
(outputs by inputs).  The loss is the mean squared error of the standardised
outputs over n x 5 values.

``bf16=True`` rounds what the GPU kernels hold in BF16 (the standardised
inputs, the weights, the hidden activations and the backward pass's dY/dZ) to
BF16 with round-to-nearest-even, so that the GPU results can be compared with
a bound set by FP32 accumulation order alone.

Pinned by: finite differences of ``loss`` against ``loss_grad`` (tests), a
network reduced by hand to one hidden unit per layer, and Adam's first step
(test_oracle_mlp.py).
"""
from __future__ import annotations

import numpy as np

NIN, NOUT, WIDTH, HIDDEN = 8, 5, 128, 4
LAYERS = HIDDEN + 1


def shapes():
    rows = [WIDTH] * HIDDEN + [NOUT]
    cols = [NIN] + [WIDTH] * HIDDEN
    return list(zip(rows, cols))


def nparams() -> int:
    return sum(r * c + r for r, c in shapes())


def unpack(p):
    out, o = [], 0
    for r, c in shapes():
        w = p[o:o + r * c].reshape(r, c)
        o += r * c
        b = p[o:o + r]
        o += r
        out.append((w, b))
    return out


def bf16(a):
    """Round float values to BF16 (nearest even), returned as float64."""
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def init_params(seed: int = 0) -> np.ndarray:
    """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights, zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for r, c in shapes():
        lim = 1.0 / np.sqrt(c)
        parts += [rng.uniform(-lim, lim, r * c), np.zeros(r)]
    return np.concatenate(parts).astype(np.float32)


def _standardise(x, scaling, q):
    x = np.asarray(x, np.float64)
    s = np.asarray(scaling, np.float64)
    xs = (x - s[:NIN]) / s[NIN:2 * NIN]
    return q(xs.astype(np.float32).astype(np.float64)) if q is not None else xs


def forward_std(p, scaling, x, bf16_mode=False):
    """Standardised outputs and the per-layer pre-activations / activations."""
    q = bf16 if bf16_mode else None
    layers = unpack(np.asarray(p, np.float64))
    h = _standardise(x, scaling, q)
    hs, zs = [h], []
    for l, (w, b) in enumerate(layers):
        wq = q(w) if q else w
        z = h @ wq.T + b
        if l < LAYERS - 1:
            zs.append(z)
            h = silu(z)
            if q:
                h = q(h)
            hs.append(h)
        else:
            return z, hs, zs


def forward(p, scaling, x, bf16_mode=False):
    """Outputs in the targets' own units."""
    s = np.asarray(scaling, np.float64)
    y, _, _ = forward_std(p, scaling, x, bf16_mode)
    return y * s[2 * NIN + NOUT:] + s[2 * NIN:2 * NIN + NOUT]


def loss(p, scaling, x, t):
    s = np.asarray(scaling, np.float64)
    y, _, _ = forward_std(p, scaling, x)
    ts = (np.asarray(t, np.float64) - s[2 * NIN:2 * NIN + NOUT]) / s[2 * NIN + NOUT:]
    return float(np.mean((y - ts) ** 2))


def loss_grad(p, scaling, x, t, bf16_mode=False):
    """(loss, gradient) by backpropagation."""
    q = bf16 if bf16_mode else None
    s = np.asarray(scaling, np.float64)
    layers = unpack(np.asarray(p, np.float64))
    y, hs, zs = forward_std(p, scaling, x, bf16_mode)
    ts = (np.asarray(t, np.float64) - s[2 * NIN:2 * NIN + NOUT]) / s[2 * NIN + NOUT:]
    n = y.shape[0]
    r = y - ts
    ls = float(np.mean(r ** 2))
    dz = 2.0 * r / (NOUT * n)
    grads = [None] * LAYERS
    for l in range(LAYERS - 1, -1, -1):
        w, _ = layers[l]
        gb = dz.sum(axis=0)
        dzq = q(dz) if q else dz
        gw = dzq.T @ hs[l]
        grads[l] = (gw, gb)
        if l == 0:
            break
        wq = q(w) if q else w
        dh = dzq @ wq
        dz = dh * silu_grad(zs[l - 1])
    flat = np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in grads])
    return ls, flat


def adam_step(p, g, m1, m2, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (step counts from 1); returns (p, m1, m2)."""
    m1 = b1 * m1 + (1 - b1) * g
    m2 = b2 * m2 + (1 - b2) * g * g
    c1, c2 = 1 - b1 ** step, 1 - b2 ** step
    return p - lr * (m1 / c1) / (np.sqrt(m2 / c2) + eps), m1, m2


def fit_scaling(x, y):
    """Per-feature mean and standard deviation (a floor keeps constant features finite)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    sx = np.maximum(x.std(axis=0), 1e-6)
    sy = np.maximum(y.std(axis=0), 1e-12)
    return np.concatenate([x.mean(axis=0), sx, y.mean(axis=0), sy]).astype(np.float32)
