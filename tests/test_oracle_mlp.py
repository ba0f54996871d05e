"""Pins of the CPU FP64 network (oracle/mlp.py), which the GPU closure
kernels are checked against: backpropagation against central finite
differences, a network reduced by hand to a closed form, BF16 rounding against known
values, Adam's first step.  Each check fails for a planted mistake in the
oracle (a dropped bias, a transposed weight, a wrong SiLU derivative)."""
import importlib.util
import pathlib

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]


def _mlp():
    spec = importlib.util.spec_from_file_location("oracle_mlp", ROOT / "oracle" / "mlp.py")
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


M = _mlp()


def _data(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, M.NIN)) * 2.0 + 1.0
    t = rng.normal(size=(n, M.NOUT))
    s = np.concatenate([np.full(M.NIN, 1.0), np.full(M.NIN, 2.0), np.zeros(M.NOUT), np.full(M.NOUT, 0.5)])
    return x, t, s


def test_nparams():
    assert M.nparams() == 128 * 8 + 128 + 3 * (128 * 128 + 128) + 5 * 128 + 5 == 51333


def test_backprop_matches_central_differences():
    x, t, s = _data(37, 1)
    p = M.init_params(3).astype(np.float64)
    p[M.nparams() - 5:] = 0.3  # non-zero output bias
    _, g = M.loss_grad(p, s, x, t)
    rng = np.random.default_rng(5)
    for k in rng.choice(M.nparams(), 40, replace=False):
        h = 1e-6
        e = np.zeros_like(p)
        e[k] = h
        fd = (M.loss(p + e, s, x, t) - M.loss(p - e, s, x, t)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 + 1e-5 * abs(fd), k
    v = rng.normal(size=p.size)
    h = 1e-6
    fd = (M.loss(p + h * v, s, x, t) - M.loss(p - h * v, s, x, t)) / (2 * h)
    assert fd == pytest.approx(float(g @ v), rel=1e-6)


def test_one_unit_chain_closed_form():
    # one active unit per layer: y_k = w5_k silu(w4 silu(w3 silu(w2 silu(w1 . xs + b1) + b2) + b3) + b4) + b5_k
    p = np.zeros(M.nparams())
    layers = M.unpack(p)  # views into p
    (w1, b1), (w2, b2), (w3, b3), (w4, b4), (w5, b5) = layers
    w1[0, :] = np.linspace(-0.5, 0.5, M.NIN)
    b1[0] = 0.1
    w2[0, 0], b2[0] = 1.5, -0.2
    w3[0, 0], b3[0] = -0.7, 0.3
    w4[0, 0], b4[0] = 2.0, 0.05
    w5[:, 0] = [1.0, -2.0, 0.5, 3.0, -1.0]
    b5[:] = [0.1, 0.2, 0.3, 0.4, 0.5]
    x, _, s = _data(4, 2)
    xs = (x - s[:8]) / s[8:16]
    sl = lambda z: z / (1 + np.exp(-z))
    a = sl(sl(sl(sl(xs @ w1[0] + 0.1) * 1.5 - 0.2) * -0.7 + 0.3) * 2.0 + 0.05)
    ys = a[:, None] * w5[:, 0] + b5
    assert np.allclose(M.forward(p, s, x), ys * s[21:] + s[16:21], rtol=1e-14, atol=1e-14)


def test_bf16_rounding_known_values():
    # BF16 keeps 7 fraction bits: the spacing above 1 is 2^-7
    v = np.array([1.0, 1.0 + 2.0 ** -7, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 1.1 * 2.0 ** -8,
                  -3.14159265, 0.0])
    got = M.bf16(v)
    # 1+2^-8 ties between 1 and 1+2^-7 (even: 1); 1+3*2^-8 ties between
    # 1+2^-7 and 1+2^-6 (even: 1+2^-6); just above a tie rounds up
    assert list(got[:5]) == [1.0, 1.0 + 2.0 ** -7, 1.0, 1.0 + 2.0 ** -6, 1.0 + 2.0 ** -7]
    assert got[5] == -3.140625 and got[6] == 0.0


def test_bf16_mode_differs_by_rounding_only():
    x, t, s = _data(64, 4)
    p = M.init_params(9)
    y = M.forward(p, s, x)
    yq = M.forward(p, s, x, bf16_mode=True)
    rel = np.linalg.norm(yq - y) / np.linalg.norm(y)
    assert 1e-5 < rel < 3e-2


def test_adam_first_step_is_signed_lr():
    g = np.array([1e-3, -2.0, 5.0, 0.0])
    p, m1, m2 = M.adam_step(np.zeros(4), g, np.zeros(4), np.zeros(4), 1, 0.01)
    assert np.allclose(p[:3], -0.01 * np.sign(g[:3]), rtol=1e-4)
    assert p[3] == 0.0 and np.allclose(m1, 0.1 * g) and np.allclose(m2, 0.001 * g * g)
