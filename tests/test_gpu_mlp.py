"""The closure network's BF16-MFMA kernels against the CPU FP64 network
(oracle/mlp.py): forward pass and gradient against the oracle with the same
BF16 roundings (bound set by FP32 accumulation order) and against plain FP64
(bound set by BF16), the gradient against finite differences of the FP64
loss, ragged batch sizes, Adam's first step, and training that lowers the
loss on a smooth target."""
import numpy as np
import pytest

from helpers import load_pkg
from test_oracle_mlp import M


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def _case(n, seed):
    rng = np.random.default_rng(seed)
    x = (rng.normal(size=(n, 8)) * np.linspace(0.5, 3.0, 8) + np.linspace(-1, 1, 8)).astype(np.float32)
    t = (np.sin(x[:, :5]) + 0.3 * x[:, 5:6] * x[:, 6:7]).astype(np.float32)
    s = np.concatenate([np.linspace(-1, 1, 8), np.linspace(0.5, 3.0, 8), np.full(5, 0.1),
                        np.linspace(0.6, 1.2, 5)]).astype(np.float32)
    return x, t, s


def _net(seed):
    g = load_pkg()
    from gsis_amd import closure
    p = M.init_params(seed)
    p[-5:] = np.float32(0.25)
    rng = np.random.default_rng(seed + 100)
    for l in range(M.LAYERS):  # non-zero hidden biases
        r, c = M.shapes()[l]
        off = sum(rr * cc + rr for rr, cc in M.shapes()[:l]) + r * c
        p[off:off + r] = rng.uniform(-0.2, 0.2, r).astype(np.float32)
    return closure, p


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 100, 4096, 8192, 10000])
def test_forward_matches_oracle(n):
    closure, p = _net(1)
    x, _, s = _case(n, 2)
    print("forward n =", n, flush=True)
    net = closure.Mlp(p, s)
    y = net.forward(x)
    net.close()
    yq = M.forward(p, s, x, bf16_mode=True)
    y64 = M.forward(p, s, x)
    # same roundings: FP32 accumulation order and SiLU's exp (relative ~1e-6)
    # plus the rare BF16 rounding decision that those flip
    assert _rel(y, yq) < 2e-3, _rel(y, yq)
    assert _rel(y, y64) < 3e-2
    assert np.isfinite(y).all()


@pytest.mark.gpu
@pytest.mark.parametrize("n", [64, 333, 8192])
def test_gradient_matches_oracle(n):
    closure, p = _net(4)
    x, t, s = _case(n, 5)
    print("loss_grad n =", n, flush=True)
    net = closure.Mlp(p, s)
    loss, g = net.loss_grad(x, t)
    net.close()
    lq, gq = M.loss_grad(p, s, x, t, bf16_mode=True)
    l64, g64 = M.loss_grad(p, s, x, t)
    assert loss == pytest.approx(lq, rel=2e-3)
    assert loss == pytest.approx(l64, rel=3e-2)
    assert _rel(g, gq) < 1e-2, _rel(g, gq)
    assert _rel(g, g64) < 5e-2, _rel(g, g64)
    # every layer's weight and bias block on its own
    off = 0
    for r, c in M.shapes():
        for sl in (slice(off, off + r * c), slice(off + r * c, off + r * c + r)):
            assert _rel(g[sl], gq[sl]) < 2e-2, sl
        off += r * c + r


@pytest.mark.gpu
def test_gradient_against_finite_differences_of_fp64_loss():
    closure, p = _net(7)
    x, t, s = _case(512, 8)
    net = closure.Mlp(p, s)
    _, g = net.loss_grad(x, t)
    net.close()
    rng = np.random.default_rng(9)
    p64 = p.astype(np.float64)
    for trial in range(4):
        v = rng.normal(size=p.size)
        h = 1e-5
        fd = (M.loss(p64 + h * v, s, x, t) - M.loss(p64 - h * v, s, x, t)) / (2 * h)
        assert float(g.astype(np.float64) @ v) == pytest.approx(fd, rel=5e-2), trial
    # single coordinates with large gradients, one per layer
    off = 0
    for r, c in M.shapes():
        k = off + int(np.argmax(np.abs(g[off:off + r * c])))
        e = np.zeros(p.size)
        e[k] = 1e-4
        fd = (M.loss(p64 + e, s, x, t) - M.loss(p64 - e, s, x, t)) / 2e-4
        assert g[k] == pytest.approx(fd, rel=5e-2), k
        off += r * c + r


@pytest.mark.gpu
def test_adam_first_step_and_training_lowers_loss():
    closure, p = _net(11)
    x, t, s = _case(4096, 12)
    net = closure.Mlp(p, s)
    _, g = net.loss_grad(x, t)
    net.train(x, t, epochs=1, batch=4096, lr=1e-3, seed=1)  # one full batch: one Adam step
    p1 = net.params()
    step = p1.astype(np.float64) - p
    big = np.abs(g) > 1e-5
    # first Adam step from zero moments is -lr * g/(|g| + eps): -lr sign(g) where |g| >> eps
    assert np.allclose(step[big], -1e-3 * np.sign(g[big]), rtol=2e-3, atol=1e-7)
    net.set_params(p)
    hist = net.train(x, t, epochs=40, batch=256, lr=2e-3, seed=3)
    y = net.forward(x)
    net.close()
    assert hist[-1] < 0.1 * hist[0], hist
    ys = (t - s[16:21]) / s[21:]
    fit = np.mean(((y - s[16:21]) / s[21:] - ys) ** 2)
    assert fit == pytest.approx(hist[-1], rel=0.5)


@pytest.mark.gpu
def test_bad_arguments_are_refused():
    g = load_pkg()
    from gsis_amd import closure
    net = closure.Mlp()
    with pytest.raises(g.capi.GsisError):
        net.set_scaling(np.zeros(26, np.float32))  # zero standard deviations
    with pytest.raises(g.capi.GsisError):
        net.loss_grad(np.zeros((8193, 8), np.float32), np.zeros((8193, 5), np.float32))
    with pytest.raises(g.capi.GsisError):
        net.train(np.zeros((10, 8), np.float32), np.zeros((10, 5), np.float32), epochs=1, batch=100)
    net.close()
