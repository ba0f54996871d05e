"""The learned HoT closure inside the synthetic NS solve (gsis_set_closure,
gsis_closure_samples): the inner loop run in chunks with the HoT updated in
between is the plain loop when the network's update is zero, the samples are
the Convert_NS HoT at the documented face order, and the network inputs are
the face gradients and averages of the Convert_NS state."""
import math

import numpy as np
import pytest

from helpers import load_oracle, load_pkg, rel_l2, solver_mod

pytestmark = pytest.mark.gpu


def _m1(case, **kw):
    s = solver_mod().GsisSolver(case, mode=1, **kw)
    s.initialize()
    return s


def test_chunked_loop_with_zero_update_equals_plain_m1():
    g = load_pkg()
    from gsis_amd import closure
    case = g.cases.cavity_case(16, 24, kn=0.1, lid=0.15, dt=1.0)
    plain = _m1(case)
    chunked = _m1(case)
    net = closure.Mlp()  # all parameters zero: the same output for every input
    closure.set_closure(chunked, net, gate=1e30, every=7)
    for k in range(4):
        print("step", k + 1, flush=True)
        plain.step(1)
        chunked.step(1)
        ip, ic = plain.counters()[2], chunked.counters()[2]
        st = closure.stats(chunked)
        assert ip == ic, (ip, ic)
        assert st["refreshes"] == (ic - 1) // 7, (st, ic)
        assert st["accepted"] == st["faces"] == (17 - 2) * (17 - 3) * 2
    a, b = plain.macro(), chunked.macro()
    for key in ("s", "s0", "cf", "u", "taoi", "taoj", "qi", "qj"):
        assert np.array_equal(a[key], b[key]), key
    # the oracle's plain M1 agrees as before
    cpu = load_oracle().OracleSolver(case, mode=1)
    cpu.initialize()
    cpu.step(4)
    assert rel_l2(b["u"], cpu.macro()["u"]) <= 1e-10
    cpu.close()
    plain.close()
    chunked.close()


def test_closure_samples_are_convert_ns_hot_and_face_gradients():
    g = load_pkg()
    from gsis_amd import closure
    n = 20
    case = g.cases.cavity_case(n, 24, kn=0.3, lid=0.2, dt=0.5)
    s = solver_mod().GsisSolver(case)
    s.initialize()
    s.step(6)
    x, y = closure.samples(s)
    m = s.macro()
    s.close()
    nx = ny = n + 1
    nfi = (nx - 2) * (ny - 3)
    assert x.shape == (nfi + (nx - 3) * (ny - 2), 8)
    # face order: i-faces i = 2..nx-1 fastest, j = 2..ny-2; then j-faces
    faces = [(0, i, j) for j in range(2, ny - 1) for i in range(2, nx)]
    faces += [(1, i, j) for j in range(2, ny) for i in range(2, nx - 1)]
    U = m["u"]  # U(4, -1:nx+1, -1:ny+1)
    rho = U[0]
    u, v = U[1] / rho, U[2] / rho
    cv = case.const["cv"]
    T = (U[3] - 0.5 * rho * (u * u + v * v)) / (cv * rho)
    h = 1.0 / n
    kn = math.sqrt(math.pi) / (2.0 * case.const["cfc"])
    for f, (d, i, j) in enumerate(faces):
        tg, qg = (m["taoi"], m["qi"]) if d == 0 else (m["taoj"], m["qj"])
        want = [tg[0, i + 1, j + 1], tg[1, i + 1, j + 1], tg[2, i + 1, j + 1], qg[0, i + 1, j + 1], qg[1, i + 1, j + 1]]
        assert np.array_equal(y[f], np.float32(want)), (d, i, j)
        # the kinetic and NS stresses are traceless: tau_33 = -(tau_11 + tau_22)
        # up to the rounding of the O(1) moments the HoT are differences of
        assert abs(tg[3, i + 1, j + 1] + tg[0, i + 1, j + 1] + tg[1, i + 1, j + 1]) <= 1e-13
        c = lambda a, ii, jj: a[ii + 1, jj + 1]
        if d == 0:  # uniform Cartesian cells: central differences across and along the face
            ux = (c(u, i, j) - c(u, i - 1, j)) / h
            vy = (c(v, i - 1, j + 1) + c(v, i, j + 1) - c(v, i - 1, j - 1) - c(v, i, j - 1)) / (4 * h)
            uy = (c(u, i - 1, j + 1) + c(u, i, j + 1) - c(u, i - 1, j - 1) - c(u, i, j - 1)) / (4 * h)
            vx = (c(v, i, j) - c(v, i - 1, j)) / h
            tx = (c(T, i, j) - c(T, i - 1, j)) / h
            ty = (c(T, i - 1, j + 1) + c(T, i, j + 1) - c(T, i - 1, j - 1) - c(T, i, j - 1)) / (4 * h)
            r0, t0 = (c(rho, i, j) + c(rho, i - 1, j)) / 2, (c(T, i, j) + c(T, i - 1, j)) / 2
        else:
            ux = (c(u, i + 1, j - 1) + c(u, i + 1, j) - c(u, i - 1, j - 1) - c(u, i - 1, j)) / (4 * h)
            vy = (c(v, i, j) - c(v, i, j - 1)) / h
            uy = (c(u, i, j) - c(u, i, j - 1)) / h
            vx = (c(v, i + 1, j - 1) + c(v, i + 1, j) - c(v, i - 1, j - 1) - c(v, i - 1, j)) / (4 * h)
            tx = (c(T, i + 1, j - 1) + c(T, i + 1, j) - c(T, i - 1, j - 1) - c(T, i - 1, j)) / (4 * h)
            ty = (c(T, i, j) - c(T, i, j - 1)) / h
            r0, t0 = (c(rho, i, j) + c(rho, i, j - 1)) / 2, (c(T, i, j) + c(T, i, j - 1)) / 2
        want = np.array([ux, vy, uy + vx, tx, ty, r0, t0, math.log10(kn)])
        assert np.allclose(x[f], want, rtol=1e-5, atol=1e-6 * (1 + np.abs(want).max())), (d, i, j, x[f], want)


def test_wrong_network_is_rejected_and_step_equals_plain():
    """A network whose HoT updates wreck the inner loop trips the residual
    check on every step: the step falls back to the DVM HoT and the loop's
    starting state, and the fields equal those of the plain step."""
    g = load_pkg()
    from gsis_amd import closure
    case = g.cases.cavity_case(16, 24, kn=0.1, lid=0.15, dt=1.0)
    plain = _m1(case)
    guarded = _m1(case)
    rng = np.random.default_rng(3)
    p = (closure.init_params(5) * 4.0).astype(np.float32)
    scaling = np.concatenate([np.zeros(8), np.full(8, 1e-3), np.zeros(5), np.full(5, 50.0)]).astype(np.float32)
    net = closure.Mlp(p, scaling)
    closure.set_closure(guarded, net, gate=1e30, every=5)
    rejected = 0
    for k in range(3):
        plain.step(1)
        guarded.step(1)
        st = closure.stats(guarded)
        print("step", k + 1, st, "plain its", plain.counters()[2], "guarded its", guarded.counters()[2], flush=True)
        rejected += st["rejected"]
    assert rejected >= 1
    a, b = plain.macro(), guarded.macro()
    for key in ("s", "cf", "u", "taoi", "taoj", "qi", "qj"):
        assert np.array_equal(a[key], b[key]), key
    plain.close()
    guarded.close()


def test_warmup_uses_the_network_then_the_plain_step():
    """gsis_set_closure_warmup: the first steps' synthetic solves run with the
    network's HoT (refreshed between launches), later steps are plain; a
    network whose predictions are not finite is refused on every face, and
    those steps equal the plain ones bit for bit."""
    g = load_pkg()
    from gsis_amd import closure
    case = g.cases.cavity_case(16, 24, kn=0.1, lid=0.15, dt=1.0)
    plain, warm, inf = _m1(case), _m1(case), _m1(case)
    net = closure.trained_network()
    closure.set_warmup(warm, net, 2, every=5)
    p = np.zeros(closure.Mlp.nparams(), np.float32)
    p[-5:] = 10.0  # output bias: times the huge output scale below, +inf
    scaling = np.concatenate([np.zeros(8), np.ones(8), np.zeros(5), np.full(5, 3e38)]).astype(np.float32)
    bad = closure.Mlp(p, scaling)
    closure.set_warmup(inf, bad, 3, every=5)
    for k in range(3):
        for s in (plain, warm, inf):
            s.step(1)
        st_w, st_i = closure.stats(warm), closure.stats(inf)
        print("step", k + 1, "warm-up", st_w, "non-finite", st_i, flush=True)
        if k < 2:  # the network's HoT were in use (and kept, or refused by the residual check)
            assert st_w["accepted"] > 0.9 * st_w["faces"] and st_w["refreshes"] + st_w["rejected"] > 0
        assert st_i["accepted"] == 0
    a, c = plain.macro(), inf.macro()
    for key in ("s", "u", "taoi", "taoj"):
        assert np.array_equal(a[key], c[key]), key
    assert np.all(np.isfinite(warm.macro()["u"]))
    for s in (plain, warm, inf):
        s.close()
