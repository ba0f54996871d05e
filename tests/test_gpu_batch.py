"""Batched instances (gsis_batch_*): every instance of a heterogeneous batch
(Couette and cavity, different mesh, velocity grid, Kn and wall speed),
stepped by graph replays of one multi-instance kinetic launch, equals the
same instance stepped on its own; the autotuner's trial steps are real steps
of the same computation; the device-side uniform initial fill equals the
upload of the host-built FRD array."""
import numpy as np
import pytest

from helpers import load_pkg, rel_l2, solver_mod

pytestmark = pytest.mark.gpu
KEYS = ("s", "cf", "u", "taoi", "taoj", "qi", "qj", "swall")


def _make(g, cases, prec):
    out = []
    for c in cases:
        s = solver_mod().GsisSolver(c, precision=prec, kstep_show=1000000)
        s.initialize()
        out.append(s)
    return out


def test_device_fill_equals_frd_upload():
    g = load_pkg()
    for prec in (g.capi.PREC_FP64, g.capi.PREC_FP32):
        case = g.cases.cavity_case(10, 12, kn=0.2, lid=0.1, dt=0.5)
        a = solver_mod().GsisSolver(case, precision=prec)
        b = solver_mod().GsisSolver(case, precision=prec)
        a.initialize(device_fill=True)
        b.initialize(device_fill=False)
        assert np.array_equal(a.frd(), b.frd())
        ma, mb = a.macro(), b.macro()
        for k in KEYS:
            assert np.array_equal(ma[k], mb[k]), k
        a.close()
        b.close()


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_batch_equals_single_instances(prec):
    g = load_pkg()
    from gsis_amd import batch
    p = g.capi.PREC_FP32 if prec == "fp32" else g.capi.PREC_FP64
    cases = batch.instances(7, seed=20261202, side=(12, 40), nv=(8, 24))
    print([c.name for c in cases], flush=True)
    single = _make(g, cases, p)
    grouped = _make(g, cases, p)
    b = batch.Batch(grouped)
    for s in single:
        s.step(3)
    b.step(3)
    for s, t in zip(single, grouped):
        assert s.counters()[:2] == t.counters()[:2]
        assert np.array_equal(s.frd(), t.frd()), s.case.name
        ms, mt = s.macro(), t.macro()
        for k in KEYS:
            assert np.array_equal(ms[k], mt[k]), (s.case.name, k)
    # autotuner: 4 candidates x 1 step, every one a real step
    nw, order, ms = b.tune(1)
    print("tuned: nw", nw, "order", order, f"{ms:.3f} ms/step", flush=True)
    assert nw in (8, 16) and order in (0, 1) and ms > 0
    b.step(2)
    for s in single:
        s.step(6)
    for s, t in zip(single, grouped):
        assert s.counters()[0] == t.counters()[0] == 9
        e = rel_l2(t.frd(), s.frd())
        assert e <= (1e-6 if prec == "fp32" else 1e-12), (s.case.name, e)
    # the solvers stay usable on their own
    grouped[0].step(1)
    b.close()
    for s in single + grouped:
        s.close()


@pytest.mark.parametrize("ns_graph", [-1, 0, 1])
def test_batch_runs_m1_instances_like_single_steps(ns_graph):
    """M1 instances in a batch: the captured part of the step, then each
    instance's synthetic solve (inside the graph, one workgroup each, or after
    it, one cooperative launch each); the same fields and inner iteration
    counts as stepping each instance on its own."""
    g = load_pkg()
    from gsis_amd import batch
    cases = batch.instances(3, seed=7, side=(12, 24), nv=(8, 16), precision_scheme=g.capi.SCHEME_UP1)
    single, grouped = [], []
    for c in cases:
        for lst in (single, grouped):
            s = solver_mod().GsisSolver(c, mode=g.capi.MODE_M1, kstep_show=1000000)
            s.initialize()
            lst.append(s)
    b = batch.Batch(grouped)
    b.set_ns_graph(ns_graph)
    # automatic: three instances are fewer than the 32 that put the solves in the graph
    assert b.ns_graph() == (3 if ns_graph == 1 else 0)
    for k in range(3):
        b.step(1)
        for s, t in zip(single, grouped):
            s.step(1)
            assert s.counters() == t.counters(), (s.case.name, k)
    for s, t in zip(single, grouped):
        ms, mt = s.macro(), t.macro()
        for key in KEYS:
            assert np.array_equal(ms[key], mt[key]), (s.case.name, key)
    b.close()
    for s in single + grouped:
        s.close()
