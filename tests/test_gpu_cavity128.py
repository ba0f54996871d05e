"""The lid-driven cavity configuration at full size (128^2 cells, 64^2
velocities, lid 0.15; Kn 0.01, 0.1 and 1) in GSIS with the reference's
MassCorrect and first-order upwind fluxes, against the oracle's fixtures
(tests/golden/cavity128_fixtures.npz, written by
tests/golden/make_cavity128_fixtures.py): the residual every 50 steps and
the final density, velocity, temperature and collision frequency fields.
At Kn 0.1 also the FP32-sweep run's relative L2 error of rho, |u| and T
against those FP64 fields."""
import json

import numpy as np
import pytest

from helpers import GOLDEN, load_pkg, rel_l2, solver_mod

pytestmark = pytest.mark.gpu
FIX = GOLDEN / "cavity128_fixtures.npz"


def _case(g, kn):
    c = g.cases.cavity_case(128, 64, kn=kn, lid=0.15, dt=1.0)
    c.const["scheme"] = g.capi.SCHEME_UP1
    return c


@pytest.mark.skipif(not FIX.exists(), reason="fixture not generated")
@pytest.mark.parametrize("kn", [0.01, 0.1, 1.0])
def test_cavity128_gsis_follows_the_oracle(kn):
    g = load_pkg()
    with np.load(FIX, allow_pickle=False) as z:
        meta = json.loads(bytes(z["meta"]).decode())
        s_ref, cf_ref = z[f"kn{kn}_s"], z[f"kn{kn}_cf"]
    run = meta["runs"][f"kn{kn}"]
    every = meta["every"]
    errs = {}
    precs = (g.capi.PREC_FP64, g.capi.PREC_FP32) if kn == 0.1 else (g.capi.PREC_FP64,)
    for prec in precs:
        s = solver_mod().GsisSolver(_case(g, kn), mode=1, precision=prec, ns_maxit=meta["ns_maxit"], kstep_show=every)
        s.initialize()
        s.set_mass_correct(True)
        for k, r_ref in run["history"]:
            s.step(every)
            r = s.residual()
            print(f"kn {kn} prec {prec} step {k}: residual {r} oracle {r_ref}", flush=True)
            if prec == g.capi.PREC_FP64:
                assert np.allclose(r, r_ref, rtol=1e-6, atol=0), (k, r, r_ref)
        m = s.macro()
        s.close()
        if prec == g.capi.PREC_FP64:
            e_s = rel_l2(m["s"][0:4, 4:132, 4:132], s_ref)
            e_cf = rel_l2(m["cf"], cf_ref)
            print(f"kn {kn} FP64 fields vs oracle: S {e_s:.2e} CF {e_cf:.2e}", flush=True)
            assert e_s <= 1e-10 and e_cf <= 1e-10, (e_s, e_cf)
        else:
            f = m["s"][0:4, 4:132, 4:132]
            speed = np.hypot(f[1], f[2])
            errs = {"rho": rel_l2(f[0], s_ref[0]), "u": rel_l2(speed, np.hypot(s_ref[1], s_ref[2])),
                    "T": rel_l2(f[3], s_ref[3])}
            print(f"kn {kn} FP32 sweeps vs FP64 oracle fields: {errs}", flush=True)
            assert all(v < 1e-3 for v in errs.values()), errs
