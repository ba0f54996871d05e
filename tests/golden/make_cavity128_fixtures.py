"""Full-size fixtures of the lid-driven cavity configuration (128^2 cells,
64^2 sinh velocity grid, lid 0.15) at Kn 0.01, 0.1 and 1, computed by the
oracle: GSIS (mode M1, at most 1000 inner iterations per step) with the
reference's MassCorrect, first-order upwind fluxes (with NND the cavity's
residual stalls), dt = 1, run until every residual is below 1e-8 (checked
every 50 steps) or 200 steps.  Writes cavity128_fixtures.npz beside this
script: per Kn the residual every 50 steps, the number of steps run, and
the final S_GKUA(1:4) of the interior cells and CF.  tests/test_gpu_cavity128.py
checks the GPU against it.

    python tests/golden/make_cavity128_fixtures.py
"""
import json
import pathlib
import sys
import time

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
from helpers import load_oracle, load_pkg  # noqa: E402

KNS = (0.01, 0.1, 1.0)
TOL, EVERY, MAXSTEPS, MAXIT = 1e-8, 50, 200, 1000


def case_for(kn):
    g = load_pkg()
    c = g.cases.cavity_case(128, 64, kn=kn, lid=0.15, dt=1.0)
    c.const["scheme"] = g.capi.SCHEME_UP1
    return c


def main():
    orc = load_oracle()
    out, meta = {}, {}
    for kn in KNS:
        c = case_for(kn)
        o = orc.OracleSolver(c, mode=1, ns_maxit=MAXIT)
        o.initialize()
        o.set_mass_correct(True)
        hist, t0, k = [], time.time(), 0
        while k < MAXSTEPS:
            o.step(EVERY)
            k += EVERY
            r = [float(x) for x in o.residual()]
            hist.append([k, r])
            print(f"kn {kn} step {k}: residual {r} ({time.time() - t0:.0f} s)", flush=True)
            if max(r) < TOL:
                break
        m = o.macro()
        key = f"kn{kn}"
        out[key + "_s"] = m["s"][0:4, 4:132, 4:132].copy()
        out[key + "_cf"] = m["cf"].copy()
        meta[key] = {"steps": k, "converged": max(hist[-1][1]) < TOL, "history": hist}
        o.close()
    out["meta"] = np.frombuffer(json.dumps({"tol": TOL, "every": EVERY, "max_steps": MAXSTEPS, "ns_maxit": MAXIT,
                                            "runs": meta}).encode(), dtype=np.uint8)
    np.savez_compressed(HERE / "cavity128_fixtures.npz", **out)


if __name__ == "__main__":
    main()
