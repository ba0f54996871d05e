"""Residual histories of the reference's shipped decks (cylinder and channel,
NND, dt_global = 0.01, mode M0) on the oracle: the residual every 50 steps up
to 2150, or up to the step at which the run stops with T < 0.1.  Writes
deck_histories.json beside this script; tests/test_gpu_decks.py checks the GPU
against it.

    python tests/golden/make_deck_histories.py
"""
import json
import pathlib
import sys

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
from helpers import channel_case, cylinder_case, load_oracle  # noqa: E402

NSTEPS, EVERY = 2150, 50


def main():
    orc = load_oracle()
    out = {"steps": NSTEPS, "every": EVERY, "decks": {}}
    for name, make in (("channel", channel_case), ("cylinder", cylinder_case)):
        o = orc.OracleSolver(make())
        o.initialize()
        hist, stop = [], None
        for k in range(EVERY, NSTEPS + 1, EVERY):
            try:
                o.step(EVERY)
            except orc.OracleError as e:
                stop = {"within_steps": [k - EVERY + 1, k], "error": str(e)}
                break
            hist.append([k, [float(x) for x in o.residual()]])
            print(name, k, hist[-1][1], flush=True)
        o.close()
        out["decks"][name] = {"history": hist, "stopped": stop}
    (HERE / "deck_histories.json").write_text(json.dumps(out, indent=0))


if __name__ == "__main__":
    main()
