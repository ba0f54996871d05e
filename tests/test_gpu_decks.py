"""The reference's shipped decks as shipped (NND, dt_global = 0.01, mode M0)
over the long run: the GPU's residual every 50 steps follows the oracle's
history (tests/golden/deck_histories.json, written by
tests/golden/make_deck_histories.py) for 2150 steps of the channel, and the
cylinder stops with T < 0.1 in the same 50-step window as the oracle."""
import json

import numpy as np
import pytest

from helpers import GOLDEN, channel_case, cylinder_case, load_pkg, solver_mod

pytestmark = pytest.mark.gpu
RTOL = 1e-6


@pytest.mark.parametrize("name,make", [("channel", channel_case), ("cylinder", cylinder_case)])
def test_deck_residual_history_follows_the_oracle(name, make):
    g = load_pkg()
    ref = json.loads((GOLDEN / "deck_histories.json").read_text())["decks"][name]
    s = solver_mod().GsisSolver(make(), kstep_show=50)
    s.initialize()
    worst = 0.0
    for k, r_ref in ref["history"]:
        s.step(50)
        r = s.residual()
        worst = max(worst, float(np.max(np.abs(r - r_ref) / np.abs(r_ref))))
        assert np.allclose(r, r_ref, rtol=RTOL, atol=0), (name, k, r, r_ref)
    print(name, "max relative residual difference", worst, flush=True)
    if ref["stopped"] is not None:
        with pytest.raises(g.capi.GsisError) as err:
            s.step(50)
            s.counters()
        assert err.value.code == g.capi.GSIS_ETEMP
    s.close()
