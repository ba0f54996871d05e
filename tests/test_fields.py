"""Host logic of the FP32-vs-FP64 report: field extraction and relative L2."""
import numpy as np

from helpers import load_pkg


def test_field_errors_pick_the_right_slots():
    fields = load_pkg().fields
    nx = ny = 6
    rng = np.random.default_rng(5)
    s = rng.standard_normal((11, nx + 7, ny + 7))
    assert all(v == 0.0 for v in fields.field_errors(s, s, nx, ny).values())
    t = s.copy()
    t[6, 4:nx + 3, 4:ny + 3] *= 1.0 + 1e-3   # p_xy of interior cells only
    t[0, 0, 0] += 5.0                          # outside the interior: ignored
    e = fields.field_errors(t, s, nx, ny)
    assert abs(e["sigma_xy"] - 1e-3) < 1e-12
    assert e["rho"] == e["T"] == e["u"] == e["q"] == 0.0
    t = s.copy()
    t[7:9, 4:nx + 3, 4:ny + 3] *= 1.0 - 2e-4  # heat flux vector
    assert abs(fields.field_errors(t, s, nx, ny)["q"] - 2e-4) < 1e-12
