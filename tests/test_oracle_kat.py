"""Known-answer tests that pin the oracle and the host-side producers against
data recorded by the reference itself (tests/golden, written by
make_fixtures.py from the reference's output files and sources)."""
import json

import numpy as np
import pytest

from helpers import GOLDEN, cylinder_case, load_oracle, load_pkg


def test_gauleg_table_is_float32_of_reference_literals():
    """GAULEG constants (src/GKUA_lisansudu.f90:305-351) are REAL(4) literals;
    the float32 roundings of numpy's Legendre roots must equal them."""
    g = load_pkg()
    tab = json.loads((GOLDEN / "gauleg_table.json").read_text())
    for kgln, t in tab.items():
        x, w = g.cases.gauss_legendre_f32(int(kgln))
        xs = sorted(set(np.abs(x)))
        ws = {}
        for xi, wi in zip(x, w):
            ws[abs(xi)] = wi
        agl = sorted(float(np.float32(float(v))) for v in t["AGL"].values())
        assert np.array_equal(np.array(xs), np.array(agl)), kgln
        for k, a in t["AGL"].items():
            assert ws[float(np.float32(float(a)))] == float(np.float32(float(t["HGL"][k]))), (kgln, k)


def test_rcon_matches_fort21():
    """init prints RCON = Boltz/(Matom*Mole_Mass) with REAL(4) Boltz/Matom
    (src/sub_init.f90:17); fort.21 recorded 207.861758824218 for argon."""
    g = load_pkg()
    ref = json.loads((GOLDEN / "fort21.json").read_text())["Rcon"]
    rcon = g.cases.BOLTZ_F32 / (g.cases.MATOM_F32 * 40.0)
    assert abs(rcon - ref) < 5e-13 * ref
    # the FP64 constants would miss it by ~3e-10 relative: the literals matter
    assert abs(1.38065e-23 / (1.66053873e-27 * 40.0) - ref) > 1e-10 * ref


@pytest.fixture(scope="module")
def cyl_init():
    case = cylinder_case()
    orc = load_oracle().OracleSolver(case)
    orc.initialize()
    m = orc.macro()
    yield case, m
    orc.close()


def test_cell_centres_match_flow2d_gkua(cyl_init):
    """x1,y1 of cells 0..40 (incl. the extrapolated ghost ring) against the
    x,y columns of output/flow2d-GKUA.dat (f15.10)."""
    case, _ = cyl_init
    kat = json.loads((GOLDEN / "kat_flow2d_gkua.json").read_text())
    x1, y1 = case.mesh.x1y1()
    # x1 has lower bound -1; rows of the file run i fastest over 0..nx, j over 0..ny
    xs = x1[1:case.nx + 2, 1:case.ny + 2].ravel(order="F")
    ys = y1[1:case.nx + 2, 1:case.ny + 2].ravel(order="F")
    assert np.max(np.abs(xs - np.array(kat["x1"]))) < 6e-11
    assert np.max(np.abs(ys - np.array(kat["y1"]))) < 6e-11


def test_initial_moments_match_flow2d_gkua(cyl_init):
    """Maxwellian fill + GL moments + DETFP conversion of the uniform Ma=3 state
    reproduce output/flow2d-GKUA.dat to its 10 printed digits on all 41x41 cells."""
    case, m = cyl_init
    kat = json.loads((GOLDEN / "kat_flow2d_gkua.json").read_text())["moments"]
    s = m["s"][:, 3:case.nx + 4, 3:case.ny + 4]
    d, u, v, T = s[0], s[1], s[2], s[3]
    fields = {"d": d, "u": u, "v": v, "T": T, "p": d * T, "txx": s[4] - d * T, "tyy": s[5] - d * T,
              "txy": s[6], "qx": s[7], "qy": s[8], "tzz": s[9] - d * T}
    for k, val in fields.items():
        assert np.max(np.abs(val - kat[k])) < 5.01e-11, k


def test_initial_ns_state_matches_flow2d(cyl_init):
    """init_flow_zero's U on cells 3..n-3 against output/flow2d.dat (src/IO.f90:330-344)."""
    case, _ = cyl_init
    kat = json.loads((GOLDEN / "kat_flow2d.json").read_text())
    rows = np.array(kat["rows"])
    frd, s, cf, u = load_pkg().cases.initial_state(case)
    c = case.const
    x1, y1 = case.mesh.x1y1()
    got = []
    for j in range(3, case.ny - 2):
        for i in range(3, case.nx - 2):
            d1, u1, v1, tmp = u[:, i + 1, j + 1]
            T1 = (tmp - 0.5 * d1 * (u1 * u1 + v1 * v1)) / (c["cv"] * d1)
            got.append([x1[i + 1, j + 1], y1[i + 1, j + 1], d1, u1, v1, T1, d1 * T1 / (c["gamma"] * c["ma"] ** 2),
                        c["ma"] * np.sqrt(u1 ** 2 + v1 ** 2) / np.sqrt(T1)])
    got = np.array(got)
    assert got.shape == rows.shape
    assert np.max(np.abs(got - rows)) < 6e-11
