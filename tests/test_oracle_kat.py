"""Known-answer tests that pin the oracle and the host-side producers against
data recorded by the reference itself (tests/golden, written by
make_fixtures.py from the reference's output files and sources)."""
import json
import math

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


def test_gauher_table_is_float32_of_reference_literals():
    """GaussHermite's 8- and 16-point tables (src/GKUA_lisansudu.f90:384-433)
    are REAL(4) literals; the float32 roundings of the half-range rule computed
    from exp(-v^2) on [0, inf) must equal them.  The 7-point table
    (:369-383) is not that rule: it integrates v^k exp(-v^2) exactly only up
    to k = 6 (a 7-point Gauss rule reaches k = 13), so it cannot be derived
    and the host-side producer refuses NV = 7."""
    g = load_pkg()
    tab = json.loads((GOLDEN / "gauher_table.json").read_text())
    for nv in (8, 16):
        t = tab[str(nv)]
        x, w = g.cases.gauss_hermite_half_f32(nv)
        agh = np.array([float(np.float32(float(t["AGH"][str(k + 1)]))) for k in range(nv)])
        hgh = np.array([float(np.float32(float(t["HGH"][str(k + 1)]))) for k in range(nv)])
        assert np.array_equal(x, agh), nv
        assert np.array_equal(w, hgh), nv
    a7 = np.array([float(tab["7"]["AGH"][str(k + 1)]) for k in range(7)])
    h7 = np.array([float(tab["7"]["HGH"][str(k + 1)]) for k in range(7)])
    mom = [abs(np.sum(h7 * a7 ** k) / (math.gamma((k + 1) / 2) / 2) - 1.0) for k in range(14)]
    assert max(mom[:7]) < 1e-8 and min(mom[7:]) > 1e-6, mom
    with pytest.raises(ValueError):
        g.cases.gauss_hermite_half_f32(7)


def test_gauss_hermite_grid_integrates_polynomials():
    """The mirrored 2*NV-point grid with w_eff = w exp(v^2) integrates
    v^k exp(-v^2) over the real line exactly up to degree 2 NV - 1 (the
    half-range Gauss rule on each side), here to the float32 rounding of its
    table."""
    g = load_pkg()
    for nv in (8, 16):
        v, w = g.cases.gh_velocity_grid(nv)
        assert np.all(np.diff(v) > 0) and len(v) == 2 * nv
        for k in range(2 * nv):
            got = np.sum(w * np.exp(-v * v) * v ** k)
            want = 0.0 if k % 2 else math.gamma((k + 1) / 2)
            assert abs(got - want) <= 2e-7 * math.gamma((k + 1) / 2 + 0.5), (nv, k, got, want)


def test_gauss_hermite_input_read_from_quadrature_para(tmp_path):
    """Quadrature_Para.dat with method 1 and option 2 selects the 8-point
    half-range rule in both directions (src/opencfd_ec2d.f90:84-98)."""
    g = load_pkg()
    src = GOLDEN / "cylinder"
    for f in ("generic.dat", "generic.inp", "control.in"):
        (tmp_path / f).write_text((src / f).read_text())
    lines = (src / "Quadrature_Para.dat").read_text().splitlines()
    lines[2] = "1"
    (tmp_path / "Quadrature_Para.dat").write_text("\n".join(lines) + "\n")
    case = g.cases.shipped_case(tmp_path, "cylinder40-gh")
    v, w = g.cases.gh_velocity_grid(8)
    assert np.array_equal(case.vi, v) and np.array_equal(case.wj, w)


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
