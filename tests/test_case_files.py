"""Cases written in the reference's own input formats (generic.dat,
generic.inp, control.in, Quadrature_Para.dat) read back unchanged by the
readers that restate the reference's (src/sub_init.f90:102-284,
src/opencfd_ec2d.f90:79-141)."""
import math

import numpy as np
import pytest

from helpers import load_pkg


def _nodes(case):
    m = case.mesh
    return m.x[2:-2, 2:-2], m.y[2:-2, 2:-2]


def test_mesh_and_bc_files_round_trip(tmp_path):
    g = load_pkg().cases
    case = g.couette_case(ny_cells=16, nx_cells=4, nv=8, uwall=0.0)
    xn, yn = _nodes(case)
    g.write_generic_dat(tmp_path / "generic.dat", xn, yn)
    g.write_generic_inp(tmp_path / "generic.inp", case.nx, case.ny, case.bcs)
    x2, y2 = g.read_generic_dat(tmp_path / "generic.dat")
    assert np.array_equal(x2, xn) and np.array_equal(y2, yn)
    bcs = g.read_generic_inp(tmp_path / "generic.inp", case.nx, case.ny)
    assert bcs == case.bcs
    mesh = g.build_mesh(x2, y2, bcs)
    for k in ("x", "y", "vol", "si", "sj", "ni1", "ni2", "nj1", "nj2"):
        assert np.array_equal(getattr(mesh, k), getattr(case.mesh, k)), k


def test_moving_wall_is_refused(tmp_path):
    g = load_pkg().cases
    case = g.cavity_case(6, 8, kn=0.1, lid=0.2)
    with pytest.raises(ValueError):
        g.write_generic_inp(tmp_path / "generic.inp", case.nx, case.ny, case.bcs)


@pytest.mark.parametrize("kn", [0.01, 0.1, 1.0])
def test_control_for_kn_reproduces_the_constants(tmp_path, kn):
    """init's constants (constants_from_control) of the written control.in
    equal those of the case given by Kn."""
    g = load_pkg().cases
    g.write_control_in(tmp_path / "control.in", g.control_for_kn(kn, dt=0.5))
    got = g.constants_from_control(g.read_control_in(tmp_path / "control.in"))
    want = g.constants_from_kn(kn, dt=0.5)
    for k in ("cfc", "re", "pr", "cpr", "x1s", "cv", "cp", "gamma", "ma", "twall", "dt"):
        assert math.isclose(got[k], want[k], rel_tol=1e-13, abs_tol=1e-15), (k, got[k], want[k])
    assert got["scheme"] == want["scheme"] and got["kstep_show"] == want["kstep_show"]


def test_quadrature_files_round_trip(tmp_path):
    g = load_pkg().cases
    gl = dict(mode=2, dvx=1.2, nvi=15, vxdown=-9.0, vxup=9.0, dvy=0.8, nvj=10, vydown=-4.0, vyup=4.0, kgln=4)
    g.write_quadrature_para(tmp_path / "q.dat", gl)
    assert g.read_quadrature_para(tmp_path / "q.dat") == gl
    for ngh in (7, 8, 16):
        g.write_quadrature_para(tmp_path / "q.dat", dict(mode=1, ngh=ngh))
        assert g.read_quadrature_para(tmp_path / "q.dat") == dict(mode=1, ngh=ngh)


def test_whole_case_directory_reads_back(tmp_path):
    """A 12 x 12 cavity with fixed walls and a Gauss-Legendre grid, written
    as the four files and read by shipped_case, is the case built in memory."""
    g = load_pkg().cases
    case = g.cavity_case(12, 8, kn=0.1, lid=0.0, dt=0.5)
    xn, yn = _nodes(case)
    q = dict(mode=2, dvx=1.2, nvi=15, vxdown=-9.0, vxup=9.0, dvy=1.2, nvj=15, vydown=-9.0, vyup=9.0, kgln=4)
    g.write_generic_dat(tmp_path / "generic.dat", xn, yn)
    g.write_generic_inp(tmp_path / "generic.inp", case.nx, case.ny, case.bcs)
    g.write_control_in(tmp_path / "control.in", g.control_for_kn(0.1, dt=0.5))
    g.write_quadrature_para(tmp_path / "Quadrature_Para.dat", q)
    back = g.shipped_case(tmp_path, "cavity12-files")
    vi, wi = g.gl_velocity_grid(1.2, 15, -9.0, 4)
    assert np.array_equal(back.vi, vi) and np.array_equal(back.wj, wi)
    assert back.bcs == case.bcs
    assert np.array_equal(back.mesh.vol, case.mesh.vol)
    assert math.isclose(back.const["cfc"], case.const["cfc"], rel_tol=1e-13)
