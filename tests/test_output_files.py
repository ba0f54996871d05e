"""Tecplot output of the reference (output, src/IO.f90:307-553) written from
the oracle's fields of the shipped 40x40 cylinder's initial state: every file
that this state determines equals the reference's shipped file byte for byte
(SHA-256 fixtures from tests/golden/make_fixtures.py)."""
import hashlib
import json

import numpy as np
import pytest

from helpers import GOLDEN, cylinder_case, load_oracle, load_pkg

DIGESTS = json.loads((GOLDEN / "output_digests.json").read_text())


def _out():
    load_pkg()
    import importlib
    return importlib.import_module("gsis_amd.output")


def write_initial_output(solver, case, directory):
    """All files from the solver's fields.  flow2d_Dem.dat averages ghost
    cells into the boundary nodes, and the shipped one was written while they
    still held the initial state (backup_Mesco_read calls output before the NS
    ghost BCs, src/Backup_write_read.f90:109 and :50), so it gets that U."""
    out = _out()
    m = solver.macro()
    out.write_output(directory, [(case, m)])
    m2 = dict(m)
    m2["u"] = load_pkg().cases.initial_state(case)[3]
    out.write_flow2d_dem(directory / "flow2d_Dem.dat", [(case, m2)])


def check_digests(directory):
    bad = []
    for name, d in DIGESTS.items():
        raw = (directory / name).read_bytes()
        if hashlib.sha256(raw).hexdigest() != d["sha256"]:
            bad.append((name, len(raw), d["bytes"], raw.count(b"\n"), d["lines"]))
    assert not bad, bad


def test_oracle_initial_state_writes_the_shipped_output_files(tmp_path):
    case = cylinder_case()
    o = load_oracle().OracleSolver(case)
    o.initialize()
    write_initial_output(o, case, tmp_path)
    check_digests(tmp_path)
    o.close()


def test_list_directed_records():
    """The shipped VAR_wall.dat header: a character item longer than the
    80-column record continues on the next record after a blank; integers take
    12 columns and a blank separates them from a following character item."""
    out = _out()
    assert out.list_directed("variables=x,y,rho_wall,U_wall,V_wall,Ttra_wall,P_wall,P_t_wall,tau_wall,"
                             "q_wall,qwall2") == (
        " variables=x,y,rho_wall,U_wall,V_wall,Ttra_wall,P_wall,P_t_wall,tau_wall,q_wall,\n qwall2\n")
    assert out.list_directed("ZONE I=", 39, ", J=", 1, ", F=point") == \
        " ZONE I=          39 , J=           1 , F=point\n"


def test_f_edit_descriptor_rules():
    """F15.10: asterisks when the value needs more columns, the sign kept on a
    negative value that rounds to zero, none on a zero; a row longer than the
    format continues on new records."""
    out = _out()
    assert out.fmt_f([[12345.6]], 1, 15, 10) == "*" * 15 + "\n"
    assert out.fmt_f([[-1e-13, -0.0, 0.5]], 3, 15, 10) == "  -0.0000000000   0.0000000000   0.5000000000\n"
    assert out.fmt_f([[1, 2, 3], [4, 5, 6]], 2, 4, 1) == " 1.0 2.0\n 3.0\n 4.0 5.0\n 6.0\n"


def test_blocks_write_one_zone_each(tmp_path):
    """One zone per block, in block order (the DO m=1,Num_Block loops)."""
    g = load_pkg()
    out = _out()
    case = g.cases.cavity_case(8, 4, 0.1)
    o = load_oracle().OracleSolver(case)
    o.initialize()
    m = o.macro()
    out.write_output(tmp_path, [(case, m), (case, m)], hybrid=True)
    txt = (tmp_path / "flow2d.dat").read_text().splitlines()
    zones = [k for k, ln in enumerate(txt) if "zone" in ln]
    n = (case.nx - 5) * (case.ny - 5)
    assert zones == [1, 2 + n, 3 + 2 * n, 4 + 2 * n + case.nx * case.ny]
    gk = (tmp_path / "flow2d-GKUA.dat").read_text().splitlines()
    assert len(gk) == 1 + 2 * (1 + (case.nx + 1) * (case.ny + 1))
    qj = (tmp_path / "q_g_j.dat").read_text().splitlines()
    assert len(qj) == 1 + 2 * (1 + 2 * (case.nx - 1) * case.ny)
    o.close()
