"""Tecplot files of the reference's ``output`` (src/IO.f90:307-553), written
from the fields the C-ABI downloads (gsis_download_macro), in the reference's
formats to the byte.

Each file has a header record written list-directed, one zone per block and
one record per point written with an explicit F edit descriptor.  The
list-directed records follow the compiler that wrote the reference's shipped
output files: every record starts with a blank, a character item that follows
a numeric one gets a separating blank, default integers take 12 columns and a
record wraps at 80 columns.  F fields too narrow for a value are filled with
asterisks, and a format with fewer descriptors than values starts a new record
and reuses itself (format reversion).

Files written, for each block given as (case, macro) with macro the dict of
GsisSolver.macro():
  flow2d.dat       U on cells 3..n-3 (:329-346); with ``hybrid`` also the zone
                   of S_GKUA averaged to the nodes that the reference adds in
                   MODE == Hybrid_Mode (:348-364)
  flow2d-GKUA.dat  S_GKUA and CF on cells 0..n (:369-389)
  flow2d-S0.dat    S0_GKUA on cells 3..n-3 (:391-410)
  Tao_g_i.dat, Tao_g_j.dat, q_g_i.dat, q_g_j.dat
                   the NS-side and kinetic stresses and heat fluxes on the
                   faces (:454-527); the NS side (TaoNS*_in, Q*_in) comes from
                   the synthetic solve, which mode M0 never runs, so it is 0
                   unless given
  flow2d_Dem.dat   U averaged to the nodes, in dimensional units (:532-550),
                   for cases whose constants come from control.in
The wall files (VAR_wall.dat, Force_wall.dat) need the reference's Swall0 and
F_Wall, which the kinetic loop does not produce, and are not written.
"""
from __future__ import annotations

import math
import pathlib
from typing import Optional, Sequence

import numpy as np

from .cases import MATOM_F32

RECL = 80  # list-directed record length of the shipped files


def list_directed(*items) -> str:
    """A list-directed WRITE of character and integer items (one or more
    records, newline-terminated)."""
    recs = [" "]
    prev_num = False
    for it in items:
        if isinstance(it, str):
            text = (" " if prev_num else "") + it
            while len(recs[-1]) + len(text) > RECL:
                room = RECL - len(recs[-1])
                recs[-1] += text[:room]
                text = text[room:]
                recs.append(" ")
            recs[-1] += text
            prev_num = False
        else:
            text = f"{int(it):12d}"
            if len(recs[-1]) + len(text) > RECL:
                recs.append(" ")
            recs[-1] += text
            prev_num = True
    return "\n".join(recs) + "\n"


def fmt_f(rows: Sequence[Sequence[float]], n: int, w: int, d: int) -> str:
    """One WRITE(u, "(nFw.d)") per row: n fields per record (a longer row
    continues on new records), asterisks for a value that does not fit, and
    zero of either sign written as 0."""
    out = []
    for row in rows:
        for k in range(0, len(row), n):
            rec = []
            for v in row[k:k + n]:
                v = float(v)
                if v == 0.0:
                    v = 0.0
                t = f"{v:{w}.{d}f}"
                rec.append("*" * w if len(t) > w or not math.isfinite(v) else t)
            out.append("".join(rec))
    return "\n".join(out) + "\n"


def _zone(ni: int, nj: int) -> str:
    return list_directed(" zone ", "i= ", ni, " j= ", nj)


def write_flow2d(path, blocks, hybrid: bool = False) -> None:
    """flow2d.dat (src/IO.f90:329-366).  The first zone takes U(2), U(3) as
    they are (momenta) in its u, v, T and Ma columns, as the reference does."""
    parts = [list_directed("variables=x,y,d,u,v,T,p,Ma")]
    for case, m in blocks:
        c = case.const
        g, ma, cv = c["gamma"], c["ma"], c["cv"]
        x1, y1 = case.mesh.x1y1()
        u = m["u"]
        nx, ny = case.nx, case.ny
        parts.append(_zone(nx - 5, ny - 5))
        vals = []
        for j in range(3, ny - 2):
            for i in range(3, nx - 2):
                d1, u1, v1, e = (u[k, i + 1, j + 1] for k in range(4))
                t1 = (e - 0.5 * d1 * (u1 * u1 + v1 * v1)) / (cv * d1)
                vals.append([x1[i + 1, j + 1], y1[i + 1, j + 1], d1, u1, v1, t1, d1 * t1 / (g * ma * ma),
                             ma * math.sqrt(u1 ** 2 + v1 ** 2) / math.sqrt(t1)])
        parts.append(fmt_f(vals, 8, 20, 10))
    if hybrid:
        for case, m in blocks:
            c = case.const
            g, ma = c["gamma"], c["ma"]
            s = m["s"]
            nx, ny = case.nx, case.ny
            parts.append(_zone(nx, ny))
            vals = []
            sq = math.sqrt(2.0 / g)

            def avg(k, i, j):
                return (s[k, i + 3, j + 3] + s[k, i + 2, j + 3] + s[k, i + 3, j + 2] + s[k, i + 2, j + 2]) * 0.25

            for j in range(1, ny + 1):
                for i in range(1, nx + 1):
                    d1 = avg(0, i, j)
                    u1 = avg(1, i, j) / ma * sq
                    v1 = avg(2, i, j) / ma * sq
                    t1 = avg(3, i, j)
                    vals.append([case.mesh.x[i + 1, j + 1], case.mesh.y[i + 1, j + 1], d1, u1, v1, t1,
                                 d1 * t1 / (g * ma * ma), ma * math.sqrt(u1 ** 2 + v1 ** 2) / math.sqrt(t1)])
            parts.append(fmt_f(vals, 8, 20, 10))
    pathlib.Path(path).write_text("".join(parts))


def write_flow2d_gkua(path, blocks) -> None:
    """flow2d-GKUA.dat (src/IO.f90:369-389): S_GKUA and CF on cells 0..nx, 0..ny."""
    parts = [list_directed("variables=x,y,d,u,v,T,p,txx,tyy,txy,qx,qy,tzz,cf")]
    for case, m in blocks:
        x1, y1 = case.mesh.x1y1()
        s, cf = m["s"], m["cf"]
        nx, ny = case.nx, case.ny
        parts.append(_zone(nx + 1, ny + 1))
        vals = []
        for j in range(0, ny + 1):
            for i in range(0, nx + 1):
                S = s[:, i + 3, j + 3]
                d1, t1 = S[0], S[3]
                vals.append([x1[i + 1, j + 1], y1[i + 1, j + 1], d1, S[1], S[2], t1, d1 * t1, S[4] - d1 * t1,
                             S[5] - d1 * t1, S[6], S[7], S[8], S[9] - d1 * t1, cf[i + 1, j + 1]])
        parts.append(fmt_f(vals, 14, 15, 10))
    pathlib.Path(path).write_text("".join(parts))


def write_flow2d_s0(path, blocks) -> None:
    """flow2d-S0.dat (src/IO.f90:391-410): S0_GKUA on cells 3..n-3."""
    parts = [list_directed("variables=x,y,d,u,v,T,p,txx,tyy,txy,qx,qy,tzz")]
    for case, m in blocks:
        x1, y1 = case.mesh.x1y1()
        s0 = m["s0"]
        nx, ny = case.nx, case.ny
        parts.append(_zone(nx - 5, ny - 5))
        vals = []
        for j in range(3, ny - 2):
            for i in range(3, nx - 2):
                S = s0[:, i + 3, j + 3]
                vals.append([x1[i + 1, j + 1], y1[i + 1, j + 1], S[0], S[1], S[2], S[3], S[0] * S[3]] + list(S[4:10]))
        parts.append(fmt_f(vals, 13, 15, 10))
    pathlib.Path(path).write_text("".join(parts))


def _faces(path, header, blocks, key_in, key_g, ncomp, idir, fmt_n):
    """Tao_g_* / q_g_* (src/IO.f90:454-527): i-faces (idir 0) on i = 1..nx,
    j = 1..ny-1 at the mean of the two cell centres in x; j-faces likewise."""
    parts = [list_directed(header)]
    for case, m in blocks:
        x1, y1 = case.mesh.x1y1()
        nx, ny = case.nx, case.ny
        g = m[key_g]
        a = m.get(key_in)
        if a is None:
            a = np.zeros_like(g)
        ni, nj = (nx, ny - 1) if idir == 0 else (nx - 1, ny)
        parts.append(_zone(ni, nj))
        vals = []
        for j in range(1, nj + 1):
            for i in range(1, ni + 1):
                if idir == 0:
                    row = [(x1[i + 1, j + 1] + x1[i, j + 1]) / 2.0, y1[i + 1, j + 1]]
                else:
                    row = [x1[i + 1, j + 1], (y1[i + 1, j + 1] + y1[i + 1, j]) / 2.0]
                for k in range(ncomp):
                    row += [a[k, i + 1, j + 1], g[k, i + 1, j + 1]]
                vals.append(row)
        parts.append(fmt_f(vals, fmt_n, 20, 10))
    pathlib.Path(path).write_text("".join(parts))


def write_flow2d_dem(path, blocks, amu_t: Optional[Sequence[np.ndarray]] = None) -> None:
    """flow2d_Dem.dat (src/IO.f90:532-550): U averaged to the nodes, in SI
    units from the control.in constants.  Amu_t, the turbulent viscosity, is 0
    without a turbulence model."""
    parts = [list_directed("variables=x,y,d,u,v,T,p,Amut")]
    for kb, (case, m) in enumerate(blocks):
        c = case.const
        g, ma, cv, re = c["gamma"], c["ma"], c["cv"], c["re"]
        rcon, t_inf, nd_inf, mole, lref = c["rcon"], c["t_inf"], c["nd_inf"], c["mole_mass"], c["lref"]
        u = m["u"]
        nx, ny = case.nx, case.ny
        at = None if amu_t is None else amu_t[kb]
        parts.append(_zone(nx, ny))
        vals = []
        c_u = math.sqrt(g * rcon * t_inf)
        for j in range(1, ny + 1):
            for i in range(1, nx + 1):
                q = [(u[k, i + 1, j + 1] + u[k, i, j + 1] + u[k, i + 1, j] + u[k, i, j]) * 0.25 for k in range(4)]
                d1 = q[0]
                u1 = q[1] / d1
                v1 = q[2] / d1
                t1 = (q[3] - 0.5 * d1 * (u1 * u1 + v1 * v1)) / (cv * d1)
                amut = 0.0 if at is None else at[i + 1, j + 1]
                vals.append([case.mesh.x[i + 1, j + 1] * lref, case.mesh.y[i + 1, j + 1] * lref,
                             d1 * nd_inf * MATOM_F32 * mole, u1 * ma * c_u, v1 * ma * c_u, t1 * t_inf,
                             d1 * t1 * nd_inf * MATOM_F32 * mole * rcon * t_inf, amut * re])
        parts.append(fmt_f(vals, 8, 20, 10))
    pathlib.Path(path).write_text("".join(parts))


def write_output(directory, blocks, hybrid: bool = False) -> pathlib.Path:
    """All files above into ``directory`` (the reference's output/)."""
    d = pathlib.Path(directory)
    d.mkdir(parents=True, exist_ok=True)
    write_flow2d(d / "flow2d.dat", blocks, hybrid)
    write_flow2d_gkua(d / "flow2d-GKUA.dat", blocks)
    write_flow2d_s0(d / "flow2d-S0.dat", blocks)
    _faces(d / "Tao_g_i.dat", "variables=x,y,T1_i,T1_G_i,T2_i,T2_G_i,T12_i,T12_G_i,T3_i,T3_G_i", blocks,
           "taoi_in", "taoi", 4, 0, 10)
    _faces(d / "Tao_g_j.dat", "variables=x,y,T1_j,T1_G_j,T2_j,T2_G_j,T12_j,T12_G_j,T3_j,T3_G_j", blocks,
           "taoj_in", "taoj", 4, 1, 10)
    _faces(d / "q_g_i.dat", "variables=x,y,qx_i,qx_G_i,qy_i,qy_G_i", blocks, "qi_in", "qi", 2, 0, 6)
    _faces(d / "q_g_j.dat", "variables=x,y,qx_j,qx_G_j,qy_j,qy_G_j", blocks, "qj_in", "qj", 2, 1, 5)
    if all("t_inf" in case.const for case, _ in blocks):  # cases read from control.in
        write_flow2d_dem(d / "flow2d_Dem.dat", blocks)
    return d
