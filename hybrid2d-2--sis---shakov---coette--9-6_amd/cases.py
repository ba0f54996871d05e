"""Host-side producers of the GSIS kernel inputs.

These mirror what the Fortran host computes before the first step: the mesh
geometry (Creat_Mesh, src/sub_init.f90:88-219, and the
ghost-node extrapolation of Update_coordinate_buffer_onemesh,
src/sub_init.f90:719-823), the discrete velocity grid (GKUA_lisansudu,
src/GKUA_lisansudu.f90:4-353), the constants of ``init`` (src/sub_init.f90:6-79)
and the initial state of ``init_flow_zero``/``GKUA_sudufenbu``
(src/sub_init.f90:390-463, src/GKUA_sudufenbu.f90:22-46).  They read the
reference's own input formats (generic.dat, generic.inp, control.in,
Quadrature_Para.dat) unchanged.

Arrays that cross the C-ABI keep the reference's Fortran layout: node
coordinates x(1-2:nx+2,1-2:ny+2), geometry (nx,ny), all column-major.  In
numpy they are held as arrays with order="F" indexed [i - lower, j - lower].
"""
from __future__ import annotations

import dataclasses
import math
import pathlib
from typing import List, Optional, Sequence

import numpy as np

from . import capi

LAP = 2
F32 = np.float32

# Gauss-Legendre nodes/weights of GAULEG (src/GKUA_lisansudu.f90:298-353).  The
# reference writes them as default-REAL (4-byte) literals, so their values are
# the float32 roundings of the textbook constants; they are produced here from
# numpy's Legendre roots and rounded the same way (checked against the
# reference's 10-digit table in the tests).
_KGLN_OK = (3, 4, 5, 6, 9, 10)


def gauss_legendre_f32(kgln: int):
    """Nodes (ascending, in [-1,1]) and weights rounded to float32."""
    if kgln not in _KGLN_OK:
        raise ValueError(f"KGLN must be one of {_KGLN_OK}")
    x, w = np.polynomial.legendre.leggauss(kgln)
    return x.astype(F32).astype(np.float64), w.astype(F32).astype(np.float64)


def gl_velocity_grid(dv: float, nint: int, vdown: float, kgln: int):
    """Composite Gauss-Legendre grid: lisanv/VNODEQUAN/DISVEL
    (src/GKUA_lisansudu.f90:119-128,170-296).  Returns (v, w_eff) with
    w_eff = FBAV(interval) * w_GL, the per-direction factor of the moment
    weight ``FBAVX*wi*FBAVY*wj`` (src/GKUA_DETFP.f90:314)."""
    x, w = gauss_legendre_f32(kgln)
    # the reference builds the nodes as VAPB + VBMA*(+-AGL(k)) with AGL>=0
    vd = [vdown + float(F32(iv)) * dv for iv in range(nint + 1)]
    nodes, weights = [], []
    for k in range(nint):
        va, vb = vd[k], vd[k + 1]
        vapb = (va + vb) / 2.0
        vbma = (vb - va) / 2.0
        for q in range(kgln):
            xq = x[q]
            if xq == 0.0:
                node = vapb
            elif xq < 0.0:
                node = vapb + vbma * (-abs(xq))
            else:
                node = vapb + vbma * abs(xq)
            nodes.append(node)
            weights.append(vbma * w[q])
    return np.array(nodes), np.array(weights)


# Half-range Gauss-Hermite rules of GaussHermite (src/GKUA_lisansudu.f90:364-436):
# the nodes and weights of  int_0^inf f(v) exp(-v^2) dv  with NV points, written
# there as REAL(4) literals.  They are computed here from the weight function
# itself -- the recurrence coefficients of the polynomials orthogonal on
# [0, inf) by the Stieltjes procedure on a 96 x 48-point Gauss-Legendre
# discretisation of exp(-v^2) over [0, 12], then the Golub-Welsch eigenvalue
# problem -- and rounded to float32 (checked against the reference's table in
# the tests).  NV = 8 and 16 only: the reference's 7-point table is not this
# rule (it is exact only up to degree 6, where 7 Gauss points reach 13), so it
# cannot be derived.
_NGH_OK = (8, 16)


def gauss_hermite_half_f32(nv: int):
    """Nodes (ascending, > 0) and weights of the NV-point half-range rule,
    rounded to float32."""
    if nv not in _NGH_OK:
        raise ValueError(f"half-range Gauss-Hermite NV must be one of {_NGH_OK}")
    x, w = np.polynomial.legendre.leggauss(48)
    edges = np.linspace(0.0, 12.0, 97)
    a, b = edges[:-1, None], edges[1:, None]
    X = ((a + b) / 2 + (b - a) / 2 * x).ravel()
    W = ((b - a) / 2 * w).ravel() * np.exp(-X * X)
    alpha, beta = np.zeros(nv), np.zeros(nv)
    p_prev, p, nrm_prev = np.zeros_like(X), np.ones_like(X), 1.0
    for k in range(nv):
        nrm = float(np.sum(W * p * p))
        alpha[k] = float(np.sum(W * X * p * p)) / nrm
        beta[k] = nrm if k == 0 else nrm / nrm_prev
        p_prev, p, nrm_prev = p, (X - alpha[k]) * p - beta[k] * p_prev, nrm
    off = np.sqrt(beta[1:])
    nodes, vec = np.linalg.eigh(np.diag(alpha) + np.diag(off, 1) + np.diag(off, -1))
    weights = beta[0] * vec[0, :] ** 2
    return nodes.astype(F32).astype(np.float64), weights.astype(F32).astype(np.float64)


def gh_velocity_grid(nv: int):
    """Gauss-Hermite grid of lisanv (src/GKUA_lisansudu.f90:103-118): the
    half-range rule mirrored to 2*NV ascending nodes.  Returns (v, w_eff) with
    w_eff = w exp(v^2), the per-direction factor IntegralGaussHermite applies
    (src/GKUA_DETFP.f90:197-200; the same in GaussHermite_rs and the wall
    integrals)."""
    x, w = gauss_hermite_half_f32(nv)
    v = np.concatenate([-x[::-1], x])
    wd = np.concatenate([w[::-1], w])
    return v, wd * np.exp(v * v)


def newton_cotes_weights(n: int, h: float) -> np.ndarray:
    """Composite closed Newton-Cotes weights on n equally spaced points:
    Simpson 1/3 on an even number of intervals, plus Simpson 3/8 on the last
    three intervals when n-1 is odd."""
    if n < 3:
        raise ValueError("need at least 3 points")
    m = n - 1
    w = np.zeros(n)
    n13 = m if m % 2 == 0 else m - 3
    for k in range(0, n13, 2):
        w[k] += h / 3.0
        w[k + 1] += 4.0 * h / 3.0
        w[k + 2] += h / 3.0
    if m % 2 == 1:
        k = n13
        w[k] += 3.0 * h / 8.0
        w[k + 1] += 9.0 * h / 8.0
        w[k + 2] += 9.0 * h / 8.0
        w[k + 3] += 3.0 * h / 8.0
    return w


def nc_velocity_grid(n: int, vmax: float):
    """Uniform grid on [-vmax, vmax] with composite Newton-Cotes weights."""
    v = np.linspace(-vmax, vmax, n)
    return v, newton_cotes_weights(n, v[1] - v[0])


def sinh_velocity_grid(n: int, vmax: float, a: float = 2.0):
    """Non-uniform grid, dense near zero: v = vmax sinh(a s)/sinh(a) on a uniform
    s-grid of [-1,1]; weights = Newton-Cotes weight in s times dv/ds."""
    s = np.linspace(-1.0, 1.0, n)
    v = vmax * np.sinh(a * s) / math.sinh(a)
    jac = vmax * a * np.cosh(a * s) / math.sinh(a)
    return v, newton_cotes_weights(n, s[1] - s[0]) * jac


# ---------------------------------------------------------------------------
# mesh
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class Mesh:
    nx: int
    ny: int
    x: np.ndarray  # (nx+4, ny+4) Fortran x(-1:nx+2, -1:ny+2)
    y: np.ndarray
    vol: np.ndarray  # (nx, ny)
    si: np.ndarray
    sj: np.ndarray
    ni1: np.ndarray
    ni2: np.ndarray
    nj1: np.ndarray
    nj2: np.ndarray

    def x1y1(self):
        """Cell centres on (-1:nx+1, -1:ny+1) (src/sub_init.f90:814-819)."""
        x, y = self.x, self.y
        x1 = (x[:-1, :-1] + x[1:, :-1] + x[:-1, 1:] + x[1:, 1:]) * 0.25
        y1 = (y[:-1, :-1] + y[1:, :-1] + y[:-1, 1:] + y[1:, 1:]) * 0.25
        return x1, y1


def build_mesh(xn: np.ndarray, yn: np.ndarray, bcs: Sequence["BC"], lref: float = 1.0,
               inner: Optional[dict] = None) -> Mesh:
    """Geometry of Creat_Mesh (src/sub_init.f90:178-208) and the ghost nodes of
    Update_coordinate_buffer_onemesh (src/sub_init.f90:719-823): linear
    extrapolation on physical faces, then the corners.  xn, yn: node
    coordinates (nx, ny).  ``inner`` maps the list index of an inner subface
    to (xn_nb, yn_nb, bc_nb, orient): the neighbour block's nodes, its matching
    subface and the orientation; that face's ghost nodes are the neighbour's
    nodes next to bc_nb instead (:739-770)."""
    nx, ny = xn.shape
    x = np.zeros((nx + 4, ny + 4), order="F")
    y = np.zeros((nx + 4, ny + 4), order="F")
    o = 1  # array offset of node index 1 is 2 (lower bound -1): X[i+1]
    x[2:nx + 2, 2:ny + 2] = xn / lref
    y[2:nx + 2, 2:ny + 2] = yn / lref

    def X(a, i, j):
        return a[i + o, j + o]

    vol = np.zeros((nx, ny), order="F")
    for j in range(1, ny):
        for i in range(1, nx):
            vol[i - 1, j - 1] = abs((X(x, i, j) - X(x, i + 1, j + 1)) * (X(y, i + 1, j) - X(y, i, j + 1))
                                    - (X(x, i + 1, j) - X(x, i, j + 1)) * (X(y, i, j) - X(y, i + 1, j + 1))) * 0.5
    si = np.zeros((nx, ny), order="F"); ni1 = np.zeros_like(si); ni2 = np.zeros_like(si)
    for j in range(1, ny):
        for i in range(1, nx + 1):
            dx = X(x, i, j + 1) - X(x, i, j)
            dy = X(y, i, j + 1) - X(y, i, j)
            s = math.sqrt(dx * dx + dy * dy)
            si[i - 1, j - 1] = s
            ni1[i - 1, j - 1] = dy / s
            ni2[i - 1, j - 1] = -dx / s
    sj = np.zeros((nx, ny), order="F"); nj1 = np.zeros_like(sj); nj2 = np.zeros_like(sj)
    for j in range(1, ny + 1):
        for i in range(1, nx):
            dx = X(x, i + 1, j) - X(x, i, j)
            dy = X(y, i + 1, j) - X(y, i, j)
            s = math.sqrt(dx * dx + dy * dy)
            sj[i - 1, j - 1] = s
            nj1[i - 1, j - 1] = -dy / s
            nj2[i - 1, j - 1] = dx / s
    # ghost nodes: neighbour nodes on inner faces, linear extrapolation on
    # physical ones, in BC order
    for kb, b in enumerate(bcs):
        ib, ie, jb, je = b.ist, b.iend, b.jst, b.jend
        if inner is not None and kb in inner:
            xnb, ynb, c, orient = inner[kb]
            n1 = ie - ib + je - jb + 1
            for l in range(1, LAP + 1):
                for k in range(1, n1 + 1):
                    m = k if orient == 2 else n1 + 1 - k  # position along the neighbour's face
                    if c.face == 1:
                        sa = (c.ist + l, c.jst + m - 1)
                    elif c.face == 2:
                        sa = (c.ist + m - 1, c.jst + l)
                    elif c.face == 3:
                        sa = (c.iend - l, c.jst + m - 1)
                    else:
                        sa = (c.ist + m - 1, c.jend - l)
                    if b.face == 1:
                        da = (ib - l, jb + k - 1)
                    elif b.face == 2:
                        da = (ib + k - 1, jb - l)
                    elif b.face == 3:
                        da = (ie + l, jb + k - 1)
                    else:
                        da = (ib + k - 1, je + l)
                    x[da[0] + o, da[1] + o] = xnb[sa[0] - 1, sa[1] - 1] / lref
                    y[da[0] + o, da[1] + o] = ynb[sa[0] - 1, sa[1] - 1] / lref
            continue
        for k in range(1, LAP + 1):
            if b.face == 1:
                for a in (x, y):
                    a[ib - k + o, jb + o:je + 1 + o] = 2.0 * a[ib + o, jb + o:je + 1 + o] - a[ib + k + o, jb + o:je + 1 + o]
            elif b.face == 2:
                for a in (x, y):
                    a[ib + o:ie + 1 + o, jb - k + o] = 2.0 * a[ib + o:ie + 1 + o, jb + o] - a[ib + o:ie + 1 + o, jb + k + o]
            elif b.face == 3:
                for a in (x, y):
                    a[ie + k + o, jb + o:je + 1 + o] = 2.0 * a[ie + o, jb + o:je + 1 + o] - a[ie - k + o, jb + o:je + 1 + o]
            else:
                for a in (x, y):
                    a[ib + o:ie + 1 + o, je + k + o] = 2.0 * a[ib + o:ie + 1 + o, je + o] - a[ib + o:ie + 1 + o, je - k + o]
    for j in range(1, LAP + 1):
        for i in range(1, LAP + 1):
            i1, j1, i2, j2 = 1 - i, 1 - j, nx + i, ny + j
            for a in (x, y):
                a[i1 + o, j1 + o] = X(a, i1, 1) + X(a, 1, j1) - X(a, 1, 1)
                a[i2 + o, j2 + o] = X(a, i2, ny) + X(a, nx, j2) - X(a, nx, ny)
                a[i1 + o, j2 + o] = X(a, i1, ny) + X(a, 1, j2) - X(a, 1, ny)
                a[i2 + o, j1 + o] = X(a, i2, 1) + X(a, nx, j1) - X(a, nx, 1)
    return Mesh(nx, ny, x, y, vol, si, sj, ni1, ni2, nj1, nj2)


@dataclasses.dataclass
class BC:
    face: int
    ist: int
    iend: int
    jst: int
    jend: int
    code: int  # GSIS_BC_* or, for an inner face, the neighbour block number (>= 1)
    uw: float = 0.0
    vw: float = 0.0
    # inner face: the neighbour's node range (ist, iend, jst, jend) as given on
    # the neighbour line of generic.inp (src/sub_init.f90:282-284)
    nrange: Optional[tuple] = None


def bc_from_range(ist: int, iend: int, jst: int, jend: int, code_pos: int, nx: int, ny: int,
                  uw: float = 0.0, vw: float = 0.0) -> BC:
    """Face detection and code negation of read_bcin (src/sub_init.f90:268-278)."""
    if ist == iend and ist == 1:
        face = 1
    elif ist == iend and ist == nx:
        face = 3
    elif jst == jend and jst == 1:
        face = 2
    elif jst == jend and jst == ny:
        face = 4
    else:
        raise ValueError("BC range is not a block face")
    return BC(face, ist, iend, jst, jend, -int(code_pos), uw, vw)


# ---------------------------------------------------------------------------
# reference input files
# ---------------------------------------------------------------------------
class ListDirected:
    """Fortran list-directed READ over the records (lines) of a text file: each
    READ starts at the next record, takes values across records until it has
    what it asks for and drops the rest of its last record.  Separators are
    blanks and commas; ``r*c`` repeats c r times."""

    def __init__(self, text: str):
        self.lines = text.splitlines()
        self.pos = 0

    def read(self, n: int) -> List[str]:
        out: List[str] = []
        while len(out) < n:
            if self.pos >= len(self.lines):
                raise EOFError("end of file in a list-directed read")
            for tok in self.lines[self.pos].replace(",", " ").split():
                if "*" in tok:
                    r, c = tok.split("*", 1)
                    out.extend([c] * int(r))
                else:
                    out.append(tok)
            self.pos += 1
        return out[:n]

    def skip(self) -> None:
        """A READ with an empty list: skips one record."""
        self.pos += 1


def _real(tok: str) -> float:
    return float(tok.replace("d", "e").replace("D", "e"))


def read_generic_dat_blocks(path: str | pathlib.Path) -> List[tuple]:
    """generic.dat: NB, then per block a record with NI NJ (anything after them
    on that record, such as NK, is dropped) and the list x, y, z of NI*NJ
    values each (src/sub_init.f90:103-178).  Returns [(xn, yn)] with node
    arrays of shape (NI, NJ)."""
    f = ListDirected(pathlib.Path(path).read_text())
    nb = int(f.read(1)[0])
    out = []
    for _ in range(nb):
        ni, nj = (int(t) for t in f.read(2))
        vals = np.array([_real(t) for t in f.read(3 * ni * nj)])
        xn = vals[:ni * nj].reshape((ni, nj), order="F")
        yn = vals[ni * nj:2 * ni * nj].reshape((ni, nj), order="F")
        out.append((xn, yn))
    return out


def read_generic_dat(path: str | pathlib.Path):
    """generic.dat of one block: (xn, yn)."""
    blocks = read_generic_dat_blocks(path)
    if len(blocks) != 1:
        raise ValueError("only single-block meshes are supported here (see multiblock)")
    return blocks[0]


def read_generic_inp(path: str | pathlib.Path, nx: int, ny: int) -> List[BC]:
    """generic.inp: header, NB, NI NJ, solver, subface count, BC lines
    (src/sub_init.f90:236-284)."""
    lines = [ln for ln in pathlib.Path(path).read_text().splitlines()]
    toks = [ln.split() for ln in lines]
    toks = [t for t in toks]
    # line 0 is skipped by the reader
    nb = int(toks[1][0])
    if nb != 1:
        raise ValueError("only single-block meshes are supported")
    ni, nj = int(toks[2][0]), int(toks[2][1])
    if (ni, nj) != (nx, ny):
        raise ValueError("generic.inp block size does not match generic.dat")
    nsub = int(toks[4][0])
    out = []
    for k in range(nsub):
        ist, iend, jst, jend, code = (int(v) for v in toks[5 + k][:5])
        out.append(bc_from_range(ist, iend, jst, jend, code, nx, ny))
    return out


def read_control_in(path: str | pathlib.Path) -> dict:
    """control.in as read by read_parameter (src/opencfd_ec2d.f90:123-141).
    vis_ref and TREF are re-read on every model line, so the last line wins."""
    lines = pathlib.Path(path).read_text().splitlines()

    def nums(k):
        return [float(t.replace("d", "e").replace("D", "e")) for t in lines[k].split()]

    c = {}
    c["Mole_Mass"], c["gamma"], c["KHS"] = nums(2)[:3]
    c["ALPgas"], c["OMGgas"], c["vis_ref"], c["TREF"] = nums(4)[:4]
    c["Sgas"], c["vis_ref"], c["TREF"] = nums(6)[:3]
    c["DSLJ"], c["TSLJ"], c["vis_ref"], c["TREF"] = nums(8)[:4]
    (c["Ma"], c["ND_inf"], c["T_inf"], c["Twall"], c["P_outlet"], c["AoA"], c["Lref"]) = nums(10)[:7]
    (c["If_viscous"], c["Time_Method"], c["Iflag_Scheme"], c["Iflag_Flux"],
     c["IFlag_Reconstruction"]) = nums(13)[:5]
    (c["Iflag_init"], c["Iflag_local_dt"], c["dt_global"], c["CFL"], c["dtmax"], c["dtmin"], c["t_end"],
     c["Kstep_save"], c["Kstep_show"]) = nums(16)[:9]
    return c


def read_quadrature_para(path: str | pathlib.Path) -> dict:
    """Quadrature_Para.dat (src/opencfd_ec2d.f90:79-121)."""
    lines = pathlib.Path(path).read_text().splitlines()
    mode = int(lines[2].split()[0])
    out = {"mode": mode}
    if mode == 1:
        out["ngh"] = {1: 7, 2: 8, 3: 16}[int(lines[6].split()[0])]
    elif mode == 2:
        dvx, nvi, vxd, vxu = lines[9].split()[:4]
        dvy, nvj, vyd, vyu = lines[11].split()[:4]
        out.update(dvx=float(dvx), nvi=int(nvi), vxdown=float(vxd), vxup=float(vxu),
                   dvy=float(dvy), nvj=int(nvj), vydown=float(vyd), vyup=float(vyu),
                   kgln=int(lines[13].split()[0]))
    return out


# ---------------------------------------------------------------------------
# writers of the same files, so that a case built here can be given to the
# reference as input (the readers above read them back unchanged)
# ---------------------------------------------------------------------------
def _f(v: float) -> str:
    return repr(float(v))


def write_generic_dat(path: str | pathlib.Path, xn: np.ndarray, yn: np.ndarray) -> None:
    """generic.dat of one block: NB; NI NJ NK=1; x, y, z column-major,
    list-directed (src/sub_init.f90:102-178 reads x and y)."""
    ni, nj = xn.shape
    vals = np.concatenate([np.asarray(xn, float).ravel(order="F"), np.asarray(yn, float).ravel(order="F"),
                           np.zeros(ni * nj)])
    rows = [" ".join(f"{v:24.16E}" for v in vals[k:k + 4]) for k in range(0, len(vals), 4)]
    pathlib.Path(path).write_text(f"1\n{ni} {nj} 1\n" + "\n".join(rows) + "\n")


def write_generic_inp(path: str | pathlib.Path, nx: int, ny: int, bcs: Sequence[BC]) -> None:
    """generic.inp of one GKUA block (solver 2): the subfaces with their
    positive codes in list order (src/sub_init.f90:236-284).  The format has no
    wall velocity (the reference sets it to 0, src/GKUA_boundary.f90:66-67), so
    a moving wall is refused."""
    lines = [" generic.inp", "1", f"{nx} {ny}", "2", str(len(bcs))]
    for b in bcs:
        if b.uw != 0.0 or b.vw != 0.0:
            raise ValueError("generic.inp cannot express a moving wall")
        lines.append(f"{b.ist:8d} {b.iend:8d} {b.jst:8d} {b.jend:8d} {-b.code:8d}")
    pathlib.Path(path).write_text("\n".join(lines) + "\n")


def write_control_in(path: str | pathlib.Path, c: dict) -> None:
    """control.in with the keys read_control_in returns, on the lines
    read_parameter reads (src/opencfd_ec2d.f90:123-141)."""
    v = {k: _f(val) for k, val in c.items()}
    i = {k: str(int(c[k])) for k in ("KHS", "If_viscous", "Time_Method", "Iflag_Scheme", "Iflag_Flux",
                                     "IFlag_Reconstruction", "Iflag_init", "Iflag_local_dt", "Kstep_save",
                                     "Kstep_show")}
    text = [
        " #Inflow Condition#",
        " Mole_Mass,   gamma,  KHS,",
        f"{v['Mole_Mass']} {v['gamma']} {i['KHS']}",
        "VHS model :        Alpha       Omiga        Vis_ref            Tref",
        f"{v['ALPgas']} {v['OMGgas']} {v['vis_ref']} {v['TREF']}",
        "Sutherland model : S_k        Vis_ref       Tref",
        f"{v['Sgas']} {v['vis_ref']} {v['TREF']}",
        "LJ 12-6 model :    DSLJ,     TSLJ,      vis_ref                Tref",
        f"{v['DSLJ']} {v['TSLJ']} {v['vis_ref']} {v['TREF']}",
        " Ma           ND_inf(1)     T_inf(K)      Twall(K)    P_outlet      AoA(degree)        Lref(m)",
        f"{v['Ma']} {v['ND_inf']} {v['T_inf']} {v['Twall']} {v['P_outlet']} {v['AoA']} {v['Lref']}",
        "#Numerical Scheme for NS#",
        "Iflag_viscous    Time_Method   Iflag_scheme  Iflag_flux  IFlag_Reconstruction",
        f"{i['If_viscous']} {i['Time_Method']} {i['Iflag_Scheme']} {i['Iflag_Flux']} {i['IFlag_Reconstruction']}",
        "#Running Model Control#",
        " Iflag_init       Iflag_local_dt,     dt_global,         CFL,       dtmax,       dtmin,     tend"
        "        Kstep_save        Kstep_show",
        f"{i['Iflag_init']} {i['Iflag_local_dt']} {v['dt_global']} {v['CFL']} {v['dtmax']} {v['dtmin']} "
        f"{v['t_end']} {i['Kstep_save']} {i['Kstep_show']}",
    ]
    pathlib.Path(path).write_text("\n".join(text) + "\n")


def write_quadrature_para(path: str | pathlib.Path, q: dict) -> None:
    """Quadrature_Para.dat (src/opencfd_ec2d.f90:79-121) for method 1
    (Gauss-Hermite, ``ngh`` = 7, 8 or 16) or 2 (Gauss-Legendre, ``dvx, nvi,
    vxdown, vxup, dvy, nvj, vydown, vyup, kgln``); the other method's lines
    hold placeholders."""
    gh = {7: 1, 8: 2, 16: 3}[q.get("ngh", 8)]
    gl = q if q["mode"] == 2 else dict(dvx=1.2, nvi=15, vxdown=-9.0, vxup=9.0, dvy=1.2, nvj=15, vydown=-9.0,
                                      vyup=9.0, kgln=4)
    text = [
        "##Quadrature method & parameters specify",
        "Quadrature method:1:--Gauss Hermite ;2:--Gauss Legendre;3:--Newton-Cotes",
        str(int(q["mode"])),
        "#parameter for various quadrature method.",
        "###############Gauss Hermite#################",
        "#.1:--7 GH points;2:--8 GH points; 3--16 GH points",
        str(gh),
        "###############Gauss Legendre ################",
        "#Dvx    Intervals  Vx-lower-bound  Vx upper bound",
        f"{_f(gl['dvx'])} {int(gl['nvi'])} {_f(gl['vxdown'])} {_f(gl['vxup'])}",
        "#Dvy    Intervals  Vy-lower-bound  Vy upper bound",
        f"{_f(gl['dvy'])} {int(gl['nvj'])} {_f(gl['vydown'])} {_f(gl['vyup'])}",
        "#Inner division:KGLN  --3,4,5,6,9,10 are available",
        str(int(gl["kgln"])),
    ]
    pathlib.Path(path).write_text("\n".join(text) + "\n")


def control_for_kn(kn: float, omega: float = 0.81, gamma: float = 5.0 / 3.0, ma: float = 1.0, dt: float = 0.01,
                   t_inf: float = 273.15, mole_mass: float = 40.0, vis_ref: float = 2.117e-5,
                   scheme: int = capi.SCHEME_NND, kstep_show: int = 50) -> dict:
    """A control.in (VHS model) whose constants, as init derives them
    (constants_from_control), are those of constants_from_kn(kn): ND_inf is
    solved from Re = cfc Ma sqrt(2 gamma) with cfc = sqrt(pi)/(2 Kn).  The
    Prandtl number follows from gamma in init (2/3 for gamma = 5/3)."""
    rcon = BOLTZ_F32 / (MATOM_F32 * mole_mass)
    re = math.sqrt(math.pi) / (2.0 * kn) * ma * math.sqrt(2.0 * gamma)
    vis_inf = vis_ref  # T_inf = TREF
    nd = re * vis_inf / (MATOM_F32 * mole_mass * ma * math.sqrt(gamma * rcon * t_inf) * 1.0)
    return dict(Mole_Mass=mole_mass, gamma=gamma, KHS=0, ALPgas=1.0, OMGgas=omega, vis_ref=vis_ref, TREF=t_inf,
                Sgas=148.0, DSLJ=3.798e-10, TSLJ=71.4, Ma=ma, ND_inf=nd, T_inf=t_inf, Twall=t_inf, P_outlet=-1.0,
                AoA=0.0, Lref=1.0, If_viscous=1, Time_Method=0, Iflag_Scheme=scheme, Iflag_Flux=5,
                IFlag_Reconstruction=0, Iflag_init=0, Iflag_local_dt=1, dt_global=dt, CFL=20.0, dtmax=10.0,
                dtmin=1e-7, t_end=1000.0, Kstep_save=100, Kstep_show=kstep_show)


# ---------------------------------------------------------------------------
# constants of init (src/sub_init.f90:6-79)
# ---------------------------------------------------------------------------
BOLTZ_F32 = float(F32(1.38065e-23))  # Boltz, REAL(4) literal (src/COMMONDATA.f90:9)
MATOM_F32 = float(F32(1.66053873e-27))  # Matom, REAL(4) literal


def constants_from_control(c: dict) -> dict:
    """Kinetic and NS constants as computed by ``init`` for the VHS model."""
    g = c["gamma"]
    omg = c["OMGgas"]
    x1s = 1.0 - omg
    pr = g / (0.25 * (9.0 * g - 5.0))
    rcon = BOLTZ_F32 / (MATOM_F32 * c["Mole_Mass"])
    if int(c["KHS"]) != 0:
        raise NotImplementedError("only the VHS viscosity law (KHS=0) is implemented")
    vis_inf = c["vis_ref"] * (c["T_inf"] / c["TREF"]) ** omg
    ma = c["Ma"]
    re = c["ND_inf"] * MATOM_F32 * c["Mole_Mass"] * ma * math.sqrt(g * rcon * c["T_inf"]) * c["Lref"] / vis_inf
    aoa = c["AoA"] * math.pi / 180.0  # PI of the reference and math.pi coincide in FP64
    cv = 1.0 / (g * (g - 1.0) * ma * ma)
    cp = cv * g
    cfc = re / (ma * math.sqrt(2.0 * g))
    cpr = (1.0 - pr) / (5.0 / 2.0)
    return dict(gamma=g, ma=ma, re=re, pr=pr, cv=cv, cp=cp, aoa=aoa, cfc=cfc, cpr=cpr, x1s=x1s,
                twall=c["Twall"] / c["T_inf"], dt=c["dt_global"], rcon=rcon,
                if_viscous=int(c["If_viscous"]), scheme=int(c["Iflag_Scheme"]),
                local_dt=int(c["Iflag_local_dt"]), cfl=c["CFL"], dtmax=c["dtmax"], dtmin=c["dtmin"],
                kstep_show=int(c["Kstep_show"]),
                # dimensional scales of the output files (src/IO.f90:532-550)
                t_inf=c["T_inf"], nd_inf=c["ND_inf"], mole_mass=c["Mole_Mass"], lref=c["Lref"])


def constants_from_kn(kn: float, omega: float = 0.81, pr: float = 2.0 / 3.0, gamma: float = 5.0 / 3.0,
                      ma: float = 1.0, twall: float = 1.0, dt: float = 0.01) -> dict:
    """Constants for a case given by its Knudsen number.  The rarefaction
    parameter cfc = p L / (mu sqrt(2RT)) relates to Kn = mu sqrt(pi R T/2)/(p L)
    by cfc = sqrt(pi)/(2 Kn).  cpr = (1-Pr)/2.5 as in init (src/sub_init.f90:66).
    ``ma`` only scales the NS-side variables (Convert_NS); the kinetic solution
    of wall/periodic cases does not depend on it."""
    cfc = math.sqrt(math.pi) / (2.0 * kn)
    cv = 1.0 / (gamma * (gamma - 1.0) * ma * ma)
    re = cfc * ma * math.sqrt(2.0 * gamma)
    return dict(gamma=gamma, ma=ma, re=re, pr=pr, cv=cv, cp=cv * gamma, aoa=0.0, cfc=cfc,
                cpr=(1.0 - pr) / 2.5, x1s=1.0 - omega, twall=twall, dt=dt, if_viscous=1,
                scheme=capi.SCHEME_NND, local_dt=1, cfl=20.0, dtmax=10.0, dtmin=1e-7, kstep_show=50)


# ---------------------------------------------------------------------------
# a case = mesh + BCs + velocity grid + constants
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class Case:
    name: str
    mesh: Mesh
    bcs: List[BC]
    vi: np.ndarray
    wi: np.ndarray
    vj: np.ndarray
    wj: np.ndarray
    const: dict
    init: str = "freestream"  # or "rest"

    @property
    def nx(self):
        return self.mesh.nx

    @property
    def ny(self):
        return self.mesh.ny

    @property
    def nvi(self):
        return len(self.vi)

    @property
    def nvj(self):
        return len(self.vj)

    @property
    def ncell(self):
        return (self.nx - 1) * (self.ny - 1)

    @property
    def nv(self):
        return self.nvi * self.nvj


def shipped_case(input_dir: str | pathlib.Path, name: str = "shipped") -> Case:
    """A case read from the reference's own input files."""
    d = pathlib.Path(input_dir)
    ctl = read_control_in(d / "control.in")
    xn, yn = read_generic_dat(d / "generic.dat")
    nx, ny = xn.shape
    bcs = read_generic_inp(d / "generic.inp", nx, ny)
    mesh = build_mesh(xn, yn, bcs, ctl["Lref"])
    q = read_quadrature_para(d / "Quadrature_Para.dat")
    if q["mode"] == 1:
        vi, wi = gh_velocity_grid(q["ngh"])
        vj, wj = vi.copy(), wi.copy()
    elif q["mode"] == 2:
        vi, wi = gl_velocity_grid(q["dvx"], q["nvi"], q["vxdown"], q["kgln"])
        vj, wj = gl_velocity_grid(q["dvy"], q["nvj"], q["vydown"], q["kgln"])
    else:
        raise NotImplementedError("Newton-Cotes input (mode 3) stops in the reference (src/GKUA_DETFP.f90:40-43)")
    return Case(name, mesh, bcs, vi, wi, vj, wj, constants_from_control(ctl), "freestream")


def cavity_case(n: int, nv: int, kn: float, lid: float = 0.15, vmax: float = 6.0, grid: str = "sinh",
                dt: float = 0.1, name: Optional[str] = None) -> Case:
    """Lid-driven cavity on [0,1]^2 with n x n uniform cells, diffuse walls at
    T=1 and the lid (j = ny) moving with velocity ``lid`` (sqrt(2RT) units)."""
    nx = ny = n + 1
    s = np.linspace(0.0, 1.0, n + 1)
    xn, yn = np.meshgrid(s, s, indexing="ij")
    bcs = [BC(2, 1, nx, 1, 1, capi.BC_WALL), BC(3, nx, nx, 1, ny, capi.BC_WALL),
           BC(4, 1, nx, ny, ny, capi.BC_WALL, lid, 0.0), BC(1, 1, 1, 1, ny, capi.BC_WALL)]
    mesh = build_mesh(xn, yn, bcs)
    if grid == "sinh":
        v, w = sinh_velocity_grid(nv, vmax)
    elif grid == "gh":
        v, w = gh_velocity_grid(nv // 2)  # nv = 2 NV nodes per direction
    else:
        v, w = nc_velocity_grid(nv, vmax)
    const = constants_from_kn(kn, dt=dt)
    return Case(name or f"cavity{n}_v{nv}_kn{kn}", mesh, bcs, v, w, v.copy(), w.copy(), const, "rest")


def couette_case(ny_cells: int = 128, nx_cells: int = 4, nv: int = 32, kn: float = 0.1, uwall: float = 0.1,
                 beta: float = 2.0, vmax: float = 6.0, dt: float = 0.1, name: Optional[str] = None) -> Case:
    """Planar Couette flow: x periodic over nx_cells cells of width 1/128,
    tanh-clustered y in [0,1], diffuse walls at y=0 (u=-uwall) and y=1 (u=+uwall)."""
    nx, ny = nx_cells + 1, ny_cells + 1
    h = 1.0 / 128.0
    xs = np.arange(nx) * h
    jj = np.arange(ny) / ny_cells
    ys = 0.5 * (1.0 + np.tanh(beta * (2.0 * jj - 1.0)) / math.tanh(beta))
    xn, yn = np.meshgrid(xs, ys, indexing="ij")
    bcs = [BC(2, 1, nx, 1, 1, capi.BC_WALL, -uwall, 0.0), BC(3, nx, nx, 1, ny, capi.BC_PERIODIC),
           BC(4, 1, nx, ny, ny, capi.BC_WALL, uwall, 0.0), BC(1, 1, 1, 1, ny, capi.BC_PERIODIC)]
    mesh = build_mesh(xn, yn, bcs)
    v, w = nc_velocity_grid(nv, vmax)
    const = constants_from_kn(kn, dt=dt)
    return Case(name or f"couette{ny_cells}_v{nv}_kn{kn}", mesh, bcs, v, w, v.copy(), w.copy(), const, "rest")


# ---------------------------------------------------------------------------
# parameters struct and initial state
# ---------------------------------------------------------------------------
def _dp(a: np.ndarray):
    import ctypes
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class ParamPack:
    """Holds a filled GsisParams and keeps every array it points to alive."""

    def __init__(self, case: Case, precision: int = capi.PREC_FP64, mode: int = capi.MODE_M0,
                 ns_tol: float = float(F32(1e-6)), ns_maxit: int = 0, kstep_show: int = 0,
                 split_kernels: int = 0, walk_tiles: int = 0):
        import ctypes
        m, c = case.mesh, case.const
        self.arrays = [np.ascontiguousarray(a, dtype=np.float64) for a in (case.vi, case.vj, case.wi, case.wj)]
        self.farrays = [np.asfortranarray(a, dtype=np.float64) for a in
                        (m.x, m.y, m.vol, m.si, m.sj, m.ni1, m.ni2, m.nj1, m.nj2)]
        self.bc_arr = (capi.GsisBC * max(1, len(case.bcs)))()
        for k, b in enumerate(case.bcs):
            self.bc_arr[k] = capi.GsisBC(b.face, b.ist, b.iend, b.jst, b.jend, b.code, b.uw, b.vw)
        p = capi.GsisParams()
        p.nx, p.ny, p.nvi, p.nvj = case.nx, case.ny, case.nvi, case.nvj
        p.vi, p.vj, p.wi, p.wj = (_dp(a) for a in self.arrays)
        (p.x, p.y, p.vol, p.si, p.sj, p.ni1, p.ni2, p.nj1, p.nj2) = (_dp(a) for a in self.farrays)
        p.nbc = len(case.bcs)
        p.bc = ctypes.cast(self.bc_arr, ctypes.POINTER(capi.GsisBC))
        for k in ("cfc", "x1s", "cpr", "twall", "dt", "ma", "gamma", "re", "pr", "cv", "cp", "aoa"):
            setattr(p, k, float(c[k]))
        p.scheme = int(c.get("scheme", capi.SCHEME_NND))
        p.mode = int(mode)
        p.precision = int(precision)
        p.if_viscous = int(c.get("if_viscous", 1))
        p.local_dt = int(c.get("local_dt", 1))
        p.cfl = float(c.get("cfl", 20.0))
        p.dtmax = float(c.get("dtmax", 10.0))
        p.dtmin = float(c.get("dtmin", 1e-7))
        p.ns_tol = float(ns_tol)
        p.ns_maxit = int(ns_maxit)
        p.kstep_show = int(kstep_show)
        p.split_kernels = int(split_kernels)
        p.walk_tiles = int(walk_tiles)
        self.p = p


def frd_shape(case: Case):
    return (2, case.nx + 7, case.ny + 7, case.nvi, case.nvj)


def initial_state(case: Case, with_frd: bool = True):
    """init_flow_zero (src/sub_init.f90:400-457) + the Maxwellian fill of
    GKUA_sudufenbu (src/GKUA_sudufenbu.f90:22-46).  Returns Fortran-layout
    arrays (frd, s, cf, u).  ``case.init == "rest"`` starts from the gas at
    rest (n=T=1, u=0).  The initial state is uniform, so with
    ``with_frd=False`` the first item is instead the pair (g, h) of
    per-velocity values that every cell of FRD holds."""
    c = case.const
    nx, ny = case.nx, case.ny
    g, ma, aoa = c["gamma"], c["ma"], c["aoa"]
    d0 = 1.0
    if case.init == "rest":
        u0 = v0 = 0.0
    else:
        u0, v0 = 1.0 * math.cos(aoa), 1.0 * math.sin(aoa)
    p0 = 1.0 / (g * ma * ma)
    u = np.zeros((4, nx + 3, ny + 3), order="F")
    u[0] = d0
    u[1] = d0 * u0
    u[2] = d0 * v0
    u[3] = p0 / (g - 1.0) + 0.5 * d0 * (u0 * u0 + v0 * v0)
    s = np.zeros((11, nx + 7, ny + 7), order="F")
    sl = (slice(2, nx + 5), slice(2, ny + 5))  # cells -1..n+1
    n_, uk, vk, tk = d0, u0 * math.sqrt(g / 2.0) * ma, v0 * math.sqrt(g / 2.0) * ma, 1.0
    s[0][sl] = n_
    s[1][sl] = uk
    s[2][sl] = vk
    s[3][sl] = tk
    cf = np.zeros((nx + 3, ny + 3), order="F")
    if c["x1s"] <= float(F32(1e-8)):
        cf[:, :] = c["cfc"] * n_
    else:
        cf[:, :] = c["cfc"] * n_ * tk ** c["x1s"]
    crt = n_ / (math.pi * tk)
    vi, vj = case.vi, case.vj
    viu = vi - uk
    vjv = vj - vk
    vij2 = (viu[:, None] ** 2 + vjv[None, :] ** 2) / tk
    gdv = crt * np.exp(-vij2)  # (nvi, nvj)
    if not with_frd:
        # the per-velocity values every cell -1..n+1 holds (velocity iv fastest)
        return (gdv.ravel(order="F").copy(), (tk * gdv / 2.0).ravel(order="F").copy()), s, cf, u
    frd = np.zeros(frd_shape(case), order="F")
    frd[0, 2:nx + 5, 2:ny + 5, :, :] = gdv[None, None, :, :]
    frd[1, 2:nx + 5, 2:ny + 5, :, :] = (tk * gdv / 2.0)[None, None, :, :]
    return frd, s, cf, u


def ns_boundary_init(case: Case, u: np.ndarray, swall: np.ndarray) -> np.ndarray:
    """NS ghost cells at initialisation, Boundary_condition_onemesh
    (src/sub_boundary.f90:10-39): far field = isothermal wall fed by the
    kinetic wall moments Swall (boundary_Farfield GKUA branch, :204-229),
    periodic "inlet" (:248-298) and the symmetry mirror
    (Symmetry_or_slidewall, :410-471).  ``u`` is (4,-1:nx+1,-1:ny+1) Fortran."""
    c = case.const
    g, ma = c["gamma"], c["ma"]
    nx, ny = case.nx, case.ny
    sq = math.sqrt(2.0 / g)
    u = u.copy(order="F")

    def U(m, i, j):
        return u[m, i + 1, j + 1]

    soff = 0
    for b in case.bcs:
        ib, ie, jb, je = b.ist, b.iend, b.jst, b.jend
        icheck, jcheck = max(ib, ie), max(jb, je)
        nk = (je - jb + 1) if b.face in (1, 3) else (ie - ib + 1)
        if b.code == capi.BC_FARFIELD and c["twall"] > 0:
            sw = swall[soff:soff + 10 * nk].reshape((10, nk), order="F")
            for j in range(jb, je + 1):
                for i in range(ib, ie + 1):
                    if (jb == je and i == icheck) or (ib == ie and j == jcheck):
                        continue
                    kk = (j - jb) if b.face in (1, 3) else (i - ib)
                    dwall = sw[0, kk]
                    uwall = sw[1, kk] / ma * sq
                    vwall = sw[2, kk] / ma * sq
                    twl = sw[3, kk]
                    for k in range(1, LAP + 1):
                        if b.face == 1:
                            i1, j1, i2, j2 = i + k - 1, j, i - k, j
                        elif b.face == 2:
                            i1, j1, i2, j2 = i, j + k - 1, i, j - k
                        elif b.face == 3:
                            i1, j1, i2, j2 = i - k, j, i + k - 1, j
                        else:
                            i1, j1, i2, j2 = i, j - k, i, j + k - 1
                        d1 = U(0, i1, j1)
                        u1 = U(1, i1, j1) / d1
                        v1 = U(2, i1, j1) / d1
                        p1 = (U(3, i1, j1) - 0.5 * d1 * (u1 * u1 + v1 * v1)) * (g - 1.0)
                        t1 = g * ma * ma * p1 / d1
                        t2 = 2.0 * twl - t1
                        if t2 <= 0.0:
                            t2 = twl
                        d2 = -d1 + 2.0 * dwall
                        p2 = d2 * t2 / (g * ma * ma)
                        u2 = -u1 + 2.0 * uwall
                        v2 = -v1 + 2.0 * vwall
                        u[:, i2 + 1, j2 + 1] = (d2, d2 * u2, d2 * v2, p2 / (g - 1.0) + 0.5 * d2 * (u2 * u2 + v2 * v2))
        elif b.code == capi.BC_PERIODIC:
            for j in range(jb, je + 1):
                for i in range(ib, ie + 1):
                    if (jb == je and i == icheck) or (ib == ie and j == jcheck):
                        continue
                    for k in range(1, LAP + 1):
                        if b.face == 1:
                            i1, j1, i2, j2 = nx - k, j, i - k, j
                        elif b.face == 2:
                            i1, j1, i2, j2 = i, ny - k, i, j - k
                        elif b.face == 3:
                            i1, j1, i2, j2 = k, j, i + k - 1, j
                        else:
                            i1, j1, i2, j2 = i, k, i, j + k - 1
                        u[:, i2 + 1, j2 + 1] = u[:, i1 + 1, j1 + 1]
        elif b.code == capi.BC_SYMMETRY:
            # every node of the range, the last one included; the face normal
            # at a node comes from the edge to the next node (:433-444)
            xg, yg = case.mesh.x, case.mesh.y
            for j in range(jb, je + 1):
                for i in range(ib, ie + 1):
                    if b.face in (1, 3):
                        dx = xg[i + 1, j + 2] - xg[i + 1, j + 1]
                        dy = yg[i + 1, j + 2] - yg[i + 1, j + 1]
                        si = math.sqrt(dx * dx + dy * dy)
                        n1, n2 = dy / si, -dx / si
                    else:
                        dx = xg[i + 2, j + 1] - xg[i + 1, j + 1]
                        dy = yg[i + 2, j + 1] - yg[i + 1, j + 1]
                        si = math.sqrt(dx * dx + dy * dy)
                        n1, n2 = -dy / si, dx / si
                    for k in range(1, LAP + 1):
                        if b.face == 1:
                            i1, j1, i2, j2 = i + k - 1, j, i - k, j
                        elif b.face == 2:
                            i1, j1, i2, j2 = i, j + k - 1, i, j - k
                        elif b.face == 3:
                            i1, j1, i2, j2 = i - k, j, i + k - 1, j
                        else:
                            i1, j1, i2, j2 = i, j - k, i, j + k - 1
                        vn = U(1, i1, j1) * n1 + U(2, i1, j1) * n2
                        u[0, i2 + 1, j2 + 1] = U(0, i1, j1)
                        u[1, i2 + 1, j2 + 1] = U(1, i1, j1) - 2.0 * vn * n1
                        u[2, i2 + 1, j2 + 1] = U(2, i1, j1) - 2.0 * vn * n2
                        u[3, i2 + 1, j2 + 1] = U(3, i1, j1)
        if b.code in (capi.BC_WALL, capi.BC_FARFIELD):
            soff += 10 * nk
    return u


def swall_len(case: Case) -> int:
    n = 0
    for b in case.bcs:
        if b.code in (capi.BC_WALL, capi.BC_FARFIELD):
            n += 10 * ((b.jend - b.jst + 1) if b.face in (1, 3) else (b.iend - b.ist + 1))
    return n
