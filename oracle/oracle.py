"""ctypes wrapper of the CPU oracle (gsis_oracle.c).  TEST INFRASTRUCTURE ONLY.

Loaded by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg as
the reference values; the product path (the GsisSolver over libgsis.so) never imports
it.  OracleSolver offers the same methods as the product's GsisSolver so that
both can be driven from the same inputs.
"""
from __future__ import annotations

import ctypes
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB = HERE / "libgsis_oracle.so"
SOURCES = [HERE / "gsis_oracle.c", HERE / "gsis_oracle_ns.c"]


def build(force: bool = False) -> pathlib.Path:
    """Compile the oracle with gcc (FP contraction off, like the reference's
    separate multiply/add)."""
    if force or not LIB.exists() or any(s.stat().st_mtime > LIB.stat().st_mtime for s in SOURCES):
        cmd = ["gcc", "-O2", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared", "-o", str(LIB)]
        cmd += [str(s) for s in SOURCES] + ["-lm"]
        subprocess.run(cmd, check=True)
    return LIB


_lib = None


def load():
    global _lib
    if _lib is None:
        build()
        lib = ctypes.CDLL(str(LIB))
        dp = ctypes.POINTER(ctypes.c_double)
        vp = ctypes.c_void_p
        lib.orc_create.restype = vp
        lib.orc_create.argtypes = [ctypes.c_void_p]
        lib.orc_destroy.argtypes = [vp]
        lib.orc_frd_len.restype = ctypes.c_size_t
        lib.orc_frd_len.argtypes = [vp]
        lib.orc_set_frd.argtypes = [vp, dp]
        lib.orc_get_frd.argtypes = [vp, dp]
        lib.orc_get_frd00.argtypes = [vp, dp]
        lib.orc_get_frd0.argtypes = [vp, dp]
        lib.orc_get_prims.argtypes = [vp, dp]
        lib.orc_feq.argtypes = [vp]
        lib.orc_van_leer.argtypes = [dp, dp, ctypes.c_double, dp]
        lib.orc_set_macro.argtypes = [vp, dp, dp, dp, dp]
        lib.orc_get_macro.argtypes = [vp] + [dp] * 9
        lib.orc_counters.argtypes = [vp, ctypes.POINTER(ctypes.c_int), dp, ctypes.POINTER(ctypes.c_int)]
        lib.orc_set_counters.argtypes = [vp, ctypes.c_int, ctypes.c_double]
        lib.orc_set_counters.restype = None
        lib.orc_phy_boundary.argtypes = [vp]
        lib.orc_detfp.argtypes = [vp]
        lib.orc_convert_ns.argtypes = [vp]
        lib.orc_convert_ns.restype = ctypes.c_int
        lib.orc_step.argtypes = [vp, ctypes.c_int]
        lib.orc_step.restype = ctypes.c_int
        lib.orc_residual.argtypes = [vp, dp, dp]
        lib.orc_residual.restype = ctypes.c_int
        lib.orc_out_mass.argtypes = [vp]
        lib.orc_out_mass.restype = ctypes.c_double
        lib.orc_set_mass_correct.argtypes = [vp, ctypes.c_int]
        lib.orc_set_mass_correct.restype = None
        lib.orc_ns_advance.argtypes = [vp, ctypes.c_int]
        lib.orc_ns_advance.restype = ctypes.c_double
        lib.orc_ns_work.argtypes = [vp, dp, dp, dp]
        lib.orc_ns_work.restype = None
        _lib = lib
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def van_leer(ql, qr, gamma):
    """Van Leer flux of the oracle for rotated states (d, un, ut, p)."""
    lib = load()
    a = np.ascontiguousarray(ql, dtype=np.float64)
    b = np.ascontiguousarray(qr, dtype=np.float64)
    f = np.zeros(4)
    lib.orc_van_leer(_dp(a), _dp(b), float(gamma), _dp(f))
    return f


# ---------------------------------------------------------------------------
# inner faces between blocks, on Fortran-layout arrays
# ---------------------------------------------------------------------------
LAP = 2  # COMMONDATA.f90:20


def _inner_src_cell(b, m: int, l: int):
    """Neighbour cell that fills buffer position m (1..n1), layer l (1..LAP):
    the Ua fill of GKUA_inner_mesco_boundary (src/GKUA_boundary.f90:437-467)."""
    if b.face == 1:
        return b.ist + l - 1, b.jst + m - 1
    if b.face == 2:
        return b.ist + m - 1, b.jst + l - 1
    if b.face == 3:
        return b.iend - l, b.jst + m - 1
    return b.ist + m - 1, b.jend - l


def _inner_ghost_cell(b, k: int, l: int):
    """Ghost cell written from buffer position k, layer l (src/GKUA_boundary.f90:470-494)."""
    if b.face == 1:
        return b.ist - l, b.jst + k - 1
    if b.face == 2:
        return b.ist + k - 1, b.jst - l
    if b.face == 3:
        return b.iend + l - 1, b.jst + k - 1
    return b.ist + k - 1, b.jend + l - 1


def inner_exchange_arrays(frd_d, s_d, bd, frd_s, s_s, bs, orient: int) -> None:
    """GKUA_inner_mesco_boundary + GKUA_inner_macro_boundary
    (src/GKUA_boundary.f90:400-633) for one subface, in place on frd_d, s_d:
    FRD(2,-3:nx+3,-3:ny+3,nvi,nvj) and S_GKUA(11,-3:nx+3,-3:ny+3) of the
    receiving block from those of the neighbour.  n1 is the receiving face's
    cell count; orient == 2 keeps the direction, anything else reverses it."""
    n1 = bd.iend - bd.ist + bd.jend - bd.jst
    if n1 != bs.iend - bs.ist + bs.jend - bs.jst:
        raise ValueError("the two subfaces have different cell counts")
    for k in range(1, n1 + 1):
        k1 = k if orient == 2 else n1 + 1 - k
        for l in range(1, LAP + 1):
            ia, ja = _inner_src_cell(bs, k1, l)
            ib, jb = _inner_ghost_cell(bd, k, l)
            frd_d[:, ib + 3, jb + 3] = frd_s[:, ia + 3, ja + 3]
            s_d[0:10, ib + 3, jb + 3] = s_s[0:10, ia + 3, ja + 3]


class OracleError(RuntimeError):
    pass


class OracleSolver:
    """CPU restatement of one GSIS case (same interface as GsisSolver)."""

    def __init__(self, case, mode: int = 0, **kw):
        from importlib import import_module

        self.case = case
        pkg = import_module(type(case).__module__.rsplit(".", 1)[0])
        self.cases = import_module(pkg.__name__ + ".cases")
        self.pack = self.cases.ParamPack(case, precision=0, mode=mode, **kw)
        self.mode = mode
        self.lib = load()
        self.h = self.lib.orc_create(ctypes.addressof(self.pack.p))

    def upload(self, frd=None, s=None, cf=None, u=None, u0=None):
        if frd is not None:
            frd = np.asfortranarray(frd, dtype=np.float64)
            self.lib.orc_set_frd(self.h, _dp(frd))
        arrs = [None if a is None else np.asfortranarray(a, dtype=np.float64) for a in (s, cf, u, u0)]
        if any(a is not None for a in arrs):
            self.lib.orc_set_macro(self.h, *[None if a is None else _dp(a) for a in arrs])

    def refresh_moments(self):
        self.lib.orc_phy_boundary(self.h)
        self.lib.orc_detfp(self.h)

    def initialize(self):
        frd, s, cf, u = self.cases.initial_state(self.case)
        self.upload(frd=frd, s=s, cf=cf, u=u, u0=u if self.mode == 1 else None)
        del frd
        self.refresh_moments()
        m = self.macro()
        u = self.cases.ns_boundary_init(self.case, m["u"], m["swall"])
        self.upload(u=u, u0=u if self.mode == 1 else None)

    def macro(self) -> dict:
        c = self.case
        nx, ny = c.nx, c.ny
        out = {
            "s": np.zeros((11, nx + 7, ny + 7), order="F"),
            "s0": np.zeros((11, nx + 7, ny + 7), order="F"),
            "cf": np.zeros((nx + 3, ny + 3), order="F"),
            "u": np.zeros((4, nx + 3, ny + 3), order="F"),
            "taoi": np.zeros((4, nx + 3, ny + 3), order="F"),
            "taoj": np.zeros((4, nx + 3, ny + 3), order="F"),
            "qi": np.zeros((2, nx + 3, ny + 3), order="F"),
            "qj": np.zeros((2, nx + 3, ny + 3), order="F"),
            "swall": np.zeros(max(1, self.cases.swall_len(c))),
        }
        keys = ("s", "s0", "cf", "u", "taoi", "taoj", "qi", "qj", "swall")
        self.lib.orc_get_macro(self.h, *[_dp(out[k]) for k in keys])
        return out

    def frd(self) -> np.ndarray:
        a = np.zeros(self.cases.frd_shape(self.case), order="F")
        self.lib.orc_get_frd(self.h, _dp(a))
        return a

    def frd00(self) -> np.ndarray:
        a = np.zeros(self.cases.frd_shape(self.case), order="F")
        self.lib.orc_get_frd00(self.h, _dp(a))
        return a

    def feq(self) -> np.ndarray:
        """Shakhov equilibrium of the current S on interior cells (FRD0)."""
        self.lib.orc_feq(self.h)
        a = np.zeros(self.cases.frd_shape(self.case), order="F")
        self.lib.orc_get_frd0(self.h, _dp(a))
        return a

    def prims(self) -> np.ndarray:
        c = self.case
        a = np.zeros((5, c.nx + 3, c.ny + 3), order="F")
        self.lib.orc_get_prims(self.h, _dp(a))
        return a

    def inner_exchange(self, ksub: int, src: "OracleSolver", msub: int, orient: int) -> None:
        """The exchange of subface ksub (1-based) from subface msub of ``src``."""
        if (src.case.nvi, src.case.nvj) != (self.case.nvi, self.case.nvj):
            raise ValueError("the two blocks have different velocity grids")
        frd_d, s_d = self.frd(), self.macro()["s"]
        if src is self:
            frd_s, s_s = frd_d, s_d
        else:
            frd_s, s_s = src.frd(), src.macro()["s"]
        inner_exchange_arrays(frd_d, s_d, self.case.bcs[ksub - 1], frd_s, s_s, src.case.bcs[msub - 1], orient)
        self.upload(frd=frd_d, s=s_d)

    def convert_ns(self) -> int:
        return self.lib.orc_convert_ns(self.h)

    def ns_advance(self, n: int = 1) -> float:
        """n inner iterations of the synthetic NS solve on the current U and
        HoT fields; the max |Res_4| of the last one (negative on T < 0.1)."""
        return self.lib.orc_ns_advance(self.h, int(n))

    def ns_work(self) -> dict:
        """Res, dU, Lci, Lcj, Lvi, Lvj and dt of the last inner iteration."""
        c = self.case
        sh = (c.nx + 3, c.ny + 3)
        res = np.zeros((4,) + sh, order="F")
        du = np.zeros((4,) + sh, order="F")
        lc = np.zeros((5,) + sh, order="C")
        self.lib.orc_ns_work(self.h, _dp(res), _dp(du), _dp(lc))
        # each of the five is a Fortran (-1:nx+1,-1:ny+1) array
        lc = np.stack([np.asfortranarray(lc[k].ravel().reshape(sh, order="F")) for k in range(5)])
        return {"res": res, "du": du, "lci": lc[0], "lcj": lc[1], "lvi": lc[2], "lvj": lc[3], "dt": lc[4]}

    def out_mass(self) -> float:
        return float(self.lib.orc_out_mass(self.h))

    def set_mass_correct(self, on: bool = True):
        self.lib.orc_set_mass_correct(self.h, int(bool(on)))

    def step(self, n: int = 1):
        rc = self.lib.orc_step(self.h, int(n))
        if rc:
            raise OracleError(f"oracle step failed with code {rc}")

    def residual(self, with_max: bool = False):
        r = np.zeros(4)
        m = np.zeros(4)
        rc = self.lib.orc_residual(self.h, _dp(r), _dp(m))
        if rc:
            raise OracleError(f"oracle residual code {rc}")
        return (r, m) if with_max else r

    def counters(self):
        k = ctypes.c_int()
        t = ctypes.c_double()
        it = ctypes.c_int()
        self.lib.orc_counters(self.h, ctypes.byref(k), ctypes.byref(t), ctypes.byref(it))
        return k.value, t.value, it.value

    def set_counters(self, kstep: int, tt: float) -> None:
        self.lib.orc_set_counters(self.h, int(kstep), float(tt))

    def run(self, max_steps: int, tol: float = 1e-8, show: int = 50):
        hist = []
        done = 0
        while done < max_steps:
            n = min(show, max_steps - done)
            self.step(n)
            done += n
            r = self.residual()
            hist.append((done, r.copy()))
            if np.all(r < tol):
                break
        return done, hist

    def close(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
