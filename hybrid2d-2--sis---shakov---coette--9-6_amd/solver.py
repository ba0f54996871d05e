"""High-level driver over the C-ABI: the loop of ``program main``
(src/opencfd_ec2d.f90:45-56) around Hybrid_Time_advance_LU_SGS, with the
residual monitor every Kstep_show steps and a convergence stop."""
from __future__ import annotations

import ctypes

import numpy as np

from . import capi, restart
from .cases import Case, ParamPack, frd_shape, initial_state, ns_boundary_init, swall_len


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class GsisSolver:
    """One GSIS case resident on the GPU."""

    def __init__(self, case: Case, precision: int = capi.PREC_FP64, mode: int = capi.MODE_M0, **kw):
        self.case = case
        self.lib = capi.load_library()
        self.pack = ParamPack(case, precision=precision, mode=mode, **kw)
        self.mode = mode
        self.kstep_show = int(kw.get("kstep_show", 0))
        h = ctypes.c_void_p()
        capi.check(self.lib, self.lib.gsis_create(ctypes.byref(self.pack.p), ctypes.byref(h)))
        self.h = h

    # -- state transfer ---------------------------------------------------
    def upload(self, frd=None, s=None, cf=None, u=None, u0=None):
        if frd is not None:
            frd = np.asfortranarray(frd, dtype=np.float64)
            capi.check(self.lib, self.lib.gsis_upload_frd(self.h, _dp(frd)))
        arrs = [None if a is None else np.asfortranarray(a, dtype=np.float64) for a in (s, cf, u, u0)]
        if any(a is not None for a in arrs):
            ptrs = [None if a is None else _dp(a) for a in arrs]
            capi.check(self.lib, self.lib.gsis_upload_macro(self.h, *ptrs))

    def refresh_moments(self):
        capi.check(self.lib, self.lib.gsis_refresh_moments(self.h))

    def initialize(self, device_fill: bool = True):
        """Initial state as init_flow_zero + GKUA_sudufenbu + the NS ghost BCs.
        The initial distribution is the same Maxwellian in every cell: with
        ``device_fill`` the device writes it (gsis_fill_frd_uniform) instead of
        uploading the full FRD array."""
        gh, s, cf, u = initial_state(self.case, with_frd=not device_fill)
        if device_fill:
            g, hv = (np.ascontiguousarray(a, dtype=np.float64) for a in gh)
            capi.check(self.lib, self.lib.gsis_fill_frd_uniform(self.h, _dp(g), _dp(hv)))
            self.upload(s=s, cf=cf, u=u, u0=u if self.mode == capi.MODE_M1 else None)
        else:
            self.upload(frd=gh, s=s, cf=cf, u=u, u0=u if self.mode == capi.MODE_M1 else None)
        del gh
        self.refresh_moments()
        m = self.macro()
        u = ns_boundary_init(self.case, m["u"], m["swall"])
        self.upload(u=u, u0=u if self.mode == capi.MODE_M1 else None)

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
            "swall": np.zeros(max(1, swall_len(c))),
        }
        keys = ("s", "s0", "cf", "u", "taoi", "taoj", "qi", "qj", "swall")
        capi.check(self.lib, self.lib.gsis_download_macro(self.h, *[_dp(out[k]) for k in keys]))
        return out

    def frd(self) -> np.ndarray:
        a = np.zeros(frd_shape(self.case), order="F")
        capi.check(self.lib, self.lib.gsis_download_frd(self.h, _dp(a)))
        return a

    # -- restart (src/Backup_write_read.f90) -------------------------------
    def save_restart(self, root, flag: int):
        """backup_Macro_write(flag): the distribution and flag.dat under
        ``root``/restart{flag}."""
        k, tt, _ = self.counters()
        return restart.write_restart(root, flag, self.frd(), tt, k)

    def resume(self, root) -> int:
        """init_flow_read: the distribution of the newest valid backup slot,
        the kinetic BCs and moments of it, and the saved step counter.  The
        macroscopic state (S, CF, U) is whatever the handle holds, as in the
        reference, where only FRD is saved.  Returns the slot read."""
        frd, tt, kstep, slot = restart.read_restart(root, frd_shape(self.case))
        self.upload(frd=frd)
        self.refresh_moments()
        self.set_counters(kstep, tt)
        return slot

    def set_mass_correct(self, on: bool = True) -> None:
        """MassCorrect (src/GKUA_DETFP.f90:153-172) after the moments of every
        later step, Init_mass taken from the current state."""
        capi.check(self.lib, self.lib.gsis_set_mass_correct(self.h, int(bool(on))))

    def set_counters(self, kstep: int, tt: float) -> None:
        capi.check(self.lib, self.lib.gsis_set_counters(self.h, int(kstep), float(tt)))

    def write_output(self, directory, hybrid: bool = False):
        """The reference's Tecplot files (output, src/IO.f90:307-553) from the
        current fields."""
        from . import output
        return output.write_output(directory, [(self.case, self.macro())], hybrid)

    # -- inner faces between blocks (src/GKUA_boundary.f90:400-633) --------
    def inner_exchange(self, ksub: int, src: "GsisSolver", msub: int, orient: int) -> None:
        """Fill the ghost layers of subface ksub (1-based) from the cells next to
        subface msub of ``src`` (another handle, or this one)."""
        capi.check(self.lib, self.lib.gsis_inner_exchange(self.h, int(ksub), src.h, int(msub), int(orient)))

    # -- stepping ---------------------------------------------------------
    def step(self, n: int = 1):
        capi.check(self.lib, self.lib.gsis_step(self.h, int(n)))

    def residual(self) -> np.ndarray:
        r = np.zeros(4)
        capi.check(self.lib, self.lib.gsis_residual(self.h, _dp(r)))
        return r

    def counters(self):
        k = ctypes.c_int32()
        t = ctypes.c_double()
        it = ctypes.c_int32()
        capi.check(self.lib, self.lib.gsis_counters(self.h, ctypes.byref(k), ctypes.byref(t), ctypes.byref(it)))
        return k.value, t.value, it.value

    def run(self, max_steps: int, tol: float = 1e-8, show: int = 50):
        """Step until all four RMS residuals are below ``tol`` (checked every
        ``show`` steps, as DETFP_res is called every Kstep_show steps).
        Returns (steps, history[(kstep, res[4])])."""
        if self.kstep_show and show % self.kstep_show:
            raise ValueError("show must be a multiple of the handle's kstep_show")
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

    def timing(self, enable: int = -1):
        """Kernel timings (ms summed over timed steps) of [rhs, sweep, moments,
        synthetic NS solve]."""
        ms = np.zeros(4)
        n = ctypes.c_int32()
        capi.check(self.lib, self.lib.gsis_timing(self.h, int(enable), _dp(ms), ctypes.byref(n)))
        return ms, n.value

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.gsis_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
