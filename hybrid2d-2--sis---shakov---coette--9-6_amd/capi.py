"""ctypes mirror of include/gsis_capi.h and the loader of the HIP library.

The C-ABI stands in for the body of ``Hybrid_Time_advance_LU_SGS``
(src/Hybrid_Sub.f90:4-160) and for ``DETFP_res`` (src/GKUA_DETFP_RES.f90:4-60).
This module only declares the structs and function prototypes; it never runs a
CPU fallback: if ``libgsis.so`` is missing the load fails loudly.
"""
from __future__ import annotations

import ctypes
import os
import pathlib

PKG_DIR = pathlib.Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libgsis.so"

GSIS_OK = 0
GSIS_EARG = 1
GSIS_EHIP = 2
GSIS_ENAN = 3
GSIS_EUNSUPPORTED = 4
GSIS_ETEMP = 5

BC_WALL = -2
BC_FARFIELD = -4
BC_PERIODIC = -5
BC_SYMMETRY = -3  # passive in the kinetic step; a code >= 1 is an inner face (neighbour block number)

SCHEME_UP1 = 0
SCHEME_NND = 1
SCHEME_WENO3 = 3

MODE_M0 = 0
MODE_M1 = 1

PREC_FP64 = 0
PREC_FP32 = 1
# kinetic kernels (gsis_params.split_kernels)
KERNEL_AUTO = 0      # one-pass column-wavefront kernel, 8 columns per workgroup
KERNEL_SPLIT = 1     # separate RHS and LU-SGS sweep kernels
KERNEL_TILES = 2     # one-pass velocity-tile walker
KERNEL_STRIP16 = 3   # column-wavefront kernel, 16 columns per workgroup

c_dp = ctypes.POINTER(ctypes.c_double)

EXPORTED = (
    "gsis_create",
    "gsis_upload_frd",
    "gsis_upload_macro",
    "gsis_refresh_moments",
    "gsis_step",
    "gsis_residual",
    "gsis_download_macro",
    "gsis_download_frd",
    "gsis_download_frd_noalloc",
    "gsis_counters",
    "gsis_set_counters",
    "gsis_inner_exchange",
    "gsis_timing",
    "gsis_destroy",
    "gsis_last_error",
    "gsis_set_mass_correct",
    "gsis_closure_samples",
    "gsis_set_closure",
    "gsis_set_closure_warmup",
    "gsis_closure_stats",
    "gsis_mlp_create",
    "gsis_mlp_nparams",
    "gsis_mlp_set_params",
    "gsis_mlp_get_params",
    "gsis_mlp_set_scaling",
    "gsis_mlp_forward",
    "gsis_mlp_loss_grad",
    "gsis_mlp_train",
    "gsis_mlp_destroy",
    "gsis_fill_frd_uniform",
    "gsis_batch_create",
    "gsis_batch_step",
    "gsis_batch_tune",
    "gsis_batch_set_ns_graph",
    "gsis_batch_ns_graph",
    "gsis_batch_destroy",
)

MLP_NIN = 8
MLP_NOUT = 5
MLP_WIDTH = 128
MLP_HIDDEN = 4


class GsisBC(ctypes.Structure):
    """gsis_bc: one boundary subface (BC_MSG_TYPE, src/COMMONDATA.f90:34-41)."""

    _fields_ = [
        ("face", ctypes.c_int32),
        ("ist", ctypes.c_int32),
        ("iend", ctypes.c_int32),
        ("jst", ctypes.c_int32),
        ("jend", ctypes.c_int32),
        ("code", ctypes.c_int32),
        ("uw", ctypes.c_double),
        ("vw", ctypes.c_double),
    ]


class GsisParams(ctypes.Structure):
    """gsis_params (field order and types identical to include/gsis_capi.h)."""

    _fields_ = [
        ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32),
        ("nvi", ctypes.c_int32),
        ("nvj", ctypes.c_int32),
        ("vi", c_dp),
        ("vj", c_dp),
        ("wi", c_dp),
        ("wj", c_dp),
        ("x", c_dp),
        ("y", c_dp),
        ("vol", c_dp),
        ("si", c_dp),
        ("sj", c_dp),
        ("ni1", c_dp),
        ("ni2", c_dp),
        ("nj1", c_dp),
        ("nj2", c_dp),
        ("nbc", ctypes.c_int32),
        ("bc", ctypes.POINTER(GsisBC)),
        ("cfc", ctypes.c_double),
        ("x1s", ctypes.c_double),
        ("cpr", ctypes.c_double),
        ("twall", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("ma", ctypes.c_double),
        ("gamma", ctypes.c_double),
        ("re", ctypes.c_double),
        ("pr", ctypes.c_double),
        ("cv", ctypes.c_double),
        ("cp", ctypes.c_double),
        ("aoa", ctypes.c_double),
        ("scheme", ctypes.c_int32),
        ("mode", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("if_viscous", ctypes.c_int32),
        ("local_dt", ctypes.c_int32),
        ("cfl", ctypes.c_double),
        ("dtmax", ctypes.c_double),
        ("dtmin", ctypes.c_double),
        ("ns_tol", ctypes.c_double),
        ("ns_maxit", ctypes.c_int32),
        ("kstep_show", ctypes.c_int32),
        ("split_kernels", ctypes.c_int32),
        ("walk_tiles", ctypes.c_int32),
    ]


class GsisError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"gsis error {code}: {msg}")
        self.code = code


_lib = None


def load_library(path: os.PathLike | str | None = None) -> ctypes.CDLL:
    """Load libgsis.so and declare the prototypes.  Raises when it is missing."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = pathlib.Path(path) if path else LIB_PATH
    if not p.exists():
        raise FileNotFoundError(
            f"{p} is not built: run __graft_entry__.build() (hipcc --offload-arch=gfx950)")
    lib = ctypes.CDLL(str(p))
    h = ctypes.c_void_p
    lib.gsis_create.argtypes = [ctypes.POINTER(GsisParams), ctypes.POINTER(h)]
    lib.gsis_upload_frd.argtypes = [h, c_dp]
    lib.gsis_upload_macro.argtypes = [h, c_dp, c_dp, c_dp, c_dp]
    lib.gsis_refresh_moments.argtypes = [h]
    lib.gsis_step.argtypes = [h, ctypes.c_int32]
    lib.gsis_residual.argtypes = [h, c_dp]
    lib.gsis_download_macro.argtypes = [h] + [c_dp] * 9
    lib.gsis_download_frd.argtypes = [h, c_dp]
    lib.gsis_download_frd_noalloc.argtypes = [h, c_dp]
    lib.gsis_counters.argtypes = [h, ctypes.POINTER(ctypes.c_int32), c_dp, ctypes.POINTER(ctypes.c_int32)]
    lib.gsis_set_counters.argtypes = [h, ctypes.c_int32, ctypes.c_double]
    lib.gsis_inner_exchange.argtypes = [h, ctypes.c_int32, h, ctypes.c_int32, ctypes.c_int32]
    lib.gsis_timing.argtypes = [h, ctypes.c_int32, c_dp, ctypes.POINTER(ctypes.c_int32)]
    lib.gsis_destroy.argtypes = [h]
    lib.gsis_last_error.restype = ctypes.c_char_p
    c_fp = ctypes.POINTER(ctypes.c_float)
    c_i64p = ctypes.POINTER(ctypes.c_int64)
    lib.gsis_set_mass_correct.argtypes = [h, ctypes.c_int32]
    lib.gsis_closure_samples.argtypes = [h, c_fp, c_fp, ctypes.c_int64, c_i64p]
    lib.gsis_set_closure.argtypes = [h, ctypes.c_void_p, ctypes.c_double, ctypes.c_int32]
    lib.gsis_set_closure_warmup.argtypes = [h, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]
    lib.gsis_closure_stats.argtypes = [h, c_i64p, c_i64p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]
    lib.gsis_mlp_create.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    lib.gsis_mlp_nparams.argtypes = []
    lib.gsis_mlp_set_params.argtypes = [ctypes.c_void_p, c_fp]
    lib.gsis_mlp_get_params.argtypes = [ctypes.c_void_p, c_fp]
    lib.gsis_mlp_set_scaling.argtypes = [ctypes.c_void_p, c_fp]
    lib.gsis_mlp_forward.argtypes = [ctypes.c_void_p, c_fp, ctypes.c_int32, c_fp]
    lib.gsis_mlp_loss_grad.argtypes = [ctypes.c_void_p, c_fp, c_fp, ctypes.c_int32, c_dp, c_fp]
    lib.gsis_mlp_train.argtypes = [ctypes.c_void_p, c_fp, c_fp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                   ctypes.c_float, ctypes.c_uint64, c_dp]
    lib.gsis_mlp_destroy.argtypes = [ctypes.c_void_p]
    i32p = ctypes.POINTER(ctypes.c_int32)
    lib.gsis_fill_frd_uniform.argtypes = [h, c_dp, c_dp]
    lib.gsis_batch_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]
    lib.gsis_batch_step.argtypes = [ctypes.c_void_p, ctypes.c_int32]
    lib.gsis_batch_tune.argtypes = [ctypes.c_void_p, ctypes.c_int32, i32p, i32p, c_dp]
    lib.gsis_batch_set_ns_graph.argtypes = [ctypes.c_void_p, ctypes.c_int32]
    lib.gsis_batch_ns_graph.argtypes = [ctypes.c_void_p, i32p]
    lib.gsis_batch_destroy.argtypes = [ctypes.c_void_p]
    for name in EXPORTED:
        if name != "gsis_last_error":
            getattr(lib, name).restype = ctypes.c_int
    if path is None:
        _lib = lib
    return lib


def check(lib: ctypes.CDLL, rc: int) -> None:
    if rc != GSIS_OK:
        raise GsisError(rc, lib.gsis_last_error().decode(errors="replace"))
