"""gsis_download_frd_noalloc: the FRD download through the library's
file-scope device stage, with no device allocation, equals gsis_download_frd
bit for bit.  Two calls in a row, the second on a smaller mesh in FP32 (the
stage keeps the first call's contents), and a mesh whose single velocity does
not fit the stage is refused before any device work."""
import ctypes

import numpy as np
import pytest

from helpers import load_pkg, solver_mod


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@pytest.mark.gpu
def test_download_frd_noalloc_equals_download_frd():
    g = load_pkg()
    sm = solver_mod()
    lib = g.capi.load_library()
    # 64^2 cells x 64^2 velocities: 340 MB of FRD, eleven chunks of the 32 MiB
    # stage; then 12^2 cells x 24^2 velocities in FP32, one chunk
    runs = [(g.cases.cavity_case(64, 64, kn=0.1, lid=0.15, dt=0.5), g.capi.PREC_FP64),
            (g.cases.cavity_case(12, 24, kn=0.5, lid=0.1, dt=0.5), g.capi.PREC_FP32)]
    for case, prec in runs:
        print("case:", case.name, "precision", prec, flush=True)
        s = sm.GsisSolver(case, precision=prec)
        s.initialize()
        s.step(3)
        ref = s.frd()
        got = np.full(g.cases.frd_shape(case), np.nan, order="F")
        g.capi.check(lib, lib.gsis_download_frd_noalloc(s.h, _dp(got)))
        s.close()
        assert np.array_equal(got, ref), case.name
        assert np.any(got != 0.0)


@pytest.mark.gpu
def test_download_frd_noalloc_refuses_a_mesh_beyond_the_stage():
    """One velocity of FRD is 2 (nx+7)(ny+7) doubles; 5 x 175 000 nodes need
    33.6 MB, more than the 32 MiB stage."""
    g = load_pkg()
    sm = solver_mod()
    lib = g.capi.load_library()
    nx, ny = 5, 175000
    mesh = g.cases.Mesh(nx, ny, np.zeros((nx + 4, ny + 4), order="F"), np.zeros((nx + 4, ny + 4), order="F"),
                        *[np.full((nx, ny), v, order="F") for v in (1e-6, 1e-3, 1e-3, 1.0, 0.0, 0.0, 1.0)])
    v = np.array([0.0])
    w = np.array([1.0])
    case = g.cases.Case("tall", mesh, [], v, w, v.copy(), w.copy(), g.cases.constants_from_kn(0.1), "rest")
    s = sm.GsisSolver(case)
    out = np.zeros(1)
    rc = lib.gsis_download_frd_noalloc(s.h, _dp(out))
    msg = lib.gsis_last_error()
    s.close()
    assert rc == g.capi.GSIS_EARG
    assert b"device stage" in msg
