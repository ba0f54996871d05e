"""Macroscopic fields of a GSIS state and their relative L2 differences, for
reporting the FP32-sweep path against the FP64 path.

After DETFP the slots of S_GKUA(1:10, i, j) hold n, U, V, T, p_xx, p_yy,
p_xy, q_x, q_y, p_zz (src/GKUA_DETFP.f90:57-84)."""
from __future__ import annotations

import numpy as np


def interior_fields(s: np.ndarray, nx: int, ny: int) -> dict:
    """rho, velocity (U,V), T, shear stress p_xy and heat flux (q_x,q_y) on the
    interior cells i=1..nx-1, j=1..ny-1 of a Fortran S_GKUA(11,-3:nx+3,-3:ny+3)."""
    c = s[:, 4:nx + 3, 4:ny + 3]
    return {"rho": c[0], "u": c[1:3], "T": c[3], "sigma_xy": c[6], "q": c[7:9]}


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    """||a - b||_2 / ||b||_2 (0 when both vanish)."""
    nb = float(np.linalg.norm(b))
    na = float(np.linalg.norm(np.asarray(a) - np.asarray(b)))
    return na / nb if nb > 0 else na


def field_errors(s_test: np.ndarray, s_ref: np.ndarray, nx: int, ny: int) -> dict:
    """Relative L2 error of each reported field of s_test against s_ref."""
    a, b = interior_fields(s_test, nx, ny), interior_fields(s_ref, nx, ny)
    return {k: rel_l2(a[k], b[k]) for k in a}
