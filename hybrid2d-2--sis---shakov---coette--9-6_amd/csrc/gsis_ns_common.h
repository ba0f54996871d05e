// gsis_ns_common.h -- device code shared by Convert_NS (gsis_kernels.hip) and
// the synthetic NS solve (gsis_synthetic.hip): viscosity law and the
// face-centred NS stress / heat flux.
#pragma once
#include "gsis_device.h"

namespace gsis {

// mu = T^0.81/Re with the REAL(4) literal 0.81 (get_viscous, src/sub_Residual.f90:553-576)
__device__ __forceinline__ double amu_of(const Dev &d, int c) {
  return 1. / d.re * pow(d.prim[3 * d.ncell + c], 0.8100000023841858);
}

// Face NS stress/heat flux (Convert_NS src/Hybrid_Sub.f90:257-309,358-407;
// the same discretisation as Residual_FVM src/sub_Residual.f90:125-165,262-298).
__device__ inline void ns_face(const Dev &d, int dir, int i, int j, double out[8]) {
  const int nx = d.nx, ny = d.ny, N = d.ncell;
  const double *uu = d.prim + N, *vv = d.prim + 2 * N, *TT = d.prim + 3 * N;
  double amu1, amk1, dix, diy, djx, djy, diu, div, dit, dju, djv, djt, u1, v1;
  const double prt = 0.9;
#define C(ii, jj) cidx(d, ii, jj)
  if (dir == 0) {
    if (i == 1) { const double a = amu_of(d, C(i, j)); amu1 = a + 0.; amk1 = d.cp * (a / d.pr + 0. / prt); }
    else if (i == nx) { const double a = amu_of(d, C(i - 1, j)); amu1 = a + 0.; amk1 = d.cp * (a / d.pr + 0. / prt); }
    else {
      const double a0 = amu_of(d, C(i - 1, j)), a1 = amu_of(d, C(i, j));
      amu1 = (a0 + a1 + 0. + 0.) * 0.5;
      amk1 = d.cp * ((a0 + a1) / d.pr + (0. + 0.) / prt) * 0.5;
    }
    dix = d.x1[C(i, j)] - d.x1[C(i - 1, j)];
    diy = d.y1[C(i, j)] - d.y1[C(i - 1, j)];
    djx = (d.x1[C(i - 1, j + 1)] + d.x1[C(i, j + 1)] - d.x1[C(i - 1, j - 1)] - d.x1[C(i, j - 1)]) * 0.25;
    djy = (d.y1[C(i - 1, j + 1)] + d.y1[C(i, j + 1)] - d.y1[C(i - 1, j - 1)] - d.y1[C(i, j - 1)]) * 0.25;
    diu = uu[C(i, j)] - uu[C(i - 1, j)];
    div = vv[C(i, j)] - vv[C(i - 1, j)];
    dit = TT[C(i, j)] - TT[C(i - 1, j)];
    dju = (uu[C(i - 1, j + 1)] + uu[C(i, j + 1)] - uu[C(i - 1, j - 1)] - uu[C(i, j - 1)]) * 0.25;
    djv = (vv[C(i - 1, j + 1)] + vv[C(i, j + 1)] - vv[C(i - 1, j - 1)] - vv[C(i, j - 1)]) * 0.25;
    djt = (TT[C(i - 1, j + 1)] + TT[C(i, j + 1)] - TT[C(i - 1, j - 1)] - TT[C(i, j - 1)]) * 0.25;
    u1 = (uu[C(i, j)] + uu[C(i - 1, j)]) * 0.5;
    v1 = (vv[C(i, j)] + vv[C(i - 1, j)]) * 0.5;
  } else {
    if (j == 1) { const double a = amu_of(d, C(i, j)); amu1 = a + 0.; amk1 = d.cp * (a / d.pr + 0. / prt); }
    else if (j == ny) { const double a = amu_of(d, C(i, j - 1)); amu1 = a + 0.; amk1 = d.cp * (a / d.pr + 0. / prt); }
    else {
      const double a0 = amu_of(d, C(i, j)), a1 = amu_of(d, C(i, j - 1));
      amu1 = (a0 + a1 + 0. + 0.) * 0.5;
      amk1 = d.cp * ((a1 + a0) / d.pr + (0. + 0.) / prt) * 0.5;
    }
    dix = (d.x1[C(i + 1, j - 1)] + d.x1[C(i + 1, j)] - d.x1[C(i - 1, j - 1)] - d.x1[C(i - 1, j)]) * 0.25;
    diy = (d.y1[C(i + 1, j - 1)] + d.y1[C(i + 1, j)] - d.y1[C(i - 1, j - 1)] - d.y1[C(i - 1, j)]) * 0.25;
    djx = d.x1[C(i, j)] - d.x1[C(i, j - 1)];
    djy = d.y1[C(i, j)] - d.y1[C(i, j - 1)];
    diu = (uu[C(i + 1, j - 1)] + uu[C(i + 1, j)] - uu[C(i - 1, j - 1)] - uu[C(i - 1, j)]) * 0.25;
    div = (vv[C(i + 1, j - 1)] + vv[C(i + 1, j)] - vv[C(i - 1, j - 1)] - vv[C(i - 1, j)]) * 0.25;
    dit = (TT[C(i + 1, j - 1)] + TT[C(i + 1, j)] - TT[C(i - 1, j - 1)] - TT[C(i - 1, j)]) * 0.25;
    dju = uu[C(i, j)] - uu[C(i, j - 1)];
    djv = vv[C(i, j)] - vv[C(i, j - 1)];
    djt = TT[C(i, j)] - TT[C(i, j - 1)];
    u1 = (uu[C(i, j)] + uu[C(i, j - 1)]) * 0.5;
    v1 = (vv[C(i, j)] + vv[C(i, j - 1)]) * 0.5;
  }
#undef C
  const double ds = 1. / (dix * djy - djx * diy);
  const double ux = (diu * djy - dju * diy) * ds;
  const double vx = (div * djy - djv * diy) * ds;
  const double tx = (dit * djy - djt * diy) * ds;
  const double uy = (-diu * djx + dju * dix) * ds;
  const double vy = (-div * djx + djv * dix) * ds;
  const double ty = (-dit * djx + djt * dix) * ds;
  out[0] = ((4. / 3.) * ux - (2. / 3.) * vy) * amu1;
  out[1] = ((4. / 3.) * vy - (2. / 3.) * ux) * amu1;
  out[2] = (uy + vx) * amu1;
  out[3] = amu1 * (-(2. / 3.) * ux - (2. / 3.) * vy);
  out[4] = amk1 * tx;
  out[5] = amk1 * ty;
  out[6] = u1;
  out[7] = v1;
}

}  // namespace gsis
