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
// P supplies the cell fields: u(i,j), v(i,j), T(i,j), mu(i,j) = amu_of, and
// the cell centres x1(i,j), y1(i,j).
template <class P>
__device__ inline void ns_face_of(const P &f, const Dev &d, int dir, int i, int j, double out[8]) {
  const int nx = d.nx, ny = d.ny;
  double amu1, amk1, dix, diy, djx, djy, diu, div, dit, dju, djv, djt, u1, v1;
  const double prt = 0.9;
  if (dir == 0) {
    if (i == 1) { const double a = f.mu(i, j); amu1 = a + 0.; amk1 = d.cp * (a / d.pr + 0. / prt); }
    else if (i == nx) { const double a = f.mu(i - 1, j); amu1 = a + 0.; amk1 = d.cp * (a / d.pr + 0. / prt); }
    else {
      const double a0 = f.mu(i - 1, j), a1 = f.mu(i, j);
      amu1 = (a0 + a1 + 0. + 0.) * 0.5;
      amk1 = d.cp * ((a0 + a1) / d.pr + (0. + 0.) / prt) * 0.5;
    }
    dix = f.x1(i, j) - f.x1(i - 1, j);
    diy = f.y1(i, j) - f.y1(i - 1, j);
    djx = (f.x1(i - 1, j + 1) + f.x1(i, j + 1) - f.x1(i - 1, j - 1) - f.x1(i, j - 1)) * 0.25;
    djy = (f.y1(i - 1, j + 1) + f.y1(i, j + 1) - f.y1(i - 1, j - 1) - f.y1(i, j - 1)) * 0.25;
    diu = f.u(i, j) - f.u(i - 1, j);
    div = f.v(i, j) - f.v(i - 1, j);
    dit = f.T(i, j) - f.T(i - 1, j);
    dju = (f.u(i - 1, j + 1) + f.u(i, j + 1) - f.u(i - 1, j - 1) - f.u(i, j - 1)) * 0.25;
    djv = (f.v(i - 1, j + 1) + f.v(i, j + 1) - f.v(i - 1, j - 1) - f.v(i, j - 1)) * 0.25;
    djt = (f.T(i - 1, j + 1) + f.T(i, j + 1) - f.T(i - 1, j - 1) - f.T(i, j - 1)) * 0.25;
    u1 = (f.u(i, j) + f.u(i - 1, j)) * 0.5;
    v1 = (f.v(i, j) + f.v(i - 1, j)) * 0.5;
  } else {
    if (j == 1) { const double a = f.mu(i, j); amu1 = a + 0.; amk1 = d.cp * (a / d.pr + 0. / prt); }
    else if (j == ny) { const double a = f.mu(i, j - 1); amu1 = a + 0.; amk1 = d.cp * (a / d.pr + 0. / prt); }
    else {
      const double a0 = f.mu(i, j), a1 = f.mu(i, j - 1);
      amu1 = (a0 + a1 + 0. + 0.) * 0.5;
      amk1 = d.cp * ((a1 + a0) / d.pr + (0. + 0.) / prt) * 0.5;
    }
    dix = (f.x1(i + 1, j - 1) + f.x1(i + 1, j) - f.x1(i - 1, j - 1) - f.x1(i - 1, j)) * 0.25;
    diy = (f.y1(i + 1, j - 1) + f.y1(i + 1, j) - f.y1(i - 1, j - 1) - f.y1(i - 1, j)) * 0.25;
    djx = f.x1(i, j) - f.x1(i, j - 1);
    djy = f.y1(i, j) - f.y1(i, j - 1);
    diu = (f.u(i + 1, j - 1) + f.u(i + 1, j) - f.u(i - 1, j - 1) - f.u(i - 1, j)) * 0.25;
    div = (f.v(i + 1, j - 1) + f.v(i + 1, j) - f.v(i - 1, j - 1) - f.v(i - 1, j)) * 0.25;
    dit = (f.T(i + 1, j - 1) + f.T(i + 1, j) - f.T(i - 1, j - 1) - f.T(i - 1, j)) * 0.25;
    dju = f.u(i, j) - f.u(i, j - 1);
    djv = f.v(i, j) - f.v(i, j - 1);
    djt = f.T(i, j) - f.T(i, j - 1);
    u1 = (f.u(i, j) + f.u(i, j - 1)) * 0.5;
    v1 = (f.v(i, j) + f.v(i, j - 1)) * 0.5;
  }
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

// the cell fields from the device arrays (prim of comput_duvtpckw)
struct GlobalCells {
  const Dev &d;
  __device__ double u(int i, int j) const { return d.prim[d.ncell + cidx(d, i, j)]; }
  __device__ double v(int i, int j) const { return d.prim[2 * d.ncell + cidx(d, i, j)]; }
  __device__ double T(int i, int j) const { return d.prim[3 * d.ncell + cidx(d, i, j)]; }
  __device__ double mu(int i, int j) const { return amu_of(d, cidx(d, i, j)); }
  __device__ double x1(int i, int j) const { return d.x1[cidx(d, i, j)]; }
  __device__ double y1(int i, int j) const { return d.y1[cidx(d, i, j)]; }
};

__device__ inline void ns_face(const Dev &d, int dir, int i, int j, double out[8]) {
  ns_face_of(GlobalCells{d}, d, dir, i, j, out);
}

}  // namespace gsis
