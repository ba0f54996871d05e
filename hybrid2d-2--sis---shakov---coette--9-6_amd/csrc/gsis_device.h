// gsis_device.h -- device-side layout of the GSIS state on one MI355X.
//
// Cells carry one ghost ring more than the reference's stencils need:
// i in [-1, nx+1], j in [-1, ny+1] (the reference allocates -3..n+3 but only
// touches -1..n+1 outside the wall slot).  Cell (i,j) has linear index
//   c = (j+1)*NI + (i+1),  NI = nx+3.
// Distributions are stored f[cell][velocity] with the reduced distributions
// G and H in separate planes, so that the 64 lanes of a wavefront read 64
// consecutive discrete velocities of one cell (one coalesced 512 B load in FP64).
// Velocity index v = (jv-1)*nvi + (iv-1), the order of the reference's
// DO jv / DO iv loops (src/GKUA_Residual.f90:28-31).
//
// The reference's ghost slot FRD(-3|n+3, ...) that carries the diffuse-wall
// distribution (src/GKUA_boundary.f90:77-92) is a separate small array:
//   slot[((face-1)*L + pos)*2 + ird][v],  pos = j+1 (faces 1,3) or i+1 (faces 2,4),
//   L = max(nx,ny)+3.
// Geometry (reference (nx,ny) arrays) is stored on a zero-padded
// (nx+2) x (ny+2) grid indexed g = j*GI + i, GI = nx+2.
// Macroscopic fields are component-major: S[k*ncell + c].
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gsis {

constexpr double kPi = 3.1415926535897932;  // const_var PI (src/COMMONDATA.f90:8)
// REAL(4) literal 1.E-30 of the positivity clamp (src/GKUA_Residual.f90:104-105)
constexpr double kClamp = 1.0000000031710769e-30;

struct Dev {
  int nx, ny, nvi, nvj, nv;
  int NI, NJ, ncell;  // padded cell grid
  int GI;             // geometry row length
  int L;              // slot length per face
  const double *vi, *vj, *wi, *wj;
  const double *rowf;  // [nvj][4]: wy, wy*vy, wy*vy^2, wy*vy^3 of each velocity row (moments)
  const double *vol, *si, *sj, *ni1, *ni2, *nj1, *nj2;
  const double *xn, *yn;  // nodes, (nx+4) x (ny+4), index (j+1)*(nx+4)+(i+1)
  const double *x1, *y1;  // cell centres, padded cell grid
  double cfc, x1s, cpr, twall, dt, ma, gamma, re, pr, cv, cp, aoa;
  int scheme;
  double *S;     // [11][ncell]
  double *S0;    // [10][ncell]
  double *CF;    // [ncell]
  double *U;     // [4][ncell]
  double *U0;    // [4][ncell]
  double *U2;    // [4][ncell]
  double *taoi, *taoj;  // [4][ncell]
  double *qi, *qj;      // [2][ncell]
  double *prim;  // [5][ncell]: d, u, v, T, p  (comput_duvtpckw)
  double *res4;  // [4][ncell]: velocity moments of the kinetic RHS per cell
  int *err;      // [0]: T<0.1 flag
  const uint8_t *ovrI1, *ovrI3;  // per row j: wall/farfield override on i-face 1 / nx
  const uint8_t *ovrJ2, *ovrJ4;  // per column i: override on j-face 1 / ny
};

__host__ __device__ inline int cidx(const Dev &d, int i, int j) { return (j + 1) * d.NI + (i + 1); }
__host__ __device__ inline int gidx(const Dev &d, int i, int j) { return j * d.GI + i; }
__host__ __device__ inline int nidx(const Dev &d, int i, int j) { return (j + 1) * (d.nx + 4) + (i + 1); }
__host__ __device__ inline size_t slot_idx(const Dev &d, int face, int pos, int ird, int v) {
  return ((((size_t)(face - 1) * d.L + pos) * 2 + ird) * (size_t)d.nv) + v;
}

// Loads and stores coherent across the whole device (all XCDs), for values
// one workgroup hands to another during a launch.
template <typename T>
__device__ __forceinline__ T ld_agent(const T *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void st_agent(T *p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One boundary subface prepared for the device.
struct BcDev {
  int face, ist, iend, jst, jend, code;
  double uw, vw;
  int swall_off;  // offset (in doubles) into the Swall buffer
};

}  // namespace gsis
