// gsis_closure.hip -- the learned HoT closure inside the synthetic NS solve:
// network inputs and targets of the closure faces, the per-face acceptance
// test against the DVM HoT, and the HoT update between inner-loop chunks.
//
// The HoT that Convert_NS builds on an NS face (src/Hybrid_Sub.f90:317-349 for
// i-faces, :419-448 for j-faces) are stored per face in taoi/taoj [4][ncell]
// (tau_11, tau_22, tau_12, tau_33) and qi/qj [2][ncell], face (i, j) at cell
// index cidx(i, j).  The kinetic parts of tau are traceless and so are the NS
// stresses, so tau_33 = -(tau_11 + tau_22) and five numbers per face carry the
// HoT; the closure updates tau_33 by minus the updates of tau_11 and tau_22.
#include <cmath>

#include "gsis_closure.h"
#include "gsis_ns_common.h"

namespace gsis {

namespace {

// the cell fields of the NS face gradients with viscosity 1, so that
// ns_face_of returns strain-rate combinations and cp/Pr grad T
template <bool PRIM>
struct UnitMuCells {
  const Dev &d;
  __device__ double rho(int i, int j) const {
    const int c = cidx(d, i, j);
    return PRIM ? d.prim[c] : d.U[c];
  }
  __device__ double u(int i, int j) const {
    const int c = cidx(d, i, j);
    return PRIM ? d.prim[d.ncell + c] : d.U[d.ncell + c] / d.U[c];
  }
  __device__ double v(int i, int j) const {
    const int c = cidx(d, i, j);
    return PRIM ? d.prim[2 * d.ncell + c] : d.U[2 * d.ncell + c] / d.U[c];
  }
  __device__ double T(int i, int j) const {
    const int c = cidx(d, i, j);
    if (PRIM) return d.prim[3 * d.ncell + c];
    const int N = d.ncell;
    const double r = d.U[c], uu = d.U[N + c] / r, vv = d.U[2 * N + c] / r;
    return (d.U[3 * N + c] - 0.5 * r * (uu * uu + vv * vv)) / (d.cv * r);  // comput_duvtpckw
  }
  __device__ double mu(int, int) const { return 1.0; }
  __device__ double x1(int i, int j) const { return d.x1[cidx(d, i, j)]; }
  __device__ double y1(int i, int j) const { return d.y1[cidx(d, i, j)]; }
};

// closure face f -> direction and face indices
__device__ __forceinline__ void face_of(const Dev &d, int nfi, int f, int &dir, int &i, int &j) {
  if (f < nfi) {
    dir = 0;
    i = 2 + f % (d.nx - 2);
    j = 2 + f / (d.nx - 2);
  } else {
    const int g = f - nfi;
    dir = 1;
    i = 2 + g % (d.nx - 3);
    j = 2 + g / (d.nx - 3);
  }
}

template <bool PRIM>
__global__ void k_cl_features(Dev d, int nfi, int nf, float lgkn, float *x) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  int dir, i, j;
  face_of(d, nfi, f, dir, i, j);
  const UnitMuCells<PRIM> cells{d};
  double o[8];
  ns_face_of(cells, d, dir, i, j, o);
  // o = (4/3 ux - 2/3 vy, 4/3 vy - 2/3 ux, uy + vx, ., k tx, k ty) with k = cp/Pr
  const double k = d.cp / d.pr;
  const int i0 = dir == 0 ? i - 1 : i, j0 = dir == 0 ? j : j - 1;
  float *xf = x + (size_t)f * kMlpIn;
  xf[0] = (float)(o[0] + 0.5 * o[1]);
  xf[1] = (float)(o[1] + 0.5 * o[0]);
  xf[2] = (float)o[2];
  xf[3] = (float)(o[4] / k);
  xf[4] = (float)(o[5] / k);
  xf[5] = (float)(0.5 * (cells.rho(i, j) + cells.rho(i0, j0)));
  xf[6] = (float)(0.5 * (cells.T(i, j) + cells.T(i0, j0)));
  xf[7] = lgkn;
}

__global__ void k_cl_targets(Dev d, int nfi, int nf, float *y) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  int dir, i, j;
  face_of(d, nfi, f, dir, i, j);
  const int N = d.ncell, c = cidx(d, i, j);
  const double *tg = dir == 0 ? d.taoi : d.taoj;
  const double *qg = dir == 0 ? d.qi : d.qj;
  float *yf = y + (size_t)f * kMlpOut;
  yf[0] = (float)tg[c];
  yf[1] = (float)tg[N + c];
  yf[2] = (float)tg[2 * N + c];
  yf[3] = (float)qg[c];
  yf[4] = (float)qg[N + c];
}

// accept[f] = 1 when every output of the network at the Convert_NS state is
// within gate standard deviations of the DVM HoT; counts the accepted faces
__global__ void k_cl_gate(int nf, const float *ybase, const float *t, const float *sc, float gate, int *accept,
                          int *count) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  int ok = 0;
  if (f < nf) {
    ok = 1;
    for (int k = 0; k < kMlpOut; k++) {
      const float e = fabsf(ybase[(size_t)f * kMlpOut + k] - t[(size_t)f * kMlpOut + k]) / sc[2 * kMlpIn + kMlpOut + k];
      if (!(e <= gate)) ok = 0;  // a NaN prediction is rejected
    }
    accept[f] = ok;
  }
  const unsigned long long b = __ballot(ok);
  if ((threadIdx.x & 63) == 0 && b) atomicAdd(count, (int)__popcll(b));
}

// HoT of the accepted faces = saved DVM HoT + N(now) - N(base), or with
// replace = N(now) itself (tau_33 = -(tau_11 + tau_22))
__global__ void k_cl_apply(Dev d, int nfi, int nf, const float *ynow, const float *ybase, const int *accept,
                           const double *saved, int replace) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf || !accept[f]) return;
  int dir, i, j;
  face_of(d, nfi, f, dir, i, j);
  const int N = d.ncell, c = cidx(d, i, j);
  double *tg = dir == 0 ? d.taoi : d.taoj;
  double *qg = dir == 0 ? d.qi : d.qj;
  const double *st = saved + (dir == 0 ? 0 : 4 * (size_t)N);
  const double *sq = saved + 8 * (size_t)N + (dir == 0 ? 0 : 2 * (size_t)N);
  double dl[kMlpOut];
  if (replace) {
    for (int k = 0; k < kMlpOut; k++) dl[k] = (double)ynow[(size_t)f * kMlpOut + k];
    tg[c] = dl[0];
    tg[N + c] = dl[1];
    tg[2 * N + c] = dl[2];
    tg[3 * N + c] = -dl[0] - dl[1];
    qg[c] = dl[3];
    qg[N + c] = dl[4];
    return;
  }
  for (int k = 0; k < kMlpOut; k++)
    dl[k] = (double)ynow[(size_t)f * kMlpOut + k] - (double)ybase[(size_t)f * kMlpOut + k];
  tg[c] = st[c] + dl[0];
  tg[N + c] = st[N + c] + dl[1];
  tg[2 * N + c] = st[2 * N + c] + dl[2];
  tg[3 * N + c] = st[3 * N + c] - dl[0] - dl[1];
  qg[c] = sq[c] + dl[3];
  qg[N + c] = sq[N + c] + dl[4];
}

__global__ void k_cl_zero(int *p) {
  if (threadIdx.x == 0) *p = 0;
}

unsigned blocks(int n) { return (unsigned)((n + 255) / 256); }

hipError_t copy_hot(const Dev &d, double *dst, bool to_saved, hipStream_t s) {
  const size_t N = d.ncell;
  double *dev[4] = {d.taoi, d.taoj, d.qi, d.qj};
  const size_t off[4] = {0, 4 * N, 8 * N, 10 * N}, len[4] = {4 * N, 4 * N, 2 * N, 2 * N};
  for (int k = 0; k < 4; k++) {
    const hipError_t e = to_saved ? hipMemcpyAsync(dst + off[k], dev[k], len[k] * 8, hipMemcpyDeviceToDevice, s)
                                  : hipMemcpyAsync(dev[k], dst + off[k], len[k] * 8, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace

int closure_faces(const Dev &d, int *nfi) {
  const int a = d.nx >= 4 && d.ny >= 4 ? (d.nx - 2) * (d.ny - 3) : 0;
  const int b = d.nx >= 4 && d.ny >= 4 ? (d.nx - 3) * (d.ny - 2) : 0;
  if (nfi) *nfi = a;
  return a + b;
}

float closure_log10_kn(const Dev &d) { return (float)std::log10(std::sqrt(M_PI) / (2.0 * d.cfc)); }

int closure_alloc(const Dev &d, ClosureBufs &b) {
  if (b.mem) return 0;
  b.nf = closure_faces(d, &b.nfi);
  const size_t nf = std::max(b.nf, 1), N = d.ncell;
  b.bytes = 16 * N * 8 + nf * (kMlpIn + 3 * kMlpOut) * 4 + nf * 4 + 64;
  if (hipMalloc(&b.mem, b.bytes) != hipSuccess) {
    b.mem = nullptr;
    return GSIS_EHIP;
  }
  char *p = reinterpret_cast<char *>(b.mem);
  b.saved = reinterpret_cast<double *>(p); p += 12 * N * 8;
  b.usave = reinterpret_cast<double *>(p); p += 4 * N * 8;
  b.x = reinterpret_cast<float *>(p); p += nf * kMlpIn * 4;
  b.ybase = reinterpret_cast<float *>(p); p += nf * kMlpOut * 4;
  b.y = reinterpret_cast<float *>(p); p += nf * kMlpOut * 4;
  b.t = reinterpret_cast<float *>(p); p += nf * kMlpOut * 4;
  b.accept = reinterpret_cast<int *>(p); p += nf * 4;
  b.count = reinterpret_cast<int *>(p);
  return 0;
}

void closure_free(ClosureBufs &b) {
  if (b.mem) (void)hipFree(b.mem);
  b = ClosureBufs{};
}

void launch_cl_features(const Dev &d, int nfi, int nf, bool from_prim, float *x, hipStream_t s) {
  if (nf <= 0) return;
  const float lk = closure_log10_kn(d);
  if (from_prim) hipLaunchKernelGGL(k_cl_features<true>, dim3(blocks(nf)), dim3(256), 0, s, d, nfi, nf, lk, x);
  else hipLaunchKernelGGL(k_cl_features<false>, dim3(blocks(nf)), dim3(256), 0, s, d, nfi, nf, lk, x);
}

void launch_cl_targets(const Dev &d, int nfi, int nf, float *y, hipStream_t s) {
  if (nf > 0) hipLaunchKernelGGL(k_cl_targets, dim3(blocks(nf)), dim3(256), 0, s, d, nfi, nf, y);
}

int closure_begin(const Dev &d, ClosureBufs &b, MlpDev &net, float gate, bool replace, hipStream_t s) {
  if (copy_hot(d, b.saved, true, s) != hipSuccess ||
      hipMemcpyAsync(b.usave, d.U, 4 * (size_t)d.ncell * 8, hipMemcpyDeviceToDevice, s) != hipSuccess)
    return GSIS_EHIP;
  b.refreshes = 0;
  b.rejected = 0;
  if (b.nf <= 0) return 0;
  launch_cl_features(d, b.nfi, b.nf, true, b.x, s);
  mlp_forward_dev(net, b.x, b.nf, b.ybase, s);
  launch_cl_targets(d, b.nfi, b.nf, b.t, s);
  hipLaunchKernelGGL(k_cl_zero, dim3(1), dim3(64), 0, s, b.count);
  // replacing: every face whose prediction is finite
  hipLaunchKernelGGL(k_cl_gate, dim3(blocks(b.nf)), dim3(256), 0, s, b.nf, b.ybase, b.t, net.dscale,
                     replace ? 3.0e38f : gate, b.accept, b.count);
  b.replace = replace;
  if (replace)  // the first launch already runs with the network's HoT
    hipLaunchKernelGGL(k_cl_apply, dim3(blocks(b.nf)), dim3(256), 0, s, d, b.nfi, b.nf, b.ybase, b.ybase, b.accept,
                       b.saved, 1);
  return hipGetLastError() == hipSuccess ? 0 : GSIS_EHIP;
}

int closure_refresh(const Dev &d, ClosureBufs &b, MlpDev &net, hipStream_t s) {
  b.refreshes += 1;
  if (b.nf <= 0) return 0;
  launch_cl_features(d, b.nfi, b.nf, false, b.x, s);
  mlp_forward_dev(net, b.x, b.nf, b.y, s);
  hipLaunchKernelGGL(k_cl_apply, dim3(blocks(b.nf)), dim3(256), 0, s, d, b.nfi, b.nf, b.y, b.ybase, b.accept,
                     b.saved, b.replace ? 1 : 0);
  return hipGetLastError() == hipSuccess ? 0 : GSIS_EHIP;
}

int closure_end(const Dev &d, ClosureBufs &b, hipStream_t s) {
  return copy_hot(d, b.saved, false, s) == hipSuccess ? 0 : GSIS_EHIP;
}

int closure_reject(const Dev &d, ClosureBufs &b, hipStream_t s) {
  b.rejected = 1;
  if (copy_hot(d, b.saved, false, s) != hipSuccess ||
      hipMemcpyAsync(d.U, b.usave, 4 * (size_t)d.ncell * 8, hipMemcpyDeviceToDevice, s) != hipSuccess)
    return GSIS_EHIP;
  return 0;
}

}  // namespace gsis
