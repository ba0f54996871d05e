// gsis_synthetic.hip -- the GSIS synthetic NS solve (mode M1) on one MI355X.
//
// Restates the block of Hybrid_Time_advance_LU_SGS that the reference ships
// commented out (src/Hybrid_Sub.f90:45-153) and its inner loop
// NS_Time_advance_LU_SGS (src/sub_NS_singlegrid.f90:4-40) with the options of
// the shipped control.in: original-variable NND2 reconstruction
// (src/sub_Residual.f90:387-397,476-546), Van Leer splitting
// (src/sub_flux_split.f90:141-192), viscous terms plus the kinetic HoT
// correction (src/sub_Residual.f90:176-193,309-326), local time step
// (src/sub_NS_singlegrid.f90:160-221) and the plane-by-plane LU-SGS
// (src/sub_LU_SGS.f90:4-143).
//
// The O(cells) inner loop is a chain of dependent sweeps, so it runs as one
// persistent workgroup (1024 threads) that iterates on the device until the
// max |energy residual| drops to ns_tol or ns_maxit is reached -- every
// iteration of the loop ends, so the kernel always drains.  The prepare and
// feedback stages are ordinary grid-wide kernels.
#include <string>

#include "gsis_ns_common.h"
#include "gsis_synthetic.h"

namespace gsis {

namespace {
std::string g_syn_err;
}

const char *synthetic_error() { return g_syn_err.c_str(); }

// S0 <- S - S0, S00 <- S on interior cells; the kinetic increment added to the
// synthetic state of the previous step on cells 3..n-3 (src/Hybrid_Sub.f90:45-94)
__global__ void k_ns_prepare(Dev d, SyntheticState st) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nc = (d.nx - 1) * (d.ny - 1);
  if (t >= nc) return;
  const int i = t % (d.nx - 1) + 1, j = t / (d.nx - 1) + 1;
  const int N = d.ncell, c = cidx(d, i, j);
  for (int k = 0; k < 10; k++) d.S0[k * N + c] = d.S[k * N + c] - d.S0[k * N + c];
  for (int k = 0; k < 4; k++) st.s00[k * N + c] = d.S[k * N + c];
  if (i < 3 || i > d.nx - 3 || j < 3 || j > d.ny - 3) return;
  const double sq = sqrt(2.0 / d.gamma), cv = d.cv;
  const double trho = d.S0[c], tu = d.S0[N + c] / d.ma * sq, tv = d.S0[2 * N + c] / d.ma * sq, tt = d.S0[3 * N + c];
  const double *U0 = d.U0;
  double rns = U0[c];
  double uns = U0[N + c] / U0[c], vns = U0[2 * N + c] / U0[c];
  double tns = (U0[3 * N + c] - 0.5 * U0[c] * (uns * uns + vns * vns)) / (cv * U0[c]);
  rns = rns + trho; uns = uns + tu; vns = vns + tv; tns = tns + tt;
  d.U[c] = rns;
  d.U[N + c] = rns * uns;
  d.U[2 * N + c] = rns * vns;
  d.U[3 * N + c] = cv * tns * rns + 0.5 * rns * (uns * uns + vns * vns);
  for (int m = 0; m < 4; m++) d.U2[m * N + c] = d.U[m * N + c];
}

__device__ __forceinline__ double minmod_ns(double a, double b) {  // src/sub_Residual.f90:357-368
  if (a * b <= 0.0) return 0.0;
  return (fabs(a) <= fabs(b)) ? a : b;
}

// Flux_Van_Leer_1Da, src/sub_flux_split.f90:141-192
__device__ void van_leer(const double ql[4], const double qr[4], double g, double f[4]) {
  const double dl = ql[0], uul = ql[1], vvl = ql[2], pl = ql[3];
  const double dr = qr[0], uur = qr[1], vvr = qr[2], pr = qr[3];
  const double al = sqrt(g * pl / dl), ar = sqrt(g * pr / dr);
  const double ml = uul / al, mr = uur / ar;
  double fp[4] = {0, 0, 0, 0}, fm[4] = {0, 0, 0, 0};
  if (ml >= 1.0) {
    fp[0] = dl * uul;
    fp[1] = dl * uul * uul + pl;
    fp[2] = dl * uul * vvl;
    fp[3] = uul * (g * pl / (g - 1.0) + 0.5 * dl * (uul * uul + vvl * vvl));
  } else if (fabs(ml) < 1.0) {
    const double mp = 0.25 * (1.0 + ml) * (1.0 + ml);
    const double t0 = dl * al * mp;
    fp[0] = t0;
    fp[1] = t0 * ((g - 1.0) * uul + 2.0 * al) / g;
    fp[2] = t0 * vvl;
    fp[3] = t0 * (((g - 1.0) * uul + 2.0 * al) * ((g - 1.0) * uul + 2.0 * al) * 0.5 / (g * g - 1.0) + 0.5 * (vvl * vvl));
  }
  if (mr >= 1.0) {
  } else if (fabs(mr) < 1.0) {
    const double mm = -0.25 * (mr - 1.0) * (mr - 1.0);
    const double t0 = dr * ar * mm;
    fm[0] = t0;
    fm[1] = t0 * ((g - 1.0) * uur - 2.0 * ar) / g;
    fm[2] = t0 * vvr;
    fm[3] = t0 * (((g - 1.0) * uur - 2.0 * ar) * ((g - 1.0) * uur - 2.0 * ar) * 0.5 / (g * g - 1.0) + 0.5 * (vvr * vvr));
  } else {
    fm[0] = dr * uur;
    fm[1] = dr * uur * uur + pr;
    fm[2] = dr * uur * vvr;
    fm[3] = uur * (g * pr / (g - 1.0) + 0.5 * dr * (uur * uur + vvr * vvr));
  }
  for (int m = 0; m < 4; m++) f[m] = fp[m] + fm[m];
}

// Residual_FVM face flux (src/sub_Residual.f90:68-202, :210-338)
__device__ void ns_face_flux(const Dev &d, int dir, int i, int j, bool visc, double fl[4]) {
  const int N = d.ncell;
  double ul[4], ur[4];
  for (int m = 0; m < 4; m++) {
    const double *q = d.prim + (m < 3 ? m : 4) * N;  // d, u, v, p
    double u1, u2, u3, u4;
    if (dir == 0) { u1 = q[cidx(d, i - 2, j)]; u2 = q[cidx(d, i - 1, j)]; u3 = q[cidx(d, i, j)]; u4 = q[cidx(d, i + 1, j)]; }
    else { u1 = q[cidx(d, i, j - 2)]; u2 = q[cidx(d, i, j - 1)]; u3 = q[cidx(d, i, j)]; u4 = q[cidx(d, i, j + 1)]; }
    ul[m] = u2 + 0.5 * minmod_ns(u2 - u1, u3 - u2);
    ur[m] = u3 - 0.5 * minmod_ns(u3 - u2, u4 - u3);
  }
  const int g = gidx(d, i, j);
  double s, n1, n2;
  if (dir == 0) { s = d.si[g]; n1 = d.ni1[g]; n2 = d.ni2[g]; }
  else { s = d.sj[g]; n1 = d.nj1[g]; n2 = d.nj2[g]; }
  const double ql[4] = {ul[0], ul[1] * n1 + ul[2] * n2, -ul[1] * n2 + ul[2] * n1, ul[3]};
  const double qr[4] = {ur[0], ur[1] * n1 + ur[2] * n2, -ur[1] * n2 + ur[2] * n1, ur[3]};
  double f0[4];
  van_leer(ql, qr, d.gamma, f0);
  fl[0] = -f0[0] * s;
  fl[1] = -(f0[1] * n1 - f0[2] * n2) * s;
  fl[2] = -(f0[1] * n2 + f0[2] * n1) * s;
  fl[3] = -f0[3] * s;
  if (!visc) return;
  double o[8];
  ns_face(d, dir, i, j, o);
  const int c = cidx(d, i, j);
  const double *tg = dir == 0 ? d.taoi : d.taoj;
  const double *qg = dir == 0 ? d.qi : d.qj;
  const double t11 = o[0] + tg[c], t22 = o[1] + tg[N + c], t12 = o[2] + tg[2 * N + c];
  const double qx = o[4] + qg[c], qy = o[5] + qg[N + c];
  const double u1 = o[6], v1 = o[7];
  const double e1 = u1 * t11 + v1 * t12 + qx;
  const double e2 = u1 * t12 + v1 * t22 + qy;
  fl[1] = fl[1] + (t11 * n1 + t12 * n2) * s;
  fl[2] = fl[2] + (t12 * n1 + t22 * n2) * s;
  fl[3] = fl[3] + (e1 * n1 + e2 * n2) * s;
}

// comput_DFn, src/sub_LU_SGS.f90:122-143 (U and dU component-major)
__device__ void comput_dfn(const double u[4], const double du[4], double n1, double n2, double g, double df[4]) {
  double u2[4];
  for (int m = 0; m < 4; m++) u2[m] = u[m] + du[m];
  const double un1 = (u[1] * n1 + u[2] * n2) / u[0];
  const double p1 = (g - 1.0) * (u[3] - 0.5 * (u[1] * u[1] + u[2] * u[2]) / u[0]);
  const double un2 = (u2[1] * n1 + u2[2] * n2) / u2[0];
  const double p2 = (g - 1.0) * (u2[3] - 0.5 * (u2[1] * u2[1] + u2[2] * u2[2]) / u2[0]);
  df[0] = u2[0] * un2 - u[0] * un1;
  df[1] = (u2[1] * un2 + p2 * n1) - (u[1] * un1 + p1 * n1);
  df[2] = (u2[2] * un2 + p2 * n2) - (u[2] * un1 + p1 * n2);
  df[3] = (u2[3] + p2) * un2 - (u[3] + p1) * un1;
}

__device__ double block_max(double x) {
  __shared__ double shm[16];
  for (int off = 32; off > 0; off >>= 1) x = fmax(x, __shfl_down(x, off, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if (lane == 0) shm[wid] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = shm[0];
    for (int w = 1; w < nw; w++) m = fmax(m, shm[w]);
    shm[0] = m;
  }
  __syncthreads();
  const double r = shm[0];
  __syncthreads();
  return r;
}

__device__ __forceinline__ void ld4(const double *a, int N, int c, double v[4]) {
  for (int m = 0; m < 4; m++) v[m] = a[m * N + c];
}

// The inner loop (src/Hybrid_Sub.f90:99-104) as one persistent workgroup.
__global__ void __launch_bounds__(1024) k_ns_loop(Dev d, SyntheticState st) {
  const int nx = d.nx, ny = d.ny, N = d.ncell, T = blockDim.x, tid = threadIdx.x;
  const double g = d.gamma, w = 1.0;
  const bool visc = st.if_viscous != 0;
  int it = 0;
  double res = 1.0;
  while (res > st.tol && it < st.maxit) {
    it++;
    // comput_duvtpckw over the padded grid, T<0.1 stop on interior cells
    int bad = 0;
    for (int c = tid; c < N; c += T) {
      const double dd = d.U[c], u = d.U[N + c] / dd, v = d.U[2 * N + c] / dd;
      const double tt = (d.U[3 * N + c] - 0.5 * dd * (u * u + v * v)) / (d.cv * dd);
      d.prim[c] = dd;
      d.prim[N + c] = u;
      d.prim[2 * N + c] = v;
      d.prim[3 * N + c] = tt;
      d.prim[4 * N + c] = (1. / (g * d.ma * d.ma)) * dd * tt;
      const int i = c % d.NI - 1, j = c / d.NI - 1;
      if (i >= 1 && i <= nx - 1 && j >= 1 && j <= ny - 1 && tt < 0.10000000149011612) bad = 1;
    }
    bad = __syncthreads_or(bad);
    if (bad) {
      if (tid == 0) st.iters[1] = 1;
      break;
    }
    // face fluxes of Residual_FVM
    const int nfi = nx * (ny - 1), nfj = (nx - 1) * ny;
    for (int t = tid; t < nfi + nfj; t += T) {
      double fl[4];
      int i, j;
      if (t < nfi) {
        i = t % nx + 1; j = t / nx + 1;
        ns_face_flux(d, 0, i, j, visc, fl);
        for (int m = 0; m < 4; m++) st.fi[m * N + cidx(d, i, j)] = fl[m];
      } else {
        const int u = t - nfi;
        i = u % (nx - 1) + 1; j = u / (nx - 1) + 1;
        ns_face_flux(d, 1, i, j, visc, fl);
        for (int m = 0; m < 4; m++) st.fj[m * N + cidx(d, i, j)] = fl[m];
      }
    }
    __syncthreads();
    // residual, its max on cells 3..n-3, comput_Lvc and comput_dt
    double rmax = 0.0;
    const int ncl = (nx - 1) * (ny - 1);
    for (int t = tid; t < ncl; t += T) {
      const int i = t % (nx - 1) + 1, j = t / (nx - 1) + 1;
      const int c = cidx(d, i, j), ci = cidx(d, i + 1, j), cj = cidx(d, i, j + 1);
      for (int m = 0; m < 4; m++)
        st.res[m * N + c] = st.fi[m * N + ci] - st.fi[m * N + c] + st.fj[m * N + cj] - st.fj[m * N + c];
      if (i >= 3 && i <= nx - 3 && j >= 3 && j <= ny - 3) rmax = fmax(rmax, fabs(st.res[3 * N + c]));
      const int g0 = gidx(d, i, j), gi = gidx(d, i + 1, j), gj = gidx(d, i, j + 1);
      const double s0 = d.vol[g0];
      const double si = 0.5 * (d.si[g0] + d.si[gi]);
      const double ni1 = 0.5 * (d.ni1[g0] + d.ni1[gi]), ni2 = 0.5 * (d.ni2[g0] + d.ni2[gi]);
      const double sj = 0.5 * (d.sj[g0] + d.sj[gj]);
      const double nj1 = 0.5 * (d.nj1[g0] + d.nj1[gj]), nj2 = 0.5 * (d.nj2[g0] + d.nj2[gj]);
      const double uu = d.prim[N + c], vv = d.prim[2 * N + c];
      const double uni = uu * ni1 + vv * ni2, unj = uu * nj1 + vv * nj2;
      const double cc = sqrt(d.prim[3 * N + c]) / d.ma;
      const double lci = (fabs(uni) + cc) * si, lcj = (fabs(unj) + cc) * sj;
      const double amu = visc ? amu_of(d, c) : 0.0;
      const double tmp1 = g / d.prim[c] * (amu / d.pr + 0.0 / 0.9);
      const double lvi = tmp1 * si * si / s0, lvj = tmp1 * sj * sj / s0;
      st.lci[c] = lci; st.lcj[c] = lcj; st.lvi[c] = lvi; st.lvj[c] = lvj;
      double dt;
      if (!st.local_dt) dt = d.dt;
      else {
        dt = st.cfl * d.vol[g0] / (lci + lcj + 1.0 * (lvi + lvj));
        if (dt > st.dtmax) dt = st.dtmax;
        if (dt < st.dtmin) dt = st.dtmin;
      }
      st.dtl[c] = dt;
    }
    res = block_max(rmax);  // includes the barrier that publishes res/L*/dt
    // du_LU_SGS_2D: forward planes (recursive), then backward planes (recursive)
    for (int plane = 6; plane <= nx + ny - 6; plane++) {
      const int jlo = max(3, plane - (nx - 3)), jhi = min(ny - 3, plane - 3);
      for (int j = jlo + tid; j <= jhi; j += T) {
        const int i = plane - j;
        const int c = cidx(d, i, j), g0 = gidx(d, i, j);
        double alfa = d.vol[g0] / st.dtl[c] + w * (st.lci[c] + st.lcj[c]);
        if (visc) alfa = alfa + 2.0 * (st.lvi[c] + st.lvj[c]);
        double dui[4] = {0, 0, 0, 0}, duj[4] = {0, 0, 0, 0}, df[4], uu[4], dd[4];
        if (i != 3) {
          const int cm = cidx(d, i - 1, j);
          ld4(d.U, N, cm, uu); ld4(st.du, N, cm, dd);
          comput_dfn(uu, dd, d.ni1[g0], d.ni2[g0], g, df);
          for (int m = 0; m < 4; m++) {
            dui[m] = 0.5 * (df[m] * d.si[g0] + w * st.lci[cm] * dd[m]);
            if (visc) dui[m] = dui[m] + st.lvi[cm] * dd[m];
          }
        }
        if (j != 3) {
          const int cm = cidx(d, i, j - 1);
          ld4(d.U, N, cm, uu); ld4(st.du, N, cm, dd);
          comput_dfn(uu, dd, d.nj1[g0], d.nj2[g0], g, df);
          for (int m = 0; m < 4; m++) {
            duj[m] = 0.5 * (df[m] * d.sj[g0] + w * st.lcj[cm] * dd[m]);
            if (visc) duj[m] = duj[m] + st.lvj[cm] * dd[m];
          }
        }
        for (int m = 0; m < 4; m++) st.du[m * N + c] = (st.res[m * N + c] + dui[m] + duj[m]) / alfa;
      }
      __syncthreads();
    }
    for (int plane = nx + ny - 6; plane >= 6; plane--) {
      const int jlo = max(3, plane - (nx - 3)), jhi = min(ny - 3, plane - 3);
      for (int j = jlo + tid; j <= jhi; j += T) {
        const int i = plane - j;
        const int c = cidx(d, i, j), g0 = gidx(d, i, j);
        double alfa = d.vol[g0] / st.dtl[c] + w * (st.lci[c] + st.lcj[c]);
        if (visc) alfa = alfa + 2.0 * (st.lvi[c] + st.lvj[c]);
        double dui[4] = {0, 0, 0, 0}, duj[4] = {0, 0, 0, 0}, df[4], uu[4], dd[4];
        if (i != nx - 3) {
          const int cp = cidx(d, i + 1, j), gp = gidx(d, i + 1, j);
          ld4(d.U, N, cp, uu); ld4(st.du, N, cp, dd);
          comput_dfn(uu, dd, d.ni1[g0], d.ni2[g0], g, df);
          for (int m = 0; m < 4; m++) {
            dui[m] = -0.5 * (df[m] * d.si[gp] - w * st.lci[cp] * dd[m]);
            if (visc) dui[m] = dui[m] + st.lvi[cp] * dd[m];
          }
        }
        if (j != ny - 3) {
          const int cp = cidx(d, i, j + 1), gp = gidx(d, i, j + 1);
          ld4(d.U, N, cp, uu); ld4(st.du, N, cp, dd);
          comput_dfn(uu, dd, d.nj1[g0], d.nj2[g0], g, df);
          for (int m = 0; m < 4; m++) {
            duj[m] = -0.5 * (df[m] * d.sj[gp] - w * st.lcj[cp] * dd[m]);
            if (visc) duj[m] = duj[m] + st.lvj[cp] * dd[m];
          }
        }
        for (int m = 0; m < 4; m++) st.du[m * N + c] = st.du[m * N + c] + (dui[m] + duj[m]) / alfa;
      }
      __syncthreads();
    }
    // U <- U + dU on cells 3..n-3
    const int nci = (nx - 5) * (ny - 5);
    for (int t = tid; t < nci; t += T) {
      const int i = t % (nx - 5) + 3, j = t / (nx - 5) + 3;
      const int c = cidx(d, i, j);
      for (int m = 0; m < 4; m++) d.U[m * N + c] = d.U[m * N + c] + st.du[m * N + c];
    }
    __syncthreads();
  }
  if (tid == 0) st.iters[0] = it;
}

// U0 <- U and the synthetic correction of S on cells 3..n-3 (src/Hybrid_Sub.f90:108-153)
__global__ void k_ns_feedback(Dev d, SyntheticState st) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nci = (d.nx - 5) * (d.ny - 5);
  if (t >= nci) return;
  const int i = t % (d.nx - 5) + 3, j = t / (d.nx - 5) + 3;
  const int N = d.ncell, c = cidx(d, i, j);
  const double sq = sqrt(2.0 / d.gamma), cv = d.cv;
  for (int m = 0; m < 4; m++) d.U0[m * N + c] = d.U[m * N + c];
  const double *U = d.U, *U2 = d.U2;
  const double rns = U[c], uns = U[N + c] / U[c], vns = U[2 * N + c] / U[c];
  const double tns = (U[3 * N + c] - 0.5 * U[c] * (uns * uns + vns * vns)) / (cv * U[c]);
  const double rk = U2[c], uk = U2[N + c] / U2[c], vk = U2[2 * N + c] / U2[c];
  const double tk = (U2[3 * N + c] - 0.5 * U2[c] * (uk * uk + vk * vk)) / (cv * U2[c]);
  double s0[4];
  s0[0] = rns - rk;
  s0[1] = (uns - uk) * d.ma / sq;
  s0[2] = (vns - vk) * d.ma / sq;
  s0[3] = tns - tk;
  for (int k = 0; k < 4; k++) {
    d.S[k * N + c] = d.S[k * N + c] + s0[k];
    d.S0[k * N + c] = s0[k] + st.s00[k * N + c];
  }
}

int synthetic_init(Dev &d, const gsis_params &p, SyntheticState &st) {
  const size_t N = d.ncell;
  const size_t n_dbl = 4 * N * 5 + 5 * N;
  void *mem = nullptr;
  if (hipMalloc(&mem, n_dbl * 8 + 64) != hipSuccess || hipMemset(mem, 0, n_dbl * 8 + 64) != hipSuccess) {
    g_syn_err = "synthetic solve: device allocation failed";
    return GSIS_EHIP;
  }
  double *b = reinterpret_cast<double *>(mem);
  st.s00 = b; b += 4 * N;
  st.res = b; b += 4 * N;
  st.du = b; b += 4 * N;
  st.fi = b; b += 4 * N;
  st.fj = b; b += 4 * N;
  st.dtl = b; b += N;
  st.lci = b; b += N;
  st.lcj = b; b += N;
  st.lvi = b; b += N;
  st.lvj = b; b += N;
  st.iters = reinterpret_cast<int *>(b);
  st.block = mem;
  st.tol = p.ns_tol;
  st.maxit = p.ns_maxit > 0 ? p.ns_maxit : 100000;
  st.local_dt = p.local_dt;
  st.if_viscous = p.if_viscous;
  st.cfl = p.cfl;
  st.dtmax = p.dtmax;
  st.dtmin = p.dtmin;
  st.on = true;
  (void)d;
  return 0;
}

int synthetic_step(const Dev &d, SyntheticState &st, hipStream_t s, int *iters) {
  if (!st.on) {
    g_syn_err = "synthetic solve not initialised";
    return GSIS_EARG;
  }
  const int nc = (d.nx - 1) * (d.ny - 1), nci = (d.nx - 5) * (d.ny - 5);
  hipLaunchKernelGGL(k_ns_prepare, dim3((nc + 255) / 256), dim3(256), 0, s, d, st);
  hipLaunchKernelGGL(k_ns_loop, dim3(1), dim3(1024), 0, s, d, st);
  if (nci > 0) hipLaunchKernelGGL(k_ns_feedback, dim3((nci + 255) / 256), dim3(256), 0, s, d, st);
  (void)iters;
  return 0;
}

int synthetic_read_iters(SyntheticState &st, hipStream_t s, int *iters, int *bad_t) {
  int h[2] = {0, 0};
  if (!st.on) {
    if (iters) *iters = 0;
    if (bad_t) *bad_t = 0;
    return 0;
  }
  if (hipMemcpyAsync(h, st.iters, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    g_syn_err = "synthetic solve: reading the iteration counter failed";
    return GSIS_EHIP;
  }
  if (iters) *iters = h[0];
  if (bad_t) *bad_t = h[1];
  return 0;
}

void synthetic_free(SyntheticState &st) {
  if (st.block) (void)hipFree(st.block);
  st.block = nullptr;
  st.on = false;
}

}  // namespace gsis
