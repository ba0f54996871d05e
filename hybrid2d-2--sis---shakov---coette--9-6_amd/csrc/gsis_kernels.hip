// gsis_kernels.hip -- hand-written gfx950 kernels of one GSIS outer iteration.
//
// Every kernel names the reference routine whose arithmetic it carries; the
// layout is described in gsis_device.h.  Storage/compute type T is double for
// the FP64 path and float for the FP32-sweep path (moments, wall sums and the
// macroscopic fields stay FP64 in both).
#include <type_traits>

#include "gsis_kernels.h"
#include "gsis_ns_common.h"

namespace gsis {

// ------------------------------------------------------------------------
// small helpers
// ------------------------------------------------------------------------
// FMINMD (src/Functions.f90:10-26): the argument of smaller magnitude when
// both have the same sign, else 0 -- the median of (x, y, 0), written without
// branches (v_med3_f32 in FP32).
__device__ __forceinline__ double fminmd(double x, double y) { return fmax(fmin(x, y), fmin(fmax(x, y), 0.0)); }
__device__ __forceinline__ float fminmd(float x, float y) { return __builtin_amdgcn_fmed3f(x, y, 0.0f); }
// positive / negative parts (a+|a|)/2 and (a-|a|)/2 of the reference's
// flux splitting, i.e. max(a,0) and min(a,0)
template <typename T>
__device__ __forceinline__ T posp(T a) { return fmax(a, T(0)); }
template <typename T>
__device__ __forceinline__ T negp(T a) { return fmin(a, T(0)); }
// 1/a for a > 0 (the LU-SGS diagonal, the WENO3 weight sum): hardware reciprocal estimate and two
// Newton-Raphson steps r <- r + r(1 - a r), accurate to about 1 ulp
__device__ __forceinline__ double rcp_pos(double a) {
  double r = __builtin_amdgcn_rcp(a);
  double e = fma(-a, r, 1.0);
  r = fma(r, e, r);
  e = fma(-a, r, 1.0);
  return fma(r, e, r);
}
__device__ __forceinline__ float rcp_pos(float a) {
  float r = __builtin_amdgcn_rcpf(a);
  const float e = fmaf(-a, r, 1.0f);
  return fmaf(r, e, r);
}

// WENO3 reconstruction (RHSFDVsub, src/GKUA_Residual.f90:167-185 and the
// wall-adjacent faces :202-250, :333-380).  With D_k the cell differences of
// the stencil, IS_k = D_k^2, alpha_k = C_k / (IS_k + eps)^2 and the
// normalised weights omega_k, the reference's two states are
//   UL = F(f-1) + 1/2 (omega_1 D(f-1) + omega_2 D(f))    (C_1 = 1/3, C_2 = 2/3)
//   UR = F(f)   - 1/2 (omega_1 D(f)   + omega_2 D(f+1))  (C_1 = 2/3, C_2 = 1/3)
// i.e. the NND form with the minmod slope replaced by a weighted mean of the
// difference across the face (dc, weight 2/3) and the outer one (dn, 1/3).
// Clearing the denominators (the REAL(4) literals 2./3. and 1./3. differ by
// exactly a factor 2) leaves one division:
//   slope = (bc dn + 2 bn dc) / (bc + 2 bn),  b = (D^2 + eps)^2.
// eps is the REAL(4) literal 1.E-15 (src/GKUA_Residual.f90:143).
constexpr int kSchemeWeno3 = 3;  // Iflag_scheme of WENO3 (src/COMMONDATA.f90:18)
// Sum of x over the 64 lanes of a wavefront, valid in lane 63, by DPP moves
// (quad_perm, half-row and row mirrors, then row_bcast:15 and row_bcast:31
// into the upper rows) instead of lane shuffles through LDS.
__device__ __forceinline__ double dpp_mov(double x, int ctrl, int row_mask) {
  const unsigned long long b = __double_as_longlong(x);
  int lo = (int)(unsigned)b, hi = (int)(unsigned)(b >> 32);
  switch (ctrl) {  // the DPP control must be a constant
    case 0xB1: lo = __builtin_amdgcn_update_dpp(0, lo, 0xB1, 0xF, 0xF, false);
               hi = __builtin_amdgcn_update_dpp(0, hi, 0xB1, 0xF, 0xF, false); break;
    case 0x4E: lo = __builtin_amdgcn_update_dpp(0, lo, 0x4E, 0xF, 0xF, false);
               hi = __builtin_amdgcn_update_dpp(0, hi, 0x4E, 0xF, 0xF, false); break;
    case 0x141: lo = __builtin_amdgcn_update_dpp(0, lo, 0x141, 0xF, 0xF, false);
                hi = __builtin_amdgcn_update_dpp(0, hi, 0x141, 0xF, 0xF, false); break;
    case 0x140: lo = __builtin_amdgcn_update_dpp(0, lo, 0x140, 0xF, 0xF, false);
                hi = __builtin_amdgcn_update_dpp(0, hi, 0x140, 0xF, 0xF, false); break;
    case 0x142: lo = __builtin_amdgcn_update_dpp(0, lo, 0x142, 0xA, 0xF, false);
                hi = __builtin_amdgcn_update_dpp(0, hi, 0x142, 0xA, 0xF, false); break;
    default: lo = __builtin_amdgcn_update_dpp(0, lo, 0x143, 0xC, 0xF, false);
             hi = __builtin_amdgcn_update_dpp(0, hi, 0x143, 0xC, 0xF, false); break;
  }
  (void)row_mask;
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ double wave_sum_last(double x) {
  x += dpp_mov(x, 0xB1, 0xF);   // quad_perm [1,0,3,2]
  x += dpp_mov(x, 0x4E, 0xF);   // quad_perm [2,3,0,1]: quad sums
  x += dpp_mov(x, 0x141, 0xF);  // row_half_mirror: sums of 8
  x += dpp_mov(x, 0x140, 0xF);  // row_mirror: row sums of 16
  x += dpp_mov(x, 0x142, 0xA);  // row_bcast:15 into rows 1, 3
  x += dpp_mov(x, 0x143, 0xC);  // row_bcast:31 into rows 2, 3
  return x;
}

constexpr double kEpsWeno = 1.0000000036274937e-15;  // (double)1.0e-15f
template <typename T>
__device__ __forceinline__ T wslope(T dc, T dn) {
  const T sc = fma(dc, dc, T(kEpsWeno)), sn = fma(dn, dn, T(kEpsWeno));
  const T bc = sc * sc, bn2 = T(2) * (sn * sn);
  return fma(bc, dn, bn2 * dc) * rcp_pos(bc + bn2);
}
// limited slope of the NND (SL = 0: minmod) or WENO3 (SL = 1) state
template <int SL, typename T>
__device__ __forceinline__ T slope(T dc, T dn) {
  if constexpr (SL == 0) return fminmd(dc, dn);
  else return wslope(dc, dn);
}

// Block-wide sum of N doubles (wave64 shuffles, then one LDS pass).  All
// threads receive the totals.  blockDim.x must be a multiple of 64, <= 1024.
template <int N>
__device__ void block_sum(double (&v)[N]) {
  __shared__ double sh[N][16];
#pragma unroll
  for (int k = 0; k < N; k++)
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < N; k++) sh[k][wid] = v[k];
  __syncthreads();
  if (threadIdx.x < N) {
    double s = 0.0;
    for (int w = 0; w < nw; w++) s += sh[threadIdx.x][w];
    sh[threadIdx.x][0] = s;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; k++) v[k] = sh[k][0];
  __syncthreads();
}

__device__ int block_min_int(int x) {
  __shared__ int shm[16];
  for (int off = 32; off > 0; off >>= 1) x = min(x, __shfl_down(x, off, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if (lane == 0) shm[wid] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    int m = shm[0];
    for (int w = 1; w < nw; w++) m = min(m, shm[w]);
    shm[0] = m;
  }
  __syncthreads();
  int r = shm[0];
  __syncthreads();
  return r;
}

// ------------------------------------------------------------------------
// S0 <- S at the start of the step (src/Hybrid_Sub.f90:29-38)
// ------------------------------------------------------------------------
__global__ void k_save_s0(Dev d) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int ncl = (d.nx - 1) * (d.ny - 1);
  if (t >= ncl) return;
  int i = t % (d.nx - 1) + 1, j = t / (d.nx - 1) + 1;
  int c = cidx(d, i, j);
  for (int k = 0; k < 10; k++) d.S0[k * d.ncell + c] = d.S[k * d.ncell + c];
}

// ------------------------------------------------------------------------
// Diffuse wall: ghost-slot extrapolation (GKUA_boundary_wall,
// src/GKUA_boundary.f90:69-97), wall density (NWGaussLegendre,
// src/GKUA_wall_functions.f90:106-199) and incoming Maxwellian
// (GKUA_wall_boundary, src/GKUA_wall_functions.f90:27-63).
// One block per face node (the last node of the face is skipped, as in the
// reference's icheck/jcheck test).
// ------------------------------------------------------------------------
template <typename T>
__device__ void wall_node(const Dev &d, const BcDev &b, int kk, T *FG, T *FH, T *slot) {
  int i, j, px, py, px1, py1, pos;
  double swnx, swny, st;
  if (b.face == 1) { i = b.ist; j = b.jst + kk; px = 1; py = j; px1 = 2; py1 = j; pos = j + 1; }
  else if (b.face == 2) { i = b.ist + kk; j = b.jst; px = i; py = 1; px1 = i; py1 = 2; pos = i + 1; }
  else if (b.face == 3) { i = b.iend; j = b.jst + kk; px = d.nx - 1; py = j; px1 = d.nx - 2; py1 = j; pos = j + 1; }
  else { i = b.ist + kk; j = b.jend; px = i; py = d.ny - 1; px1 = i; py1 = d.ny - 2; pos = i + 1; }
  if (b.face == 1 || b.face == 3) { swnx = d.ni1[gidx(d, i, j)]; swny = d.ni2[gidx(d, i, j)]; }
  else { swnx = d.nj1[gidx(d, i, j)]; swny = d.nj2[gidx(d, i, j)]; }
  st = (b.face <= 2) ? 1.0 : -1.0;
  const size_t c0 = (size_t)cidx(d, px, py) * d.nv, c1 = (size_t)cidx(d, px1, py1) * d.nv;
  const double tw = d.twall;
  // exp(-|c|^2/Tw) = exp(-cx^2/Tw) exp(-cy^2/Tw): nvi + nvj exponentials per
  // face node, in LDS
  __shared__ double ex[kMaxV1], ey[kMaxV1];
  for (int k = threadIdx.x; k < d.nvi; k += blockDim.x) {
    const double cx = d.vi[k] - b.uw;
    ex[k] = exp(-(cx * cx) / tw);
  }
  for (int k = threadIdx.x; k < d.nvj; k += blockDim.x) {
    const double cy = d.vj[k] - b.vw;
    ey[k] = exp(-(cy * cy) / tw);
  }
  __syncthreads();
  // velocity v = jv*nvi + iv, stepped by blockDim.x without a division
  const int dj = (int)blockDim.x / d.nvi, di = (int)blockDim.x - dj * d.nvi;
  double acc[2] = {0.0, 0.0};
  int first_in = 0x7fffffff;
  {
    int iv = (int)threadIdx.x % d.nvi, jv = (int)threadIdx.x / d.nvi;
    for (int v = threadIdx.x; v < d.nv; v += blockDim.x) {
      T outg, outh;
      {
        T f1 = FG[c0 + v], f2 = FG[c1 + v];
        T t = (-f2 + T(3) * f1) * T(0.5);
        if (t <= T(0)) t = f1;
        if (t >= T(1.5) * f1) t = T(1.5) * f1;
        outg = t;
        f1 = FH[c0 + v]; f2 = FH[c1 + v];
        t = (-f2 + T(3) * f1) * T(0.5);
        if (t <= T(0)) t = f1;
        if (t >= T(1.5) * f1) t = T(1.5) * f1;
        outh = t;
      }
      slot[slot_idx(d, b.face, pos, 0, v)] = outg;
      slot[slot_idx(d, b.face, pos, 1, v)] = outh;
      const double w = d.wi[iv] * d.wj[jv];
      const double cx = d.vi[iv] - b.uw, cy = d.vj[jv] - b.vw;
      const double vnw = (cx * swnx + cy * swny) * st;
      const double vnn = (vnw - fabs(vnw)) / 2.;
      acc[0] += w * (vnn * (double)outg);
      acc[1] += w * ((vnw + fabs(vnw)) / 2. * (ex[iv] * ey[jv]) / (kPi * tw));
      if (vnw >= 0.0) first_in = min(first_in, v);
      iv += di; jv += dj;
      if (iv >= d.nvi) { iv -= d.nvi; jv += 1; }
    }
  }
  block_sum<2>(acc);
  first_in = block_min_int(first_in);
  const double rho = -acc[0] / acc[1];
  {
    int iv = (int)threadIdx.x % d.nvi, jv = (int)threadIdx.x / d.nvi;
    for (int v = threadIdx.x; v < d.nv; v += blockDim.x) {
      const double cx = d.vi[iv] - b.uw, cy = d.vj[jv] - b.vw;
      const double vnw = (cx * swnx + cy * swny) * st;
      if (vnw >= 0.0) {
        // the reference clamps a negative wall density only after its first use
        // (src/GKUA_wall_functions.f90:52-54)
        const double r = (v == first_in || rho >= 0.0) ? rho : 0.0;
        const double g = r * (ex[iv] * ey[jv]) / (kPi * tw);
        slot[slot_idx(d, b.face, pos, 0, v)] = T(g);
        slot[slot_idx(d, b.face, pos, 1, v)] = T(tw * g / 2.);
      }
      iv += di; jv += dj;
      if (iv >= d.nvi) { iv -= d.nvi; jv += 1; }
    }
  }
}

template <typename T>
__global__ void k_wall(Dev d, BcDev b, T *FG, T *FH, T *slot) {
  wall_node(d, b, blockIdx.x, FG, FH, slot);
}
// All diffuse walls of the block in one launch: a wall reads only interior
// cells and writes only its own face's slot, so the walls are independent of
// each other and of the far-field / periodic ghost fills.
template <typename T>
__global__ void k_walls(Dev d, WallSet ws, T *FG, T *FH, T *slot) {
  int w = 0;
  while (w + 1 < ws.n && (int)blockIdx.x >= ws.start[w + 1]) w++;
  wall_node(d, ws.b[w], (int)blockIdx.x - ws.start[w], FG, FH, slot);
}

// ------------------------------------------------------------------------
// Kinetic far field: Maxwellian ghosts from the NS state
// (GKUA_boundary_farfield, src/GKUA_boundary.f90:103-215).
// One block per (face node, layer k).
// ------------------------------------------------------------------------
template <typename T>
__global__ void k_farfield(Dev d, BcDev b, T *FG, T *FH) {
  const int nn = (b.face == 1 || b.face == 3) ? (b.jend - b.jst + 1) : (b.iend - b.ist + 1);
  const int m = blockIdx.x % nn, k = blockIdx.x / nn + 1;
  int i, j;
  if (b.face == 1 || b.face == 3) { i = b.ist; j = b.jst + m; } else { i = b.ist + m; j = b.jst; }
  int i1, j1, i2, j2;
  if (b.face == 1 || b.face == 2) { i1 = i; j1 = j; }
  else if (b.face == 3) { i1 = i - 1; j1 = j; }
  else { i1 = i; j1 = j - 1; }
  if (b.face == 1) { i2 = i - k; j2 = j; }
  else if (b.face == 2) { i2 = i; j2 = j - k; }
  else if (b.face == 3) { i2 = i + k - 1; j2 = j; }
  else { i2 = i; j2 = j + k - 1; }
  double dx, dy, s, n1, n2;
  if (b.face == 1 || b.face == 3) {
    dx = d.xn[nidx(d, i, j + 1)] - d.xn[nidx(d, i, j)];
    dy = d.yn[nidx(d, i, j + 1)] - d.yn[nidx(d, i, j)];
    s = sqrt(dx * dx + dy * dy);
    if (b.face == 1) { n1 = -dy / s; n2 = dx / s; } else { n1 = dy / s; n2 = -dx / s; }
  } else {
    dx = d.xn[nidx(d, i + 1, j)] - d.xn[nidx(d, i, j)];
    dy = d.yn[nidx(d, i + 1, j)] - d.yn[nidx(d, i, j)];
    s = sqrt(dx * dx + dy * dy);
    if (b.face == 2) { n1 = dy / s; n2 = -dx / s; } else { n1 = -dy / s; n2 = dx / s; }
  }
  const double g = d.gamma, ma = d.ma;
  const double d_inf = 1., u_inf = cos(d.aoa), v_inf = sin(d.aoa), p_inf = 1. / (g * ma * ma);
  const double ma_n = (u_inf * n1 + v_inf * n2) * ma;
  const int ca = cidx(d, i1, j1), cb = cidx(d, i2, j2);
  double u4[4];
  if (ma_n > 0.) {
    for (int q = 0; q < 4; q++) u4[q] = d.U[q * d.ncell + ca];
  } else {
    u4[0] = d_inf; u4[1] = d_inf * u_inf; u4[2] = d_inf * v_inf;
    u4[3] = p_inf / (g - 1.) + 0.5 * d_inf * (u_inf * u_inf + v_inf * v_inf);
  }
  const double rho = u4[0], tu = u4[1] / rho, tv = u4[2] / rho;
  const double sq = sqrt(2.0 / g);
  const double n0 = u4[0];
  const double uu0 = (u4[1] / rho) * ma / sq, vv0 = (u4[2] / rho) * ma / sq;
  const double tt = (u4[3] - 0.5 * rho * (tu * tu + tv * tv)) / (d.cv * rho);
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 0; q < 4; q++) d.U[q * d.ncell + cb] = u4[q];
    d.S[0 * d.ncell + cb] = n0;
    d.S[1 * d.ncell + cb] = uu0;
    d.S[2 * d.ncell + cb] = vv0;
    d.S[3 * d.ncell + cb] = tt;
    d.S[4 * d.ncell + cb] = tt;
  }
  const size_t cf = (size_t)cb * d.nv;
  for (int v = threadIdx.x; v < d.nv; v += blockDim.x) {
    const int iv = v % d.nvi, jv = v / d.nvi;
    const double viu = d.vi[iv] - uu0, vjv = d.vj[jv] - vv0;
    const double vij2 = (viu * viu + vjv * vjv) / tt;
    const double gd = n0 / (kPi * tt) * exp(-vij2);
    FG[cf + v] = T(gd);
    FH[cf + v] = T(gd * tt / 2.);
  }
}

// ------------------------------------------------------------------------
// Periodic copy ("inlet", src/GKUA_boundary.f90:219-273): one block per
// (face node, layer).
// ------------------------------------------------------------------------
template <typename T>
__global__ void k_periodic(Dev d, BcDev b, T *FG, T *FH) {
  const int nn = (b.face == 1 || b.face == 3) ? (b.jend - b.jst + 1) : (b.iend - b.ist + 1);
  const int m = blockIdx.x % nn, k = blockIdx.x / nn + 1;
  int i, j;
  if (b.face == 1 || b.face == 3) { i = b.ist; j = b.jst + m; } else { i = b.ist + m; j = b.jst; }
  int i1, j1, i2, j2;
  if (b.face == 1) { i1 = d.nx - k; j1 = j; i2 = i - k; j2 = j; }
  else if (b.face == 2) { i1 = i; j1 = d.ny - k; i2 = i; j2 = j - k; }
  else if (b.face == 3) { i1 = k; j1 = j; i2 = i + k - 1; j2 = j; }
  else { i1 = i; j1 = k; i2 = i; j2 = j + k - 1; }
  const int ca = cidx(d, i1, j1), cb = cidx(d, i2, j2);
  if (threadIdx.x < 11) d.S[threadIdx.x * d.ncell + cb] = d.S[threadIdx.x * d.ncell + ca];
  const size_t sa = (size_t)ca * d.nv, sb = (size_t)cb * d.nv;
  for (int v = threadIdx.x; v < d.nv; v += blockDim.x) {
    FG[sb + v] = FG[sa + v];
    FH[sb + v] = FH[sa + v];
  }
}

// ------------------------------------------------------------------------
// Inner face between two blocks: GKUA_inner_mesco_boundary and
// GKUA_inner_macro_boundary (src/GKUA_boundary.f90:400-633).  The LAP = 2 cell
// layers next to the neighbour's subface are copied into this block's ghost
// layers: FRD(1:2) for every velocity and S_GKUA(1:10).  Position k = 1..n1
// along this block's face takes the neighbour's position k (orient == 2) or
// n1+1-k (any other orient).  One workgroup per (position, layer); threads
// stride over velocities.  The neighbour's cells are interior cells and this
// block's are ghost cells, so a block may be its own neighbour.
// ------------------------------------------------------------------------
__host__ __device__ inline void inner_src_cell(const BcDev &b, int m, int l, int &i, int &j) {
  if (b.face == 1) { i = b.ist + l - 1; j = b.jst + m - 1; }
  else if (b.face == 2) { i = b.ist + m - 1; j = b.jst + l - 1; }
  else if (b.face == 3) { i = b.iend - l; j = b.jst + m - 1; }
  else { i = b.ist + m - 1; j = b.jend - l; }
}
__host__ __device__ inline void inner_ghost_cell(const BcDev &b, int k, int l, int &i, int &j) {
  if (b.face == 1) { i = b.ist - l; j = b.jst + k - 1; }
  else if (b.face == 2) { i = b.ist + k - 1; j = b.jst - l; }
  else if (b.face == 3) { i = b.iend + l - 1; j = b.jst + k - 1; }
  else { i = b.ist + k - 1; j = b.jend + l - 1; }
}

template <typename T>
__global__ void k_inner(Dev dd, BcDev bd, T *FGd, T *FHd, Dev ds, BcDev bs, const T *FGs, const T *FHs, int n1,
                        int orient) {
  const int k = blockIdx.x % n1 + 1, l = blockIdx.x / n1 + 1;
  const int k1 = orient == 2 ? k : n1 + 1 - k;
  int ia, ja, ib, jb;
  inner_src_cell(bs, k1, l, ia, ja);
  inner_ghost_cell(bd, k, l, ib, jb);
  const int ca = cidx(ds, ia, ja), cb = cidx(dd, ib, jb);
  if (threadIdx.x < 10) dd.S[threadIdx.x * dd.ncell + cb] = ds.S[threadIdx.x * ds.ncell + ca];
  const size_t sa = (size_t)ca * ds.nv, sb = (size_t)cb * dd.nv;
  for (int v = threadIdx.x; v < dd.nv; v += blockDim.x) {
    FGd[sb + v] = FGs[sa + v];
    FHd[sb + v] = FHs[sa + v];
  }
}

// ------------------------------------------------------------------------
// Kinetic RHS for every interior cell and velocity: reduced Shakhov
// equilibrium (Get_Feq, src/GKUA_Residual.f90:420-469), NND upwind fluxes with
// the wall/farfield face overrides and the BGK-Shakhov source (RHSFDVsub,
// src/GKUA_Residual.f90:117-418).  Also reduces the four conserved moments
// of the RHS per cell (GaussLegendre_rs, src/GKUA_DETFP_RES.f90:81-120).
// One block per interior cell; threads stride over velocities.
// ------------------------------------------------------------------------
// sch: GSIS scheme code (0 UP1, 1 NND4A2, 3 WENO3), uniform over the launch
template <typename T>
__device__ __forceinline__ T flux_nnd(T a, T fm2, T fm1, T f0, T fp1, int sch) {
  const T asp = posp(a), asn = negp(a);
  if (sch == 0) return asp * fm1 + asn * f0;
  const T d0 = f0 - fm1;
  const T sl = sch == 3 ? slope<1>(d0, fm1 - fm2) : slope<0>(d0, fm1 - fm2);
  const T sr = sch == 3 ? slope<1>(d0, fp1 - f0) : slope<0>(d0, fp1 - f0);
  return asp * (fm1 + T(0.5) * sl) + asn * (f0 - T(0.5) * sr);
}
// face next to a wall/farfield face on the low side: unlimited (central)
// left state (src/GKUA_Residual.f90:202-220)
template <typename T>
__device__ __forceinline__ T flux_lowadj(T a, T f1, T f2, T f3, int sch) {
  const T asp = posp(a), asn = negp(a);
  if (sch == 0) return asp * f1 + asn * f2;
  const T ulg = f1 + T(0.5) * (f2 - f1);
  const T sr = sch == 3 ? slope<1>(f2 - f1, f3 - f2) : slope<0>(f2 - f1, f3 - f2);
  return asp * ulg + asn * (f2 - T(0.5) * sr);
}
// face next to a wall/farfield face on the high side: unlimited (central)
// right state (src/GKUA_Residual.f90:239-257); f1 = F(n-1), f2 = F(n-2),
// f3 = F(n-3)
template <typename T>
__device__ __forceinline__ T flux_highadj(T a, T f1, T f2, T f3, int sch) {
  const T asp = posp(a), asn = negp(a);
  if (sch == 0) return asp * f2 + asn * f1;
  const T sl = sch == 3 ? slope<1>(f1 - f2, f2 - f3) : slope<0>(f1 - f2, f2 - f3);
  const T urg = f1 - T(0.5) * (f1 - f2);
  return asp * (f2 + T(0.5) * sl) + asn * urg;
}

// flux through i-face fi of row j (between cells fi-1 and fi); stencil
// F(fi-2), F(fi-1), F(fi), F(fi+1)
template <typename T>
__device__ __forceinline__ T iface_flux(const Dev &d, int fi, int j, T a, T fm2, T fm1, T f0, T fp1,
                                        const T *slot, int ird, int v, int sch) {
  if (d.ovrI1[j]) {
    if (fi == 1) return a * slot[slot_idx(d, 1, j + 1, ird, v)];
    if (fi == 2) return flux_lowadj(a, fm1, f0, fp1, sch);
  }
  if (d.ovrI3[j]) {
    if (fi == d.nx) return a * slot[slot_idx(d, 3, j + 1, ird, v)];
    if (fi == d.nx - 1) return flux_highadj(a, f0, fm1, fm2, sch);
  }
  return flux_nnd(a, fm2, fm1, f0, fp1, sch);
}
template <typename T>
__device__ __forceinline__ T jface_flux(const Dev &d, int i, int fj, T a, T fm2, T fm1, T f0, T fp1,
                                        const T *slot, int ird, int v, int sch) {
  if (d.ovrJ2[i]) {
    if (fj == 1) return a * slot[slot_idx(d, 2, i + 1, ird, v)];
    if (fj == 2) return flux_lowadj(a, fm1, f0, fp1, sch);
  }
  if (d.ovrJ4[i]) {
    if (fj == d.ny) return a * slot[slot_idx(d, 4, i + 1, ird, v)];
    if (fj == d.ny - 1) return flux_highadj(a, f0, fm1, fm2, sch);
  }
  return flux_nnd(a, fm2, fm1, f0, fp1, sch);
}

template <typename T>
__global__ void __launch_bounds__(256) k_rhs(Dev d, const T *FG, const T *FH, const T *slot, T *RG, T *RH) {
  __shared__ T ex[kMaxV1], ey[kMaxV1];
  const int t = blockIdx.x;
  const int i = t % (d.nx - 1) + 1, j = t / (d.nx - 1) + 1;
  const int c = cidx(d, i, j);
  const double n = d.S[0 * d.ncell + c], uu = d.S[1 * d.ncell + c], vv = d.S[2 * d.ncell + c];
  const double tt = d.S[3 * d.ncell + c], qx = d.S[7 * d.ncell + c], qy = d.S[8 * d.ncell + c];
  // exp(-|c|^2/T) = exp(-cx^2/T) exp(-cy^2/T): nvi+nvj exponentials per cell
  for (int q = threadIdx.x; q < d.nvi; q += blockDim.x) {
    const double cx = d.vi[q] - uu;
    ex[q] = T(exp(-(cx * cx) / tt));
  }
  for (int q = threadIdx.x; q < d.nvj; q += blockDim.x) {
    const double cy = d.vj[q] - vv;
    ey[q] = T(exp(-(cy * cy) / tt));
  }
  __syncthreads();
  const T crt = T(n / (kPi * tt));
  const T cpt = T(d.cpr / (n * tt * tt));
  const T Tt = T(tt), Uu = T(uu), Vv = T(vv), Qx = T(qx), Qy = T(qy);
  const T vol = T(d.vol[gidx(d, i, j)]), cf = T(d.CF[c]);
  const double sil = d.si[gidx(d, i, j)], sir = d.si[gidx(d, i + 1, j)];
  const double nil1 = d.ni1[gidx(d, i, j)], nil2 = d.ni2[gidx(d, i, j)];
  const double nir1 = d.ni1[gidx(d, i + 1, j)], nir2 = d.ni2[gidx(d, i + 1, j)];
  const double sjb = d.sj[gidx(d, i, j)], sjt = d.sj[gidx(d, i, j + 1)];
  const double njb1 = d.nj1[gidx(d, i, j)], njb2 = d.nj2[gidx(d, i, j)];
  const double njt1 = d.nj1[gidx(d, i, j + 1)], njt2 = d.nj2[gidx(d, i, j + 1)];
  const int sch = d.scheme;
  const size_t nv = d.nv;
  const size_t o0 = (size_t)c * nv;
  const size_t oim2 = (size_t)cidx(d, i - 2, j) * nv, oim1 = (size_t)cidx(d, i - 1, j) * nv;
  const size_t oip1 = (size_t)cidx(d, i + 1, j) * nv, oip2 = (size_t)cidx(d, i + 2, j) * nv;
  const size_t ojm2 = (size_t)cidx(d, i, j - 2) * nv, ojm1 = (size_t)cidx(d, i, j - 1) * nv;
  const size_t ojp1 = (size_t)cidx(d, i, j + 1) * nv, ojp2 = (size_t)cidx(d, i, j + 2) * nv;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int v = threadIdx.x; v < d.nv; v += blockDim.x) {
    const int iv = v % d.nvi, jv = v / d.nvi;
    const double vxd = d.vi[iv], vyd = d.vj[jv];
    const T axl = T((vxd * nil1 + vyd * nil2) * sil), axr = T((vxd * nir1 + vyd * nir2) * sir);
    const T ayb = T((vxd * njb1 + vyd * njb2) * sjb), ayt = T((vxd * njt1 + vyd * njt2) * sjt);
    // Shakhov equilibrium of this cell (Get_Feq)
    const T viu = T(vxd) - Uu, vjv = T(vyd) - Vv;
    const T vij2 = (viu * viu + vjv * vjv) / Tt;
    const T fvx = cpt * (viu * Qx + vjv * Qy);
    const T gd = crt * (ex[iv] * ey[jv]);
    const T feqg = gd * (T(1) + fvx * (T(2) * vij2 - T(4)));
    const T feqh = (Tt * gd / T(2)) * (T(1) + fvx * (T(2) * vij2 - T(2)));
    T rhs[2];
#pragma unroll
    for (int ird = 0; ird < 2; ird++) {
      const T *F = ird == 0 ? FG : FH;
      const T f0 = F[o0 + v];
      const T fim2 = F[oim2 + v], fim1 = F[oim1 + v], fip1 = F[oip1 + v], fip2 = F[oip2 + v];
      const T fjm2 = F[ojm2 + v], fjm1 = F[ojm1 + v], fjp1 = F[ojp1 + v], fjp2 = F[ojp2 + v];
      const T flx_l = iface_flux(d, i, j, axl, fim2, fim1, f0, fip1, slot, ird, v, sch);
      const T flx_r = iface_flux(d, i + 1, j, axr, fim1, f0, fip1, fip2, slot, ird, v, sch);
      const T fly_b = jface_flux(d, i, j, ayb, fjm2, fjm1, f0, fjp1, slot, ird, v, sch);
      const T fly_t = jface_flux(d, i, j + 1, ayt, fjm1, f0, fjp1, fjp2, slot, ird, v, sch);
      const T net = ((flx_r - flx_l) + fly_t) - fly_b;
      const T feq = ird == 0 ? feqg : feqh;
      rhs[ird] = -net / vol + cf * (feq - f0);
    }
    RG[o0 + v] = rhs[0];
    RH[o0 + v] = rhs[1];
    const double w = d.wi[iv] * d.wj[jv];
    const double g = (double)rhs[0], h = (double)rhs[1];
    const double vx1 = vxd * g, vy1 = vyd * g;
    acc[0] += w * g;
    acc[1] += w * vx1;
    acc[2] += w * vy1;
    acc[3] += w * (h + vxd * vx1 + vyd * vy1);
  }
  block_sum<4>(acc);
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; k++) d.res4[k * d.ncell + c] = acc[k];
}

// ------------------------------------------------------------------------
// LU-SGS sweep of the kinetic equation (src/GKUA_Residual.f90:34-109):
// forward recursion on anti-diagonal planes i+j=p, the non-recursive
// backward pass (it reads forward values only, :82-86), update F+=dF and
// the positivity clamp of interior cells.  One block per tile of VT
// consecutive velocities walks all planes; the forward values of the last
// three planes live in LDS.  F is updated in place (the RHS kernel already
// consumed the old values).
// ------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_sweep(Dev d, int VT, T *FG, T *FH, const T *RG, const T *RH) {
  extern __shared__ unsigned char lds_raw[];
  T *D = reinterpret_cast<T *>(lds_raw);  // [3][nx+1][VT][2]
  const int lane = threadIdx.x % VT, slot_s = threadIdx.x / VT, NS = blockDim.x / VT;
  const int v = blockIdx.x * VT + lane;
  const bool act = v < d.nv;
  const int vv = act ? v : 0;
  const double vxd = d.vi[vv % d.nvi], vyd = d.vj[vv / d.nvi];
  const int nx = d.nx, ny = d.ny;
  const T dt = T(d.dt);
  const int plast = (nx - 1) + (ny - 1);
  auto didx = [&](int ring, int col, int ird) { return ((ring * (nx + 1) + col) * VT + lane) * 2 + ird; };
  auto ax = [&](int i, int j) { const int g = gidx(d, i, j); return T((vxd * d.ni1[g] + vyd * d.ni2[g]) * d.si[g]); };
  auto ay = [&](int i, int j) { const int g = gidx(d, i, j); return T((vxd * d.nj1[g] + vyd * d.nj2[g]) * d.sj[g]); };
  for (int p = 2; p <= plast + 1; p++) {
    if (p <= plast) {
      const int ilo = max(1, p - (ny - 1)), ihi = min(nx - 1, p - 1);
      const int ring = p % 3, rprev = (p - 1) % 3;
      for (int k = slot_s; k <= ihi - ilo; k += NS) {
        const int i = ilo + k, j = p - i;
        const T axi = ax(i, j), axi1 = ax(i + 1, j), ayj = ay(i, j), ayj1 = ay(i, j + 1);
        const T vol = T(d.vol[gidx(d, i, j)]);
        const int c = cidx(d, i, j);
        const T alfa = vol / dt + (posp(axi1) - negp(axi) + posp(ayj1) - negp(ayj)) + vol * T(d.CF[c]);
        const T axp1 = posp(axi), ayp1 = posp(ayj);
        const size_t o = (size_t)c * d.nv + vv;
#pragma unroll
        for (int ird = 0; ird < 2; ird++) {
          const T dl = (i - 1 >= 1) ? D[didx(rprev, i - 1, ird)] : T(0);
          const T db = (j - 1 >= 1) ? D[didx(rprev, i, ird)] : T(0);
          const T r = act ? (ird == 0 ? RG[o] : RH[o]) : T(0);
          D[didx(ring, i, ird)] = (vol * r + axp1 * dl + ayp1 * db) / alfa;
        }
      }
    }
    __syncthreads();
    const int q = p - 1;
    if (q >= 2) {
      const int ilo = max(1, q - (ny - 1)), ihi = min(nx - 1, q - 1);
      const int rq = q % 3, rn = p % 3;
      for (int k = slot_s; k <= ihi - ilo; k += NS) {
        const int i = ilo + k, j = q - i;
        const T axi = ax(i, j), axi1 = ax(i + 1, j), ayj = ay(i, j), ayj1 = ay(i, j + 1);
        const T vol = T(d.vol[gidx(d, i, j)]);
        const int c = cidx(d, i, j);
        const T alfa = vol / dt + (posp(axi1) - negp(axi) + posp(ayj1) - negp(ayj)) + vol * T(d.CF[c]);
        const T axn1 = negp(axi1), ayn1 = negp(ayj1);
        const size_t o = (size_t)c * d.nv + vv;
#pragma unroll
        for (int ird = 0; ird < 2; ird++) {
          const T dr = (i + 1 <= nx - 1) ? D[didx(rn, i + 1, ird)] : T(0);
          const T dtp = (j + 1 <= ny - 1) ? D[didx(rn, i, ird)] : T(0);
          const T du = D[didx(rq, i, ird)] + (-axn1 * dr + -ayn1 * dtp) / alfa;
          if (act) {
            T *F = ird == 0 ? FG : FH;
            T f = F[o] + du;
            if ((double)f < kClamp) f = T(kClamp);
            F[o] = f;
          }
        }
      }
    }
  }
}

// ------------------------------------------------------------------------
// Per-cell constants of the kinetic step, computed once per step so that the
// per-velocity work needs no division, sorted by "wrapped anti-diagonal":
// cell (i,j) goes to plane p = (i+j+2) mod (ny+3), slot i-1, as
// cpar[p*pstride + (i-1)*kNW + k] (the fused kernel's position of row j in a tile is j+3,
// after three virtual rows, and a tile has ny+3 positions).  Each plane then
// holds exactly one cell per mesh column: the cells the fused kernel visits
// in one step.  Values (k): 0 n/(pi T), 1 cpr/(n T^2) (Get_Feq,
// src/GKUA_Residual.f90:449-450), 2 1/T, 3 T, 4 u, 5 v, 6 qx, 7 qy, 8 CF,
// 9 1/vol, 10 vol, 11 vol/dt + vol*CF (the velocity-independent part of the
// LU-SGS diagonal, src/GKUA_Residual.f90:43), then s*(n1,n2) of the faces:
// 12-13 i-face i, 14-15 i-face i+1, 16-17 j-face j, 18-19 j-face j+1.
// ------------------------------------------------------------------------
constexpr int kNW = kNCellPar;
// With pstride = 0 the constants are stored by cell instead, cpar[((j-1)*(nx-1) + i-1)*kNW + k]
// (the layout of the column-wavefront kernel).
template <typename P>
__device__ __forceinline__ void cellpar_body(const Dev &d, P *cpar, int pstride, int t) {
  const int ncol = d.nx - 1, nrow = d.ny - 1;
  if (t >= ncol * nrow) return;
  const int i = t % ncol + 1, j = t / ncol + 1;
  const int c = cidx(d, i, j), N = d.ncell;
  const int r = (i + j + 2) % (nrow + 4);
  P *o = pstride ? cpar + (size_t)r * pstride + (size_t)(i - 1) * kNW : cpar + (size_t)t * kNW;
  const double n = d.S[c], T = d.S[3 * N + c];
  const int g0 = gidx(d, i, j), gi = gidx(d, i + 1, j), gj = gidx(d, i, j + 1);
  const double vol = d.vol[g0], cf = d.CF[c];
  const double vals[kNW] = {n / (kPi * T), d.cpr / (n * T * T), 1.0 / T, T, d.S[N + c], d.S[2 * N + c],
                            d.S[7 * N + c], d.S[8 * N + c], cf, 1.0 / vol, vol, vol / d.dt + vol * cf,
                            d.ni1[g0] * d.si[g0], d.ni2[g0] * d.si[g0], d.ni1[gi] * d.si[gi], d.ni2[gi] * d.si[gi],
                            d.nj1[g0] * d.sj[g0], d.nj2[g0] * d.sj[g0], d.nj1[gj] * d.sj[gj], d.nj2[gj] * d.sj[gj]};
#pragma unroll
  for (int k = 0; k < kNW; k++) o[k] = P(vals[k]);
}
template <typename P>
__global__ void k_cellpar(Dev d, P *cpar, int pstride) {
  cellpar_body(d, cpar, pstride, (int)(blockIdx.x * blockDim.x + threadIdx.x));
}

// Flux through one face in the fused kernel: the rules of iface_flux /
// jface_flux written without branches.  fl holds the override flags of the
// row (i-faces) or column (j-faces): bit 0 low boundary face, bit 1 high one;
// sv is the wall-slot value of the boundary face.  Next to an overridden face
// the limiter on that side is dropped (flux_lowadj / flux_highadj), on it the
// flux is a * slot (src/GKUA_Residual.f90:200-262).
// hs is 1/2 for the NND and WENO3 schemes and 0 for first-order upwind (the
// limited slopes then drop out exactly); SL picks the slope rule (0 minmod,
// 1 WENO3, 2 none).
// The two limited slopes of a face, s(d0, dl) and s(d0, dr).  For minmod in
// FP64, minmod(d0, x) = clamp(x, min(d0, 0), max(d0, 0)) (the same value: x
// when it lies between 0 and d0, else the nearer bound), so the bounds are
// shared by the two slopes: 6 instructions instead of 8.  FP32 keeps
// v_med3_f32, one instruction per slope.
template <int SL, typename T>
__device__ __forceinline__ void slopes(T d0, T dl, T dr, T &sl, T &sr) {
  if constexpr (SL == 2) {  // first-order upwind: no slopes
    sl = T(0);
    sr = T(0);
  } else if constexpr (SL == 0 && std::is_same<T, double>::value) {
    const T lo = fmin(d0, T(0)), hi = fmax(d0, T(0));
    sl = fmin(fmax(dl, lo), hi);
    sr = fmin(fmax(dr, lo), hi);
  } else {
    sl = slope<SL>(d0, dl);
    sr = slope<SL>(d0, dr);
  }
}
template <int SL, typename T>
__device__ __forceinline__ T wflux(int f, int n, int fl, T a, T fm2, T fm1, T f0, T fp1, T sv, T hs) {
  const T asp = posp(a), asn = negp(a);
  const T d0 = f0 - fm1;
  T sl, sr;
  slopes<SL>(d0, fm1 - fm2, fp1 - f0, sl, sr);
  const T liml = ((fl & 1) && f == 2) ? d0 : sl;
  const T limr = ((fl & 2) && f == n - 1) ? d0 : sr;
  const T fx = asp * (fm1 + hs * liml) + asn * (f0 - hs * limr);
  const bool wall = ((fl & 1) && f == 1) || ((fl & 2) && f == n);
  return wall ? a * sv : fx;
}
// wflux for the face at index 1 of its direction.  With n >= 5 (checked by
// gsis_create) neither of wflux's n-1 / n rules can apply, and leaving their
// comparisons out keeps loop-invariant masks of n out of scalar registers.
template <int SL, typename T>
__device__ __forceinline__ T wflux1(int fl, T a, T fm2, T fm1, T f0, T fp1, T sv, T hs) {
  const T asp = posp(a), asn = negp(a);
  const T d0 = f0 - fm1;
  T sl, sr;
  slopes<SL>(d0, fm1 - fm2, fp1 - f0, sl, sr);
  const T fx = asp * (fm1 + hs * sl) + asn * (f0 - hs * sr);
  return (fl & 1) ? a * sv : fx;
}
// interior face (no override) of the fused kernel's fast path
template <int SL, typename T>
__device__ __forceinline__ T iflux(T a, T fm2, T fm1, T f0, T fp1, T hs) {
  if constexpr (SL == 2) return posp(a) * fm1 + negp(a) * f0;
  const T d0 = f0 - fm1;
  T sl, sr;
  slopes<SL>(d0, fm1 - fm2, fp1 - f0, sl, sr);
  return posp(a) * (fm1 + hs * sl) + negp(a) * (f0 - hs * sr);
}
// exp(x) for the equilibrium, x <= 0.  Cody-Waite reduction x = k ln2 + r
// (|r| <= ln2/2, ln2 split into the fdlibm high/low parts so that k*ln2_hi is
// exact) and the degree-13 Taylor polynomial of e^r in Horner form
// (truncation error < 5e-18 relative); exp(x) = 2^k p(r).  Written out here
// rather than calling the library exp because its coefficients stay in
// scalar registers, loaded per call from constant memory through a pointer
// the compiler cannot hoist: the library routine's constants, hoisted out of
// the fused kernel's loop, cost ~40 vector registers and forced spills.
// Accuracy is checked end to end: the fused kernel's fields agree with the
// oracle (libm exp) to <= 1e-10 relative L2.
__constant__ double kExpC[16] = {1.0 / 6227020800.0, 1.0 / 479001600.0, 1.0 / 39916800.0, 1.0 / 3628800.0,
                                 1.0 / 362880.0,     1.0 / 40320.0,     1.0 / 5040.0,     1.0 / 720.0,
                                 1.0 / 120.0,        1.0 / 24.0,        1.0 / 6.0,        0.5,
                                 1.4426950408889634, -6.93147180369123816490e-01, -1.90821492927058770002e-10,
                                 -746.0};
typedef const __attribute__((address_space(4))) double *cdptr;
__device__ __forceinline__ double wexp(double x) {
  cdptr kc = (cdptr)kExpC;
  asm volatile("" : "+s"(kc));
  x = x < kc[15] ? kc[15] : x;  // below -746 the result underflows to 0 anyway
  const double k = rint(x * kc[12]);
  double r = fma(k, kc[13], x);
  r = fma(k, kc[14], r);
  // Horner steps p <- p*r + c with the coefficient read straight from its
  // scalar register (the compiler would copy each one into a vector register
  // pair first to use the accumulating fma form)
  double p = kc[0];
#pragma unroll
  for (int m = 1; m < 12; m++) {
    const double c = kc[m];
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(p) : "v"(p), "v"(r), "s"(c));
  }
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, (int)k);
}
__device__ __forceinline__ float wexp(float x) { return expf(x); }
// element at a 32-bit byte offset from a uniform base (global_load with an
// SGPR base and a VGPR offset)
template <typename T>
__device__ __forceinline__ T ldo(const T *base, unsigned off) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + off);
}
template <typename T>
__device__ __forceinline__ void sto(T *base, unsigned off, T v) {
  *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + off) = v;
}

// ------------------------------------------------------------------------
// Fused kinetic step ("column walker"): Shakhov equilibrium, NND fluxes,
// source, forward LU-SGS recursion, non-recursive backward pass, update and
// interior clamp in ONE pass over F (src/GKUA_Residual.f90:26-109,117-469).
//
// G and H are independent transport problems in this step, so a tile is one
// of them for VT consecutive velocities; F holds the G plane followed by the
// H plane.  Thread (column i, lane) of a block owns mesh column i and one
// velocity of the tile.  A block walks K tiles back to back: at step t column
// i is at position q = t - (i-1) of its sequence (a tile is ny+3 positions:
// three window-filling virtual rows, then rows 1..ny-1), so all active
// columns sit on the same wrapped anti-diagonal and the forward recursion's
// inputs from column i-1 (its face flux and Delta*) were produced one step
// earlier and come through LDS, while the j-stencil F(i, j-2..j+2) slides in
// registers.  Column i finalises position q-2 at the top of step t (backward
// pass + update): it needs Delta*(i+1, .), published by column i+1 at step
// t-1, and its own Delta* of the next row, kept in registers with the
// backward coefficients.  F is updated in place: every other reader of a
// value has consumed it at an earlier step (column i+1 reads it at t+1 and
// waits for it at the top of t+2, before the write).  All global loads of a
// position are issued one step ahead, and the velocity-independent constants
// of the cells two steps ahead are staged cooperatively into LDS.
// Offsets are 32-bit byte offsets from F, which holds the G plane, the H
// plane and the wall slots (slot at byte offset slotb) in one allocation of
// less than 4 GiB.
// ------------------------------------------------------------------------
template <typename T, int SL>
__global__ void __launch_bounds__(1024) k_walk(Dev d, const double *cpar, int VT, int K, T *F, unsigned planeb,
                                               unsigned slotb, T *R, int store_rhs) {
  using T2 = typename std::conditional<std::is_same<T, double>::value, double2, float2>::type;
  constexpr unsigned ES = sizeof(T);
  extern __shared__ double lds_d[];
  const int ncol = d.nx - 1, nrow = d.ny - 1, nv = d.nv, nx = d.nx, ny = d.ny;
  const int nthr = blockDim.x;
  // one plane of cell constants, padded to 3 values per thread so that the
  // cooperative copy needs no bounds test (kNW*ncol <= 3*nthr since VT >= 8)
  const int nconst = 3 * nthr;
  double *C = lds_d;                                            // [2][nconst]: [ncol][kNW] + padding
  double2 *V = reinterpret_cast<double2 *>(C + 2 * nconst);     // [K][VT] velocities of the tiles
  T2 *X = reinterpret_cast<T2 *>(V + K * VT);                   // [2][ncol][VT] (face flux, Delta*)
  uint8_t *OF = reinterpret_cast<uint8_t *>(X + 2 * ncol * VT);  // override flags, rows and columns
  const int ntiles = 2 * ((nv + VT - 1) / VT);
  const int tile0 = blockIdx.x * K;
  const int ktiles = min(K, ntiles - tile0);
  const T hs = d.scheme != 0 ? T(0.5) : T(0);  // NND slope factor (0: upwind)
  const unsigned cb = (unsigned)nv * ES;       // one cell
  const unsigned rb = (unsigned)d.NI * cb;     // one row of cells
  const unsigned L2 = 2u * d.L;
  const int vshift = 31 - __builtin_clz(VT);   // VT is a power of two
  {
    const int tid = threadIdx.x;
    for (int e = tid; e < ktiles * VT; e += nthr) {
      const int tau = tile0 + e / VT, v = (tau >> 1) * VT + e % VT;
      V[e] = v < nv ? make_double2(d.vi[v % d.nvi], d.vj[v / d.nvi]) : make_double2(0.0, 0.0);
    }
    // bits 0-1: i-face overrides of row e; bits 2-3: j-face overrides of column e
    const int nof = max(nrow + 2, d.nx + 2);
    for (int e = tid; e < nof; e += nthr)
      OF[e] = (e <= nrow + 1 ? ((d.ovrI1[e] ? 1 : 0) | (d.ovrI3[e] ? 2 : 0)) : 0) |
              (e < d.nx + 2 ? ((d.ovrJ2[e] ? 4 : 0) | (d.ovrJ4[e] ? 8 : 0)) : 0);
    for (int e = tid; e < 2 * nconst; e += nthr) C[e] = cpar[e];  // planes 0 and 1
  }
  // Per-thread indices are recomputed inside the loop from a laundered copy
  // of threadIdx.x: keeping them live across the loop costs more registers
  // than the few integer instructions that recompute them.
  auto tile_of = [&](int kt, int lane, unsigned &tb, int &ird, int &vv) {
    const int tau = tile0 + kt, v = (tau >> 1) * VT + lane;
    ird = tau & 1;
    const bool act = v < nv;
    vv = act ? v : 0;
    tb = (unsigned)ird * planeb + (unsigned)vv * ES;
    return act;
  };
  // byte offset (from F) of slot[((face-1)*L + pos)*2 + ird][v]
  auto soff = [&](unsigned face, unsigned pos, unsigned ird, unsigned v) {
    return slotb + (((face - 1) * L2 + 2 * pos + ird) * (unsigned)nv + v) * ES;
  };
  // Each tile's sequence has nrow+4 positions: four "virtual" rows -3..0,
  // which only shift F(i,-1..2) of the tile into the register window (and
  // finalise the previous tile's last two rows), then rows 1..nrow.  The
  // window shift is then the same at every step and the value to finalise is
  // always the oldest window entry.
  const int LT = nrow + 4;
  const int nseq = ktiles * LT;
  const int tsteps = (ncol - 1) + nseq + 2;
  // values of the next position, loaded one step ahead
  T pw = T(0), pm1 = T(0), pp1 = T(0), pp2 = T(0), px = T(0), py = T(0);
  __syncthreads();  // OF, V and the first two constants planes
  {
    const int tl = threadIdx.x, lane = tl & (VT - 1), i = (tl >> vshift) + 1;
    unsigned tb;
    int ird, vv;
    tile_of(0, lane, tb, ird, vv);
    pw = ldo(F, tb + (unsigned)(i + 1) * cb);  // F(i, -1) of the first tile
  }
  T w0 = T(0), w1 = T(0), w2 = T(0), w3 = T(0), w4 = T(0);  // F(i, r-2..r+2)
  T fly_prev = T(0);
  // forward values of positions q-1 (n) and q-2 (o): Delta* and the
  // backward-pass coefficients -A^-_x/alfa, -A^-_y/alfa
  T hd_n = T(0), bx_n = T(0), by_n = T(0), hd_o = T(0), bx_o = T(0), by_o = T(0);
  int r = -4, kc = 0;  // row and tile of the current position
  // The constants of step s are loaded at step s-2 and written to LDS buffer
  // s&1 at the top of step s-1, after that step's wait for the loads issued
  // one step earlier; planes 0 and 1 are in place before the loop.
  int rstage = 2 % LT;
  double stg[3] = {0.0, 0.0, 0.0};
  for (int t = 0; t < tsteps; t++) {
    // Everything issued in the previous step (its loads for this step, the
    // stage loads, its store) is complete from here on: one explicit wait,
    // so that no later instruction of the step waits for memory.
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    int tl = threadIdx.x;
    asm volatile("" : "+v"(tl));
    const int lane = tl & (VT - 1), col = tl >> vshift, i = col + 1;
    const int buf = t & 1, pbuf = buf ^ 1;
    const int q = t - col;
    if (t > 0) {
#pragma unroll
      for (int k = 0; k < 3; k++) C[pbuf * nconst + k * nthr + tl] = stg[k];
    }
#pragma unroll
    for (int k = 0; k < 3; k++) stg[k] = ldo(cpar + (size_t)rstage * nconst + k * nthr, (unsigned)tl * 8u);
    rstage = rstage + 1 == LT ? 0 : rstage + 1;
    const bool active = q >= 0 && q < nseq;
    if (active) {
      if (++r > nrow) {
        r = -3;
        kc++;
      }
    }
    // (before a column starts and after it ends the window holds unused values)
    w0 = w1;
    w1 = w2;
    w2 = w3;
    w3 = w4;
    w4 = pw;
    // Backward pass + update of position q-2 (src/GKUA_Residual.f90:64-109),
    // first thing in the step: everything it reads was written before the
    // last barrier, and its store then has the whole step to complete.
    if (q >= 2 && q < nseq + 2) {
      int jf, kf;
      if (q < nseq) {
        jf = r - 2;
        kf = kc;
        if (jf < -3) {
          jf += LT;
          kf = kc - 1;
        }
      } else {
        jf = nrow - 1 + (q - nseq);
        kf = ktiles - 1;
      }
      if (jf >= 1) {
        const T dr = (i < ncol) ? X[(pbuf * ncol + col + 1) * VT + lane].y : T(0);
        // Delta*(i, jf+1) is the value of position q-1, 0 past the last row
        const T dn = (jf == nrow) ? T(0) : hd_n;
        const T du = hd_o + (bx_o * dr + by_o * dn);
        unsigned tb;
        int ird, vv;
        if (tile_of(kf, lane, tb, ird, vv)) {
          T f = w0 + du;
          if ((double)f < kClamp) f = T(kClamp);
          sto(F, tb + (unsigned)(jf + 1) * rb + (unsigned)(i + 1) * cb, f);
        }
      }
    }
    // Interior fast path, taken when every lane of the wavefront is at an
    // interior position (no face override, no wall slot or ghost value
    // within reach, finalised row in the same tile): plain NND (or WENO3) fluxes and no
    // boundary loads.  Everything else takes the general path.
    const bool inner = active && r >= 3 && r <= nrow - 2 && i >= 2 && i <= ncol - 2;
    if (__all(inner)) {
      const T fm1 = pm1, fp1 = pp1, fp2 = pp2;
      unsigned tb;
      int ird, vv;
      const bool act = tile_of(kc, lane, tb, ird, vv);
      const unsigned oc = tb + (unsigned)(r + 1) * rb + (unsigned)(i + 1) * cb;  // cell (i, r)
      pw = ldo(F, oc + 3 * rb);
      pm1 = ldo(F, oc + rb - cb);
      pp1 = ldo(F, oc + rb + cb);
      pp2 = ldo(F, oc + rb + 2 * cb);
      __builtin_amdgcn_sched_barrier(0);
      const double2 vxy = V[kc * VT + lane];
      const T vx = T(vxy.x), vy = T(vxy.y);
      const double *cc = C + buf * nconst + col * kNW;
      T feq;
      {  // Shakhov equilibrium of cell (i,r) (Get_Feq, src/GKUA_Residual.f90:420-469)
        const T viu = vx - T(cc[4]), vjv = vy - T(cc[5]);
        const T vij2 = (viu * viu + vjv * vjv) * T(cc[2]);
        const T fvx = T(cc[1]) * (viu * T(cc[6]) + vjv * T(cc[7]));
        const T gd = T(cc[0]) * wexp(-vij2);
        // G: gd (1 + phi (2c^2 - 4)); H: (T gd / 2)(1 + phi (2c^2 - 2))
        const T pre = ird == 0 ? gd : (T(cc[3]) * gd) * T(0.5);
        feq = pre * (T(1) + fvx * (T(2) * vij2 - (ird == 0 ? T(4) : T(2))));
      }
      const T axl = vx * T(cc[12]) + vy * T(cc[13]);
      const T axr = vx * T(cc[14]) + vy * T(cc[15]);
      const T ayb = vx * T(cc[16]) + vy * T(cc[17]);
      const T ayt = vx * T(cc[18]) + vy * T(cc[19]);
      const T2 xl = X[(pbuf * ncol + col - 1) * VT + lane];
      const T flxr = iflux<SL>(axr, fm1, w2, fp1, fp2, hs);
      const T flyt = iflux<SL>(ayt, w1, w2, w3, w4, hs);
      const T rr = -(((flxr - xl.x) + flyt) - fly_prev) * T(cc[9]) + T(cc[8]) * (feq - w2);
      fly_prev = flyt;
      // forward recursion (src/GKUA_Residual.f90:39-57)
      const T alfa = T(cc[11]) + (posp(axr) - negp(axl) + posp(ayt) - negp(ayb));
      const T ra = rcp_pos(alfa);
      const T ds = (T(cc[10]) * rr + posp(axl) * xl.y + posp(ayb) * hd_n) * ra;
      T2 xo;
      xo.x = flxr;
      xo.y = ds;
      X[(buf * ncol + col) * VT + lane] = xo;
      hd_o = hd_n; bx_o = bx_n; by_o = by_n;
      hd_n = ds; bx_n = -negp(axr) * ra; by_n = -negp(ayt) * ra;
      if (store_rhs && act) sto(R, oc, rr);
    } else {
      const T fm1 = pm1, fp1 = pp1, fp2 = pp2, xs = px, ys = py;
      // Loads of the next position, issued unconditionally (addresses that
      // are not needed point at the cell itself) so that the only wait for
      // them is at the top of the next step.  Before a column starts the next
      // position is its first one; after its last the loads are not used.
      {
        const bool started = q >= 0;
        const bool nvalid = active && q + 1 < nseq;
        const int rn = !started ? -3 : (nvalid ? (r == nrow ? -3 : r + 1) : nrow);
        const int kn = !started ? 0 : (nvalid && r == nrow ? kc + 1 : kc);
        unsigned tb;
        int ird, vv;
        tile_of(kn, lane, tb, ird, vv);
        // cell (i, rn) (row rn >= -3 may lie below the padded grid; its
        // offset is only used for rows >= 1) and cell (i, rn+2) (row >= -1)
        const unsigned oc = tb + (unsigned)(rn + 1) * rb + (unsigned)(i + 1) * cb;
        const unsigned ow = oc + 2 * rb;
        const bool rreal = rn >= 1;
        const int rf = OF[rreal ? rn : 1], cf = OF[i] >> 2;
        const unsigned ax = !rreal ? ow
                                   : (i == 1 ? ((rf & 1) ? soff(1, rn + 1, ird, vv) : oc - 2 * cb)
                                             : ((i == ncol && (rf & 2)) ? soff(3, rn + 1, ird, vv) : ow));
        const unsigned ay = (rn == 1 && (cf & 1)) ? soff(2, i + 1, ird, vv)
                                                  : ((rn == nrow && (cf & 2)) ? soff(4, i + 1, ird, vv) : ow);
        pw = ldo(F, ow);
        pm1 = ldo(F, rreal ? oc - cb : ow);
        pp1 = ldo(F, rreal ? oc + cb : ow);
        pp2 = ldo(F, rreal ? oc + 2 * cb : ow);
        px = ldo(F, ax);
        py = ldo(F, ay);
      }
      __builtin_amdgcn_sched_barrier(0);
      T hd_new = T(0), bx_new = T(0), by_new = T(0);
      if (active && r >= 1) {
        const bool first = r == 1;
        unsigned tb;
        int ird, vv;
        const bool act = tile_of(kc, lane, tb, ird, vv);
        const double2 vxy = V[kc * VT + lane];
        const T vx = T(vxy.x), vy = T(vxy.y);
        const double *cc = C + buf * nconst + col * kNW;
        T feq;
        {  // Shakhov equilibrium of cell (i,r) (Get_Feq, src/GKUA_Residual.f90:420-469)
          const T viu = vx - T(cc[4]), vjv = vy - T(cc[5]);
          const T vij2 = (viu * viu + vjv * vjv) * T(cc[2]);
          const T fvx = T(cc[1]) * (viu * T(cc[6]) + vjv * T(cc[7]));
          const T gd = T(cc[0]) * wexp(-vij2);
          const T pre = ird == 0 ? gd : (T(cc[3]) * gd) * T(0.5);
          feq = pre * (T(1) + fvx * (T(2) * vij2 - (ird == 0 ? T(4) : T(2))));
        }
        const int rf = OF[r], cf = OF[i] >> 2;
        const T axl = vx * T(cc[12]) + vy * T(cc[13]);
        const T axr = vx * T(cc[14]) + vy * T(cc[15]);
        T flxl, dl = T(0);
        if (i == 1) {
          flxl = wflux<SL>(1, nx, rf, axl, xs, fm1, w2, fp1, xs, hs);
        } else {
          const T2 xl = X[(pbuf * ncol + col - 1) * VT + lane];
          flxl = xl.x;
          dl = xl.y;
        }
        const T flxr = wflux<SL>(i + 1, nx, rf, axr, fm1, w2, fp1, fp2, xs, hs);
        const T ayb = vx * T(cc[16]) + vy * T(cc[17]);
        const T ayt = vx * T(cc[18]) + vy * T(cc[19]);
        const T flyb = first ? wflux<SL>(1, ny, cf, ayb, w0, w1, w2, w3, ys, hs) : fly_prev;
        const T flyt = wflux<SL>(r + 1, ny, cf, ayt, w1, w2, w3, w4, ys, hs);
        fly_prev = flyt;
        const T rr = -(((flxr - flxl) + flyt) - flyb) * T(cc[9]) + T(cc[8]) * (feq - w2);
        if (store_rhs && act) sto(R, tb + (unsigned)(r + 1) * rb + (unsigned)(i + 1) * cb, rr);
        // forward recursion (src/GKUA_Residual.f90:39-57); Delta*(i,0) = Delta*(0,j) = 0
        const T alfa = T(cc[11]) + (posp(axr) - negp(axl) + posp(ayt) - negp(ayb));
        const T ra = rcp_pos(alfa);
        const T db = first ? T(0) : hd_n;
        const T ds = (T(cc[10]) * rr + posp(axl) * dl + posp(ayb) * db) * ra;
        T2 xo;
        xo.x = flxr;
        xo.y = ds;
        X[(buf * ncol + col) * VT + lane] = xo;
        hd_new = ds; bx_new = -negp(axr) * ra; by_new = -negp(ayt) * ra;
      }
      hd_o = hd_n; bx_o = bx_n; by_o = by_n;
      hd_n = hd_new; bx_n = bx_new; by_n = by_new;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------
// Column-wavefront kinetic step ("strip walker"): the same single pass as
// k_walk (Shakhov equilibrium, NND fluxes, forward LU-SGS recursion,
// non-recursive backward pass, update and clamp; src/GKUA_Residual.f90:26-109,
// 117-469), mapped so that one wavefront owns one mesh column and 64
// consecutive velocities, with G and H of a velocity in the same lane.  Row,
// column, wall rules and the 20 cell constants are then wave-uniform (scalar
// registers and scalar loads), every access to F is one contiguous line per
// wavefront, and the work that depends only on the cell and the velocity
// (exp, Shakhov factor, face coefficients, LU-SGS diagonal) is shared by G
// and H.
//
// A workgroup holds NW consecutive columns (a strip) of one velocity tile.
// Wavefront w is at position q = s - w of the tile's sequence at step s
// (three window-filling virtual rows, then rows 1..ny-1), so the forward
// recursion's inputs from column i-1 were published one step earlier in LDS;
// column i finalises position q-2 at the top of step s, using Delta*(i+1, .)
// published by column i+1 one step earlier.
//
// The strips of a tile form a pipeline across workgroups.  The last column
// of strip g hands the face flux and Delta* of each row (G and H) to the
// first column of strip g+1 through a global seam buffer, then raises that
// row's flag to the launch's epoch; the consumer checks the flag one step
// before it loads the row.  The grid is strip-major, so a workgroup only
// waits for one with a lower index, which was dispatched before it: the
// waits cannot form a cycle.  The waits are bounded; a timeout sets *werr.
// The last column of a strip is not updated here: its backward pass needs
// Delta* of the next strip's first column, which k_seam applies after this
// kernel (the first column stores it in dright).  Until then the column keeps
// its old values, which are the ones its neighbours read.
// ------------------------------------------------------------------------
constexpr int kSpinMax = 1 << 20;  // about a second of polling

// element of a wavefront-uniform base plus a per-lane byte offset (a global
// access with a scalar base address and one vector offset register)
template <typename T>
__device__ __forceinline__ T ldv(const T *base, unsigned voff) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + voff);
}
template <typename T>
__device__ __forceinline__ void stv(T *base, unsigned voff, T v) {
  *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + voff) = v;
}

// Workgroup barrier that orders LDS accesses only.  __syncthreads() also
// waits for every outstanding vector memory operation (its fence covers the
// LDS writes of global_load_lds); here the staged copies are read only by the
// wavefront that issued them, after its own wait, so they stay in flight.
__device__ __forceinline__ void barrier_lds() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// A wavefront-uniform value the compiler must treat as new at this point:
// everything derived from it is recomputed (a few scalar instructions) rather
// than hoisted out of the loop and kept in registers.
template <typename X>
__device__ __forceinline__ X fresh(X x) {
  asm volatile("" : "+s"(x));
  return x;
}

// 16 bytes per lane from global memory straight into LDS at dst + 16*lane
// (gfx950 global_load_lds_dwordx4; dst is wavefront-uniform).  The builtin
// exists only for the device target; the host pass compiles the kernel's
// launch stub and never this body.
__device__ __forceinline__ void lds_copy16(const void *src, void *dst) {
#if defined(__HIP_DEVICE_COMPILE__)
  __builtin_amdgcn_global_load_lds(src, dst, 16, 0, 0);
#endif
}

// SL: slope rule of the fluxes (0 minmod for NND, 1 WENO3, 2 none: UP1).  At least
// 4 wavefronts per SIMD (at most 128 VGPRs): with 80 KB of LDS per 8-column
// workgroup two workgroups share a CU, which the register file must allow.
template <typename T>
using Vec2 = typename std::conditional<std::is_same<T, double>::value, double2, float2>::type;

// LDS of one workgroup of the column-wavefront kernel (declared by each
// kernel, so that its addresses are constants of that kernel)
template <typename T, int NW>
struct StripLds {
  Vec2<T> XG[2][NW][64], XH[2][NW][64];  // (face flux, Delta*) of each column
  // Values of the next row, copied from global memory straight into LDS
  // (global_load_lds, 16 bytes per lane) one step ahead, so that no vector
  // register holds them while they are in flight: per column, slots
  // 0 F(i, r+2), 1 F(i-1, r), 2 F(i+1, r), 3 F(i+2, r), 4 x-boundary value,
  // 5 y-boundary value, each the tile's 64 velocities of G then of H.
  T STG[NW][6][128];
};

template <typename T, int NW, int SL>
__device__ __forceinline__ void strip_body(StripLds<T, NW> &lds, const Dev &d, const T *__restrict__ cpar,
                                           T *__restrict__ F, const T *__restrict__ slot, T *__restrict__ R,
                                           int store_rhs, int ntiles, T *msg, T *dright, int *flag, int epoch,
                                           int *werr, const int *__restrict__ oflags, int bid) {
  using T2 = Vec2<T>;
  constexpr unsigned ES = sizeof(T);
  auto &XG = lds.XG;
  auto &XH = lds.XH;
  auto &STG = lds.STG;
  constexpr int kPer = 16 / sizeof(T);  // velocities per lane and copy
  constexpr int kLanes = 64 / kPer;     // lanes per component
  const int nx = d.nx, ny = d.ny, ncol = nx - 1, nrow = ny - 1, nv = d.nv;
  const int ngroups = (ncol + NW - 1) / NW;
  const int LT = nrow + 4;                 // positions: virtual rows -3..0, rows 1..ny-1
  const int nsteps = (NW - 1) + LT + 2;
  const int w0 = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int g0w = bid / ntiles, tile = bid - g0w * ntiles;
  const int i0 = g0w * NW + w0 + 1;  // mesh column of this wavefront
  // Strips early in the pipeline feed all later ones: their wavefronts get
  // the higher issue priority on a CU they share with later strips.
  if (g0w == 0) __builtin_amdgcn_s_setprio(3);
  else if (g0w == 1) __builtin_amdgcn_s_setprio(2);
  else if (g0w == 2) __builtin_amdgcn_s_setprio(1);
  if (i0 > ncol) {  // past the last column: keep the workgroup's barriers
    for (int s = 0; s < nsteps; s++) barrier_lds();
    return;
  }
  T vx, vy;
  {
    const int vv = min(tile * 64 + (int)(threadIdx.x & 63), nv - 1);
    vx = T(d.vi[vv % d.nvi]);
    vy = T(d.vj[vv / d.nvi]);
  }
  constexpr T hs = SL == 2 ? T(0) : T(0.5);  // slope factor (0: first-order upwind)
  // (override words are wavefront-uniform: scalar registers also where the
  // table's address was itself read from memory, as in the batched kernel)
  const int cfl0 = __builtin_amdgcn_readfirstlane(oflags[ny + 2 + i0]);  // j-face overrides of column i
  // Roles, re-derived every step: the first column of strip g > 0 reads the
  // seam message of strip g-1 (consumer); the last column of a strip that is
  // not the last writes it (producer) and is updated by k_seam; xlo / xhi:
  // the mesh's first / last column; xrule: an overridden i-face within reach
  // of the right face.
  // byte strides: one cell's velocities, one row of cells, one plane
  const long long cb = (long long)nv * ES, rb = (long long)d.NI * cb, pb = (long long)d.ncell * cb;
  char *const Fb0 = reinterpret_cast<char *>(F);
  char *const Rb0 = reinterpret_cast<char *>(R);
  const char *const Sb0 = reinterpret_cast<const char *>(slot);
  // Per-lane offsets are re-derived where they are used, from a copy of the
  // thread index the compiler cannot see through: keeping them live across
  // the loop costs more vector registers than the few instructions that
  // recompute them.
  // Staged copies: lane l < kLanes copies velocities tile*64 + l*kPer .. +kPer-1
  // of G, the next kLanes lanes the same of H (one plane, or one slot row,
  // further).  The last copy of a partial tile is moved back into the
  // velocity range (those lanes' velocities are not used).
  auto lane_of = [&]() {
    unsigned t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return (int)(t & 63);
  };
  // per-lane byte offsets of a copy: into F (G or H plane) and into the wall slots
  auto copy_offsets = [&](int l, long long &stF, long long &stS) {
    const int comp = l / kLanes;
    const long long pvb = (long long)min(tile * 64 + (l % kLanes) * kPer, nv - kPer) * ES;
    stF = (comp ? pb : 0ll) + pvb;
    stS = (comp ? cb : 0ll) + pvb;
  };
  auto copy = [&](int w, int k, const char *base, long long st, int l) {
    if (l < 2 * kLanes) lds_copy16(base + st, &STG[w][k][0]);
  };
  T mfg = T(0), mdg = T(0), mfh = T(0), mdh = T(0);  // seam message of row r+1
  int mflag = 0;                                     // flag of row r+2
  long long oc = -2 * rb + (long long)(i0 + 1) * cb;  // byte offset of cell (i, r), r = -3
  constexpr long long kRec = kNW * (long long)sizeof(T);        // bytes of one cell's constants
  // 32-bit offsets and indices where they fit, to keep scalar registers free
  int ocp = (-4 * ncol + (i0 - 1)) * (int)kRec;  // cpar byte offset of cell (i, r); used for r >= 1
  const int ccstep = ncol * (int)kRec;
  const unsigned sbc = ((unsigned)(g0w > 0 ? g0w - 1 : 0) * ntiles + tile) * (unsigned)nrow;  // seam rows read (row r at sbc + r - 1)
  const unsigned sbp = ((unsigned)g0w * ntiles + tile) * (unsigned)nrow;                    // seam rows written
  {
    const int l = (int)(threadIdx.x & 63);
    long long stF, stS;
    copy_offsets(l, stF, stS);
    copy(w0, 0, Fb0 + oc + 2 * rb, stF, l);  // F(i, -1), the first window value
  }
  T g0 = T(0), g1 = T(0), g2 = T(0), g3 = T(0), g4 = T(0);  // G(i, r-2..r+2)
  T h0 = T(0), h1 = T(0), h2 = T(0), h3 = T(0), h4 = T(0);
  T gfly = T(0), hfly = T(0);  // flux through the bottom face of the next row
  T gd_n = T(0), hd_n = T(0), bx_n = T(0), by_n = T(0), gd_o = T(0), hd_o = T(0), bx_o = T(0), by_o = T(0);
  int flag_due = 0;  // a seam row written in the previous step waits for its flag
  // The cell constants of the next row are touched one step ahead (one
  // dword from each 64-byte line of the 160-byte record), so that the
  // scalar loads of the row's 20 constants hit the scalar cache.
  int pf0 = 0, pf1 = 0, pf2 = 0, pf3 = 0;
  // Steps s < w0 and s >= w0 + LT + 2 only keep the workgroup's barriers.
  // The active steps q = s - w0 = 0 .. LT+1 run five at a time, each given
  // the five-row windows in rotated order: moving a window down one row
  // then renames registers instead of copying four values.
  auto step = [&](int q, T &gA, T &gB, T &gC, T &gD, T &gE, T &hA, T &hB, T &hC, T &hD, T &hE)
                  __attribute__((always_inline)) {
    T &g0 = gB, &g1 = gC, &g2 = gD, &g3 = gE, &g4 = gA;  // G(i, r-2..r+2); gA takes the new row
    T &h0 = hB, &h1 = hC, &h2 = hD, &h3 = hE, &h4 = hA;
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the previous step's copies, loads and stores are complete
    // per-wavefront values, re-derived every step (see fresh())
    const int w = fresh(w0), i = fresh(i0), g = fresh(g0w), cfl = fresh(cfl0);
    char *const Fb = fresh(Fb0);
    char *const Rb = fresh(Rb0);
    const char *const Sb = fresh(Sb0);
    const bool consumer = w == 0 && g > 0;
    const bool producer = w == NW - 1 && g < ngroups - 1;
    const bool xlo = i == 1, xhi = i == ncol;
    const bool xrule = i == 1 || i >= ncol - 1;
    const int lane = lane_of();
    const unsigned voff = (unsigned)min(tile * 64 + lane, nv - 1) * ES, loff = (unsigned)lane * ES;
    const bool act = tile * 64 + lane < nv;
    const int s = q + w, r = q - 3;
    const int buf = s & 1, pbuf = buf ^ 1;
    if (producer && flag_due) {  // its message stores are complete: publish the row
      if (lane == 0) st_agent(flag + sbp + (r - 2), epoch);
      flag_due = 0;
    }
    // backward pass + update of row r-2 (src/GKUA_Residual.f90:64-109).  Its
    // inputs are complete at the top of the step (Delta* of column i+1 was
    // published in the previous step), and storing early gives the stores a
    // whole step to complete before the next step's wait.  No wavefront reads
    // F(i, r-2) after step s-2.
    const int rf = r - 2;
    if (rf >= 1 && rf <= nrow && !producer) {
      T gr = T(0), hr = T(0);
      if (i < ncol) {
        gr = XG[pbuf][w + 1][lane].y;
        hr = XH[pbuf][w + 1][lane].y;
      }
      const bool last = rf == nrow;  // Delta*(i, ny) = 0
      const T gu = gd_o + (bx_o * gr + by_o * (last ? T(0) : gd_n));
      const T hu = hd_o + (bx_o * hr + by_o * (last ? T(0) : hd_n));
      T fg = g0 + gu, fh = h0 + hu;
      if ((double)fg < kClamp) fg = T(kClamp);
      if ((double)fh < kClamp) fh = T(kClamp);
      if (act) {
        stv(reinterpret_cast<T *>(Fb + oc - 2 * rb), voff, fg);
        stv(reinterpret_cast<T *>(Fb + oc - 2 * rb + pb), voff, fh);
      }
    }
    asm volatile("" ::"s"(pf0), "s"(pf1), "s"(pf2), "s"(pf3));  // the previous step's touches
    if (r + 1 >= 1 && r + 1 <= nrow) {
      const int *pc = reinterpret_cast<const int *>(reinterpret_cast<const char *>(cpar) + ocp + ccstep);
      pf0 = pc[0];
      pf1 = pc[16];
      pf2 = pc[kRec > 128 ? 32 : 0];
      pf3 = pc[kRec / 4 - 1];
    }
    const bool real = r >= 1 && r <= nrow;
    const int rfl = real ? __builtin_amdgcn_readfirstlane(oflags[r]) : 0;  // i-face overrides of row r
    // the values staged in the previous step for this row
    // (slots that were not copied for this row hold older values, which are
    // read here but never used)
    g4 = STG[w][0][lane];
    h4 = STG[w][0][64 + lane];
    const T cgm1 = STG[w][1][lane], chm1 = STG[w][1][64 + lane];
    const T cgp1 = STG[w][2][lane], chp1 = STG[w][2][64 + lane];
    const T cgp2 = STG[w][3][lane], chp2 = STG[w][3][64 + lane];
    const T cgxs = STG[w][4][lane], chxs = STG[w][4][64 + lane];
    const T cgys = STG[w][5][lane], chys = STG[w][5][64 + lane];
    const T cmfg = mfg, cmdg = mdg, cmfh = mfh, cmdh = mdh;
    // row r: equilibrium, fluxes, RHS, forward recursion (src/GKUA_Residual.f90:39-57,117-469)
    T gd_new = T(0), hd_new = T(0), bx_new = T(0), by_new = T(0);
    T feqg = T(0), feqh = T(0), axl = T(0), axr = T(0), ayb = T(0), ayt = T(0), ra = T(0);
    T vinv = T(0), cf = T(0), vol = T(0);
    if (real) {
      const T *cc = reinterpret_cast<const T *>(reinterpret_cast<const char *>(cpar) + ocp);
      {  // Shakhov equilibrium of cell (i,r) (Get_Feq, src/GKUA_Residual.f90:420-469)
        const T viu = vx - T(cc[4]), vjv = vy - T(cc[5]);
        const T vij2 = (viu * viu + vjv * vjv) * T(cc[2]);
        const T fvx = T(cc[1]) * (viu * T(cc[6]) + vjv * T(cc[7]));
        const T gd = T(cc[0]) * wexp(-vij2);
        const T b = fvx * (T(2) * vij2 - T(2));
        // G: gd (1 + phi (2c^2 - 4)); H: (T gd / 2)(1 + phi (2c^2 - 2))
        feqg = gd * ((T(1) - T(2) * fvx) + b);
        feqh = ((T(cc[3]) * gd) * T(0.5)) * (T(1) + b);
      }
      const T *cc2 = cc;
      axl = vx * T(cc2[12]) + vy * T(cc2[13]);
      axr = vx * T(cc2[14]) + vy * T(cc2[15]);
      ayb = vx * T(cc2[16]) + vy * T(cc2[17]);
      ayt = vx * T(cc2[18]) + vy * T(cc2[19]);
      const T alfa = T(cc2[11]) + (posp(axr) - negp(axl) + posp(ayt) - negp(ayb));
      ra = rcp_pos(alfa);
      vinv = T(cc2[9]);
      cf = T(cc2[8]);
      vol = T(cc2[10]);
    }
    // all staged values are in registers before the copies of the next row overwrite them
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
    // copies of the next row rn (one step ahead)
    const int rn = r + 1;
    long long stF, stS;
    copy_offsets(lane, stF, stS);
    auto slotrow = [&](int face, int pos) { return Sb + (((long long)(face - 1) * d.L + pos) * 2) * cb; };
    if (rn <= nrow) copy(w, 0, Fb + oc + 3 * rb, stF, lane);
    if (rn >= 1 && rn <= nrow) {
      copy(w, 1, Fb + oc + rb - cb, stF, lane);
      copy(w, 2, Fb + oc + rb + cb, stF, lane);
      copy(w, 3, Fb + oc + rb + 2 * cb, stF, lane);
      const int rfn = __builtin_amdgcn_readfirstlane(oflags[rn]);
      if (xlo) {
        if (rfn & 1) copy(w, 4, slotrow(1, rn + 1), stS, lane);
        else copy(w, 4, Fb + oc + rb - 2 * cb, stF, lane);  // ghost cell (-1, rn)
      } else if (xhi && (rfn & 2)) {
        copy(w, 4, slotrow(3, rn + 1), stS, lane);
      }
      if (rn == 1 && (cfl & 1)) copy(w, 5, slotrow(2, i + 1), stS, lane);
      else if (rn == nrow && (cfl & 2)) copy(w, 5, slotrow(4, i + 1), stS, lane);
      if (consumer) {
        // the flag of row rn was loaded one step ago; wait for it if the
        // producing strip is behind, then load the row's message
        const unsigned sr = sbc + (rn - 1);
        int fv = __builtin_amdgcn_readfirstlane(mflag);
        if (fv != epoch && __builtin_amdgcn_readfirstlane(ld_agent(werr)) == 0) {
          int it = 0;
          do {
            __builtin_amdgcn_s_sleep(2);
            fv = __builtin_amdgcn_readfirstlane(ld_agent(flag + sr));
          } while (fv != epoch && ++it < kSpinMax);
          if (fv != epoch && lane == 0) st_agent(werr, 1);
        }
        const char *m = reinterpret_cast<const char *>(msg + (size_t)sr * 256) + loff;
        mfg = ld_agent(reinterpret_cast<const T *>(m));
        mdg = ld_agent(reinterpret_cast<const T *>(m + 64 * ES));
        mfh = ld_agent(reinterpret_cast<const T *>(m + 128 * ES));
        mdh = ld_agent(reinterpret_cast<const T *>(m + 192 * ES));
      }
    }
    if (consumer && rn + 1 >= 1 && rn + 1 <= nrow) mflag = ld_agent(flag + sbc + rn);
    if (real) {
      const bool first = r == 1;
      const bool yrule = first || r >= nrow - 1;  // an overridden j-face within reach of the top face
      // left face: boundary flux, the seam message or column i-1's value
      T gfl, gdl, hfl, hdl;
      if (xlo) {
        gfl = wflux1<SL>(rfl, axl, cgxs, cgm1, g2, cgp1, cgxs, hs);
        hfl = wflux1<SL>(rfl, axl, chxs, chm1, h2, chp1, chxs, hs);
        gdl = T(0);
        hdl = T(0);
      } else if (w == 0) {
        gfl = cmfg; gdl = cmdg; hfl = cmfh; hdl = cmdh;
      } else {
        const T2 lg = XG[pbuf][w - 1][lane], lh = XH[pbuf][w - 1][lane];
        gfl = lg.x; gdl = lg.y; hfl = lh.x; hdl = lh.y;
      }
      T gfr, hfr, gft, hft, gfb, hfb;
      if (xrule) {
        gfr = wflux<SL>(i + 1, nx, rfl, axr, cgm1, g2, cgp1, cgp2, cgxs, hs);
        hfr = wflux<SL>(i + 1, nx, rfl, axr, chm1, h2, chp1, chp2, chxs, hs);
      } else {
        gfr = iflux<SL>(axr, cgm1, g2, cgp1, cgp2, hs);
        hfr = iflux<SL>(axr, chm1, h2, chp1, chp2, hs);
      }
      if (yrule) {
        gft = wflux<SL>(r + 1, ny, cfl, ayt, g1, g2, g3, g4, cgys, hs);
        hft = wflux<SL>(r + 1, ny, cfl, ayt, h1, h2, h3, h4, chys, hs);
      } else {
        gft = iflux<SL>(ayt, g1, g2, g3, g4, hs);
        hft = iflux<SL>(ayt, h1, h2, h3, h4, hs);
      }
      if (first) {
        gfb = wflux1<SL>(cfl, ayb, g0, g1, g2, g3, cgys, hs);
        hfb = wflux1<SL>(cfl, ayb, h0, h1, h2, h3, chys, hs);
      } else {
        gfb = gfly;
        hfb = hfly;
      }
      gfly = gft;
      hfly = hft;
      const T grr = -(((gfr - gfl) + gft) - gfb) * vinv + cf * (feqg - g2);
      const T hrr = -(((hfr - hfl) + hft) - hfb) * vinv + cf * (feqh - h2);
      if (store_rhs && act) {
        stv(reinterpret_cast<T *>(Rb + oc), voff, grr);
        stv(reinterpret_cast<T *>(Rb + oc + pb), voff, hrr);
      }
      gd_new = (vol * grr + posp(axl) * gdl + posp(ayb) * (first ? T(0) : gd_n)) * ra;
      hd_new = (vol * hrr + posp(axl) * hdl + posp(ayb) * (first ? T(0) : hd_n)) * ra;
      XG[buf][w][lane] = T2{gfr, gd_new};
      XH[buf][w][lane] = T2{hfr, hd_new};
      bx_new = -negp(axr) * ra;
      by_new = -negp(ayt) * ra;
      if (producer) {
        char *m = reinterpret_cast<char *>(msg + (size_t)(sbp + (r - 1)) * 256) + loff;
        st_agent(reinterpret_cast<T *>(m), gfr);
        st_agent(reinterpret_cast<T *>(m + 64 * ES), gd_new);
        st_agent(reinterpret_cast<T *>(m + 128 * ES), hfr);
        st_agent(reinterpret_cast<T *>(m + 192 * ES), hd_new);
        flag_due = 1;
      }
      if (consumer) {
        T *m = dright + (size_t)(sbc + (r - 1)) * 128;
        stv(m, loff, gd_new);
        stv(m + 64, loff, hd_new);
      }
    }
    gd_o = gd_n; hd_o = hd_n; bx_o = bx_n; by_o = by_n;
    gd_n = gd_new; hd_n = hd_new; bx_n = bx_new; by_n = by_new;
    oc += rb;
    ocp += ccstep;
    barrier_lds();
  };
  for (int s = 0; s < w0; s++) barrier_lds();
  for (int q = 0; q < LT + 2; q += 5) {
    step(q, g0, g1, g2, g3, g4, h0, h1, h2, h3, h4);
    if (q + 1 < LT + 2) step(q + 1, g1, g2, g3, g4, g0, h1, h2, h3, h4, h0);
    if (q + 2 < LT + 2) step(q + 2, g2, g3, g4, g0, g1, h2, h3, h4, h0, h1);
    if (q + 3 < LT + 2) step(q + 3, g3, g4, g0, g1, g2, h3, h4, h0, h1, h2);
    if (q + 4 < LT + 2) step(q + 4, g4, g0, g1, g2, g3, h4, h0, h1, h2, h3);
  }
  for (int s = w0 + LT + 2; s < nsteps; s++) barrier_lds();
  if (w0 == NW - 1 && g0w < ngroups - 1 && flag_due) {
    __builtin_amdgcn_s_waitcnt(0x0F70);
    if ((threadIdx.x & 63) == 0) st_agent(flag + sbp + (nrow - 1), epoch);
  }
}

template <typename T, int NW, int SL>
__global__ void __launch_bounds__(64 * NW, 4) k_strip(Dev d, const T *__restrict__ cpar, T *__restrict__ F,
                                                 const T *__restrict__ slot, T *__restrict__ R, int store_rhs,
                                                 int ntiles, T *msg, T *dright, int *flag, int epoch, int *werr,
                                                 const int *__restrict__ oflags) {
  __shared__ StripLds<T, NW> lds;
  strip_body<T, NW, SL>(lds, d, cpar, F, slot, R, store_rhs, ntiles, msg, dright, flag, epoch, werr, oflags,
                        (int)blockIdx.x);
}

// wavefront-uniform copies (scalar registers) of values read from memory
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
template <typename P>
__device__ __forceinline__ P *uni(P *p) {
  const unsigned long long v = reinterpret_cast<unsigned long long>(p);
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v), hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return reinterpret_cast<P *>(((unsigned long long)hi << 32) | lo);
}

// the job of workgroup b of a batched launch: the last job whose first
// workgroup (member K of the jobs) is <= b (wavefront-uniform)
template <typename J, int J::*K>
__device__ __forceinline__ int job_of(const J *jobs, int njobs, int b) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].*K <= b) lo = mid;
    else hi = mid - 1;
  }
  return __builtin_amdgcn_readfirstlane(lo);
}

// One launch of the column-wavefront kernel over the strips of every job of a
// batch (instances of different mesh and velocity grid).  Each job's
// workgroups are consecutive and strip-major, so a workgroup still waits only
// for lower-numbered workgroups of its own job.  The launch epoch is read
// from device memory (a replayed graph increments it in k_batch_epoch).
template <typename T, int NW, int SL>
__global__ void __launch_bounds__(64 * NW, 4) k_strip_batch(const StripJob<T> *__restrict__ jobs, int njobs,
                                                       const int *__restrict__ epoch, int *werr) {
  const int k = job_of<StripJob<T>, &StripJob<T>::strip0>(jobs, njobs, (int)blockIdx.x);
  const StripJob<T> &J = jobs[k];
  // the fields the kernel uses, in scalar registers like kernel arguments
  Dev d{};
  d.nx = uni(J.d.nx);
  d.ny = uni(J.d.ny);
  d.nvi = uni(J.d.nvi);
  d.nv = uni(J.d.nv);
  d.NI = uni(J.d.NI);
  d.ncell = uni(J.d.ncell);
  d.L = uni(J.d.L);
  d.vi = uni(J.d.vi);
  d.vj = uni(J.d.vj);
  __shared__ StripLds<T, NW> lds;
  strip_body<T, NW, SL>(lds, d, uni(J.cp), uni(J.F), uni(J.slot), uni(J.R), 0, uni(J.ntiles), uni(J.msg), uni(J.dright),
                        uni(J.flag), uni(*epoch), werr, uni(J.oflags), (int)blockIdx.x - uni(J.strip0));
}

template <typename T>
__global__ void __launch_bounds__(256) k_cellpar_batch(const StripJob<T> *__restrict__ jobs, int njobs) {
  const int k = job_of<StripJob<T>, &StripJob<T>::cell0>(jobs, njobs, (int)blockIdx.x);
  const StripJob<T> &J = jobs[k];
  cellpar_body(J.d, const_cast<T *>(J.cp), 0, ((int)blockIdx.x - J.cell0) * 256 + (int)threadIdx.x);
}

__global__ void k_batch_epoch(int *epoch) {
  if (threadIdx.x == 0) *epoch += 1;
}

// Backward pass + update of the last column of every strip but the last
// (src/GKUA_Residual.f90:64-109): Delta*(i, j) and Delta*(i, j+1) from the
// seam messages, Delta*(i+1, j) from dright, the coefficients
// -A^-_x/alfa, -A^-_y/alfa recomputed from the cell constants as k_strip
// computes them.  One thread per (seam, tile, row, velocity lane).
template <typename T>
__device__ __forceinline__ void seam_body(const Dev &d, const T *__restrict__ cpar, T *F, const T *msg,
                                          const T *dright, int ntiles, int nw, size_t t) {
  const int lane = (int)(t & 63);
  const size_t row_id = t >> 6;  // (seam * ntiles + tile) * nrow + r - 1
  const int ncol = d.nx - 1, nrow = d.ny - 1, nv = d.nv;
  const int ngroups = (ncol + nw - 1) / nw;
  if (row_id >= (size_t)(ngroups - 1) * ntiles * nrow) return;
  const int r = (int)(row_id % nrow) + 1;
  const int st = (int)(row_id / nrow);
  const int sg = st / ntiles, tile = st - sg * ntiles;
  const int i = (sg + 1) * nw;  // last column of strip sg
  const int v = tile * 64 + lane;
  if (v >= nv) return;
  const size_t nf = (size_t)d.ncell * nv;
  const T vx = T(d.vi[v % d.nvi]), vy = T(d.vj[v / d.nvi]);
  const T gds = msg[row_id * 256 + 64 + lane], hds = msg[row_id * 256 + 192 + lane];
  const T gdn = r < nrow ? msg[(row_id + 1) * 256 + 64 + lane] : T(0);  // Delta*(i, j+1), 0 past the last row
  const T hdn = r < nrow ? msg[(row_id + 1) * 256 + 192 + lane] : T(0);
  const T gdr = dright[row_id * 128 + lane], hdr = dright[row_id * 128 + 64 + lane];
  const T *cc = cpar + ((size_t)(r - 1) * ncol + (i - 1)) * kNW;
  const T axl = vx * T(cc[12]) + vy * T(cc[13]);
  const T axr = vx * T(cc[14]) + vy * T(cc[15]);
  const T ayb = vx * T(cc[16]) + vy * T(cc[17]);
  const T ayt = vx * T(cc[18]) + vy * T(cc[19]);
  const T alfa = T(cc[11]) + (posp(axr) - negp(axl) + posp(ayt) - negp(ayb));
  const T ra = rcp_pos(alfa);
  const T bx = -negp(axr) * ra, by = -negp(ayt) * ra;
  const size_t o = ((size_t)(r + 1) * d.NI + (size_t)(i + 1)) * nv + v;
  T fg = F[o] + (gds + (bx * gdr + by * gdn));
  T fh = F[nf + o] + (hds + (bx * hdr + by * hdn));
  if ((double)fg < kClamp) fg = T(kClamp);
  if ((double)fh < kClamp) fh = T(kClamp);
  F[o] = fg;
  F[nf + o] = fh;
}
template <typename T>
__global__ void __launch_bounds__(256) k_seam(Dev d, const T *__restrict__ cpar, T *F, const T *msg,
                                              const T *dright, int ntiles, int nw) {
  seam_body(d, cpar, F, msg, dright, ntiles, nw, (size_t)blockIdx.x * blockDim.x + threadIdx.x);
}

template <typename T>
__global__ void __launch_bounds__(256) k_seam_batch(const StripJob<T> *__restrict__ jobs, int njobs, int nw) {
  const int k = job_of<StripJob<T>, &StripJob<T>::seam0>(jobs, njobs, (int)blockIdx.x);
  const StripJob<T> &J = jobs[k];
  seam_body(J.d, J.cp, J.F, J.msg, J.dright, J.ntiles, nw, (size_t)((int)blockIdx.x - J.seam0) * 256 + threadIdx.x);
}

// Conserved moments of the stored RHS per interior cell (GaussLegendre_rs,
// src/GKUA_DETFP_RES.f90:81-120); one block per interior cell.
template <typename T>
__global__ void __launch_bounds__(256) k_res4(Dev d, const T *RG, const T *RH) {
  const int t = blockIdx.x;
  const int i = t % (d.nx - 1) + 1, j = t / (d.nx - 1) + 1;
  const int c = cidx(d, i, j);
  const size_t o = (size_t)c * d.nv;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int v = threadIdx.x; v < d.nv; v += blockDim.x) {
    const int iv = v % d.nvi, jv = v / d.nvi;
    const double vxd = d.vi[iv], vyd = d.vj[jv], w = d.wi[iv] * d.wj[jv];
    const double g = (double)RG[o + v], h = (double)RH[o + v];
    const double vx1 = vxd * g, vy1 = vyd * g;
    acc[0] += w * g;
    acc[1] += w * vx1;
    acc[2] += w * vy1;
    acc[3] += w * (h + vxd * vx1 + vyd * vy1);
  }
  block_sum<4>(acc);
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; k++) d.res4[k * d.ncell + c] = acc[k];
}

// ------------------------------------------------------------------------
// Positivity clamp of the ghost ring (src/GKUA_Residual.f90:101-109) and the
// velocity moments with the DETFP conversion and collision frequency
// (GaussLegendre src/GKUA_DETFP.f90:288-363, DETFP :52-105).  The reference
// zeroes all of S_GKUA before the moments (:303), so cells outside
// [0,nx]x[0,ny] get 0.
// ------------------------------------------------------------------------
// One wavefront per padded cell.  Lane l covers velocity column iv = l mod nvi
// of rows jv = l div nvi + R*k (R = 64 div nvi rows per pass when nvi <= 64),
// or iv = l + 64*m for nvi > 64, so that vi/wi stay in registers and a pass
// reads consecutive velocities.  Wave-level reduction only (no barrier).
// Raw buffer access (gfx9-family descriptor: word 3 = 0x00020000, offsets in
// bytes, stride 0; a load past num_records returns 0).
constexpr int kBufWord3 = 0x00020000;
template <typename T>
__device__ __forceinline__ T buf_ld(__amdgpu_buffer_rsrc_t r, unsigned voff, int soff) {
  if constexpr (sizeof(T) == 8)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
template <typename T>
__device__ __forceinline__ void buf_st(T v, __amdgpu_buffer_rsrc_t r, unsigned voff, int soff) {
  if constexpr (sizeof(T) == 8)
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned int, v), r,
                                          voff, soff, 0);
  else
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, voff, soff, 0);
}

// ONE: 33 <= nvi <= 64, so a lane holds one column and each pass one whole
// row; only the six row sums stay live through the loop, which leaves
// registers for more loads in flight.  Two wavefronts share a cell, each over
// half of its rows, and combine through LDS: twice the wavefronts shortens the
// last partial round of the grid.
template <typename T, bool ONE>
__global__ void __launch_bounds__(256) k_moments(Dev d, T *__restrict__ FG, T *__restrict__ FH) {
  // the wavefront index through readfirstlane, so that the row range below is
  // known to be uniform (scalar row offsets and factors)
  const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int c = ONE ? blockIdx.x * 2 + (wv >> 1) : blockIdx.x * 4 + wv;
  const int half = ONE ? (wv & 1) : 0;
  // whole wavefronts: the reduction below stays wave-uniform; the ONE path
  // returns only after its barrier
  if (!ONE && c >= d.ncell) return;
  const bool valid = c < d.ncell;
  const int i = c % d.NI - 1, j = c / d.NI - 1;
  const bool interior = (i >= 1 && i <= d.nx - 1 && j >= 1 && j <= d.ny - 1);
  const bool ring = (i >= 0 && i <= d.nx && j >= 0 && j <= d.ny);
  const size_t o = (size_t)c * d.nv;
  const int nvi = d.nvi, nvj = d.nvj;
  const int R = nvi <= 64 ? 64 / nvi : 1;  // rows per pass
  const int nch = nvi <= 64 ? 1 : (nvi + 63) / 64;
  // A lane keeps one velocity column iv, so vx and wx are fixed along its
  // rows and the ten moments factor into six sums over the rows:
  //   A0 = sum wy G, A1 = sum wy vy G, A2 = sum wy vy^2 G, A5 = sum wy vy^3 G,
  //   A3 = sum wy H, A4 = sum wy vy H,
  // one FMA each per velocity (the row factors wy vy^k are per row).
  double acc[10];
#pragma unroll
  for (int k = 0; k < 10; k++) acc[k] = 0.0;
  if constexpr (ONE) {
    double A0 = 0.0, A1 = 0.0, A2 = 0.0, A3 = 0.0, A4 = 0.0, A5 = 0.0;
    const int iv = min(lane, nvi - 1);
    const bool on = lane < nvi;
    const double vx = d.vi[iv], wx = d.wi[iv];
    // Buffer loads over the cell's velocities: the lane offset in a vector
    // register, the row offset in a scalar one, so the unrolled rows need no
    // 64-bit vector addresses (which took 32 registers as plain loads).
    // Offsets stay below num_records = the cell's bytes.
    const unsigned loff = (unsigned)iv * (unsigned)sizeof(T);
    const int cell_bytes = nvi * nvj * (int)sizeof(T);
    const __amdgpu_buffer_rsrc_t rg = __builtin_amdgcn_make_buffer_rsrc(FG + o, (short)0, cell_bytes, kBufWord3);
    const __amdgpu_buffer_rsrc_t rh = __builtin_amdgcn_make_buffer_rsrc(FH + o, (short)0, cell_bytes, kBufWord3);
    const int hn = (nvj + 1) >> 1;
    const int j0 = valid ? half * hn : 0, j1 = valid ? min(nvj, half * hn + hn) : 0;
#pragma unroll 8
    for (int jv = j0; jv < j1; jv++) {
      const int soff = jv * nvi * (int)sizeof(T);
      T g = buf_ld<T>(rg, loff, soff), h = buf_ld<T>(rh, loff, soff);
      if (!interior && on) {
        if ((double)g < kClamp) { g = T(kClamp); buf_st<T>(g, rg, loff, soff); }
        if ((double)h < kClamp) { h = T(kClamp); buf_st<T>(h, rh, loff, soff); }
      }
      // row factors from the table (scalar loads, scalar operands): formed
      // here they would take vector registers for every unrolled row
      const double wy = d.rowf[4 * jv], w1 = d.rowf[4 * jv + 1], w2 = d.rowf[4 * jv + 2],
                   w3 = d.rowf[4 * jv + 3];
      const double gd = (double)g, hd = (double)h;
      A0 += wy * gd;
      A1 += w1 * gd;
      A2 += w2 * gd;
      A5 += w3 * gd;
      A3 += wy * hd;
      A4 += w1 * hd;
    }
    const double vx2 = vx * vx;
    const double e4 = A3 + vx2 * A0 + A2;
    acc[0] = wx * A0;
    acc[1] = wx * (vx * A0);
    acc[2] = wx * A1;
    acc[3] = wx * e4;
    acc[4] = wx * (vx2 * A0);
    acc[5] = wx * A2;
    acc[6] = wx * (vx * A1);
    acc[7] = wx * (vx * e4);
    acc[8] = wx * (A4 + vx2 * A1 + A5);
    acc[9] = wx * A3;
    if (!on)  // lanes past the row add nothing
#pragma unroll
      for (int k = 0; k < 10; k++) acc[k] = 0.0;
    __shared__ double part[2][10];
#pragma unroll
    for (int k = 0; k < 10; k++) acc[k] = wave_sum_last(acc[k]);
    if (half == 1 && lane == 63)
#pragma unroll
      for (int k = 0; k < 10; k++) part[wv >> 1][k] = acc[k];
    __syncthreads();
    if (!valid || half == 1) return;
    if (lane == 63)
#pragma unroll
      for (int k = 0; k < 10; k++) acc[k] += part[wv >> 1][k];
  }
  // R = 1 (nvi > 32): one row per pass, so the row index and its factors are
  // wavefront-uniform (scalar loads)
  const bool uni = R == 1;
  for (int ch = 0; ch < (ONE ? 0 : nch); ch++) {
    const int iv = nvi <= 64 ? lane % nvi : ch * 64 + lane;
    const int jr = uni ? 0 : lane / nvi;
    if (iv >= nvi || (uni ? (nvi <= 64 && lane >= nvi) : jr >= R)) continue;
    const double vx = d.vi[iv], wx = d.wi[iv];
    double A0 = 0.0, A1 = 0.0, A2 = 0.0, A3 = 0.0, A4 = 0.0, A5 = 0.0;
    auto add = [&](int jv) {
      const size_t ov = o + (size_t)jv * nvi + iv;
      T g = FG[ov], h = FH[ov];
      if (!interior) {
        if ((double)g < kClamp) { g = T(kClamp); FG[ov] = g; }
        if ((double)h < kClamp) { h = T(kClamp); FH[ov] = h; }
      }
      const double vy = d.vj[jv], wy = d.wj[jv];
      const double w1 = wy * vy, w2 = w1 * vy, w3 = w2 * vy;
      const double gd = (double)g, hd = (double)h;
      A0 += wy * gd;
      A1 += w1 * gd;
      A2 += w2 * gd;
      A5 += w3 * gd;
      A3 += wy * hd;
      A4 += w1 * hd;
    };
    if (uni) {
#pragma unroll 4
      for (int jv = 0; jv < nvj; jv++) add(jv);
    } else {
#pragma unroll 4
      for (int jv = jr; jv < nvj; jv += R) add(jv);
    }
    // sum_v w {G, vx G, vy G, H + |v|^2 G, vx^2 G, vy^2 G, vx vy G,
    //          vx (H + |v|^2 G), vy (H + |v|^2 G), H} of this column
    const double vx2 = vx * vx;
    const double e4 = A3 + vx2 * A0 + A2;  // sum wy (H + |v|^2 G)
    acc[0] += wx * A0;
    acc[1] += wx * (vx * A0);
    acc[2] += wx * A1;
    acc[3] += wx * e4;
    acc[4] += wx * (vx2 * A0);
    acc[5] += wx * A2;
    acc[6] += wx * (vx * A1);
    acc[7] += wx * (vx * e4);
    acc[8] += wx * (A4 + vx2 * A1 + A5);
    acc[9] += wx * A3;
  }
  if (!ring) {
    if (lane < 11) d.S[lane * d.ncell + c] = 0.0;
    return;
  }
  if constexpr (!ONE)
#pragma unroll
    for (int k = 0; k < 10; k++) acc[k] = wave_sum_last(acc[k]);
  if (lane == 63) {
    double ron = acc[0];
    if (ron <= 0.) ron = 1.0;  // R00
    const double u = acc[1] / ron, v = acc[2] / ron;
    const double us1 = u * u, vs1 = v * v, sum_uv = us1 + vs1;
    double ttr = (acc[3] - ron * sum_uv) / (1.5 * ron);
    if (ttr <= 0.) ttr = 1.0;  // T00
    double *S = d.S;
    const int N = d.ncell;
    S[0 * N + c] = acc[0];
    S[1 * N + c] = u;
    S[2 * N + c] = v;
    S[3 * N + c] = ttr;
    S[4 * N + c] = (acc[4] - ron * us1) * 2.;
    S[5 * N + c] = (acc[5] - ron * vs1) * 2.;
    S[6 * N + c] = (acc[6] - ron * u * v) * 2.;
    S[7 * N + c] = acc[7] - 2. * u * acc[4] - 2. * v * acc[6] + ron * u * (sum_uv - 1.5 * ttr);
    S[8 * N + c] = acc[8] - 2. * v * acc[5] - 2. * u * acc[6] + ron * v * (sum_uv - 1.5 * ttr);
    S[9 * N + c] = acc[9] * 2.;
    S[10 * N + c] = 0.0;
    if (d.x1s <= 9.99999993922529e-09) d.CF[c] = d.cfc * acc[0];  // REAL(4) 1.e-8
    else d.CF[c] = d.cfc * acc[0] * pow(ttr, d.x1s);
  }
}

// DETFPwall for every wall/farfield face node (src/GKUA_DETFP_WALL.f90:34-95).
// list[4*k..]: ghost cell, first cell, second cell, Swall offset.
__global__ void k_det_wall(Dev d, const int *list, int n, double *swall) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int cg = list[4 * t], ca = list[4 * t + 1], cb = list[4 * t + 2], so = list[4 * t + 3];
  for (int k = 0; k < 10; k++) {
    const double sw = (-d.S[k * d.ncell + cb] + 3.0 * d.S[k * d.ncell + ca]) * 0.5;
    swall[so + k] = sw;
    d.S[k * d.ncell + cg] = 2.0 * sw - d.S[k * d.ncell + ca];
  }
}

// Convert_NS part 1 (src/Hybrid_Sub.f90:221-235): corner extrapolation of S
// and the NS conservative variables on [0,nx]x[0,ny].
__global__ void k_convert_u(Dev d) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int n1 = d.nx + 1;
  if (t >= n1 * (d.ny + 1)) return;
  const int i = t % n1, j = t / n1;
  const int nx = d.nx, ny = d.ny, N = d.ncell;
  const int c = cidx(d, i, j);
  double *S = d.S;
  int ca = -1, cb = -1, cc = -1;
  if (i == 0 && j == 0) { ca = cidx(d, 1, 0); cb = cidx(d, 0, 1); cc = cidx(d, 1, 1); }
  else if (i == nx && j == 0) { ca = cidx(d, nx, 1); cb = cidx(d, nx - 1, 0); cc = cidx(d, nx - 1, 1); }
  else if (i == 0 && j == ny) { ca = cidx(d, 1, ny); cb = cidx(d, 0, ny - 1); cc = cidx(d, 1, ny - 1); }
  else if (i == nx && j == ny) { ca = cidx(d, nx - 1, ny); cb = cidx(d, nx, ny - 1); cc = cidx(d, nx - 1, ny - 1); }
  if (ca >= 0)
    for (int k = 0; k < 11; k++) S[k * N + c] = S[k * N + ca] + S[k * N + cb] - S[k * N + cc];
  const double sq = sqrt(2.0 / d.gamma);
  const double u1 = S[c];
  const double tu = S[1 * N + c] / d.ma * sq, tv = S[2 * N + c] / d.ma * sq, tt = S[3 * N + c];
  d.U[0 * N + c] = u1;
  d.U[1 * N + c] = u1 * tu;
  d.U[2 * N + c] = u1 * tv;
  d.U[3 * N + c] = u1 * d.cv * tt + 0.5 * u1 * (tu * tu + tv * tv);
}

// Convert_NS part 2 (src/Hybrid_Sub.f90:236-239): the four whole-array ghost
// extrapolations, in order.  Single block.
__global__ void k_convert_ghost(Dev d) {
  const int nx = d.nx, ny = d.ny, N = d.ncell;
  double *U = d.U;
  for (int ph = 0; ph < 4; ph++) {
    const int len = (ph < 2) ? (ny + 3) : (nx + 3);
    for (int t = threadIdx.x; t < 4 * len; t += blockDim.x) {
      const int m = t % 4, r = t / 4 - 1;  // r in [-1, n+1]
      int cg, c1, c2;
      if (ph == 0) { cg = cidx(d, 0, r); c1 = cidx(d, 1, r); c2 = cidx(d, 2, r); }
      else if (ph == 1) { cg = cidx(d, nx, r); c1 = cidx(d, nx - 1, r); c2 = cidx(d, nx - 2, r); }
      else if (ph == 2) { cg = cidx(d, r, 0); c1 = cidx(d, r, 1); c2 = cidx(d, r, 2); }
      else { cg = cidx(d, r, ny); c1 = cidx(d, r, ny - 1); c2 = cidx(d, r, ny - 2); }
      U[m * N + cg] = 2.0 * U[m * N + c1] - U[m * N + c2];
    }
    __syncthreads();
  }
}

// comput_duvtpckw (src/sub_NS_singlegrid.f90:121-157) over the padded grid,
// with the T<0.1 stop of interior cells turned into an error flag.
__global__ void k_prims(Dev d) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.ncell) return;
  const int N = d.ncell;
  const double dd = d.U[c], u = d.U[N + c] / dd, v = d.U[2 * N + c] / dd;
  const double T = (d.U[3 * N + c] - 0.5 * dd * (u * u + v * v)) / (d.cv * dd);
  d.prim[c] = dd;
  d.prim[N + c] = u;
  d.prim[2 * N + c] = v;
  d.prim[3 * N + c] = T;
  d.prim[4 * N + c] = (1. / (d.gamma * d.ma * d.ma)) * dd * T;
  const int i = c % d.NI - 1, j = c / d.NI - 1;
  if (i >= 1 && i <= d.nx - 1 && j >= 1 && j <= d.ny - 1 && T < 0.10000000149011612) atomicOr(d.err, 1);
}

// HoT of Convert_NS (src/Hybrid_Sub.f90:317-349 and :419-448): i-faces
// (i=1..nx, j=1..ny-1) then j-faces (i=1..nx-1, j=1..ny), one thread each.
__global__ void k_hot(Dev d) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = d.nx, ny = d.ny, N = d.ncell;
  const int ni = nx * (ny - 1), nj = (nx - 1) * ny;
  if (t >= ni + nj) return;
  const double *S = d.S;
  const double gm2 = d.gamma * (d.ma * d.ma);
  // 1.4142135381698608 is sqrt(2.0) evaluated in REAL(4), as the reference does
  double o[8];
  if (t < ni) {
    const int i = t % nx + 1, j = t / nx + 1;
    ns_face(d, 0, i, j, o);
    const int c = cidx(d, i, j), cm = cidx(d, i - 1, j);
    double *tg = d.taoi, *qg = d.qi;
    if (i != 1 && i != nx) {
      const double tmp = (S[4 * N + c] + S[5 * N + c] + S[9 * N + c] + S[4 * N + cm] + S[5 * N + cm] + S[9 * N + cm]) / 6.0;
      tg[0 * N + c] = -((S[4 * N + c] + S[4 * N + cm]) / 2.0 - tmp) / gm2 - o[0];
      tg[1 * N + c] = -((S[5 * N + c] + S[5 * N + cm]) / 2.0 - tmp) / gm2 - o[1];
      tg[2 * N + c] = -(S[6 * N + c] + S[6 * N + cm]) / 2.0 / gm2 - o[2];
      tg[3 * N + c] = -((S[9 * N + c] + S[9 * N + cm]) / 2.0 - tmp) / gm2 - o[3];
      qg[0 * N + c] = -(S[7 * N + cm] + S[7 * N + c]) / 2.0 * 1.4142135381698608 / ((d.ma * d.ma * d.ma) * pow(d.gamma, 1.5)) - o[4];
      qg[1 * N + c] = -(S[8 * N + cm] + S[8 * N + c]) / 2.0 * 1.4142135381698608 / ((d.ma * d.ma * d.ma) * pow(d.gamma, 1.5)) - o[5];
    } else {
      const int cs = (i == 1) ? c : cm;
      const double tmp = (S[4 * N + cs] + S[5 * N + cs] + S[9 * N + cs]) / 3.0;
      tg[0 * N + c] = -(S[4 * N + cs] - tmp) / gm2 - o[0];
      tg[1 * N + c] = -(S[5 * N + cs] - tmp) / gm2 - o[1];
      tg[2 * N + c] = -S[6 * N + cs] / gm2 - o[2];
      tg[3 * N + c] = -(S[9 * N + cs] - tmp) / gm2 - o[3];
      qg[0 * N + c] = -S[7 * N + cs] * 1.4142135381698608 / ((d.ma * d.ma * d.ma) * pow(d.gamma, 1.5)) - o[4];
      qg[1 * N + c] = -S[8 * N + cs] * 1.4142135381698608 / ((d.ma * d.ma * d.ma) * pow(d.gamma, 1.5)) - o[5];
    }
  } else {
    const int tt = t - ni;
    const int i = tt % (nx - 1) + 1, j = tt / (nx - 1) + 1;
    ns_face(d, 1, i, j, o);
    const int c = cidx(d, i, j), cm = cidx(d, i, j - 1);
    double *tg = d.taoj, *qg = d.qj;
    if (j != 1 && j != ny) {
      const double tmp = (S[4 * N + c] + S[5 * N + c] + S[9 * N + c] + S[4 * N + cm] + S[5 * N + cm] + S[9 * N + cm]) / 6.0;
      tg[0 * N + c] = -((S[4 * N + c] + S[4 * N + cm]) / 2.0 - tmp) / gm2 - o[0];
      tg[1 * N + c] = -((S[5 * N + c] + S[5 * N + cm]) / 2.0 - tmp) / gm2 - o[1];
      tg[2 * N + c] = -(S[6 * N + c] + S[6 * N + cm]) / 2.0 / gm2 - o[2];
      tg[3 * N + c] = -((S[9 * N + c] + S[9 * N + cm]) / 2.0 - tmp) / gm2 - o[3];
      qg[0 * N + c] = -(S[7 * N + cm] + S[7 * N + c]) / 2.0 * 1.4142135381698608 / ((d.ma * d.ma * d.ma) * pow(d.gamma, 1.5)) - o[4];
      qg[1 * N + c] = -(S[8 * N + cm] + S[8 * N + c]) / 2.0 * 1.4142135381698608 / ((d.ma * d.ma * d.ma) * pow(d.gamma, 1.5)) - o[5];
    } else {
      const int cs = (j == 1) ? c : cm;
      const double tmp = (S[4 * N + cs] + S[5 * N + cs] + S[9 * N + cs]) / 3.0;
      tg[0 * N + c] = -(S[4 * N + cs] - tmp) / gm2 - o[0];
      tg[1 * N + c] = -(S[5 * N + cs] - tmp) / gm2 - o[1];
      tg[2 * N + c] = -S[6 * N + cs] / gm2 - o[2];
      tg[3 * N + c] = -(S[9 * N + cs] - tmp) / gm2 - o[3];
      qg[0 * N + c] = -S[7 * N + cs] * 1.4142135381698608 / ((d.ma * d.ma * d.ma) * pow(d.gamma, 1.5)) - o[4];
      qg[1 * N + c] = -S[8 * N + cs] * 1.4142135381698608 / ((d.ma * d.ma * d.ma) * pow(d.gamma, 1.5)) - o[5];
    }
  }
}

// RMS and max over interior cells of the per-cell RHS moments
// (DETFP_res, src/GKUA_DETFP_RES.f90:45-54).  Single block.
__global__ void k_residual(Dev d, double *out) {
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const int ncl = (d.nx - 1) * (d.ny - 1);
  for (int t = threadIdx.x; t < ncl; t += blockDim.x) {
    const int i = t % (d.nx - 1) + 1, j = t / (d.nx - 1) + 1;
    const int c = cidx(d, i, j);
    for (int k = 0; k < 4; k++) {
      const double r = d.res4[k * d.ncell + c];
      acc[k] += r * r;
      acc[4 + k] = fmax(acc[4 + k], fabs(r));
    }
  }
  // sums and maxima reduced separately
  double s4[4] = {acc[0], acc[1], acc[2], acc[3]};
  block_sum<4>(s4);
  __shared__ double shm[4][16];
  for (int k = 0; k < 4; k++) {
    double m = acc[4 + k];
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) shm[k][threadIdx.x >> 6] = m;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 0; k < 4; k++) {
      double m = 0.0;
      for (int w = 0; w < (int)(blockDim.x >> 6); w++) m = fmax(m, shm[k][w]);
      out[k] = sqrt(s4[k] / (1.0 * ncl));
      out[4 + k] = m;
    }
  }
}

// ------------------------------------------------------------------------
// Layout transposition between the reference's FRD(2,-3:nx+3,-3:ny+3,nvi,nvj)
// and the device's f[cell][velocity] planes + wall slots.
// ------------------------------------------------------------------------

// A chunk of consecutive velocities v0..v0+nvc-1 at a time.  The Fortran
// array holds each velocity's values as one contiguous block of
// 2(nx+7)(ny+7) doubles, so a chunk is a contiguous slice of it; the stage
// holds that slice, element ird + 2((i+3) + (nx+7)((j+3) + (ny+7)(v-v0))).
// One thread per stage element.  Every stage element is written (zero where
// the device keeps no value), so the stage needs no clearing.
template <typename T>
__global__ void k_frd_out_chunk(Dev d, double *stage, int v0, int nvc, const T *FG, const T *FH, const T *slot) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t nxp = d.nx + 7, nyp = d.ny + 7;
  if (t >= 2 * nxp * nyp * (size_t)nvc) return;
  const int ird = (int)(t & 1);
  const size_t q = t >> 1;
  const int i = (int)(q % nxp) - 3, j = (int)((q / nxp) % nyp) - 3;
  const int v = v0 + (int)(q / (nxp * nyp));
  double val = 0.0;
  if (i >= -1 && i <= d.nx + 1 && j >= -1 && j <= d.ny + 1) {
    const size_t f = (size_t)cidx(d, i, j) * d.nv + v;
    val = (double)(ird == 0 ? FG[f] : FH[f]);
  } else if ((i == -3 || i == d.nx + 3) && j >= -1 && j <= d.ny + 1) {
    val = (double)slot[slot_idx(d, i < 0 ? 1 : 3, j + 1, ird, v)];
  } else if ((j == -3 || j == d.ny + 3) && i >= -1 && i <= d.nx + 1) {
    val = (double)slot[slot_idx(d, j < 0 ? 2 : 4, i + 1, ird, v)];
  }
  stage[t] = val;
}

template <typename T>
__global__ void k_frd_in_chunk(Dev d, const double *stage, int v0, int nvc, T *FG, T *FH, T *slot) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t nxp = d.nx + 7, nyp = d.ny + 7;
  if (t >= 2 * nxp * nyp * (size_t)nvc) return;
  const int ird = (int)(t & 1);
  const size_t q = t >> 1;
  const int i = (int)(q % nxp) - 3, j = (int)((q / nxp) % nyp) - 3;
  const int v = v0 + (int)(q / (nxp * nyp));
  const T val = T(stage[t]);
  if (i >= -1 && i <= d.nx + 1 && j >= -1 && j <= d.ny + 1) {
    const size_t f = (size_t)cidx(d, i, j) * d.nv + v;
    if (ird == 0) FG[f] = val;
    else FH[f] = val;
  } else if ((i == -3 || i == d.nx + 3) && j >= -1 && j <= d.ny + 1) {
    slot[slot_idx(d, i < 0 ? 1 : 3, j + 1, ird, v)] = val;
  } else if ((j == -3 || j == d.ny + 3) && i >= -1 && i <= d.nx + 1) {
    slot[slot_idx(d, j < 0 ? 2 : 4, i + 1, ird, v)] = val;
  }
}

// ------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------
template <typename T>
void launch_wall(const Dev &d, const BcDev &b, T *FG, T *FH, T *slot, hipStream_t s) {
  const int nn = (b.face == 1 || b.face == 3) ? (b.jend - b.jst) : (b.iend - b.ist);
  if (nn <= 0) return;
  hipLaunchKernelGGL(k_wall<T>, dim3(nn), dim3(256), 0, s, d, b, FG, FH, slot);
}
template <typename T>
void launch_walls(const Dev &d, const WallSet &ws, T *FG, T *FH, T *slot, hipStream_t s) {
  if (ws.n <= 0 || ws.start[ws.n] <= 0) return;
  hipLaunchKernelGGL(k_walls<T>, dim3(ws.start[ws.n]), dim3(256), 0, s, d, ws, FG, FH, slot);
}
template <typename T>
void launch_farfield(const Dev &d, const BcDev &b, T *FG, T *FH, hipStream_t s) {
  const int nn = (b.face == 1 || b.face == 3) ? (b.jend - b.jst + 1) : (b.iend - b.ist + 1);
  hipLaunchKernelGGL(k_farfield<T>, dim3(2 * nn), dim3(256), 0, s, d, b, FG, FH);
}
template <typename T>
void launch_periodic(const Dev &d, const BcDev &b, T *FG, T *FH, hipStream_t s) {
  const int nn = (b.face == 1 || b.face == 3) ? (b.jend - b.jst + 1) : (b.iend - b.ist + 1);
  hipLaunchKernelGGL(k_periodic<T>, dim3(2 * nn), dim3(256), 0, s, d, b, FG, FH);
}
template <typename T>
void launch_inner(const Dev &dd, const BcDev &bd, T *FGd, T *FHd, const Dev &ds, const BcDev &bs, const T *FGs,
                  const T *FHs, int n1, int orient, hipStream_t s) {
  hipLaunchKernelGGL(k_inner<T>, dim3(2 * n1), dim3(256), 0, s, dd, bd, FGd, FHd, ds, bs, FGs, FHs, n1, orient);
}
template <typename T>
void launch_rhs(const Dev &d, const T *FG, const T *FH, const T *slot, T *RG, T *RH, hipStream_t s) {
  hipLaunchKernelGGL(k_rhs<T>, dim3((d.nx - 1) * (d.ny - 1)), dim3(256), 0, s, d, FG, FH, slot, RG, RH);
}
template <typename T>
int sweep_vt(const Dev &d) {
  // largest power-of-two tile (<=64) whose 3-plane LDS ring fits in 64 KiB
  int vt = 64;
  while (vt > 1 && (size_t)3 * (d.nx + 1) * vt * 2 * sizeof(T) > 64 * 1024) vt >>= 1;
  return vt;
}
template <typename T>
void launch_sweep(const Dev &d, T *FG, T *FH, const T *RG, const T *RH, hipStream_t s) {
  const int vt = sweep_vt<T>(d);
  const size_t lds = (size_t)3 * (d.nx + 1) * vt * 2 * sizeof(T);
  const int nb = (d.nv + vt - 1) / vt;
  hipLaunchKernelGGL(k_sweep<T>, dim3(nb), dim3(256), lds, s, d, vt, FG, FH, RG, RH);
}
template <typename T>
void launch_res4(const Dev &d, const T *RG, const T *RH, hipStream_t s) {
  hipLaunchKernelGGL(k_res4<T>, dim3((d.nx - 1) * (d.ny - 1)), dim3(256), 0, s, d, RG, RH);
}
// Launch shape of the fused kernel: VT velocities per column (a power of two,
// at least 8 so that the staged constants are at most 3 per thread), K tiles
// per block so that the grid covers the CUs about once.  False when the mesh
// or the dynamic LDS does not fit (the caller then uses the split kernels).
template <typename T>
bool walk_config(const Dev &d, int tiles, int *vt_out, int *k_out, size_t *lds_out) {
  static int ncu = 0, max_lds = 0;
  if (!ncu) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) {
      ncu = 0;
      return false;
    }
  }
  const int ncol = d.nx - 1, nrow = d.ny - 1;
  if (ncol * 8 > 1024) return false;
  // 32-bit byte offsets into F (two planes) and the wall slots
  if (2.0 * d.ncell * d.nv * sizeof(T) >= 4294967295.0 || 8.0 * d.L * d.nv * sizeof(T) >= 4294967295.0) return false;
  int vt = 64;
  while (vt > 8 && vt * ncol > 1024) vt >>= 1;
  const int ntiles = 2 * ((d.nv + vt - 1) / vt);
  const int K = tiles > 0 ? (tiles < ntiles ? tiles : ntiles) : (ntiles / ncu > 1 ? ntiles / ncu : 1);
  const size_t nthr = (size_t)vt * ncol;
  const size_t lds = 8 * (2 * 3 * nthr + 2 * (size_t)K * vt) + 2 * (size_t)ncol * vt * 2 * sizeof(T) +
                     (size_t)std::max(nrow + 2, d.nx + 2);
  if (lds > (size_t)max_lds) return false;
  static bool attr_set = false;
  if (lds > 65536 && !attr_set) {
    if (hipFuncSetAttribute((const void *)k_walk<T, 0>, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds) !=
            hipSuccess ||
        hipFuncSetAttribute((const void *)k_walk<T, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds) !=
            hipSuccess)
      return false;
    attr_set = true;
  }
  *vt_out = vt;
  *k_out = K;
  *lds_out = lds;
  return true;
}
// Column-wavefront kernel with NW columns per workgroup, then the seam update.
template <typename T, int NW>
bool launch_strip_nw(const Dev &d, double *cpar, T *F, const T *slot, T *R, int store_rhs, StripBufs &b,
                     hipStream_t s) {
  const int ncol = d.nx - 1, nrow = d.ny - 1;
  if (ncol < 2 || nrow < 2) return false;
  const int ngroups = (ncol + NW - 1) / NW, ntiles = (d.nv + 63) / 64;
  const size_t rows = (size_t)(ngroups - 1) * ntiles * nrow;
  if (rows * 256 > b.msg_elems || rows * 128 > b.dright_elems || rows > b.flag_elems) return false;
  const int ncl = ncol * nrow;
  // the constants in T: the FP32 kernels read them as single precision
  // without a conversion per use (the same values: each is rounded once)
  T *cp = reinterpret_cast<T *>(cpar);
  hipLaunchKernelGGL(k_cellpar<T>, dim3((ncl + 255) / 256), dim3(256), 0, s, d, cp, 0);
  b.epoch += 1;
  if (b.ev_before) (void)hipEventRecord(b.ev_before, s);
  // the WENO3 scheme is a separate instantiation so that the NND / UP1
  // kernel's register allocation is untouched by the WENO3 weights
  if (d.scheme == kSchemeWeno3)
    hipLaunchKernelGGL((k_strip<T, NW, 1>), dim3(ngroups * ntiles), dim3(64 * NW), 0, s, d, cp, F, slot, R,
                       store_rhs, ntiles, (T *)b.msg, (T *)b.dright, b.flag, b.epoch, b.err, b.oflags);
  else if (d.scheme == 0)  // first-order upwind
    hipLaunchKernelGGL((k_strip<T, NW, 2>), dim3(ngroups * ntiles), dim3(64 * NW), 0, s, d, cp, F, slot, R,
                       store_rhs, ntiles, (T *)b.msg, (T *)b.dright, b.flag, b.epoch, b.err, b.oflags);
  else
    hipLaunchKernelGGL((k_strip<T, NW, 0>), dim3(ngroups * ntiles), dim3(64 * NW), 0, s, d, cp, F, slot, R,
                       store_rhs, ntiles, (T *)b.msg, (T *)b.dright, b.flag, b.epoch, b.err, b.oflags);
  if (b.ev_after) (void)hipEventRecord(b.ev_after, s);
  if (ngroups > 1) {
    const size_t nthr = rows * 64;
    hipLaunchKernelGGL(k_seam<T>, dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0, s, d, cp, F,
                       (const T *)b.msg, (const T *)b.dright, ntiles, NW);
  }
  return true;
}
int strip_blocks(const Dev &d, int nw) {
  const int ncol = d.nx - 1;
  return (ncol + nw - 1) / nw * ((d.nv + 63) / 64);
}
int seam_blocks(const Dev &d, int nw) { return (int)((strip_rows(d, nw) * 64 + 255) / 256); }
int cell_blocks(const Dev &d) { return ((d.nx - 1) * (d.ny - 1) + 255) / 256; }

template <typename T, int NW>
void strip_batch_nw(const StripJob<T> *djobs, int njobs, int scheme, int nstrip, int *epoch, int *werr,
                    hipStream_t s) {
  if (scheme == kSchemeWeno3)
    hipLaunchKernelGGL((k_strip_batch<T, NW, 1>), dim3(nstrip), dim3(64 * NW), 0, s, djobs, njobs, epoch, werr);
  else if (scheme == 0)
    hipLaunchKernelGGL((k_strip_batch<T, NW, 2>), dim3(nstrip), dim3(64 * NW), 0, s, djobs, njobs, epoch, werr);
  else
    hipLaunchKernelGGL((k_strip_batch<T, NW, 0>), dim3(nstrip), dim3(64 * NW), 0, s, djobs, njobs, epoch, werr);
}

template <typename T>
void launch_strip_batch(const StripJob<T> *djobs, int njobs, int nw, int scheme, int nstrip, int nseam, int ncellb,
                        int *epoch, int *werr, hipStream_t s) {
  hipLaunchKernelGGL(k_batch_epoch, dim3(1), dim3(64), 0, s, epoch);
  hipLaunchKernelGGL(k_cellpar_batch<T>, dim3(ncellb), dim3(256), 0, s, djobs, njobs);
  if (nw == 16) strip_batch_nw<T, 16>(djobs, njobs, scheme, nstrip, epoch, werr, s);
  else strip_batch_nw<T, 8>(djobs, njobs, scheme, nstrip, epoch, werr, s);
  if (nseam > 0) hipLaunchKernelGGL(k_seam_batch<T>, dim3(nseam), dim3(256), 0, s, djobs, njobs, nw);
}

template <typename T>
__global__ void k_fill_uniform(Dev d, const double *g, const double *h, T *FG, T *FH, T *slot, size_t nslot) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t nf = (size_t)d.ncell * d.nv;
  if (t < nf) {
    const int v = (int)(t % d.nv);
    FG[t] = T(g[v]);
    FH[t] = T(h[v]);
  }
  if (t < nslot) slot[t] = T(0);
}

template <typename T>
void launch_fill_uniform(const Dev &d, const double *g, const double *h, T *FG, T *FH, T *slot, hipStream_t s) {
  const size_t nf = (size_t)d.ncell * d.nv, ns = (size_t)4 * d.L * 2 * d.nv;
  const size_t n = std::max(nf, ns);
  hipLaunchKernelGGL(k_fill_uniform<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d, g, h, FG, FH, slot, ns);
}

size_t strip_rows(const Dev &d, int nw) {
  const int ncol = d.nx - 1, nrow = d.ny - 1;
  if (ncol < 2 || nrow < 2) return 0;
  return (size_t)((ncol + nw - 1) / nw - 1) * ((d.nv + 63) / 64) * nrow;
}
template <typename T>
bool launch_strip(const Dev &d, double *cpar, T *F, const T *slot, T *R, int store_rhs, int nw, StripBufs &b,
                  hipStream_t s) {
  if (nw == 16) return launch_strip_nw<T, 16>(d, cpar, F, slot, R, store_rhs, b, s);
  if (nw == 8) return launch_strip_nw<T, 8>(d, cpar, F, slot, R, store_rhs, b, s);
  return false;
}

template <typename T>
bool launch_walk(const Dev &d, double *cpar, T *F, const T *slot, T *R, int store_rhs, int tiles, hipStream_t s) {
  int vt = 0, K = 0;
  size_t lds = 0;
  if (!walk_config<T>(d, tiles, &vt, &K, &lds)) return false;
  // the wall slots must follow the two F planes in the same allocation
  const ptrdiff_t so = reinterpret_cast<const char *>(slot) - reinterpret_cast<const char *>(F);
  const size_t slot_bytes = (size_t)8 * d.L * d.nv * sizeof(T);
  if (so < (ptrdiff_t)(2 * (size_t)d.ncell * d.nv * sizeof(T)) || (size_t)so + slot_bytes > 4294967295u) return false;
  const int ncl = (d.nx - 1) * (d.ny - 1);
  hipLaunchKernelGGL(k_cellpar<double>, dim3((ncl + 255) / 256), dim3(256), 0, s, d, cpar, 3 * vt * (d.nx - 1));
  const int ntiles = 2 * ((d.nv + vt - 1) / vt);
  const int nb = (ntiles + K - 1) / K;
  const unsigned planeb = (unsigned)((size_t)d.ncell * d.nv * sizeof(T));
  if (d.scheme == kSchemeWeno3)
    hipLaunchKernelGGL((k_walk<T, 1>), dim3(nb), dim3(vt * (d.nx - 1)), lds, s, d, cpar, vt, K, F, planeb,
                       (unsigned)so, R, store_rhs);
  else
    hipLaunchKernelGGL((k_walk<T, 0>), dim3(nb), dim3(vt * (d.nx - 1)), lds, s, d, cpar, vt, K, F, planeb,
                       (unsigned)so, R, store_rhs);
  return true;
}

template <typename T>
void launch_moments(const Dev &d, T *FG, T *FH, hipStream_t s) {
  if (d.nvi > 32 && d.nvi <= 64)
    hipLaunchKernelGGL((k_moments<T, true>), dim3((d.ncell + 1) / 2), dim3(256), 0, s, d, FG, FH);
  else
    hipLaunchKernelGGL((k_moments<T, false>), dim3((d.ncell + 3) / 4), dim3(256), 0, s, d, FG, FH);
}
template <typename T>
void launch_frd_out_chunk(const Dev &d, double *stage, int v0, int nvc, const T *FG, const T *FH, const T *slot,
                          hipStream_t s) {
  const size_t n = (size_t)2 * (d.nx + 7) * (d.ny + 7) * nvc;
  hipLaunchKernelGGL(k_frd_out_chunk<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d, stage, v0, nvc, FG,
                     FH, slot);
}
template <typename T>
void launch_frd_in_chunk(const Dev &d, const double *stage, int v0, int nvc, T *FG, T *FH, T *slot, hipStream_t s) {
  const size_t n = (size_t)2 * (d.nx + 7) * (d.ny + 7) * nvc;
  hipLaunchKernelGGL(k_frd_in_chunk<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d, stage, v0, nvc, FG,
                     FH, slot);
}

void launch_save_s0(const Dev &d, hipStream_t s) {
  const int n = (d.nx - 1) * (d.ny - 1);
  hipLaunchKernelGGL(k_save_s0, dim3((n + 255) / 256), dim3(256), 0, s, d);
}
void launch_det_wall(const Dev &d, const int *list, int n, double *swall, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_det_wall, dim3((n + 127) / 128), dim3(128), 0, s, d, list, n, swall);
}
void launch_convert(const Dev &d, hipStream_t s) {
  const int n = (d.nx + 1) * (d.ny + 1);
  hipLaunchKernelGGL(k_convert_u, dim3((n + 255) / 256), dim3(256), 0, s, d);
  hipLaunchKernelGGL(k_convert_ghost, dim3(1), dim3(256), 0, s, d);
  hipLaunchKernelGGL(k_prims, dim3((d.ncell + 255) / 256), dim3(256), 0, s, d);
}
void launch_hot(const Dev &d, hipStream_t s) {
  const int n = d.nx * (d.ny - 1) + (d.nx - 1) * d.ny;
  hipLaunchKernelGGL(k_hot, dim3((n + 127) / 128), dim3(128), 0, s, d);
}
// out_mass of DETFP (src/GKUA_DETFP.f90:134-138): the sum of S_GKUA(1)*vol
// over the interior cells.  One workgroup, a fixed summation order: the same
// value on every run.
__global__ void __launch_bounds__(1024) k_out_mass(Dev d, double *out) {
  __shared__ double red[1024];
  const int ncol = d.nx - 1, n = ncol * (d.ny - 1);
  double acc = 0.0;
  for (int t = threadIdx.x; t < n; t += 1024) {
    const int i = t % ncol + 1, j = t / ncol + 1;
    acc += d.S[cidx(d, i, j)] * d.vol[gidx(d, i, j)];
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int k = 512; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// MassCorrect (src/GKUA_DETFP.f90:153-172): S_GKUA(1) of every interior cell
// times Init_mass/out_mass, evaluated as (S*Init_mass)/out_mass
__global__ void k_mass_scale(Dev d, const double *out, double init) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int ncol = d.nx - 1;
  if (t >= ncol * (d.ny - 1)) return;
  const int c = cidx(d, t % ncol + 1, t / ncol + 1);
  d.S[c] = d.S[c] * init / out[0];
}

void launch_out_mass(const Dev &d, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_out_mass, dim3(1), dim3(1024), 0, s, d, out);
}
void launch_mass_correct(const Dev &d, double *work, double init, hipStream_t s) {
  launch_out_mass(d, work, s);
  const int n = (d.nx - 1) * (d.ny - 1);
  hipLaunchKernelGGL(k_mass_scale, dim3((n + 255) / 256), dim3(256), 0, s, d, (const double *)work, init);
}
void launch_residual(const Dev &d, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_residual, dim3(1), dim3(1024), 0, s, d, out);
}

#define INST(T)                                                                                   \
  template void launch_wall<T>(const Dev &, const BcDev &, T *, T *, T *, hipStream_t);           \
  template void launch_walls<T>(const Dev &, const WallSet &, T *, T *, T *, hipStream_t);        \
  template void launch_farfield<T>(const Dev &, const BcDev &, T *, T *, hipStream_t);            \
  template void launch_periodic<T>(const Dev &, const BcDev &, T *, T *, hipStream_t);            \
  template void launch_inner<T>(const Dev &, const BcDev &, T *, T *, const Dev &, const BcDev &, const T *, \
                                const T *, int, int, hipStream_t);                                \
  template void launch_rhs<T>(const Dev &, const T *, const T *, const T *, T *, T *, hipStream_t); \
  template void launch_sweep<T>(const Dev &, T *, T *, const T *, const T *, hipStream_t);        \
  template void launch_moments<T>(const Dev &, T *, T *, hipStream_t);                            \
  template bool launch_walk<T>(const Dev &, double *, T *, const T *, T *, int, int, hipStream_t);      \
  template bool launch_strip<T>(const Dev &, double *, T *, const T *, T *, int, int, StripBufs &, hipStream_t); \
  template void launch_res4<T>(const Dev &, const T *, const T *, hipStream_t);                  \
  template void launch_frd_out_chunk<T>(const Dev &, double *, int, int, const T *, const T *, const T *,       \
                                        hipStream_t);                                                           \
  template void launch_frd_in_chunk<T>(const Dev &, const double *, int, int, T *, T *, T *, hipStream_t);    \
  template void launch_strip_batch<T>(const StripJob<T> *, int, int, int, int, int, int, int *, int *, hipStream_t); \
  template void launch_fill_uniform<T>(const Dev &, const double *, const double *, T *, T *, T *, hipStream_t);
INST(double)
INST(float)
#undef INST

}  // namespace gsis
