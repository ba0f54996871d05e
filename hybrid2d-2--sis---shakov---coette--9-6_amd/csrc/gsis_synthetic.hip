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
// The inner loop is one cooperative launch that iterates on the device until
// the max |energy residual| drops to ns_tol or ns_maxit is reached:
//  * flux workgroups (1..nflux) own 16x16-cell tiles of the swept cells
//    3..n-3.  Per iteration each stages U of its tile plus two halo rings in
//    LDS, forms the primitives and viscosity, the face fluxes (inviscid,
//    viscous, HoT), then per cell Res, Lc, Lv, dt and everything the LU-SGS
//    sweeps need that does not depend on dU, stored by anti-diagonal plane;
//  * workgroup 0 runs the two sweeps.  Lane j of its wavefronts owns row j,
//    so one sweep step is one anti-diagonal plane: the forward recursion takes
//    dU(i-1,j) from the lane's own previous step and dU(i,j-1) from the lane
//    below (a cross-lane shift; between wavefronts a hand-off through an LDS
//    ring, no barrier per plane).  The backward pass writes U + dU of each
//    cell as soon as it is final; the sweeps read the pre-update U from the
//    records, so the update is the reference's "after both sweeps".  On a
//    mesh whose swept cells all have the normals (1,0) and (0,1) (checked
//    on the host at set-up) the sweep skips the normals: DF across an
//    i-face is dFx, across a j-face dFy.
// Any order of the cells that respects the sweep dependences gives the same
// dU, and the planes are visited in the reference's order, so the inner
// iteration counts are the reference's.  The loop ends by the reference's
// rule: iteration k+1 runs when max |Res_4| of iteration k exceeds ns_tol and
// k < ns_maxit.  Every wait between workgroups is bounded; a timeout ends the
// loop and is reported by the host.
#include <cstring>
#include <string>

#include "gsis_ns_common.h"
#include "gsis_synthetic.h"

namespace gsis {

namespace {
std::string g_syn_err;
}

const char *synthetic_error() { return g_syn_err.c_str(); }

constexpr int kTile = 16;                    // flux tile: kTile x kTile cells
constexpr int kTW = kTile + 4;               // with two halo rings each side
constexpr int kTC = kTW * kTW;
constexpr int kNFace = (kTile + 1) * kTile;  // i-faces (or j-faces) of a tile
constexpr int kMaxSweepWaves = 16;           // rows <= 1024
constexpr int kRingMax = 10;                 // sweep hand-off slots per wavefront (at most)
constexpr int kSlot = 14;                    // doubles per hand-off slot: dU, dFx, dFy, Cbj, Hsj
constexpr long long kWaitTicks = 200000000;  // bounded waits: 2 s of the 100 MHz wall clock

enum { kCtlGo = 1, kCtlErr = 3 };

// Sweep records are blocked by (plane, wavefront): block ((p-6)*nw + w) holds
// the 64 rows 3+64w .. 3+64w+63, field k of row 3+64w+l at (k/2)*128 + 2l + k%2,
// so each pair of fields of a plane is one contiguous 1 KiB load per
// wavefront.  Offsets in doubles; the record array stays below 4 GiB
// (synthetic_init), so byte offsets fit the 32-bit buffer offsets.
constexpr int kBlk = kNRec * 64;  // doubles per block
static_assert(kRecRes == 0 && kRecU == 4 && kRecRa == 8 && kRecCbi == 9 && kRecCbj == 10 && kRecHsi == 11 &&
                  kRecHsj == 12 && kRecRr == 13 && kRecNi == 14 && kRecNj == 16 && kNRec == 18,
              "the sweep's pair loads and the flux workgroups' pair stores assume this field order");
__device__ __forceinline__ size_t rec_blk(int p, int w, int nw) { return ((size_t)(p - 6) * nw + w) * kBlk; }
__device__ __forceinline__ size_t rec_cell(int i, int j, int nw) {  // the cell's first pair
  const int r = j - 3;
  return rec_blk(i + j, r >> 6, nw) + 2 * (r & 63);
}

// The record array as a raw buffer; loads and stores through it carry sc1,
// the cache policy of ld_agent / st_agent, so they are coherent across the
// device's XCDs.
constexpr int kBufWord3 = 0x00020000;  // gfx9-family descriptor word 3 (raw buffer, stride 0)
constexpr int kCoherent = 16;          // buffer instruction cache policy: sc1
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// Global-memory views of the mesh arrays: loads through them are global
// loads also where the compiler cannot infer the address space of a Dev
// field (the flux workgroups' function, the batched kernel's job table)
typedef const __attribute__((address_space(1))) double gdbl;
__device__ __forceinline__ gdbl *gp(const double *p) { return (gdbl *)p; }
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rec_rsrc(const SyntheticState &st) {
  return __builtin_amdgcn_make_buffer_rsrc(st.rec, (short)0, (int)((size_t)st.np * st.nw * kBlk * sizeof(double)),
                                           kBufWord3);
}
__device__ __forceinline__ void st_pair(__amdgpu_buffer_rsrc_t r, unsigned byte_off, double a, double b) {
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, double2{a, b}), r, byte_off, 0, kCoherent);
}
__device__ __forceinline__ double2 ld_pair(__amdgpu_buffer_rsrc_t r, unsigned voff, int soff) {
  return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, kCoherent));
}

// Every record slot's density field and its reciprocal to 1 (the rest stays
// zero): the slots of cells that are not swept keep this finite state for good.
__global__ void k_rec_init(SyntheticState st) {
  const size_t nslot = (size_t)st.np * st.nw * 64;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nslot) return;
  const size_t blk = t / 64, l = t % 64;
  st.rec[blk * kBlk + (kRecU / 2) * 128 + 2 * l] = 1.0;
  st.rec[blk * kBlk + (kRecRr / 2) * 128 + 2 * l + 1] = 1.0;
}

// S0 <- S - S0, S00 <- S on interior cells; the kinetic increment added to the
// synthetic state of the previous step on cells 3..n-3 (src/Hybrid_Sub.f90:45-94).
// Also resets the inner loop's control words.
__global__ void k_ns_prepare(Dev d, SyntheticState st) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) {
    st.ctl[kCtlGo] = (1.0 > st.tol && st.maxit > 0) ? 1 : -1;  // the reference starts from res = 1
    st.ctl[kCtlErr] = 0;
    st.iters[0] = 0;
    st.iters[1] = 0;
    st.iters[2] = 0;
  }
  if (t < st.nflux) {  // per flux workgroup: no iteration done, no T<0.1
    reinterpret_cast<int *>(st.part + 8 * t)[0] = 0;
    reinterpret_cast<int *>(st.part + 8 * t)[1] = 0;
  }
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

// comput_DFn (src/sub_LU_SGS.f90:122-143): DF = F.n(U + dU) - F.n(U) on the
// unit normal (n1, n2).  F.n is linear in n: F.n = Fx n1 + Fy n2 with the
// Euler fluxes Fx, Fy of the state, so a cell's DF on any normal is
// (Fx(U+dU) - Fx(U)) n1 + (Fy(U+dU) - Fy(U)) n2.  The sweep forms the two
// differences once per cell, when its dU is known, and both cells that take
// it as a neighbour (its i+1 and its j+1 forward, i-1 and j-1 backward)
// project them on their own normals.  1/rho is shared by the velocities and
// the pressure (hardware reciprocal and two Newton steps, about 1 ulp).
__device__ __forceinline__ double rcp_nr(double a) {
  double r = __builtin_amdgcn_rcp(a);
  double e = fma(-a, r, 1.0);
  r = fma(r, e, r);
  e = fma(-a, r, 1.0);
  return fma(r, e, r);
}
// Fx(u), Fy(u) with r = 1/u[0] as rcp_nr forms it
__device__ __forceinline__ void fxy_r(const double u[4], double r, double g, double fx[4], double fy[4]) {
  const double vx = u[1] * r, vy = u[2] * r;
  const double p = (g - 1.0) * (u[3] - 0.5 * (u[1] * u[1] + u[2] * u[2]) * r);
  const double e = u[3] + p;
  fx[0] = u[1];
  fx[1] = u[1] * vx + p;
  fx[2] = u[2] * vx;
  fx[3] = e * vx;
  fy[0] = u[2];
  fy[1] = u[1] * vy;
  fy[2] = u[2] * vy + p;
  fy[3] = e * vy;
}
__device__ __forceinline__ void fxy(const double u[4], double g, double fx[4], double fy[4]) {
  fxy_r(u, rcp_nr(u[0]), g, fx, fy);
}
// DF of state u and increment du on (n1, n2) in the reference's form (the
// sweep of the launch for more than 256 swept rows carries states)
__device__ __forceinline__ void dfn(const double u[4], const double du[4], double n1, double n2, double g,
                                    double df[4]) {
  double u2[4];
  for (int m = 0; m < 4; m++) u2[m] = u[m] + du[m];
  const double r1 = rcp_nr(u[0]), r2 = rcp_nr(u2[0]);
  const double un1 = (u[1] * n1 + u[2] * n2) * r1;
  const double p1 = (g - 1.0) * (u[3] - 0.5 * (u[1] * u[1] + u[2] * u[2]) * r1);
  const double un2 = (u2[1] * n1 + u2[2] * n2) * r2;
  const double p2 = (g - 1.0) * (u2[3] - 0.5 * (u2[1] * u2[1] + u2[2] * u2[2]) * r2);
  df[0] = u2[0] * un2 - u[0] * un1;
  df[1] = (u2[1] * un2 + p2 * n1) - (u[1] * un1 + p1 * n1);
  df[2] = (u2[2] * un2 + p2 * n2) - (u[2] * un1 + p1 * n2);
  df[3] = (u2[3] + p2) * un2 - (u[3] + p1) * un1;
}

struct NsShared {
  // sweep: lane 63 (forward) or lane 0 (backward) of wavefront w hands its
  // row's values of step s to the neighbouring wavefront in slot s % kRing,
  // at xw[(w * kRingMax + slot) * kSlot] (first in the struct: slot
  // addresses then fit the LDS instructions' offset fields)
  double xw[kMaxSweepWaves * kRingMax * kSlot];
  double U[4][kTC];    // tile + halo
  double pr[5][kTC];   // d, u, v, T, p (comput_duvtpckw)
  double mu[kTC];      // amu_of
  double fi[4][kNFace];  // i-face fluxes, face (i0+a, j0+b) at b*(kTile+1)+a
  double fj[4][kNFace];  // j-face fluxes, face (i0+a, j0+b) at b*kTile+a
  int prog[kMaxSweepWaves];  // steps a wavefront has handed on
  int cons[kMaxSweepWaves];  // steps whose hand-off a wavefront has taken
  int abort;                 // a hand-off wait timed out
  double red[kMaxSweepWaves];
  int word;
};

// the tile's cell fields for ns_face_of
struct TileCells {
  const NsShared &sh;
  const Dev &d;
  int ib, jb;  // cell (ib, jb) is LDS cell 0
  __device__ int l(int i, int j) const { return (j - jb) * kTW + (i - ib); }
  __device__ double u(int i, int j) const { return sh.pr[1][l(i, j)]; }
  __device__ double v(int i, int j) const { return sh.pr[2][l(i, j)]; }
  __device__ double T(int i, int j) const { return sh.pr[3][l(i, j)]; }
  __device__ double mu(int i, int j) const { return sh.mu[l(i, j)]; }
  __device__ double x1(int i, int j) const { return gp(d.x1)[cidx(d, i, j)]; }
  __device__ double y1(int i, int j) const { return gp(d.y1)[cidx(d, i, j)]; }
};

// Residual_FVM face flux (src/sub_Residual.f90:68-202, :210-338) of face
// (i,j), primitives from the tile in LDS
__device__ void face_flux(const Dev &d, const TileCells &tc, int dir, int i, int j, bool visc, double fl[4]) {
  const NsShared &sh = tc.sh;
  double ul[4], ur[4];
  for (int m = 0; m < 4; m++) {
    const double *q = sh.pr[m < 3 ? m : 4];  // d, u, v, p
    double u1, u2, u3, u4;
    if (dir == 0) { u1 = q[tc.l(i - 2, j)]; u2 = q[tc.l(i - 1, j)]; u3 = q[tc.l(i, j)]; u4 = q[tc.l(i + 1, j)]; }
    else { u1 = q[tc.l(i, j - 2)]; u2 = q[tc.l(i, j - 1)]; u3 = q[tc.l(i, j)]; u4 = q[tc.l(i, j + 1)]; }
    ul[m] = u2 + 0.5 * minmod_ns(u2 - u1, u3 - u2);
    ur[m] = u3 - 0.5 * minmod_ns(u3 - u2, u4 - u3);
  }
  const int g = gidx(d, i, j);
  double s, n1, n2;
  if (dir == 0) { s = gp(d.si)[g]; n1 = gp(d.ni1)[g]; n2 = gp(d.ni2)[g]; }
  else { s = gp(d.sj)[g]; n1 = gp(d.nj1)[g]; n2 = gp(d.nj2)[g]; }
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
  ns_face_of(tc, d, dir, i, j, o);
  const int N = d.ncell, c = cidx(d, i, j);
  gdbl *tg = gp(dir == 0 ? d.taoi : d.taoj);
  gdbl *qg = gp(dir == 0 ? d.qi : d.qj);
  const double t11 = o[0] + tg[c], t22 = o[1] + tg[N + c], t12 = o[2] + tg[2 * N + c];
  const double qx = o[4] + qg[c], qy = o[5] + qg[N + c];
  const double u1 = o[6], v1 = o[7];
  const double e1 = u1 * t11 + v1 * t12 + qx;
  const double e2 = u1 * t12 + v1 * t22 + qy;
  fl[1] = fl[1] + (t11 * n1 + t12 * n2) * s;
  fl[2] = fl[2] + (t12 * n1 + t22 * n2) * s;
  fl[3] = fl[3] + (e1 * n1 + e2 * n2) * s;
}

// One tile of Comput_Residual_one_mesh (src/sub_NS_singlegrid.f90:45-114):
// comput_duvtpckw with its T<0.1 stop, get_viscous, the face fluxes, Res,
// comput_Lvc, comput_dt and the LU-SGS diagonal (src/sub_LU_SGS.f90:34-36),
// stored as sweep records.  Returns this tile's max |Res_4| and T<0.1 flag.
__device__ void flux_tile(const Dev &d, const SyntheticState &st, NsShared &sh, int tile, double &rmax, int &bad) {
  const int nx = d.nx, ny = d.ny, N = d.ncell;
  const int tid = threadIdx.x, nt = blockDim.x;
  const double g = d.gamma;
  const bool visc = st.if_viscous != 0;
  const int i0 = 3 + (tile % st.ntx) * kTile, j0 = 3 + (tile / st.ntx) * kTile;
  const int ilast = min(i0 + kTile - 1, nx - 3), jlast = min(j0 + kTile - 1, ny - 3);
  const int nci = ilast - i0 + 1, ncj = jlast - j0 + 1;
  const TileCells tc{sh, d, i0 - 2, j0 - 2};
  for (int t = tid; t < kTC; t += nt) {
    const int i = i0 - 2 + t % kTW, j = j0 - 2 + t / kTW;
    if (i > nx + 1 || j > ny + 1) continue;
    const int c = cidx(d, i, j);
    double u[4];
    for (int m = 0; m < 4; m++) {
      u[m] = __hip_atomic_load(gp(d.U) + m * N + c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      sh.U[m][t] = u[m];
    }
    const double dd = u[0], uu = u[1] / dd, vv = u[2] / dd;
    const double T = (u[3] - 0.5 * dd * (uu * uu + vv * vv)) / (d.cv * dd);
    sh.pr[0][t] = dd;
    sh.pr[1][t] = uu;
    sh.pr[2][t] = vv;
    sh.pr[3][t] = T;
    sh.pr[4][t] = (1. / (g * d.ma * d.ma)) * dd * T;
    sh.mu[t] = visc ? 1. / d.re * pow(T, 0.8100000023841858) : 0.0;
    if (i >= 1 && i <= nx - 1 && j >= 1 && j <= ny - 1 && T < 0.10000000149011612) bad = 1;
  }
  __syncthreads();
  const int nfi = (nci + 1) * ncj, nfj = nci * (ncj + 1);
  for (int t = tid; t < nfi + nfj; t += nt) {
    double fl[4];
    if (t < nfi) {
      const int a = t % (nci + 1), b = t / (nci + 1);
      face_flux(d, tc, 0, i0 + a, j0 + b, visc, fl);
      for (int m = 0; m < 4; m++) sh.fi[m][b * (kTile + 1) + a] = fl[m];
    } else {
      const int a = (t - nfi) % nci, b = (t - nfi) / nci;
      face_flux(d, tc, 1, i0 + a, j0 + b, visc, fl);
      for (int m = 0; m < 4; m++) sh.fj[m][b * kTile + a] = fl[m];
    }
  }
  __syncthreads();
  const int nw = st.nw;
  const __amdgpu_buffer_rsrc_t rr = rec_rsrc(st);
  for (int t = tid; t < nci * ncj; t += nt) {
    const int a = t % nci, b = t / nci, i = i0 + a, j = j0 + b;
    double res[4];
    for (int m = 0; m < 4; m++)
      res[m] = sh.fi[m][b * (kTile + 1) + a + 1] - sh.fi[m][b * (kTile + 1) + a] + sh.fj[m][(b + 1) * kTile + a] -
               sh.fj[m][b * kTile + a];
    rmax = fmax(rmax, fabs(res[3]));
    // comput_Lvc and comput_dt (src/sub_NS_singlegrid.f90:160-221)
    const int g0 = gidx(d, i, j), gi = gidx(d, i + 1, j), gj = gidx(d, i, j + 1);
    gdbl *gvol = gp(d.vol), *gsi = gp(d.si), *gsj = gp(d.sj);
    gdbl *gni1 = gp(d.ni1), *gni2 = gp(d.ni2), *gnj1 = gp(d.nj1), *gnj2 = gp(d.nj2);
    const double s0 = gvol[g0];
    const double si = 0.5 * (gsi[g0] + gsi[gi]);
    const double ni1 = 0.5 * (gni1[g0] + gni1[gi]), ni2 = 0.5 * (gni2[g0] + gni2[gi]);
    const double sj = 0.5 * (gsj[g0] + gsj[gj]);
    const double nj1 = 0.5 * (gnj1[g0] + gnj1[gj]), nj2 = 0.5 * (gnj2[g0] + gnj2[gj]);
    const int lc = tc.l(i, j);
    const double uu = sh.pr[1][lc], vv = sh.pr[2][lc];
    const double uni = uu * ni1 + vv * ni2, unj = uu * nj1 + vv * nj2;
    const double cc = sqrt(sh.pr[3][lc]) / d.ma;
    const double lci = (fabs(uni) + cc) * si, lcj = (fabs(unj) + cc) * sj;
    const double tmp1 = g / sh.pr[0][lc] * (sh.mu[lc] / d.pr + 0.0 / 0.9);
    const double lvi = visc ? tmp1 * si * si / s0 : 0.0, lvj = visc ? tmp1 * sj * sj / s0 : 0.0;
    double dt;
    if (!st.local_dt) dt = d.dt;
    else {
      dt = st.cfl * gvol[g0] / (lci + lcj + 1.0 * (lvi + lvj));
      if (dt > st.dtmax) dt = st.dtmax;
      if (dt < st.dtmin) dt = st.dtmin;
    }
    double alfa = gvol[g0] / dt + 1.0 * (lci + lcj);
    if (visc) alfa = alfa + 2.0 * (lvi + lvj);
    const unsigned o = (unsigned)(rec_cell(i, j, nw) * sizeof(double));
    st_pair(rr, o + (kRecRes / 2) * 1024, res[0], res[1]);
    st_pair(rr, o + (kRecRes / 2 + 1) * 1024, res[2], res[3]);
    st_pair(rr, o + (kRecU / 2) * 1024, sh.U[0][lc], sh.U[1][lc]);
    st_pair(rr, o + (kRecU / 2 + 1) * 1024, sh.U[2][lc], sh.U[3][lc]);
    st_pair(rr, o + (kRecRa / 2) * 1024, 1.0 / alfa, 0.5 * lci + lvi);                // ra, cbi
    // si/2 and sj/2 weigh the forward sweep's terms across the cell's lower
    // faces; on the first swept column (row) the cell across is not swept and
    // the term is left out (src/sub_LU_SGS.f90:44-56): a zero weight leaves it
    // out without a test in the sweep (the backward sweep reads these of cells
    // i+1, j+1 >= 4 only)
    const double hsi = i == 3 ? 0.0 : 0.5 * gsi[g0], hsj = j == 3 ? 0.0 : 0.5 * gsj[g0];
    st_pair(rr, o + (kRecCbj / 2) * 1024, 0.5 * lcj + lvj, hsi);  // cbj, hsi
    st_pair(rr, o + (kRecHsj / 2) * 1024, hsj, rcp_nr(sh.U[0][lc]));  // hsj, 1/rho
    st_pair(rr, o + (kRecNi / 2) * 1024, gni1[g0], gni2[g0]);     // ni1, ni2
    st_pair(rr, o + (kRecNj / 2) * 1024, gnj1[g0], gnj2[g0]);     // nj1, nj2
  }
  __syncthreads();  // the next tile reuses the LDS
}

__device__ __forceinline__ void barrier_lds() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// x of lane l-1 (wave_shr:1) or l+1 (wave_shl:1) of the wavefront; 0 in the
// lane that has no such neighbour (lane 0, or lane 63)
__device__ __forceinline__ double lane_shift(double x, bool up) {
  const unsigned long long b = __double_as_longlong(x);
  int lo = (int)(unsigned)b, hi = (int)(unsigned)(b >> 32);
  if (up) {
    lo = __builtin_amdgcn_mov_dpp(lo, 0x138, 0xf, 0xf, true);
    hi = __builtin_amdgcn_mov_dpp(hi, 0x138, 0xf, 0xf, true);
  } else {
    lo = __builtin_amdgcn_mov_dpp(lo, 0x130, 0xf, 0xf, true);
    hi = __builtin_amdgcn_mov_dpp(hi, 0x130, 0xf, 0xf, true);
  }
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
// the same, with 1.0 in the lane that has no such neighbour (a density)
__device__ __forceinline__ double lane_shift_rho(double x, bool up) {
  const unsigned long long b = __double_as_longlong(x), o = __double_as_longlong(1.0);
  int lo = (int)(unsigned)b, hi = (int)(unsigned)(b >> 32);
  const int olo = (int)(unsigned)o, ohi = (int)(unsigned)(o >> 32);
  if (up) {
    lo = __builtin_amdgcn_update_dpp(olo, lo, 0x138, 0xf, 0xf, false);
    hi = __builtin_amdgcn_update_dpp(ohi, hi, 0x138, 0xf, 0xf, false);
  } else {
    lo = __builtin_amdgcn_update_dpp(olo, lo, 0x130, 0xf, 0xf, false);
    hi = __builtin_amdgcn_update_dpp(ohi, hi, 0x130, 0xf, 0xf, false);
  }
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// One sweep step's record of a lane's own cell (all of it in the forward
// sweep; the backward sweep leaves res and reads dU* instead).
struct Rec {
  double res[4], ra, ni[2], nj[2], hsi, hsj, u[4], cbi, cbj, rr;
};

// Every lane loads its row's slot of the plane, whether or not the row has a
// cell there (the slot exists for every row of a sweep wavefront): nine
// 16-byte loads, one scalar block offset and immediate pair offsets.  On a
// mesh whose normals are (1,0) and (0,1) (kAxis) the last two pairs, the
// normals, are not loaded: seven loads.
template <bool FWD, bool kAxis>
__device__ __forceinline__ void load_rec(__amdgpu_buffer_rsrc_t rr, int blk_bytes, __amdgpu_buffer_rsrc_t rd,
                                         int ds_bytes, int lane, Rec &x) {
  const unsigned v = (unsigned)lane * 16u;
  auto pr = [&](int q) { return ld_pair(rr, v + (unsigned)q * 1024u, blk_bytes); };
  double2 r01, r23;
  if (FWD) {
    r01 = pr(kRecRes / 2);
    r23 = pr(kRecRes / 2 + 1);
  } else {  // dU* was stored by this lane
    r01 = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(rd, v, ds_bytes, 0));
    r23 = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(rd, v + 1024u, ds_bytes, 0));
  }
  const double2 u01 = pr(kRecU / 2), u23 = pr(kRecU / 2 + 1), p4 = pr(kRecRa / 2), p5 = pr(kRecCbj / 2),
                p6 = pr(kRecHsj / 2);
  x.res[0] = r01.x; x.res[1] = r01.y; x.res[2] = r23.x; x.res[3] = r23.y;
  x.u[0] = u01.x; x.u[1] = u01.y; x.u[2] = u23.x; x.u[3] = u23.y;
  x.ra = p4.x; x.cbi = p4.y;
  x.cbj = p5.x; x.hsi = p5.y;
  x.hsj = p6.x;
  x.rr = p6.y;
  if (!kAxis) {
    const double2 p7 = pr(kRecNi / 2), p8 = pr(kRecNj / 2);
    x.ni[0] = p7.x; x.ni[1] = p7.y;
    x.nj[0] = p8.x; x.nj[1] = p8.y;
  }
}

// LDS words read and written by several wavefronts of the workgroup
__device__ __forceinline__ int lds_get(int *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); }
__device__ __forceinline__ void lds_put(int *p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// x in a vector register: LDS addresses formed from it keep their constant
// parts in the instructions' offset fields
__device__ __forceinline__ int in_vgpr(int x) {
  asm volatile("" : "+v"(x));
  return x;
}
__device__ __forceinline__ void lgkm_wait() { __builtin_amdgcn_s_waitcnt(0xC07F); }  // lgkmcnt(0)
__device__ __forceinline__ void mem_order() { asm volatile("" ::: "memory"); }  // no compiler reordering

// Waits until the LDS word *p reaches want; bounded by kWaitTicks, after
// which (or once another wavefront has timed out) it gives up, sets the
// workgroup's abort word and the error word, and the sweep runs to its end on
// whatever values it has, so that every wavefront still reaches the barriers.
__device__ __forceinline__ void wait_lds(int *p, int want, int *abort, int *err) {
  if (lds_get(p) >= want || lds_get(abort)) return;
  const long long t0 = wall_clock64();
  while (lds_get(p) < want && !lds_get(abort)) {
    __builtin_amdgcn_s_sleep(1);
    if (wall_clock64() - t0 > kWaitTicks) {
      lds_put(abort, 1);
      st_agent(err, 1);
    }
  }
}

// du_LU_SGS_2D (src/sub_LU_SGS.f90:4-113) and the update of cells 3..n-3.
// Lane j owns row j (j = 3 + thread index); step p visits plane i + j = p.
// Each step is straight-line code: every lane issues the same loads and
// stores (a lane without a cell stores out of the buffer's range, which
// discards the store), so the waits for a record loaded kDepth - 1 steps
// earlier do not wait for the later ones.  The loop runs a multiple of kDepth steps;
// the steps after the last plane have no cell in any row.
// kDepth records in registers: each step's record is loaded at the start of
// the step kDepth - 1 steps before it is used (the loads come from the
// device's coherence point, written by other CUs, and go out among the
// step's arithmetic), which takes (kDepth - 1) * 34 more VGPRs than one
// record.
//
// What a lane carries from one step to the next: its own cell's dU, flux
// differences dFx, dFy, Cbi and Hsi (the cell is the i-neighbour of the
// lane's next one), and the j-neighbour row's dU, dFx, dFy, Cbj and Hsj,
// shifted in from the next lane at the end of the step (in the edge lane:
// from the neighbouring wavefront, below).  The launch for more than 256
// swept rows (kDiff false) carries the states U instead and forms DF in the
// reference's form: with flux differences that launch, held to 128 VGPRs and
// spilling, gave results off by about 1e-6 on meshes of 294-302 swept rows,
// varying from run to run, for a reason not found; with states it matches
// the oracle on every tall mesh tried.
//
// Hand-offs between wavefronts (LDS ring of kRing slots per wavefront): the
// edge lane (63 forward, 0 backward) of wavefront w writes its values of step
// s in slot s % kRing; prog[w] = s + 1 says slots up to s are written,
// cons[w] = s + 1 that wavefront w has taken the slots up to s of its
// neighbour.  No step waits for an LDS round trip in the steady state:
//  * the words and the neighbour's next slot are read at the end of a step,
//    one step before they are needed, the slot straight into the edge lane of
//    the shifted row (first the word, then the slot: the LDS performs a
//    wavefront's operations in order, so a slot read after a word that says
//    "written" holds the written values);
//  * the next step first works on its own i-neighbour, then waits for those
//    reads (lgkmcnt(0)), publishes prog and cons, and only if the word it read
//    was too old waits for the neighbour and reads the slot again.  A
//    wavefront so settles one to two steps behind its neighbour;
//  * a slot is written only when the word cons read one step earlier says it
//    has been taken (cons only grows), else after waiting for it.
//
// kAxis (with kDiff): every swept cell's normals are ni = (1,0), nj = (0,1),
// so DF across an i-face is dFx and across a j-face dFy, exactly the values
// fma(dFy, ni2, dFx ni1) and fma(dFy, nj2, dFx nj1) take there (up to the
// sign of a zero): the normals are not loaded, a lane hands on dFy of its row
// and keeps dFx.
template <bool FWD, int kDepth, bool kDiff, bool kAxis>
__device__ void sweep_pass(const Dev &d, const SyntheticState &st, NsShared &sh) {
  static_assert(kDiff || !kAxis, "the axis-aligned form carries flux differences");
  const int nx = d.nx, ny = d.ny, N = d.ncell, nwr = st.nw;
  const double g = d.gamma;
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int j = 3 + tid;
  const int pmin = 6, pmax = nx + ny - 6;
  const int nstep = pmax - pmin + 1, nstepd = (nstep + kDepth - 1) / kDepth * kDepth;
  if (w >= nwr) return;  // a wavefront without rows
  // the wavefront this one takes the neighbouring row from, and the one it
  // hands its edge row to
  const int from = FWD ? w - 1 : w + 1, to = FWD ? w + 1 : w - 1;
  const bool has_from = from >= 0 && from < nwr, has_to = to >= 0 && to < nwr;
  const bool edge_in = FWD ? lane == 0 : lane == 63, edge_out = FWD ? lane == 63 : lane == 0;
  constexpr int kRing = 2 * kDepth;  // a multiple of kDepth: no slot index wraps within a group of steps
  static_assert(kRing <= kRingMax, "hand-off ring");
  int *abort = &sh.abort, *err = st.ctl + kCtlErr;
  const int fromc = has_from ? from : w, toc = has_to ? to : w;  // (in range also where unused)
  int *pw = &sh.prog[in_vgpr(w)], *pfrom = &sh.prog[in_vgpr(fromc)];  // prog of this wavefront, of from
  int *cw = &sh.cons[in_vgpr(w)], *cto = &sh.cons[in_vgpr(toc)];      // cons of this wavefront, of to
  // records, dU* (pairs of components, 256 doubles per block) and U as raw buffers
  const __amdgpu_buffer_rsrc_t rr = rec_rsrc(st);
  const __amdgpu_buffer_rsrc_t rd = __builtin_amdgcn_make_buffer_rsrc(
      st.dstar, (short)0, (int)((size_t)st.np * nwr * 256 * sizeof(double)), kBufWord3);
  const __amdgpu_buffer_rsrc_t ru =
      __builtin_amdgcn_make_buffer_rsrc(d.U, (short)0, (int)((size_t)4 * N * sizeof(double)), kBufWord3);
  constexpr unsigned kOff = 0x80000000u;  // past every buffer's range: the store is discarded
  // the plane of step s (clamped to the last one: the steps past it reload it)
  // as a block index (p - 6) * nw + w
  auto bidx = [&](int s) {
    const int q = min(s, nstep - 1);
    return (FWD ? q : nstep - 1 - q) * nwr + w;
  };
  auto pof = [&](int s) { return FWD ? pmin + s : pmax - s; };
  auto act = [&](int s) {
    const int p = pof(s);
    return s < nstep && j <= ny - 3 && p - j >= 3 && p - j <= nx - 3;
  };
  // byte offset of cell (p - j, j) in a component of U, less 8 p
  const unsigned ucell = (unsigned)(((j + 1) * d.NI + 1 - j) * (int)sizeof(double));
  // own previous cell (the i-neighbour) and the j-neighbour row, both of the
  // previous step: dU and the weights, and either the flux differences dFx,
  // dFy (kDiff) or the state U (cfx, nfx hold U; cfy, nfy unused)
  double cdu[4], cfx[4], cfy[4], ccbi = 0.0, chsi = 0.0;
  double ndu[4], nfx[4], nfy[4], ncb = 0.0, nhs = 0.0;
  for (int m = 0; m < 4; m++) {
    cdu[m] = cfy[m] = ndu[m] = nfy[m] = 0.0;
    cfx[m] = nfx[m] = kDiff ? 0.0 : 1.0;
  }
  // the neighbour's slot at xb into the edge lane of the j-neighbour row
  auto read_slot = [&](const double *xb) {
    if (edge_in) {
      for (int m = 0; m < 4; m++) {
        ndu[m] = xb[m];
        if (!kAxis) nfx[m] = xb[4 + m];
        if (kDiff) nfy[m] = xb[8 + m];
      }
      ncb = xb[12];
      nhs = xb[13];
    }
  };
  auto slot_of = [&](int v, int t) { return (v * kRingMax + t % kRing) * kSlot; };
  int fprog = 0, fcons = 0;  // prog[from] and cons[to] as read one step earlier
  Rec r[kDepth];
#pragma unroll
  for (int q = 0; q < kDepth - 1; q++) load_rec<FWD, kAxis>(rr, bidx(q) * kBlk * 8, rd, bidx(q) * 2048, lane, r[q]);
  // block of the loads issued in step s (those of step s + kDepth - 1), moving
  // by one plane per step and held at the last plane; block of step s
  const int dbl = FWD ? nwr : -nwr, blim = FWD ? (nstep - 1) * nwr + w : w;
  int bl = bidx(kDepth - 1), bs = w;
  // slots of the group of steps s0 .. s0+kDepth-1: xin + kSlot q (from's), xout + kSlot q (this wavefront's)
  const double *xin = sh.xw;
  double *xout = sh.xw;
  auto step = [&](int s, int q, Rec &x) {
    // the record of step s + kDepth - 1, into the slot step s - 1 has used:
    // issued first, so that the loads go out among this step's arithmetic
    load_rec<FWD, kAxis>(rr, bl * (kBlk * 8), rd, bl * 2048, lane, r[(q + kDepth - 1) % kDepth]);
    bl = FWD ? min(bl + dbl, blim) : max(bl + dbl, blim);
    const int p = pof(s), i = p - j;
    const bool on = act(s);
    // off-diagonal term from the i-neighbour (this lane's previous cell)
    double dfi[4], dfj[4];
    if constexpr (kAxis) {
      for (int m = 0; m < 4; m++) dfi[m] = cfx[m];
    } else if constexpr (kDiff) {
      for (int m = 0; m < 4; m++) dfi[m] = fma(cfy[m], x.ni[1], cfx[m] * x.ni[0]);
    } else {
      dfn(cfx, cdu, x.ni[0], x.ni[1], g, dfi);
    }
    if (s > 0) {
      lgkm_wait();  // the previous step's slot write and reads are complete
      if (has_to && lane == 0) lds_put(pw, s);
      if (has_from) {
        if (__builtin_amdgcn_readfirstlane(fprog) < s) {  // read too early: wait for slot s-1, read it again
          wait_lds(pfrom, s, abort, err);
          mem_order();
          read_slot(&sh.xw[slot_of(fromc, s - 1)]);
          lgkm_wait();
        }
        if (lane == 0) lds_put(cw, s);
      }
    }
    // off-diagonal term from the j-neighbour
    if constexpr (kAxis) {
      for (int m = 0; m < 4; m++) dfj[m] = nfy[m];
    } else if constexpr (kDiff) {
      for (int m = 0; m < 4; m++) dfj[m] = fma(nfy[m], x.nj[1], nfx[m] * x.nj[0]);
    } else {
      dfn(nfx, ndu, x.nj[0], x.nj[1], g, dfj);
    }
    // The terms across faces to cells that are not swept (i = 3, j = 3 forward;
    // i = nx-3, j = ny-3 backward) vanish without a test: the forward weights
    // Hsi, Hsj of those cells are zero, and a slot of a cell that is not swept
    // holds U = (1,0,0,0) and zero weights (written once by k_rec_init, never by
    // the flux workgroups), so its dU and flux differences stay zero.
    double du[4];
    for (int m = 0; m < 4; m++) {
      // forward:  dui = 0.5 (DF si(i,j) + Lci(i-1,j) dU) + Lvi(i-1,j) dU      (:44-49)
      // backward: dui = -0.5 (DF si(i+1,j) - Lci(i+1,j) dU) + Lvi(i+1,j) dU   (:82-86)
      const double ti = FWD ? fma(x.hsi, dfi[m], ccbi * cdu[m]) : fma(-chsi, dfi[m], ccbi * cdu[m]);
      const double tj = FWD ? fma(x.hsj, dfj[m], ncb * ndu[m]) : fma(-nhs, dfj[m], ncb * ndu[m]);
      double a = FWD ? x.res[m] : 0.0;
      a = a + ti;
      a = a + tj;
      du[m] = FWD ? a * x.ra : x.res[m] + a * x.ra;  // backward: x.res holds dU*
    }
    // plain stores: dU* is read back by this lane, U is written back to the
    // device's coherence point by the release fence after the sweep
    if (FWD) {
      const unsigned v = on ? (unsigned)lane * 16u : kOff;
      const int so = bs * 2048;  // bidx(s) where a lane has a cell
      __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, double2{du[0], du[1]}), rd, v, so, 0);
      __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, double2{du[2], du[3]}), rd, v + 1024u, so,
                                             0);
    } else {
      const unsigned v = on ? ucell + (unsigned)p * 8u : kOff;
      for (int m = 0; m < 4; m++)
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x.u[m] + du[m]), ru, v,
                                              m * N * (int)sizeof(double), 0);
    }
    // carry: this cell's dU and flux differences, and the j-neighbour row's
    // of the next step (zero in the edge lane of a wavefront without a
    // neighbour, where the term has a zero weight)
    if constexpr (kDiff) {
      double u2[4], fx1[4], fy1[4], fx2[4], fy2[4];
      for (int m = 0; m < 4; m++) {
        cdu[m] = on ? du[m] : 0.0;
        u2[m] = x.u[m] + cdu[m];
      }
      fxy_r(x.u, x.rr, g, fx1, fy1);  // 1/rho of the record's state from the record
      fxy(u2, g, fx2, fy2);
      for (int m = 0; m < 4; m++) {
        cfx[m] = fx2[m] - fx1[m];
        cfy[m] = fy2[m] - fy1[m];
        ndu[m] = lane_shift(cdu[m], FWD);
        if (!kAxis) nfx[m] = lane_shift(cfx[m], FWD);
        nfy[m] = lane_shift(cfy[m], FWD);
      }
    } else {  // the state (the edge lane of a wavefront without a neighbour gets (1,0,0,0), dU = 0)
      for (int m = 0; m < 4; m++) {
        cfx[m] = x.u[m];
        cdu[m] = on ? du[m] : 0.0;
        nfx[m] = m == 0 ? lane_shift_rho(cfx[m], FWD) : lane_shift(cfx[m], FWD);
        ndu[m] = lane_shift(cdu[m], FWD);
      }
    }
    ccbi = x.cbi;
    chsi = x.hsi;
    ncb = lane_shift(x.cbj, FWD);
    nhs = lane_shift(x.hsj, FWD);
    if (has_to) {  // hand the edge row on, once slot s % kRing has been taken
      if (__builtin_amdgcn_readfirstlane(fcons) < s - kRing + 1) {
        wait_lds(cto, s - kRing + 1, abort, err);
        mem_order();
      }
      if (edge_out) {
        double *xb = xout + kSlot * q;
        for (int m = 0; m < 4; m++) {
          xb[m] = cdu[m];
          if (!kAxis) xb[4 + m] = cfx[m];
          if (kDiff) xb[8 + m] = cfy[m];
        }
        xb[12] = x.cbj;
        xb[13] = x.hsj;
      }
      mem_order();
      fcons = lds_get(cto);
    }
    if (has_from) {  // for the next step: the word, then the slot of this step
      mem_order();
      fprog = lds_get(pfrom);
      mem_order();
      read_slot(xin + kSlot * q);
    }
    bs += nwr;
  };
  for (int s = 0; s < nstepd; s += kDepth) {
    xin = &sh.xw[in_vgpr(slot_of(fromc, s))];
    xout = &sh.xw[in_vgpr(slot_of(w, s))];
#pragma unroll
    for (int q = 0; q < kDepth; q++) step(s + q, q, r[q]);
  }
}

// The wavefronts of a pass are coupled only through their hand-offs: a
// wavefront may run up to 2 kDepth steps ahead of the one that takes its edge row.
// kDiff: carry flux differences (the launches of 256 threads); the launch of
// up to 1024 threads carries states (see sweep_pass).  kAxis: the mesh's
// normals are (1,0) and (0,1) on every swept cell.
template <int kDepth, bool kDiff = true, bool kAxis = false>
__device__ void sweep(const Dev &d, const SyntheticState &st, NsShared &sh) {
  const int tid = threadIdx.x;
  if (tid < kMaxSweepWaves) { sh.prog[tid] = 0; sh.cons[tid] = 0; }
  if (tid == 0) sh.abort = 0;
  barrier_lds();
  sweep_pass<true, kDepth, kDiff, kAxis>(d, st, sh);  // dU* = (Res + dui + duj)/alfa, planes 6 .. nx+ny-6 (:26-69)
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this lane's dU* stores are complete
  barrier_lds();
  if (tid < kMaxSweepWaves) { sh.prog[tid] = 0; sh.cons[tid] = 0; }
  barrier_lds();
  sweep_pass<false, kDepth, kDiff, kAxis>(d, st, sh);  // dU = dU* + (dui + duj)/alfa, planes nx+ny-6 .. 6 (:73-111), then U + dU
  // release: the U stores reach the device's coherence point before the
  // flux workgroups are told to read U (L2 write-back, then vmcnt(0))
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
}

// Polls *p until it reaches want (ge) or equals want or is negative (!ge).
// Bounded by kWaitTicks of the 100 MHz wall clock; a timeout sets *err.  The
// flux workgroups wait through a whole sweep, so they poll with long sleeps
// (SLEEP) to keep their loads off the memory the sweep reads.
template <int SLEEP>
__device__ bool wait_until(const int *p, int want, bool ge, int *err) {
  const long long t0 = wall_clock64();
  int it = 0, v = ld_agent(p);
  while (!(ge ? v >= want : (v == want || v < 0))) {
    __builtin_amdgcn_s_sleep(SLEEP);
    if ((++it & 15) == 0) {
      if (ld_agent(err) != 0) return false;
      if (wall_clock64() - t0 > kWaitTicks) {
        st_agent(err, 1);
        return false;
      }
    }
    v = ld_agent(p);
  }
  return true;
}

// A copy of x with every 32-bit word read once and made wavefront-uniform
// (scalar registers): x is a function argument passed through the stack,
// whose fields would otherwise be reloaded per use into vector registers.
template <typename T>
__device__ __forceinline__ T uniform_copy(const T &x) {
  static_assert(sizeof(T) % 4 == 0, "word copy");
  T y;
  const int *src = reinterpret_cast<const int *>(&x);
  int *dst = reinterpret_cast<int *>(&y);
#pragma unroll
  for (int k = 0; k < (int)(sizeof(T) / 4); k++) dst[k] = __builtin_amdgcn_readfirstlane(src[k]);
  return y;
}

// The flux workgroups' part of the inner loop.  A separate (not inlined)
// function: the sweep's registers are then allocated without the flux code's
// live values, which otherwise push the sweep's scalar registers into
// spills.  One copy per launch size B, so that each gets the register budget
// of its kernel's workgroup size (one shared copy is held to 1024 threads).
// Flux workgroup w reports an iteration in its own slot part[8w..]: its max
// |Res_4| and T<0.1 flag, then the iteration number once those stores are
// complete.
template <int B>
__device__ __attribute__((noinline)) void ns_flux_role(const Dev d_arg, const SyntheticState st_arg, NsShared &sh) {
  const Dev d = uniform_copy(d_arg);
  const SyntheticState st = uniform_copy(st_arg);
  int *ctl = st.ctl;
  const int tid = threadIdx.x;
  const int wg = blockIdx.x - 1, ntiles = st.ntx * st.nty;
  int *slot = reinterpret_cast<int *>(st.part + 8 * wg);
  for (int k = st.k0;; k++) {
    if (tid == 0) {
      int go = -1;
      if (wait_until<60>(ctl + kCtlGo, k, false, ctl + kCtlErr)) go = ld_agent(ctl + kCtlGo);
      sh.word = go == k ? 1 : 0;
    }
    __syncthreads();
    const int run = sh.word;
    __syncthreads();
    if (!run) return;
    double rmax = 0.0;
    int bad = 0;
    for (int tile = wg; tile < ntiles; tile += st.nflux) flux_tile(d, st, sh, tile, rmax, bad);
    // the workgroup's max |Res_4| and T<0.1 flag
    for (int off = 32; off > 0; off >>= 1) rmax = fmax(rmax, __shfl_down(rmax, off, 64));
    bad = __syncthreads_or(bad);
    if ((tid & 63) == 0) sh.red[tid >> 6] = rmax;
    __syncthreads();
    if (tid == 0) {
      double m = sh.red[0];
      for (int q = 1; q < (int)(blockDim.x >> 6); q++) m = fmax(m, sh.red[q]);
      st_agent(st.part + 8 * wg + 1, m);
      st_agent(slot + 1, bad);
    }
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this thread's stores are complete
    __syncthreads();
    if (tid == 0) st_agent(slot, k);
  }
}

// The inner loop (src/Hybrid_Sub.f90:99-104): workgroup 0 sweeps, the others
// compute the fluxes; see the file comment.  kAxis: the sweep of a mesh whose
// normals are (1,0) and (0,1) on every swept cell (launches of 256 threads).
template <int B, bool kAxis>
__global__ void __launch_bounds__(B) k_ns_loop(Dev d, SyntheticState st) {
  static_assert(B <= 256 || !kAxis, "the axis-aligned sweep carries flux differences");
  __shared__ NsShared sh;
  int *ctl = st.ctl;
  const int tid = threadIdx.x;
  if (blockIdx.x > 0) {
    ns_flux_role<B>(d, st, sh);
    return;
  }
  // workgroup 0: the LU-SGS sweeps and the loop control
  int it = st.k0 - 1, bad = 0, stopped = 0;
  double res_last = 0.0;
  if (ld_agent(ctl + kCtlGo) == st.k0) {
    for (int k = st.k0;; k++) {
      // every flux workgroup's slot holds k; one lane per workgroup polls
      int ok = 1;
      double rmax = 0.0;
      int b = 0;
      for (int q = tid; q < st.nflux; q += blockDim.x) {
        const int *slot = reinterpret_cast<const int *>(st.part + 8 * q);
        if (!wait_until<1>(slot, k, true, ctl + kCtlErr)) ok = 0;
        rmax = fmax(rmax, ld_agent(st.part + 8 * q + 1));
        b |= ld_agent(slot + 1);
      }
      ok = __syncthreads_and(ok);
      b = __syncthreads_or(b);
      if (!ok) break;  // a wait timed out: end the loop
      for (int off = 32; off > 0; off >>= 1) rmax = fmax(rmax, __shfl_down(rmax, off, 64));
      if ((tid & 63) == 0) sh.red[tid >> 6] = rmax;
      __syncthreads();
      double res = sh.red[0];
      for (int q = 1; q < (int)(blockDim.x >> 6); q++) res = fmax(res, sh.red[q]);
      bad = b;
      it = k;
      res_last = res;
      __syncthreads();
      if (bad) break;
      sweep<B <= 256 ? 4 : 3, B <= 256, kAxis>(d, st, sh);  // deeper prefetch where registers allow
      __syncthreads();
      bool more = res > st.tol && k < st.maxit;
      if (more && k >= st.kstop) {  // the launch ends here; a later one continues at k + 1
        more = false;
        stopped = 1;
      }
      if (tid == 0) st_agent(ctl + kCtlGo, more ? k + 1 : -1);
      if (!more) break;
    }
  }
  if (tid == 0) {
    st_agent(ctl + kCtlGo, -1);  // releases the flux workgroups on every exit path
    st.iters[0] = it;
    st.iters[1] = bad;
    st.iters[2] = stopped;
    reinterpret_cast<double *>(st.iters + 4)[0] = res_last;  // max |Res_4| of the launch's last iteration
  }
}

// The inner loops of a batch of instances in one launch, one workgroup per
// instance: the workgroup computes the fluxes of its instance's tiles one
// after another, then runs the same sweeps, until the reference's rule ends
// the loop.  No workgroup waits for another, so the launch needs neither
// co-residency nor host decisions and can be captured in a graph.  The same
// records and arithmetic as k_ns_loop: the same dU and iteration counts.
__global__ void __launch_bounds__(256) k_ns_batch(const NsJob *__restrict__ jobs) {
  __shared__ NsShared sh;
  const NsJob &J = jobs[blockIdx.x];
  const Dev &d = J.d;
  const SyntheticState &st = J.st;
  const int tid = threadIdx.x;
  int it = 0, bad = 0;
  double res_last = 0.0;
  const bool go = 1.0 > st.tol && st.maxit > 0;  // the reference starts from res = 1
  if (go && st.rs > 0) {
    const int ntiles = st.ntx * st.nty;
    for (int k = 1;; k++) {
      double rmax = 0.0;
      int b = 0;
      for (int tile = 0; tile < ntiles; tile++) flux_tile(d, st, sh, tile, rmax, b);
      for (int off = 32; off > 0; off >>= 1) rmax = fmax(rmax, __shfl_down(rmax, off, 64));
      b = __syncthreads_or(b);
      if ((tid & 63) == 0) sh.red[tid >> 6] = rmax;
      __syncthreads();
      double res = sh.red[0];
      for (int q = 1; q < (int)(blockDim.x >> 6); q++) res = fmax(res, sh.red[q]);
      it = k;
      res_last = res;
      bad = b;
      __syncthreads();
      if (bad) break;
      if (st.axis)
        sweep<4, true, true>(d, st, sh);
      else
        sweep<4>(d, st, sh);
      __syncthreads();
      if (!(res > st.tol && k < st.maxit)) break;
    }
  } else if (go) {
    it = 1;  // no swept cell: one iteration that changes nothing
  }
  if (tid == 0) {
    st.iters[0] = it;
    st.iters[1] = bad;
    st.iters[2] = 0;
    reinterpret_cast<double *>(st.iters + 4)[0] = res_last;
  }
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

// the next launch of the inner loop continues at iteration k0
__global__ void k_ns_resume(SyntheticState st, int k0) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    st.ctl[kCtlGo] = k0;
    st.iters[1] = 0;
    st.iters[2] = 0;
  }
}

// the degenerate loop of a mesh without swept cells: max |Res_4| over no
// cells is 0, so the reference runs one iteration that changes nothing
__global__ void k_ns_empty(SyntheticState st) {
  if (threadIdx.x == 0 && blockIdx.x == 0) st.iters[0] = st.ctl[kCtlGo] == 1 ? 1 : 0;
}

// the inner-loop kernel of an instance: 256 or 1024 threads, and for 256 the
// sweep of an axis-aligned mesh where it is one
static const void *ns_loop_kernel(const SyntheticState &st) {
  if (st.block > 256) return reinterpret_cast<const void *>(k_ns_loop<1024, false>);
  return st.axis ? reinterpret_cast<const void *>(k_ns_loop<256, true>)
                 : reinterpret_cast<const void *>(k_ns_loop<256, false>);
}

// whether every swept cell (3..nx-3, 3..ny-3) has the face normals ni = (1,0)
// and nj = (0,1) exactly (a zero of either sign), in the caller's Fortran
// (nx, ny) arrays
static bool axis_aligned(const gsis_params &p, int nx, int ny) {
  for (int j = 3; j <= ny - 3; j++)
    for (int i = 3; i <= nx - 3; i++) {
      const size_t g = (size_t)(i - 1) + (size_t)nx * (j - 1);
      if (!(p.ni1[g] == 1.0 && p.ni2[g] == 0.0 && p.nj1[g] == 0.0 && p.nj2[g] == 1.0)) return false;
    }
  return true;
}

int synthetic_init(Dev &d, const gsis_params &p, SyntheticState &st) {
  const size_t N = d.ncell;
  st.rs = d.ny - 5 > 0 && d.nx - 5 > 0 ? d.ny - 5 : 0;
  st.np = st.rs > 0 ? d.nx + d.ny - 11 : 0;
  if (st.rs > kMaxSweepWaves * 64) {
    g_syn_err = "synthetic solve: at most " + std::to_string(kMaxSweepWaves * 64 + 5) + " nodes in j";
    return GSIS_EUNSUPPORTED;
  }
  st.ntx = st.rs > 0 ? (d.nx - 5 + kTile - 1) / kTile : 0;
  st.nty = st.rs > 0 ? (st.rs + kTile - 1) / kTile : 0;
  const int ntiles = st.ntx * st.nty;
  st.block = std::max(256, 64 * ((st.rs + 63) / 64));
  st.axis = st.block <= 256 && st.rs > 0 && axis_aligned(p, d.nx, d.ny);
  if (ntiles > 0) {
    int dev = 0, ncu = 0, per = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, ns_loop_kernel(st), st.block, 0);
    if (e != hipSuccess || per < 1) {
      g_syn_err = std::string("synthetic solve: occupancy query: ") + hipGetErrorString(e) + ", " +
                  std::to_string(per) + " workgroups of " + std::to_string(st.block) + " threads per CU";
      return GSIS_EHIP;
    }
    st.nflux = std::min(ntiles, ncu * per - 1);
    if (st.nflux < 1) {
      g_syn_err = "synthetic solve: the inner-loop launch does not fit the device";
      return GSIS_EUNSUPPORTED;
    }
  }
  st.nw = (st.rs + 63) / 64;
  const size_t nq = (size_t)st.np * st.nw * 64;  // cells of the blocked sweep layout
  // 32-bit buffer offsets of the records and of U
  if ((size_t)kNRec * nq * sizeof(double) >= ((size_t)1 << 31) || (size_t)4 * N * sizeof(double) >= ((size_t)1 << 31)) {
    g_syn_err = "synthetic solve: the mesh is too large for the sweep's buffers (2 GiB)";
    return GSIS_EUNSUPPORTED;
  }
  const size_t n_dbl = 4 * N + (size_t)kNRec * nq + 4 * nq + 8 * (size_t)std::max(st.nflux, 1);
  const size_t bytes = n_dbl * 8 + 64;
  void *mem = nullptr;
  if (hipMalloc(&mem, bytes) != hipSuccess || hipMemset(mem, 0, bytes) != hipSuccess) {
    g_syn_err = "synthetic solve: device allocation failed";
    return GSIS_EHIP;
  }
  double *b = reinterpret_cast<double *>(mem);
  st.s00 = b; b += 4 * N;
  st.rec = b; b += kNRec * nq;
  st.dstar = b; b += 4 * nq;
  st.part = b; b += 8 * (size_t)std::max(st.nflux, 1);
  st.ctl = reinterpret_cast<int *>(b);
  st.iters = st.ctl + 8;
  st.mem = mem;
  if (nq > 0) {
    hipLaunchKernelGGL(k_rec_init, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, 0, st);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
      (void)hipFree(mem);
      st.mem = nullptr;
      g_syn_err = "synthetic solve: record initialisation failed";
      return GSIS_EHIP;
    }
  }
  st.tol = p.ns_tol;
  st.maxit = p.ns_maxit > 0 ? p.ns_maxit : 100000;
  st.local_dt = p.local_dt;
  st.if_viscous = p.if_viscous;
  st.cfl = p.cfl;
  st.dtmax = p.dtmax;
  st.dtmin = p.dtmin;
  st.on = true;
  return 0;
}

int synthetic_begin(const Dev &d, SyntheticState &st, hipStream_t s) {
  if (!st.on) {
    g_syn_err = "synthetic solve not initialised";
    return GSIS_EARG;
  }
  const int nc = (d.nx - 1) * (d.ny - 1);
  hipLaunchKernelGGL(k_ns_prepare, dim3((std::max(nc, st.nflux) + 255) / 256), dim3(256), 0, s, d, st);
  return 0;
}

int synthetic_run(const Dev &d, SyntheticState &st, int k0, int kstop, hipStream_t s) {
  if (!st.on || k0 < 1 || kstop < k0) {
    g_syn_err = "synthetic solve: not initialised, or an empty range of inner iterations";
    return GSIS_EARG;
  }
  if (k0 > 1) hipLaunchKernelGGL(k_ns_resume, dim3(1), dim3(64), 0, s, st, k0);
  if (st.nflux > 0) {
    Dev dd = d;
    SyntheticState ss = st;
    ss.k0 = k0;
    ss.kstop = kstop;
    void *args[] = {&dd, &ss};
    const hipError_t e = hipLaunchCooperativeKernel(ns_loop_kernel(st),
                                                    dim3(1 + st.nflux), dim3(st.block), args, 0, s);
    if (e != hipSuccess) {
      g_syn_err = std::string("synthetic solve: cooperative launch failed: ") + hipGetErrorString(e);
      return GSIS_EHIP;
    }
  } else if (k0 == 1) {
    hipLaunchKernelGGL(k_ns_empty, dim3(1), dim3(64), 0, s, st);
  }
  return 0;
}

bool synthetic_batchable(const SyntheticState &st) { return st.on && st.block <= 256; }

void launch_ns_batch(const NsJob *djobs, int n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_ns_batch, dim3(n), dim3(256), 0, s, djobs);
}

void synthetic_finish(const Dev &d, SyntheticState &st, hipStream_t s) {
  const int nci = (d.nx - 5) * (d.ny - 5);
  if (nci > 0) hipLaunchKernelGGL(k_ns_feedback, dim3((nci + 255) / 256), dim3(256), 0, s, d, st);
}

int synthetic_step(const Dev &d, SyntheticState &st, hipStream_t s) {
  int rc = synthetic_begin(d, st, s);
  if (!rc) rc = synthetic_run(d, st, 1, 0x7fffffff, s);
  if (!rc) synthetic_finish(d, st, s);
  return rc;
}

int synthetic_state(SyntheticState &st, hipStream_t s, bool *unfinished, int *iters, bool *bad_t, double *res) {
  int h[6] = {0, 0, 0, 0, 0, 0};
  if (hipMemcpyAsync(h, st.iters, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    g_syn_err = "synthetic solve: reading the loop state failed";
    return GSIS_EHIP;
  }
  *unfinished = h[2] != 0 && h[1] == 0;
  *iters = h[0];
  *bad_t = h[1] != 0;
  double r;
  memcpy(&r, h + 4, sizeof(r));
  *res = r;
  return 0;
}

int synthetic_read_iters(SyntheticState &st, hipStream_t s, int *iters, int *bad_t) {
  int h[2] = {0, 0}, c[4] = {0, 0, 0, 0};
  if (!st.on) {
    if (iters) *iters = 0;
    if (bad_t) *bad_t = 0;
    return 0;
  }
  if (hipMemcpyAsync(h, st.iters, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(c, st.ctl, sizeof(c), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    g_syn_err = "synthetic solve: reading the iteration counter failed";
    return GSIS_EHIP;
  }
  if (iters) *iters = h[0];
  if (bad_t) *bad_t = h[1];
  if (c[kCtlErr]) {
    g_syn_err = "synthetic solve: a wait between the inner loop's workgroups timed out";
    return GSIS_EHIP;
  }
  return 0;
}

void synthetic_free(SyntheticState &st) {
  if (st.mem) (void)hipFree(st.mem);
  st.mem = nullptr;
  st.on = false;
}

}  // namespace gsis
