/*
 * gsis_oracle_ns.c -- CPU oracle of the synthetic NS solve (mode M1).
 * TEST INFRASTRUCTURE ONLY (see gsis_oracle.c).
 *
 * Restates the GSIS block of Hybrid_Time_advance_LU_SGS that the reference
 * ships commented out (src/Hybrid_Sub.f90:45-153), with its inner loop
 * NS_Time_advance_LU_SGS (src/sub_NS_singlegrid.f90:4-40):
 *   Comput_Residual_one_mesh (:45-114) -> comput_duvtpckw (:121-157),
 *   Residual/get_viscous/Residual_FVM (src/sub_Residual.f90:11-354) with the
 *   options of the shipped control.in (original-variable NND2 reconstruction
 *   :387-397,476-546, Van Leer splitting src/sub_flux_split.f90:141-192,
 *   viscous terms + the kinetic HoT correction :176-193,309-326),
 *   comput_Lvc (:187-221), comput_dt (:160-183), du_LU_SGS_2D
 *   (src/sub_LU_SGS.f90:4-143) and the update of cells 3..n-3.
 * The inner loop runs while the max |energy residual| (get_res,
 * src/IO.f90:286-303) exceeds ns_tol, as in src/Hybrid_Sub.f90:99-104.
 * TaoNS_out/Q_out are never filled by the reference (getNSF is not called),
 * so the stress/heat-flux feedback of :142-147 adds zero and is omitted.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "gsis_oracle_impl.h"

static double minmod_ns(double a, double b) { /* src/sub_Residual.f90:357-368 */
  if (a * b <= 0.0) return 0.0;
  return (fabs(a) <= fabs(b)) ? a : b;
}

/* Flux_Van_Leer_1Da, src/sub_flux_split.f90:141-192 */
void orc_van_leer(const double ql[4], const double qr[4], double g, double f[4]);
static void van_leer(const double ql[4], const double qr[4], double g, double f[4]) { orc_van_leer(ql, qr, g, f); }
void orc_van_leer(const double ql[4], const double qr[4], double g, double f[4]) {
  double dl = ql[0], uul = ql[1], vvl = ql[2], pl = ql[3];
  double dr = qr[0], uur = qr[1], vvr = qr[2], pr = qr[3];
  double al = sqrt(g * pl / dl), ar = sqrt(g * pr / dr);
  double ml = uul / al, mr = uur / ar;
  double fp[4] = {0, 0, 0, 0}, fm[4] = {0, 0, 0, 0};
  if (ml >= 1.0) {
    fp[0] = dl * uul;
    fp[1] = dl * uul * uul + pl;
    fp[2] = dl * uul * vvl;
    fp[3] = uul * (g * pl / (g - 1.0) + 0.5 * dl * (uul * uul + vvl * vvl));
  } else if (fabs(ml) < 1.0) {
    double mp = 0.25 * (1.0 + ml) * (1.0 + ml);
    double t0 = dl * al * mp;
    fp[0] = t0;
    fp[1] = t0 * ((g - 1.0) * uul + 2.0 * al) / g;
    fp[2] = t0 * vvl;
    fp[3] = t0 * (((g - 1.0) * uul + 2.0 * al) * ((g - 1.0) * uul + 2.0 * al) * 0.5 / (g * g - 1.0) + 0.5 * (vvl * vvl));
  }
  if (mr >= 1.0) {
    /* zero */
  } else if (fabs(mr) < 1.0) {
    double mm = -0.25 * (mr - 1.0) * (mr - 1.0);
    double t0 = dr * ar * mm;
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

/* inviscid + viscous + HoT flux through i-face (dir 0) or j-face (dir 1) (i,j),
 * Residual_FVM src/sub_Residual.f90:68-202 and :210-338 */
static void ns_face_flux(orc_t *o, int dir, int i, int j, double fl[4]) {
  double *q[4] = {o->d, o->uu, o->vv, o->p};
  double s, n1, n2;
  double ul[4], ur[4];
  for (int m = 0; m < 4; m++) {
    double u1, u2, u3, u4;
    if (dir == 0) {
      u1 = C2(q[m], i - 2, j); u2 = C2(q[m], i - 1, j); u3 = C2(q[m], i, j); u4 = C2(q[m], i + 1, j);
    } else {
      u1 = C2(q[m], i, j - 2); u2 = C2(q[m], i, j - 1); u3 = C2(q[m], i, j); u4 = C2(q[m], i, j + 1);
    }
    ul[m] = u2 + 0.5 * minmod_ns(u2 - u1, u3 - u2); /* scheme_fP NND2 (:490-491) */
    ur[m] = u3 - 0.5 * minmod_ns(u3 - u2, u4 - u3); /* scheme_fm NND2 (:526-527) */
  }
  if (dir == 0) { s = G2(o->si, i, j); n1 = G2(o->ni1, i, j); n2 = G2(o->ni2, i, j); }
  else { s = G2(o->sj, i, j); n1 = G2(o->nj1, i, j); n2 = G2(o->nj2, i, j); }
  double ql[4] = {ul[0], ul[1] * n1 + ul[2] * n2, -ul[1] * n2 + ul[2] * n1, ul[3]};
  double qr[4] = {ur[0], ur[1] * n1 + ur[2] * n2, -ur[1] * n2 + ur[2] * n1, ur[3]};
  double f0[4];
  van_leer(ql, qr, o->gamma, f0);
  fl[0] = -f0[0] * s;
  fl[1] = -(f0[1] * n1 - f0[2] * n2) * s;
  fl[2] = -(f0[1] * n2 + f0[2] * n1) * s;
  fl[3] = -f0[3] * s;
  if (!o->if_viscous) return;
  double t11, t22, t12, t33, qx, qy, u1, v1;
  orc_ns_face_stress(o, dir, i, j, &t11, &t22, &t12, &t33, &qx, &qy, &u1, &v1);
  const double *tg = dir == 0 ? o->taonsi_g : o->taonsj_g;
  const double *qg = dir == 0 ? o->qi_g : o->qj_g;
  t11 = t11 + U4(tg, 1, i, j);
  t22 = t22 + U4(tg, 2, i, j);
  t12 = t12 + U4(tg, 3, i, j);
  t33 = t33 + U4(tg, 4, i, j);
  qx = qx + Q2(qg, 1, i, j);
  qy = qy + Q2(qg, 2, i, j);
  (void)t33;
  double e1 = u1 * t11 + v1 * t12 + qx;
  double e2 = u1 * t12 + v1 * t22 + qy;
  fl[1] = fl[1] + (t11 * n1 + t12 * n2) * s;
  fl[2] = fl[2] + (t12 * n1 + t22 * n2) * s;
  fl[3] = fl[3] + (e1 * n1 + e2 * n2) * s;
}

/* comput_DFn, src/sub_LU_SGS.f90:122-143 */
static void comput_dfn(const double *u, const double *du, double n1, double n2, double g, double df[4]) {
  double u2[4];
  for (int m = 0; m < 4; m++) u2[m] = u[m] + du[m];
  double un1 = (u[1] * n1 + u[2] * n2) / u[0];
  double p1 = (g - 1.0) * (u[3] - 0.5 * (u[1] * u[1] + u[2] * u[2]) / u[0]);
  double un2 = (u2[1] * n1 + u2[2] * n2) / u2[0];
  double p2 = (g - 1.0) * (u2[3] - 0.5 * (u2[1] * u2[1] + u2[2] * u2[2]) / u2[0]);
  df[0] = u2[0] * un2 - u[0] * un1;
  df[1] = (u2[1] * un2 + p2 * n1) - (u[1] * un1 + p1 * n1);
  df[2] = (u2[2] * un2 + p2 * n2) - (u[2] * un1 + p1 * n2);
  df[3] = (u2[3] + p2) * un2 - (u[3] + p1) * un1;
}

/* one NS_Time_advance_LU_SGS; returns the max |Res(4)| of the residual it
 * computed (before the update), or a negative code on T<0.1 */
static double ns_time_advance(orc_t *o, int *err) {
  const int nx = o->nx, ny = o->ny;
  const double g = o->gamma;
  *err = orc_comput_prims(o);
  if (*err) return -1.0;
  if (o->if_viscous) orc_get_viscous(o);
  /* face fluxes */
  const size_t nf = (size_t)(nx + 1) * (ny + 1);
  double *fi = (double *)calloc(4 * nf, sizeof(double)), *fj = (double *)calloc(4 * nf, sizeof(double));
#define FI(m, i, j) fi[(m) + 4 * ((i) + (size_t)(nx + 1) * (j))]
#define FJ(m, i, j) fj[(m) + 4 * ((i) + (size_t)(nx + 1) * (j))]
  for (int j = 1; j <= ny - 1; j++)
    for (int i = 1; i <= nx; i++) ns_face_flux(o, 0, i, j, &FI(0, i, j));
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx - 1; i++) ns_face_flux(o, 1, i, j, &FJ(0, i, j));
  for (int j = 1; j <= ny - 1; j++)
    for (int i = 1; i <= nx - 1; i++)
      for (int m = 1; m <= 4; m++)
        U4(o->res, m, i, j) = FI(m - 1, i + 1, j) - FI(m - 1, i, j) + FJ(m - 1, i, j + 1) - FJ(m - 1, i, j);
  free(fi); free(fj);
#undef FI
#undef FJ
  double resmax4 = 0.0;
  for (int j = 3; j <= ny - 3; j++)
    for (int i = 3; i <= nx - 3; i++) {
      double r = fabs(U4(o->res, 4, i, j));
      if (r > resmax4) resmax4 = r;
    }
  /* comput_Lvc and comput_dt */
  const double prt = 0.9;
  for (int j = 1; j <= ny - 1; j++)
    for (int i = 1; i <= nx - 1; i++) {
      double s0 = G2(o->vol, i, j);
      double si = 0.5 * (G2(o->si, i, j) + G2(o->si, i + 1, j));
      double ni1 = 0.5 * (G2(o->ni1, i, j) + G2(o->ni1, i + 1, j));
      double ni2 = 0.5 * (G2(o->ni2, i, j) + G2(o->ni2, i + 1, j));
      double sj = 0.5 * (G2(o->sj, i, j) + G2(o->sj, i, j + 1));
      double nj1 = 0.5 * (G2(o->nj1, i, j) + G2(o->nj1, i, j + 1));
      double nj2 = 0.5 * (G2(o->nj2, i, j) + G2(o->nj2, i, j + 1));
      double uni = C2(o->uu, i, j) * ni1 + C2(o->vv, i, j) * ni2;
      double unj = C2(o->uu, i, j) * nj1 + C2(o->vv, i, j) * nj2;
      double cc = sqrt(C2(o->T, i, j)) / o->ma;
      C2(o->lci, i, j) = (fabs(uni) + cc) * si;
      C2(o->lcj, i, j) = (fabs(unj) + cc) * sj;
      double amu = o->if_viscous ? C2(o->amu, i, j) : 0.0;
      double tmp1 = g / C2(o->d, i, j) * (amu / o->pr + 0.0 / prt);
      C2(o->lvi, i, j) = tmp1 * si * si / s0;
      C2(o->lvj, i, j) = tmp1 * sj * sj / s0;
      if (!o->local_dt) C2(o->dtl, i, j) = o->dt;
      else {
        double dt = o->cfl * G2(o->vol, i, j) /
                    (C2(o->lci, i, j) + C2(o->lcj, i, j) + 1.0 * (C2(o->lvi, i, j) + C2(o->lvj, i, j)));
        if (dt > o->dtmax) dt = o->dtmax;
        if (dt < o->dtmin) dt = o->dtmin;
        C2(o->dtl, i, j) = dt;
      }
    }
  /* du_LU_SGS_2D (w_LU = 1) */
  const double w = 1.0;
  for (int plane = 6; plane <= nx + ny - 6; plane++)
    for (int j = 3; j <= ny - 3; j++) {
      int i = plane - j;
      if (i < 3 || i > nx - 3) continue;
      double alfa = G2(o->vol, i, j) / C2(o->dtl, i, j) + w * (C2(o->lci, i, j) + C2(o->lcj, i, j));
      if (o->if_viscous) alfa = alfa + 2.0 * (C2(o->lvi, i, j) + C2(o->lvj, i, j));
      double dui[4] = {0, 0, 0, 0}, duj[4] = {0, 0, 0, 0}, df[4];
      if (i != 3) {
        comput_dfn(&U4(o->U, 1, i - 1, j), &U4(o->du, 1, i - 1, j), G2(o->ni1, i, j), G2(o->ni2, i, j), g, df);
        for (int m = 0; m < 4; m++) {
          dui[m] = 0.5 * (df[m] * G2(o->si, i, j) + w * C2(o->lci, i - 1, j) * U4(o->du, m + 1, i - 1, j));
          if (o->if_viscous) dui[m] = dui[m] + C2(o->lvi, i - 1, j) * U4(o->du, m + 1, i - 1, j);
        }
      }
      if (j != 3) {
        comput_dfn(&U4(o->U, 1, i, j - 1), &U4(o->du, 1, i, j - 1), G2(o->nj1, i, j), G2(o->nj2, i, j), g, df);
        for (int m = 0; m < 4; m++) {
          duj[m] = 0.5 * (df[m] * G2(o->sj, i, j) + w * C2(o->lcj, i, j - 1) * U4(o->du, m + 1, i, j - 1));
          if (o->if_viscous) duj[m] = duj[m] + C2(o->lvj, i, j - 1) * U4(o->du, m + 1, i, j - 1);
        }
      }
      for (int m = 0; m < 4; m++) U4(o->du, m + 1, i, j) = (U4(o->res, m + 1, i, j) + dui[m] + duj[m]) / alfa;
    }
  for (int plane = nx + ny - 6; plane >= 6; plane--)
    for (int j = ny - 3; j >= 3; j--) {
      int i = plane - j;
      if (i < 3 || i > nx - 3) continue;
      double alfa = G2(o->vol, i, j) / C2(o->dtl, i, j) + w * (C2(o->lci, i, j) + C2(o->lcj, i, j));
      if (o->if_viscous) alfa = alfa + 2.0 * (C2(o->lvi, i, j) + C2(o->lvj, i, j));
      double dui[4] = {0, 0, 0, 0}, duj[4] = {0, 0, 0, 0}, df[4];
      if (i != nx - 3) {
        comput_dfn(&U4(o->U, 1, i + 1, j), &U4(o->du, 1, i + 1, j), G2(o->ni1, i, j), G2(o->ni2, i, j), g, df);
        for (int m = 0; m < 4; m++) {
          dui[m] = -0.5 * (df[m] * G2(o->si, i + 1, j) - w * C2(o->lci, i + 1, j) * U4(o->du, m + 1, i + 1, j));
          if (o->if_viscous) dui[m] = dui[m] + C2(o->lvi, i + 1, j) * U4(o->du, m + 1, i + 1, j);
        }
      }
      if (j != ny - 3) {
        comput_dfn(&U4(o->U, 1, i, j + 1), &U4(o->du, 1, i, j + 1), G2(o->nj1, i, j), G2(o->nj2, i, j), g, df);
        for (int m = 0; m < 4; m++) {
          duj[m] = -0.5 * (df[m] * G2(o->sj, i, j + 1) - w * C2(o->lcj, i, j + 1) * U4(o->du, m + 1, i, j + 1));
          if (o->if_viscous) duj[m] = duj[m] + C2(o->lvj, i, j + 1) * U4(o->du, m + 1, i, j + 1);
        }
      }
      for (int m = 0; m < 4; m++) U4(o->du, m + 1, i, j) = U4(o->du, m + 1, i, j) + (dui[m] + duj[m]) / alfa;
    }
  for (int j = 3; j <= ny - 3; j++)
    for (int i = 3; i <= nx - 3; i++)
      for (int m = 1; m <= 4; m++) U4(o->U, m, i, j) = U4(o->U, m, i, j) + U4(o->du, m, i, j);
  return resmax4;
}

/* n inner iterations of NS_Time_advance_LU_SGS from the current U and HoT
 * fields, without the GSIS block around them; returns the max |Res(4)| of the
 * last one (a negative value on T<0.1).  For checks of the inner loop. */
double orc_ns_advance(orc_t *o, int n) {
  double r = 0.0;
  for (int k = 0; k < n; k++) {
    int err = 0;
    r = ns_time_advance(o, &err);
    if (err) return -1.0;
  }
  return r;
}

/* the work arrays of the last inner iteration, Fortran layouts; any pointer
 * may be NULL: res, du (4,-1:nx+1,-1:ny+1); lc = Lci, Lcj, Lvi, Lvj, dt as
 * five (-1:nx+1,-1:ny+1) arrays */
void orc_ns_work(const orc_t *o, double *res, double *du, double *lc) {
  const size_t nc = (size_t)(o->nx + 3) * (o->ny + 3);
  if (res) memcpy(res, o->res, 4 * nc * sizeof(double));
  if (du) memcpy(du, o->du, 4 * nc * sizeof(double));
  if (lc) {
    const double *a[5] = {o->lci, o->lcj, o->lvi, o->lvj, o->dtl};
    for (int k = 0; k < 5; k++) memcpy(lc + k * nc, a[k], nc * sizeof(double));
  }
}

/* the GSIS block of Hybrid_Time_advance_LU_SGS, src/Hybrid_Sub.f90:45-153 */
int orc_synthetic_ns(orc_t *o) {
  const int nx = o->nx, ny = o->ny;
  const double g = o->gamma, ma = o->ma, cv = o->cv;
  const double sq = sqrt(2.0 / g);
  for (int i = 1; i <= nx - 1; i++)
    for (int j = 1; j <= ny - 1; j++)
      for (int k = 1; k <= 10; k++) {
        S0_(k, i, j) = S_(k, i, j) - S0_(k, i, j);
        S00_(k, i, j) = S_(k, i, j);
      }
  for (int i = 3; i <= nx - 3; i++)
    for (int j = 3; j <= ny - 3; j++) {
      double trho = S0_(1, i, j);
      double tu = S0_(2, i, j) / ma * sq, tv = S0_(3, i, j) / ma * sq, tt = S0_(4, i, j);
      double rns = U4(o->U0, 1, i, j);
      double uns = U4(o->U0, 2, i, j) / U4(o->U0, 1, i, j);
      double vns = U4(o->U0, 3, i, j) / U4(o->U0, 1, i, j);
      double tns = (U4(o->U0, 4, i, j) - 0.5 * U4(o->U0, 1, i, j) * (uns * uns + vns * vns)) / (cv * U4(o->U0, 1, i, j));
      rns = rns + trho; uns = uns + tu; vns = vns + tv; tns = tns + tt;
      U4(o->U, 1, i, j) = rns;
      U4(o->U, 2, i, j) = rns * uns;
      U4(o->U, 3, i, j) = rns * vns;
      U4(o->U, 4, i, j) = cv * tns * rns + 0.5 * rns * (uns * uns + vns * vns);
      for (int m = 1; m <= 4; m++) U4(o->U2, m, i, j) = U4(o->U, m, i, j);
    }
  double res = 1.0;
  int it = 0;
  while (res > o->ns_tol) {
    if (it >= o->ns_maxit) break;
    it++;
    int err = 0;
    res = ns_time_advance(o, &err);
    if (err) return GSIS_ETEMP;
  }
  o->ns_iter = it;
  for (int i = 3; i <= nx - 3; i++)
    for (int j = 3; j <= ny - 3; j++)
      for (int m = 1; m <= 4; m++) U4(o->U0, m, i, j) = U4(o->U, m, i, j);
  for (int i = 3; i <= nx - 3; i++)
    for (int j = 3; j <= ny - 3; j++) {
      double rns = U4(o->U, 1, i, j);
      double uns = U4(o->U, 2, i, j) / U4(o->U, 1, i, j);
      double vns = U4(o->U, 3, i, j) / U4(o->U, 1, i, j);
      double tns = (U4(o->U, 4, i, j) - 0.5 * U4(o->U, 1, i, j) * (uns * uns + vns * vns)) / (cv * U4(o->U, 1, i, j));
      double rk = U4(o->U2, 1, i, j);
      double uk = U4(o->U2, 2, i, j) / U4(o->U2, 1, i, j);
      double vk = U4(o->U2, 3, i, j) / U4(o->U2, 1, i, j);
      double tk = (U4(o->U2, 4, i, j) - 0.5 * U4(o->U2, 1, i, j) * (uk * uk + vk * vk)) / (cv * U4(o->U2, 1, i, j));
      S0_(1, i, j) = rns - rk;
      S0_(2, i, j) = (uns - uk) * ma / sq;
      S0_(3, i, j) = (vns - vk) * ma / sq;
      S0_(4, i, j) = tns - tk;
      for (int k = 1; k <= 4; k++) S_(k, i, j) = S_(k, i, j) + S0_(k, i, j);
      for (int k = 1; k <= 4; k++) S0_(k, i, j) = S0_(k, i, j) + S00_(k, i, j);
    }
  return 0;
}
