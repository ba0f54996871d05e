/*
 * gsis_oracle.c -- CPU oracle of the GSIS kinetic loop.  TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, in plain C and in the reference's own array layout
 * (Fortran column-major with the same ghost lower bounds), the FP64 algorithm
 * of the Fortran code Hybrid2D for one outer GSIS iteration.  The HIP path is
 * compared against it.  Only tests/, __graft_entry__.smoke() and
 * bench.py's CPU-baseline leg may load it; the product path never does.
 *
 * Every routine cites the reference lines it follows.  Single-precision
 * literals of the reference (default REAL is 4 bytes: no real-size option in
 * Hybrid2D.vfproj) are reproduced where they change a value: 1.E-30 in the
 * positivity clamp (src/GKUA_Residual.f90:104-105), sqrt(2.0) in the HoT heat
 * flux (src/Hybrid_Sub.f90:325,335) and OMGgas=0.81 in get_viscous
 * (src/sub_Residual.f90:563).
 *
 * Differences from the reference that do not change results beyond 1 ulp:
 *  - quadrature weights enter as the product wi_eff*wj_eff (the reference
 *    multiplies FBAVX*wi*FBAVY*wj left to right, src/GKUA_DETFP.f90:314);
 *  - S_GKUA(11) is left 0: the reference accumulates an unassigned FBEI(11)
 *    into it (src/GKUA_DETFP.f90:337-339) and never reads it on the step path.
 */
#include "gsis_oracle_impl.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static double fminmd(double x, double y) { /* FMINMD, src/Functions.f90:10-26 */
  if ((x > 0. && y > 0.) || (x < 0. && y < 0.)) return (fabs(x) <= fabs(y)) ? x : y;
  return 0.;
}

static double *dup_arr(const double *src, size_t n) {
  double *a = (double *)malloc(n * sizeof(double));
  if (src) memcpy(a, src, n * sizeof(double));
  else memset(a, 0, n * sizeof(double));
  return a;
}
static double *zalloc(size_t n) { return (double *)calloc(n, sizeof(double)); }

static int bc_nk(const gsis_bc *b) { return (b->face == 1 || b->face == 3) ? (b->jend - b->jst + 1) : (b->iend - b->ist + 1); }

orc_t *orc_create(const gsis_params *p) {
  orc_t *o = (orc_t *)calloc(1, sizeof(orc_t));
  o->nx = p->nx; o->ny = p->ny; o->nvi = p->nvi; o->nvj = p->nvj;
  o->vi = dup_arr(p->vi, p->nvi); o->vj = dup_arr(p->vj, p->nvj);
  o->wi = dup_arr(p->wi, p->nvi); o->wj = dup_arr(p->wj, p->nvj);
  size_t nn = (size_t)(p->nx + 4) * (p->ny + 4);
  o->x = dup_arr(p->x, nn); o->y = dup_arr(p->y, nn);
  size_t ng = (size_t)p->nx * p->ny;
  o->vol = dup_arr(p->vol, ng); o->si = dup_arr(p->si, ng); o->sj = dup_arr(p->sj, ng);
  o->ni1 = dup_arr(p->ni1, ng); o->ni2 = dup_arr(p->ni2, ng);
  o->nj1 = dup_arr(p->nj1, ng); o->nj2 = dup_arr(p->nj2, ng);
  o->nbc = p->nbc;
  o->bc = (gsis_bc *)malloc(sizeof(gsis_bc) * (p->nbc > 0 ? p->nbc : 1));
  memcpy(o->bc, p->bc, sizeof(gsis_bc) * p->nbc);
  o->cfc = p->cfc; o->x1s = p->x1s; o->cpr = p->cpr; o->twall = p->twall; o->dt = p->dt;
  o->ma = p->ma; o->gamma = p->gamma; o->re = p->re; o->pr = p->pr; o->cv = p->cv; o->cp = p->cp;
  o->aoa = p->aoa; o->scheme = p->scheme; o->mode = p->mode; o->if_viscous = p->if_viscous;
  o->local_dt = p->local_dt; o->cfl = p->cfl; o->dtmax = p->dtmax; o->dtmin = p->dtmin;
  o->ns_tol = p->ns_tol; o->ns_maxit = p->ns_maxit > 0 ? p->ns_maxit : 100000;
  size_t nf = FPLANE * (size_t)o->nvi * o->nvj;
  o->F = zalloc(nf); o->F0 = zalloc(nf); o->F00 = zalloc(nf);
  size_t ns = (size_t)11 * NI7 * NJ7;
  o->S = zalloc(ns); o->S0 = zalloc(ns); o->S00 = zalloc(ns);
  size_t nc = (size_t)NI3 * (o->ny + 3);
  o->CF = zalloc(nc);
  o->U = zalloc(4 * nc); o->U0 = zalloc(4 * nc); o->U2 = zalloc(4 * nc);
  for (size_t c = 0; c < nc; c++) { /* Creat_Mesh initial U/U0/U2, src/sub_init.f90:159-162 */
    o->U[4 * c] = 1.; o->U[4 * c + 3] = 1.;
    o->U0[4 * c] = 1.; o->U0[4 * c + 3] = 1.;
    o->U2[4 * c] = 1.; o->U2[4 * c + 3] = 1.;
  }
  o->taonsi_g = zalloc(4 * nc); o->taonsj_g = zalloc(4 * nc);
  o->qi_g = zalloc(2 * nc); o->qj_g = zalloc(2 * nc);
  o->taonsi_in = zalloc(4 * nc); o->taonsj_in = zalloc(4 * nc);
  o->qi_in = zalloc(2 * nc); o->qj_in = zalloc(2 * nc);
  o->d = zalloc(nc); o->uu = zalloc(nc); o->vv = zalloc(nc); o->T = zalloc(nc); o->p = zalloc(nc);
  o->amu = zalloc(nc);
  o->res = zalloc(4 * nc); o->du = zalloc(4 * nc); o->dtl = zalloc(nc);
  o->lci = zalloc(nc); o->lcj = zalloc(nc); o->lvi = zalloc(nc); o->lvj = zalloc(nc);
  o->swall_off = (int *)calloc(p->nbc + 1, sizeof(int));
  int off = 0;
  for (int b = 0; b < p->nbc; b++) {
    o->swall_off[b] = off;
    if (p->bc[b].code == GSIS_BC_WALL || p->bc[b].code == GSIS_BC_FARFIELD) off += 10 * bc_nk(&p->bc[b]);
  }
  o->swall_off[p->nbc] = off;
  o->swall = zalloc(off > 0 ? off : 1);
  /* cell centres, src/sub_init.f90:814-819 */
  o->x1 = zalloc(nc); o->y1 = zalloc(nc);
  for (int j = -1; j <= o->ny + 1; j++)
    for (int i = -1; i <= o->nx + 1; i++) {
      C2(o->x1, i, j) = (XN(o->x, i, j) + XN(o->x, i + 1, j) + XN(o->x, i, j + 1) + XN(o->x, i + 1, j + 1)) * 0.25;
      C2(o->y1, i, j) = (XN(o->y, i, j) + XN(o->y, i + 1, j) + XN(o->y, i, j + 1) + XN(o->y, i + 1, j + 1)) * 0.25;
    }
  return o;
}

void orc_destroy(orc_t *o) {
  if (!o) return;
  double *arrs[] = {o->vi, o->vj, o->wi, o->wj, o->x, o->y, o->x1, o->y1, o->vol, o->si, o->sj,
                    o->ni1, o->ni2, o->nj1, o->nj2, o->F, o->F0, o->F00, o->S, o->S0, o->S00,
                    o->CF, o->U, o->U0, o->U2, o->taonsi_g, o->taonsj_g, o->qi_g, o->qj_g,
                    o->taonsi_in, o->taonsj_in, o->qi_in, o->qj_in, o->swall, o->d, o->uu, o->vv,
                    o->T, o->p, o->amu, o->res, o->du, o->dtl, o->lci, o->lcj, o->lvi, o->lvj};
  for (size_t k = 0; k < sizeof(arrs) / sizeof(arrs[0]); k++) free(arrs[k]);
  free(o->bc); free(o->swall_off); free(o);
}

size_t orc_frd_len(const orc_t *o) { return FPLANE * (size_t)o->nvi * o->nvj; }

void orc_set_frd(orc_t *o, const double *frd) { memcpy(o->F, frd, orc_frd_len(o) * sizeof(double)); }
void orc_get_frd(const orc_t *o, double *frd) { memcpy(frd, o->F, orc_frd_len(o) * sizeof(double)); }
void orc_get_frd00(const orc_t *o, double *frd) { memcpy(frd, o->F00, orc_frd_len(o) * sizeof(double)); }
void orc_get_frd0(const orc_t *o, double *frd) { memcpy(frd, o->F0, orc_frd_len(o) * sizeof(double)); }
/* primitives and viscosity of the last comput_duvtpckw/get_viscous:
   out[5][(nx+3)(ny+3)] = d, u, v, T, mu on (-1:nx+1,-1:ny+1) */
void orc_get_prims(const orc_t *o, double *out) {
  size_t nc = (size_t)NI3 * (o->ny + 3);
  memcpy(out, o->d, nc * sizeof(double));
  memcpy(out + nc, o->uu, nc * sizeof(double));
  memcpy(out + 2 * nc, o->vv, nc * sizeof(double));
  memcpy(out + 3 * nc, o->T, nc * sizeof(double));
  memcpy(out + 4 * nc, o->amu, nc * sizeof(double));
}
/* Get_Feq on the current S (src/GKUA_Residual.f90:420-469) into FRD0 */
static void get_feq(orc_t *o);
void orc_feq(orc_t *o) { get_feq(o); }

void orc_set_macro(orc_t *o, const double *s, const double *cf, const double *u, const double *u0) {
  size_t nc = (size_t)NI3 * (o->ny + 3);
  if (s) memcpy(o->S, s, sizeof(double) * 11 * NI7 * NJ7);
  if (cf) memcpy(o->CF, cf, sizeof(double) * nc);
  if (u) memcpy(o->U, u, sizeof(double) * 4 * nc);
  if (u0) memcpy(o->U0, u0, sizeof(double) * 4 * nc);
}

void orc_get_macro(const orc_t *o, double *s, double *s0, double *cf, double *u, double *taonsi_g,
                   double *taonsj_g, double *qi_g, double *qj_g, double *swall) {
  size_t nc = (size_t)NI3 * (o->ny + 3);
  if (s) memcpy(s, o->S, sizeof(double) * 11 * NI7 * NJ7);
  if (s0) memcpy(s0, o->S0, sizeof(double) * 11 * NI7 * NJ7);
  if (cf) memcpy(cf, o->CF, sizeof(double) * nc);
  if (u) memcpy(u, o->U, sizeof(double) * 4 * nc);
  if (taonsi_g) memcpy(taonsi_g, o->taonsi_g, sizeof(double) * 4 * nc);
  if (taonsj_g) memcpy(taonsj_g, o->taonsj_g, sizeof(double) * 4 * nc);
  if (qi_g) memcpy(qi_g, o->qi_g, sizeof(double) * 2 * nc);
  if (qj_g) memcpy(qj_g, o->qj_g, sizeof(double) * 2 * nc);
  if (swall) memcpy(swall, o->swall, sizeof(double) * o->swall_off[o->nbc]);
}

void orc_counters(const orc_t *o, int *kstep, double *tt, int *ns_iter) {
  if (kstep) *kstep = o->kstep;
  if (tt) *tt = o->tt;
  if (ns_iter) *ns_iter = o->ns_iter;
}

/* the counters init_flow_read restores from flag.dat (src/Backup_write_read.f90:20-32) */
void orc_set_counters(orc_t *o, int kstep, double tt) {
  o->kstep = kstep;
  o->tt = tt;
}

static double *fvel(const orc_t *o, double *base, int iv, int jv) { /* per-velocity plane */
  return base + FPLANE * ((size_t)(iv - 1) + (size_t)o->nvi * (jv - 1));
}

/* ------------------------------------------------------------------------ */
/* Diffuse wall: GKUA_boundary_wall (src/GKUA_boundary.f90:49-100), then     */
/* GKUA_wall_boundary (src/GKUA_wall_functions.f90:1-68) with DETNW ->       */
/* NWGaussLegendre (src/GKUA_wall_functions.f90:70-199).                     */
/* ------------------------------------------------------------------------ */
static void bc_wall(orc_t *o, const gsis_bc *b) {
  const int nx = o->nx, ny = o->ny;
  const int ib = b->ist, ie = b->iend, jb = b->jst, je = b->jend;
  const int icheck = ib > ie ? ib : ie, jcheck = jb > je ? jb : je;
  const double uw = b->uw, vw = b->vw, tw = o->twall;
  /* 1) extrapolate the wall-adjacent distribution to the ghost slot -3/n+3 */
  for (int i = ib; i <= ie; i++)
    for (int j = jb; j <= je; j++) {
      if ((jb == je && i == icheck) || (ib == ie && j == jcheck)) continue;
      int px, py, px1, py1, wx, wy;
      if (b->face == 1) { px = 1; py = j; px1 = 2; py1 = j; wx = -3; wy = j; }
      else if (b->face == 2) { px = i; py = 1; px1 = i; py1 = 2; wx = i; wy = -3; }
      else if (b->face == 3) { px = nx - 1; py = j; px1 = nx - 2; py1 = j; wx = nx + 3; wy = j; }
      else { px = i; py = ny - 1; px1 = i; py1 = ny - 2; wx = i; wy = ny + 3; }
      for (int jv = 1; jv <= o->nvj; jv++)
        for (int iv = 1; iv <= o->nvi; iv++) {
          double *Fv = fvel(o, o->F, iv, jv);
          for (int ird = 1; ird <= 2; ird++) {
            double f1 = FV(Fv, ird, px, py);
            double t = (-FV(Fv, ird, px1, py1) + 3. * f1) * 0.5;
            if (t <= 0.) t = f1;
            if (t >= 1.5 * f1) t = 1.5 * f1;
            FV(Fv, ird, wx, wy) = t;
          }
        }
    }
  /* 2) wall density per face node (NWGaussLegendre) */
  const int kmax = (b->face == 1 || b->face == 3) ? ny + 3 : nx + 3;
  double *sumnw = zalloc(kmax + 1), *sumnw1 = zalloc(kmax + 1);
  for (int i = ib; i <= ie; i++)
    for (int j = jb; j <= je; j++) {
      if ((jb == je && i == icheck) || (ib == ie && j == jcheck)) continue;
      int k, wx, wy;
      double swnx, swny, st;
      if (b->face == 1) { k = j; wx = -3; wy = j; swnx = G2(o->ni1, i, j); swny = G2(o->ni2, i, j); st = 1.; }
      else if (b->face == 2) { k = i; wx = i; wy = -3; swnx = G2(o->nj1, i, j); swny = G2(o->nj2, i, j); st = 1.; }
      else if (b->face == 3) { k = j; wx = nx + 3; wy = j; swnx = G2(o->ni1, i, j); swny = G2(o->ni2, i, j); st = -1.; }
      else { k = i; wx = i; wy = ny + 3; swnx = G2(o->nj1, i, j); swny = G2(o->nj2, i, j); st = -1.; }
      for (int jv = 1; jv <= o->nvj; jv++)
        for (int iv = 1; iv <= o->nvi; iv++) {
          double w = o->wi[iv - 1] * o->wj[jv - 1];
          double cx = o->vi[iv - 1] - uw, cy = o->vj[jv - 1] - vw;
          double vnw = (cx * swnx + cy * swny) * st;
          double vnn = (vnw - fabs(vnw)) / 2.;
          sumnw[k] += w * (vnn * FV(fvel(o, o->F, iv, jv), 1, wx, wy));
          double fb = (vnw + fabs(vnw)) / 2. * exp(-(cx * cx + cy * cy) / tw) / (PI_D * tw);
          sumnw1[k] += w * fb;
        }
    }
  for (int i = ib; i <= ie; i++)
    for (int j = jb; j <= je; j++) {
      if ((jb == je && i == icheck) || (ib == ie && j == jcheck)) continue;
      int k = (b->face == 1 || b->face == 3) ? j : i;
      sumnw[k] = -sumnw[k] / sumnw1[k];
    }
  /* 3) incoming half: diffuse Maxwellian at the wall (src/GKUA_wall_functions.f90:27-63) */
  for (int i = ib; i <= ie; i++)
    for (int j = jb; j <= je; j++) {
      if ((jb == je && i == icheck) || (ib == ie && j == jcheck)) continue;
      int k, wx, wy;
      double swnx, swny, st;
      if (b->face == 1) { k = j; wx = -3; wy = j; swnx = G2(o->ni1, i, j); swny = G2(o->ni2, i, j); st = 1.; }
      else if (b->face == 2) { k = i; wx = i; wy = -3; swnx = G2(o->nj1, i, j); swny = G2(o->nj2, i, j); st = 1.; }
      else if (b->face == 3) { k = j; wx = nx + 3; wy = j; swnx = G2(o->ni1, i, j); swny = G2(o->ni2, i, j); st = -1.; }
      else { k = i; wx = i; wy = ny + 3; swnx = G2(o->nj1, i, j); swny = G2(o->nj2, i, j); st = -1.; }
      for (int jv = 1; jv <= o->nvj; jv++)
        for (int iv = 1; iv <= o->nvi; iv++) {
          double cx = o->vi[iv - 1] - uw, cy = o->vj[jv - 1] - vw;
          double vnw = (cx * swnx + cy * swny) * st;
          if (vnw >= 0.0) {
            double cvuv = cx * cx + cy * cy;
            double g = sumnw[k] * exp(-cvuv / tw) / (PI_D * tw);
            if (sumnw[k] < 0.) sumnw[k] = 0.; /* applied after first use, src/GKUA_wall_functions.f90:52-54 */
            double *Fv = fvel(o, o->F, iv, jv);
            FV(Fv, 1, wx, wy) = g;
            FV(Fv, 2, wx, wy) = tw * g / 2.;
          }
        }
    }
  free(sumnw); free(sumnw1);
}

/* GKUA_boundary_farfield, src/GKUA_boundary.f90:103-215 */
static void bc_farfield(orc_t *o, const gsis_bc *b) {
  const double g = o->gamma, ma = o->ma;
  const double d_inf = 1., u_inf = cos(o->aoa), v_inf = sin(o->aoa), p_inf = 1. / (g * ma * ma);
  const double sq = sqrt(2.0 / g);
  for (int j = b->jst; j <= b->jend; j++)
    for (int i = b->ist; i <= b->iend; i++) {
      int i1, j1;
      if (b->face == 1 || b->face == 2) { i1 = i; j1 = j; }
      else if (b->face == 3) { i1 = i - 1; j1 = j; }
      else { i1 = i; j1 = j - 1; }
      double dx, dy, s, n1, n2;
      if (b->face == 1 || b->face == 3) {
        dx = XN(o->x, i, j + 1) - XN(o->x, i, j); dy = XN(o->y, i, j + 1) - XN(o->y, i, j);
        s = sqrt(dx * dx + dy * dy);
        if (b->face == 1) { n1 = -dy / s; n2 = dx / s; } else { n1 = dy / s; n2 = -dx / s; }
      } else {
        dx = XN(o->x, i + 1, j) - XN(o->x, i, j); dy = XN(o->y, i + 1, j) - XN(o->y, i, j);
        s = sqrt(dx * dx + dy * dy);
        if (b->face == 2) { n1 = dy / s; n2 = -dx / s; } else { n1 = -dy / s; n2 = dx / s; }
      }
      double ma_n = (u_inf * n1 + v_inf * n2) * ma;
      for (int k = 1; k <= 2; k++) {
        int i2, j2;
        if (b->face == 1) { i2 = i - k; j2 = j; }
        else if (b->face == 2) { i2 = i; j2 = j - k; }
        else if (b->face == 3) { i2 = i + k - 1; j2 = j; }
        else { i2 = i; j2 = j + k - 1; }
        if (ma_n > 0.) {
          for (int m = 1; m <= 4; m++) U4(o->U, m, i2, j2) = U4(o->U, m, i1, j1);
        } else {
          U4(o->U, 1, i2, j2) = d_inf;
          U4(o->U, 2, i2, j2) = d_inf * u_inf;
          U4(o->U, 3, i2, j2) = d_inf * v_inf;
          U4(o->U, 4, i2, j2) = p_inf / (g - 1.) + 0.5 * d_inf * (u_inf * u_inf + v_inf * v_inf);
        }
        double rho = U4(o->U, 1, i2, j2);
        double tu = U4(o->U, 2, i2, j2) / rho, tv = U4(o->U, 3, i2, j2) / rho;
        S_(1, i2, j2) = U4(o->U, 1, i2, j2);
        S_(2, i2, j2) = (U4(o->U, 2, i2, j2) / rho) * ma / sq;
        S_(3, i2, j2) = (U4(o->U, 3, i2, j2) / rho) * ma / sq;
        double tt = (U4(o->U, 4, i2, j2) - 0.5 * rho * (tu * tu + tv * tv)) / (o->cv * rho);
        S_(4, i2, j2) = tt;
        S_(5, i2, j2) = tt;
        double n0 = S_(1, i2, j2), u0 = S_(2, i2, j2), v0 = S_(3, i2, j2);
        for (int jv = 1; jv <= o->nvj; jv++)
          for (int iv = 1; iv <= o->nvi; iv++) {
            double viu = o->vi[iv - 1] - u0, vjv = o->vj[jv - 1] - v0;
            double vij2 = (viu * viu + vjv * vjv) / tt;
            double gd = n0 / (PI_D * tt) * exp(-vij2);
            double *Fv = fvel(o, o->F, iv, jv);
            FV(Fv, 1, i2, j2) = gd;
            FV(Fv, 2, i2, j2) = gd * tt / 2.;
          }
      }
    }
}

/* GKUA_boundary_inlet = periodic copy, src/GKUA_boundary.f90:219-273 */
static void bc_periodic(orc_t *o, const gsis_bc *b) {
  const int nx = o->nx, ny = o->ny;
  for (int j = b->jst; j <= b->jend; j++)
    for (int i = b->ist; i <= b->iend; i++)
      for (int k = 1; k <= 2; k++) {
        int i1, j1, i2, j2;
        if (b->face == 1) { i1 = nx - k; j1 = j; i2 = i - k; j2 = j; }
        else if (b->face == 2) { i1 = i; j1 = ny - k; i2 = i; j2 = j - k; }
        else if (b->face == 3) { i1 = k; j1 = j; i2 = i + k - 1; j2 = j; }
        else { i1 = i; j1 = k; i2 = i; j2 = j + k - 1; }
        for (int m = 1; m <= 11; m++) S_(m, i2, j2) = S_(m, i1, j1);
        for (int jv = 1; jv <= o->nvj; jv++)
          for (int iv = 1; iv <= o->nvi; iv++) {
            double *Fv = fvel(o, o->F, iv, jv);
            FV(Fv, 1, i2, j2) = FV(Fv, 1, i1, j1);
            FV(Fv, 2, i2, j2) = FV(Fv, 2, i1, j1);
          }
      }
}

/* GKUA_Phy_boundary, src/GKUA_boundary.f90:2-46 (BCs in input order) */
void orc_phy_boundary(orc_t *o) {
  for (int b = 0; b < o->nbc; b++) {
    const gsis_bc *bc = &o->bc[b];
    if (bc->code == GSIS_BC_WALL) bc_wall(o, bc);
    else if (bc->code == GSIS_BC_FARFIELD) bc_farfield(o, bc);
    else if (bc->code == GSIS_BC_PERIODIC) bc_periodic(o, bc);
  }
}

/* Get_Feq: reduced Shakhov equilibrium, src/GKUA_Residual.f90:420-469 */
static void get_feq(orc_t *o) {
  const double cpr = o->cpr;
#pragma omp parallel for schedule(static)
  for (int jv = 1; jv <= o->nvj; jv++)
    for (int iv = 1; iv <= o->nvi; iv++) {
      double *F0v = fvel(o, o->F0, iv, jv);
      for (int j = 1; j <= o->ny - 1; j++)
        for (int i = 1; i <= o->nx - 1; i++) {
          double n = S_(1, i, j), T = S_(4, i, j);
          double vjv = o->vj[jv - 1] - S_(3, i, j);
          double viu = o->vi[iv - 1] - S_(2, i, j);
          double crt = n / (PI_D * T);
          double cpt = cpr / (n * T * T);
          double viuq = viu * S_(8, i, j);
          double vq = viuq + vjv * S_(9, i, j);
          double vij2 = (viu * viu + vjv * vjv) / T;
          double fvx = cpt * vq;
          double gd = crt * exp(-vij2);
          FV(F0v, 1, i, j) = gd * (1.0 + fvx * (2. * vij2 - 4.));
          FV(F0v, 2, i, j) = (T * gd / 2.) * (1.0 + fvx * (2. * vij2 - 2.));
        }
    }
}

/* per-thread work arrays for one velocity */
typedef struct {
  double *flux; /* FluxFRD(2, 0:nx+1, 0:ny+1) */
  double *dfrd; /* DFRD(2, -1:max(nx,ny)+2) */
  double *rhs;  /* RHSFRD(2, 0:nx, 0:ny) */
  double *du;   /* DUFRD */
  double *ax;   /* ATSVX(0:nx+1, 0:ny+1) */
  double *ay;   /* ATSVY */
} work_t;

#define WNI (o->nx + 2)
#define WFL(w, ird, i, j) (w)->flux[((ird)-1) + 2 * ((i) + (size_t)WNI * (j))]
#define WRH(w, ird, i, j) (w)->rhs[((ird)-1) + 2 * ((i) + (size_t)WNI * (j))]
#define WDU(w, ird, i, j) (w)->du[((ird)-1) + 2 * ((i) + (size_t)WNI * (j))]
#define WAX(w, i, j) (w)->ax[(i) + (size_t)WNI * (j)]
#define WAY(w, i, j) (w)->ay[(i) + (size_t)WNI * (j)]
#define WD(w, ird, k) (w)->dfrd[((ird)-1) + 2 * ((k) + 1)]

/* WENO3 constants of RHSFDVsub (src/GKUA_Residual.f90:143-144): the
 * literals are REAL(4), converted to the REAL(8) variables they are stored in */
static const double EPS_WENO = (double)1.0e-15f;
static const double C_WENO1 = (double)(1.0f / 3.0f), C_WENO2 = (double)(2.0f / 3.0f);

/* WENO3 left state at a face from F(f-2), F(f-1), F(f) with smoothness
 * indicators is1 = (F(f-1)-F(f-2))^2, is2 = (F(f)-F(f-1))^2
 * (src/GKUA_Residual.f90:170-175) */
static double weno_left(double fm2, double fm1, double f0, double is1, double is2) {
  double a1 = C_WENO1 / ((is1 + EPS_WENO) * (is1 + EPS_WENO));
  double a2 = C_WENO2 / ((is2 + EPS_WENO) * (is2 + EPS_WENO));
  double as = a1 + a2, o1 = a1 / as, o2 = a2 / as;
  return o1 * (-0.5 * fm2 + 1.5 * fm1) + o2 * (0.5 * fm1 + 0.5 * f0);
}
/* WENO3 right state at a face from F(f-1), F(f), F(f+1) with
 * is2 = (F(f)-F(f-1))^2, is3 = (F(f+1)-F(f))^2 (src/GKUA_Residual.f90:177-182) */
static double weno_right(double fm1, double f0, double fp1, double is2, double is3) {
  double a1 = C_WENO2 / ((is2 + EPS_WENO) * (is2 + EPS_WENO));
  double a2 = C_WENO1 / ((is3 + EPS_WENO) * (is3 + EPS_WENO));
  double as = a1 + a2, o1 = a1 / as, o2 = a2 / as;
  return o1 * (0.5 * fm1 + 0.5 * f0) + o2 * (1.5 * f0 - 0.5 * fp1);
}

/* RHSFDVsub, src/GKUA_Residual.f90:117-418 (UP1, NND4A2 and WENO3 schemes) */
static void rhs_velocity(const orc_t *o, work_t *w, int iv, int jv) {
  const int nx = o->nx, ny = o->ny;
  const double *Fv = fvel(o, o->F, iv, jv);
  const double *F0v = fvel(o, o->F0, iv, jv);
  const double vx = o->vi[iv - 1], vy = o->vj[jv - 1];
  const int nnd = (o->scheme == GSIS_SCHEME_NND);
  const int weno = (o->scheme == GSIS_SCHEME_WENO3);
  /* i-direction (:145-187) */
  for (int j = 1; j <= ny - 1; j++) {
    for (int i = 0; i <= nx + 1; i++)
      for (int r = 1; r <= 2; r++) WD(w, r, i) = FV(Fv, r, i, j) - FV(Fv, r, i - 1, j);
    for (int i = 1; i <= nx; i++) {
      double a = (vx * G2(o->ni1, i, j) + vy * G2(o->ni2, i, j)) * G2(o->si, i, j);
      WAX(w, i, j) = a;
      double asp = (a + fabs(a)) * 0.5, asn = (a - fabs(a)) * 0.5;
      for (int r = 1; r <= 2; r++) {
        if (weno) { /* :167-185 */
          double is1 = WD(w, r, i - 1) * WD(w, r, i - 1), is2 = WD(w, r, i) * WD(w, r, i);
          double is3 = WD(w, r, i + 1) * WD(w, r, i + 1);
          double ulh = weno_left(FV(Fv, r, i - 2, j), FV(Fv, r, i - 1, j), FV(Fv, r, i, j), is1, is2);
          double urh = weno_right(FV(Fv, r, i - 1, j), FV(Fv, r, i, j), FV(Fv, r, i + 1, j), is2, is3);
          WFL(w, r, i, j) = asp * ulh + asn * urh;
        } else if (!nnd) {
          WFL(w, r, i, j) = asp * FV(Fv, r, i - 1, j) + asn * FV(Fv, r, i, j);
        } else {
          double ulg = FV(Fv, r, i - 1, j) + 0.5 * fminmd(WD(w, r, i), WD(w, r, i - 1));
          double urg = FV(Fv, r, i, j) - 0.5 * fminmd(WD(w, r, i), WD(w, r, i + 1));
          WFL(w, r, i, j) = asp * ulg + asn * urg;
        }
      }
    }
  }
  /* wall / farfield face overrides, i-faces (:189-266) */
  for (int b = 0; b < o->nbc; b++) {
    const gsis_bc *bc = &o->bc[b];
    if (!(bc->code == GSIS_BC_WALL || bc->code == GSIS_BC_FARFIELD)) continue;
    if (bc->face == 1) {
      for (int j = bc->jst; j <= bc->jend - 1; j++) {
        double a = WAX(w, 2, j), asp = (a + fabs(a)) * 0.5, asn = (a - fabs(a)) * 0.5;
        for (int r = 1; r <= 2; r++) {
          double f1 = FV(Fv, r, 1, j), f2 = FV(Fv, r, 2, j), f3 = FV(Fv, r, 3, j);
          if (weno) { /* central left state, WENO right state (:202-213) */
            double urh = weno_right(f1, f2, f3, (f2 - f1) * (f2 - f1), (f3 - f2) * (f3 - f2));
            WFL(w, r, 2, j) = asp * (0.5 * (f1 + f2)) + asn * urh;
          } else if (!nnd) WFL(w, r, 2, j) = asp * f1 + asn * f2;
          else {
            double ulg = f1 + 0.5 * (f2 - f1);
            double urg = f2 - 0.5 * fminmd(f2 - f1, f3 - f2);
            WFL(w, r, 2, j) = asp * ulg + asn * urg;
          }
        }
        for (int r = 1; r <= 2; r++) WFL(w, r, 1, j) = WAX(w, 1, j) * FV(Fv, r, -3, j);
      }
    } else if (bc->face == 3) {
      for (int j = bc->jst; j <= bc->jend - 1; j++) {
        double a = WAX(w, nx - 1, j), asp = (a + fabs(a)) * 0.5, asn = (a - fabs(a)) * 0.5;
        for (int r = 1; r <= 2; r++) {
          double f1 = FV(Fv, r, nx - 1, j), f2 = FV(Fv, r, nx - 2, j), f3 = FV(Fv, r, nx - 3, j);
          if (weno) { /* WENO left state, central right state (:239-250) */
            double ulh = weno_left(f3, f2, f1, (f2 - f3) * (f2 - f3), (f1 - f2) * (f1 - f2));
            WFL(w, r, nx - 1, j) = asp * ulh + asn * (0.5 * f2 + 0.5 * f1);
          } else if (!nnd) WFL(w, r, nx - 1, j) = asp * f2 + asn * f1;
          else {
            double ulg = f2 + 0.5 * fminmd(f1 - f2, f2 - f3);
            double urg = f1 - 0.5 * (f1 - f2);
            WFL(w, r, nx - 1, j) = asp * ulg + asn * urg;
          }
        }
        for (int r = 1; r <= 2; r++) WFL(w, r, nx, j) = WAX(w, nx, j) * FV(Fv, r, nx + 3, j);
      }
    }
  }
  for (int j = 1; j <= ny - 1; j++)
    for (int i = 1; i <= nx - 1; i++)
      for (int r = 1; r <= 2; r++) WRH(w, r, i, j) = WFL(w, r, i + 1, j) - WFL(w, r, i, j);
  /* j-direction (:278-320) */
  for (int i = 1; i <= nx - 1; i++) {
    for (int j = 0; j <= ny + 1; j++)
      for (int r = 1; r <= 2; r++) WD(w, r, j) = FV(Fv, r, i, j) - FV(Fv, r, i, j - 1);
    for (int j = 1; j <= ny; j++) {
      double a = (vx * G2(o->nj1, i, j) + vy * G2(o->nj2, i, j)) * G2(o->sj, i, j);
      WAY(w, i, j) = a;
      double asp = (a + fabs(a)) * 0.5, asn = (a - fabs(a)) * 0.5;
      for (int r = 1; r <= 2; r++) {
        if (weno) { /* :300-318 */
          double is1 = WD(w, r, j - 1) * WD(w, r, j - 1), is2 = WD(w, r, j) * WD(w, r, j);
          double is3 = WD(w, r, j + 1) * WD(w, r, j + 1);
          double ulh = weno_left(FV(Fv, r, i, j - 2), FV(Fv, r, i, j - 1), FV(Fv, r, i, j), is1, is2);
          double urh = weno_right(FV(Fv, r, i, j - 1), FV(Fv, r, i, j), FV(Fv, r, i, j + 1), is2, is3);
          WFL(w, r, i, j) = asp * ulh + asn * urh;
        } else if (!nnd) {
          WFL(w, r, i, j) = asp * FV(Fv, r, i, j - 1) + asn * FV(Fv, r, i, j);
        } else {
          double ulg = FV(Fv, r, i, j - 1) + 0.5 * fminmd(WD(w, r, j), WD(w, r, j - 1));
          double urg = FV(Fv, r, i, j) - 0.5 * fminmd(WD(w, r, j), WD(w, r, j + 1));
          WFL(w, r, i, j) = asp * ulg + asn * urg;
        }
      }
    }
  }
  /* j-face overrides (:321-396) */
  for (int b = 0; b < o->nbc; b++) {
    const gsis_bc *bc = &o->bc[b];
    if (!(bc->code == GSIS_BC_WALL || bc->code == GSIS_BC_FARFIELD)) continue;
    if (bc->face == 2) {
      for (int i = bc->ist; i <= bc->iend - 1; i++) {
        double a = WAY(w, i, 2), asp = (a + fabs(a)) * 0.5, asn = (a - fabs(a)) * 0.5;
        for (int r = 1; r <= 2; r++) {
          double f1 = FV(Fv, r, i, 1), f2 = FV(Fv, r, i, 2), f3 = FV(Fv, r, i, 3);
          if (weno) { /* :333-344 */
            double urh = weno_right(f1, f2, f3, (f2 - f1) * (f2 - f1), (f3 - f2) * (f3 - f2));
            WFL(w, r, i, 2) = asp * (0.5 * f1 + 0.5 * f2) + asn * urh;
          } else if (!nnd) WFL(w, r, i, 2) = asp * f1 + asn * f2;
          else {
            double ulg = f1 + 0.5 * (f2 - f1);
            double urg = f2 - 0.5 * fminmd(f2 - f1, f3 - f2);
            WFL(w, r, i, 2) = asp * ulg + asn * urg;
          }
        }
        for (int r = 1; r <= 2; r++) WFL(w, r, i, 1) = WAY(w, i, 1) * FV(Fv, r, i, -3);
      }
    } else if (bc->face == 4) {
      for (int i = bc->ist; i <= bc->iend - 1; i++) {
        double a = WAY(w, i, ny - 1), asp = (a + fabs(a)) * 0.5, asn = (a - fabs(a)) * 0.5;
        for (int r = 1; r <= 2; r++) {
          double f1 = FV(Fv, r, i, ny - 1), f2 = FV(Fv, r, i, ny - 2), f3 = FV(Fv, r, i, ny - 3);
          if (weno) { /* :369-380 */
            double ulh = weno_left(f3, f2, f1, (f2 - f3) * (f2 - f3), (f1 - f2) * (f1 - f2));
            WFL(w, r, i, ny - 1) = asp * ulh + asn * (0.5 * f2 + 0.5 * f1);
          } else if (!nnd) WFL(w, r, i, ny - 1) = asp * f2 + asn * f1;
          else {
            double ulg = f2 + 0.5 * fminmd(f1 - f2, f2 - f3);
            double urg = f1 - 0.5 * (f1 - f2);
            WFL(w, r, i, ny - 1) = asp * ulg + asn * urg;
          }
        }
        for (int r = 1; r <= 2; r++) WFL(w, r, i, ny) = WAY(w, i, ny) * FV(Fv, r, i, ny + 3);
      }
    }
  }
  for (int i = 1; i <= nx - 1; i++)
    for (int j = 1; j <= ny - 1; j++)
      for (int r = 1; r <= 2; r++) WRH(w, r, i, j) = WRH(w, r, i, j) + WFL(w, r, i, j + 1) - WFL(w, r, i, j);
  /* source term (:406-414) */
  for (int j = 1; j <= ny - 1; j++)
    for (int i = 1; i <= nx - 1; i++) {
      double vol = G2(o->vol, i, j), cf = C2(o->CF, i, j);
      for (int r = 1; r <= 2; r++)
        WRH(w, r, i, j) = -WRH(w, r, i, j) / vol + cf * (FV(F0v, r, i, j) - FV(Fv, r, i, j));
    }
}

static const double CLAMP_1EM30 = (double)1.0e-30f; /* 1.E-30 is a REAL(4) literal */

/* GKUA_Residual: per-velocity RHS, forward/backward LU-SGS sweeps on
 * anti-diagonal planes, update, FRD00<-RHS, positivity clamp
 * (src/GKUA_Residual.f90:4-114). */
static void kinetic_solve(orc_t *o) {
  const int nx = o->nx, ny = o->ny;
  const double dt = o->dt;
  get_feq(o);
  const int nvt = o->nvi * o->nvj;
  const size_t wn = (size_t)(nx + 2) * (ny + 2);
  const int nd = (nx > ny ? nx : ny) + 4;
#pragma omp parallel
  {
    work_t w;
    w.flux = zalloc(2 * wn); w.rhs = zalloc(2 * wn); w.du = zalloc(2 * wn);
    w.ax = zalloc(wn); w.ay = zalloc(wn); w.dfrd = zalloc(2 * nd);
#pragma omp for schedule(dynamic, 1)
    for (int v = 0; v < nvt; v++) {
      const int iv = v % o->nvi + 1, jv = v / o->nvi + 1;
      rhs_velocity(o, &w, iv, jv);
      double *Fv = fvel(o, o->F, iv, jv);
      double *Dv = fvel(o, o->F00, iv, jv);
      for (int plane = 2; plane <= nx - 1 + ny - 1; plane++)
        for (int j = 1; j <= ny - 1; j++) {
          int i = plane - j;
          if (i < 1 || i > nx - 1) continue;
          double axp = (WAX(&w, i + 1, j) + fabs(WAX(&w, i + 1, j))) * 0.5;
          double ayp = (WAY(&w, i, j + 1) + fabs(WAY(&w, i, j + 1))) * 0.5;
          double axn = (WAX(&w, i, j) - fabs(WAX(&w, i, j))) * 0.5;
          double ayn = (WAY(&w, i, j) - fabs(WAY(&w, i, j))) * 0.5;
          double vol = G2(o->vol, i, j);
          double alfa = vol / dt + (axp - axn + ayp - ayn) + vol * C2(o->CF, i, j);
          double axp1 = (WAX(&w, i, j) + fabs(WAX(&w, i, j))) * 0.5;
          double ayp1 = (WAY(&w, i, j) + fabs(WAY(&w, i, j))) * 0.5;
          for (int r = 1; r <= 2; r++) {
            double di = axp1 * FV(Dv, r, i - 1, j);
            double dj = ayp1 * FV(Dv, r, i, j - 1);
            FV(Dv, r, i, j) = (vol * WRH(&w, r, i, j) + di + dj) / alfa;
          }
        }
      for (int plane = nx - 1 + ny - 1; plane >= 2; plane--)
        for (int j = ny - 1; j >= 1; j--) {
          int i = plane - j;
          if (i < 1 || i > nx - 1) continue;
          double axp = (WAX(&w, i + 1, j) + fabs(WAX(&w, i + 1, j))) * 0.5;
          double ayp = (WAY(&w, i, j + 1) + fabs(WAY(&w, i, j + 1))) * 0.5;
          double axn = (WAX(&w, i, j) - fabs(WAX(&w, i, j))) * 0.5;
          double ayn = (WAY(&w, i, j) - fabs(WAY(&w, i, j))) * 0.5;
          double vol = G2(o->vol, i, j);
          double alfa = vol / dt + (axp - axn + ayp - ayn) + vol * C2(o->CF, i, j);
          double axn1 = (WAX(&w, i + 1, j) - fabs(WAX(&w, i + 1, j))) * 0.5;
          double ayn1 = (WAY(&w, i, j + 1) - fabs(WAY(&w, i, j + 1))) * 0.5;
          for (int r = 1; r <= 2; r++) {
            double di = -axn1 * FV(Dv, r, i + 1, j);
            double dj = -ayn1 * FV(Dv, r, i, j + 1);
            WDU(&w, r, i, j) = FV(Dv, r, i, j) + (di + dj) / alfa;
          }
        }
      for (int j = 1; j <= ny - 1; j++)
        for (int i = 1; i <= nx - 1; i++)
          for (int r = 1; r <= 2; r++) {
            FV(Fv, r, i, j) = FV(Fv, r, i, j) + WDU(&w, r, i, j);
            FV(Dv, r, i, j) = WRH(&w, r, i, j);
          }
      for (int j = -1; j <= ny + 1; j++)
        for (int i = -1; i <= nx + 1; i++)
          for (int r = 1; r <= 2; r++)
            if (FV(Fv, r, i, j) < CLAMP_1EM30) FV(Fv, r, i, j) = CLAMP_1EM30;
    }
    free(w.flux); free(w.rhs); free(w.du); free(w.ax); free(w.ay); free(w.dfrd);
  }
}

/* DETFPwall for wall and farfield faces, src/GKUA_DETFP_WALL.f90:1-95 */
static void det_wall(orc_t *o) {
  for (int b = 0; b < o->nbc; b++) {
    const gsis_bc *bc = &o->bc[b];
    if (!(bc->code == GSIS_BC_WALL || bc->code == GSIS_BC_FARFIELD)) continue;
    const int ib = bc->ist, ie = bc->iend, jb = bc->jst, je = bc->jend;
    const int icheck = ib > ie ? ib : ie, jcheck = jb > je ? jb : je;
    double *sw = o->swall + o->swall_off[b];
    for (int i = ib; i <= ie; i++)
      for (int j = jb; j <= je; j++) {
        if ((jb == je && i == icheck) || (ib == ie && j == jcheck)) continue;
        int g1, g2, a1, a2, b1, b2, kk;
        if (bc->face == 1) { g1 = ib - 1; g2 = j; a1 = ib; a2 = j; b1 = ib + 1; b2 = j; kk = j - jb; }
        else if (bc->face == 2) { g1 = i; g2 = jb - 1; a1 = i; a2 = jb; b1 = i; b2 = jb + 1; kk = i - ib; }
        else if (bc->face == 3) { g1 = ie; g2 = j; a1 = ie - 1; a2 = j; b1 = ie - 2; b2 = j; kk = j - jb; }
        else { g1 = i; g2 = je; a1 = i; a2 = je - 1; b1 = i; b2 = je - 2; kk = i - ib; }
        for (int k = 1; k <= 10; k++) {
          double swk = (-S_(k, b1, b2) + 3.0 * S_(k, a1, a2)) * 0.5;
          sw[(k - 1) + 10 * kk] = swk;
          S_(k, g1, g2) = 2.0 * swk - S_(k, a1, a2);
        }
      }
  }
}

/* DETFP with GaussLegendre moments and the primitive conversion,
 * src/GKUA_DETFP.f90:4-151 and :288-363 (VHS collision law). */
void orc_detfp(orc_t *o) {
  const int nx = o->nx, ny = o->ny;
  memset(o->S, 0, sizeof(double) * 11 * NI7 * NJ7); /* B%S_GKUA=0.0 (:303) */
#pragma omp parallel for schedule(static)
  for (int j = 0; j <= ny; j++)
    for (int i = 0; i <= nx; i++) {
      double acc[10] = {0};
      for (int jv = 1; jv <= o->nvj; jv++) {
        const double vy = o->vj[jv - 1];
        for (int iv = 1; iv <= o->nvi; iv++) {
          const double vx = o->vi[iv - 1];
          const double wgt = o->wi[iv - 1] * o->wj[jv - 1];
          const double *Fv = fvel(o, o->F, iv, jv);
          double g = FV(Fv, 1, i, j), h = FV(Fv, 2, i, j);
          double vx1 = vx * g, vy1 = vy * g, vx2 = vx * vx1, vy2 = vy * vy1;
          double f4 = h + vx2 + vy2;
          acc[0] += wgt * g;
          acc[1] += wgt * vx1;
          acc[2] += wgt * vy1;
          acc[3] += wgt * f4;
          acc[4] += wgt * vx2;
          acc[5] += wgt * vy2;
          acc[6] += wgt * (vx * vy * g);
          acc[7] += wgt * (vx * f4);
          acc[8] += wgt * (vy * f4);
          acc[9] += wgt * h;
        }
      }
      for (int k = 0; k < 10; k++) S_(k + 1, i, j) = acc[k];
    }
  /* primitives and collision frequency (:57-105) */
  for (int j = 0; j <= ny; j++)
    for (int i = 0; i <= nx; i++) {
      double s4 = S_(4, i, j), s5 = S_(5, i, j), s6 = S_(6, i, j), s7 = S_(7, i, j);
      double ron = S_(1, i, j);
      if (ron <= 0.) ron = 1.0; /* R00 = d0 = 1 (src/sub_init.f90:405) */
      double u = S_(2, i, j) / ron; S_(2, i, j) = u;
      double v = S_(3, i, j) / ron; S_(3, i, j) = v;
      double us1 = u * u, vs1 = v * v, sum_uv = us1 + vs1;
      double ttr = (s4 - ron * sum_uv) / (1.5 * ron);
      if (ttr <= 0.) ttr = 1.0; /* T00 = 1 (src/sub_init.f90:404) */
      S_(4, i, j) = ttr;
      S_(5, i, j) = (s5 - ron * us1) * 2.;
      S_(6, i, j) = (s6 - ron * vs1) * 2.;
      S_(7, i, j) = (s7 - ron * u * v) * 2.;
      S_(8, i, j) = S_(8, i, j) - 2. * u * s5 - 2. * v * s7 + ron * u * (sum_uv - 1.5 * ttr);
      S_(9, i, j) = S_(9, i, j) - 2. * v * s6 - 2. * u * s7 + ron * v * (sum_uv - 1.5 * ttr);
      S_(10, i, j) = S_(10, i, j) * 2.;
      if (o->x1s <= (double)1.e-8f) C2(o->CF, i, j) = o->cfc * S_(1, i, j);
      else C2(o->CF, i, j) = o->cfc * S_(1, i, j) * pow(S_(4, i, j), o->x1s);
    }
  det_wall(o);
}

/* comput_duvtpckw, src/sub_NS_singlegrid.f90:121-157; returns nonzero on T<0.1 */
int orc_comput_prims(orc_t *o) {
  const double ptmp = 1. / (o->gamma * o->ma * o->ma);
  int bad = 0;
  for (int j = -1; j <= o->ny + 1; j++)
    for (int i = -1; i <= o->nx + 1; i++) {
      double d = U4(o->U, 1, i, j);
      double u = U4(o->U, 2, i, j) / d, v = U4(o->U, 3, i, j) / d;
      double T = (U4(o->U, 4, i, j) - 0.5 * d * (u * u + v * v)) / (o->cv * d);
      C2(o->d, i, j) = d; C2(o->uu, i, j) = u; C2(o->vv, i, j) = v; C2(o->T, i, j) = T;
      C2(o->p, i, j) = ptmp * d * T;
    }
  for (int j = 1; j <= o->ny - 1; j++)
    for (int i = 1; i <= o->nx - 1; i++)
      if (C2(o->T, i, j) < (double)0.1f) bad = 1;
  return bad;
}

/* get_viscous: mu = T^0.81/Re with the REAL(4) literal 0.81, src/sub_Residual.f90:553-576 */
void orc_get_viscous(orc_t *o) {
  const double omg = (double)0.81f;
  for (int j = 0; j <= o->ny; j++)
    for (int i = 0; i <= o->nx; i++) C2(o->amu, i, j) = 1. / o->re * pow(C2(o->T, i, j), omg);
}

/* Face-centred NS stress and heat flux (the same discretisation is used by
 * Convert_NS src/Hybrid_Sub.f90:257-309,358-407 and by Residual_FVM
 * src/sub_Residual.f90:125-165,262-298).  dir 0: i-face (i,j), 1: j-face. */
void orc_ns_face_stress(const orc_t *o, int dir, int i, int j, double *t11, double *t22,
                           double *t12, double *t33, double *qx, double *qy, double *u1o, double *v1o) {
  const int nx = o->nx, ny = o->ny;
  double amu1, amk1, dix, diy, djx, djy, diu, div, dit, dju, djv, djt, u1, v1;
  const double prt = 0.9; /* PrT, src/COMMONDATA.f90:22 */
  if (dir == 0) {
    if (i == 1) { amu1 = C2(o->amu, i, j) + 0.; amk1 = o->cp * (C2(o->amu, i, j) / o->pr + 0. / prt); }
    else if (i == nx) { amu1 = C2(o->amu, i - 1, j) + 0.; amk1 = o->cp * (C2(o->amu, i - 1, j) / o->pr + 0. / prt); }
    else {
      amu1 = (C2(o->amu, i - 1, j) + C2(o->amu, i, j) + 0. + 0.) * 0.5;
      amk1 = o->cp * ((C2(o->amu, i - 1, j) + C2(o->amu, i, j)) / o->pr + (0. + 0.) / prt) * 0.5;
    }
    dix = C2(o->x1, i, j) - C2(o->x1, i - 1, j);
    diy = C2(o->y1, i, j) - C2(o->y1, i - 1, j);
    djx = (C2(o->x1, i - 1, j + 1) + C2(o->x1, i, j + 1) - C2(o->x1, i - 1, j - 1) - C2(o->x1, i, j - 1)) * 0.25;
    djy = (C2(o->y1, i - 1, j + 1) + C2(o->y1, i, j + 1) - C2(o->y1, i - 1, j - 1) - C2(o->y1, i, j - 1)) * 0.25;
    diu = C2(o->uu, i, j) - C2(o->uu, i - 1, j);
    div = C2(o->vv, i, j) - C2(o->vv, i - 1, j);
    dit = C2(o->T, i, j) - C2(o->T, i - 1, j);
    dju = (C2(o->uu, i - 1, j + 1) + C2(o->uu, i, j + 1) - C2(o->uu, i - 1, j - 1) - C2(o->uu, i, j - 1)) * 0.25;
    djv = (C2(o->vv, i - 1, j + 1) + C2(o->vv, i, j + 1) - C2(o->vv, i - 1, j - 1) - C2(o->vv, i, j - 1)) * 0.25;
    djt = (C2(o->T, i - 1, j + 1) + C2(o->T, i, j + 1) - C2(o->T, i - 1, j - 1) - C2(o->T, i, j - 1)) * 0.25;
    u1 = (C2(o->uu, i, j) + C2(o->uu, i - 1, j)) * 0.5;
    v1 = (C2(o->vv, i, j) + C2(o->vv, i - 1, j)) * 0.5;
  } else {
    if (j == 1) { amu1 = C2(o->amu, i, j) + 0.; amk1 = o->cp * (C2(o->amu, i, j) / o->pr + 0. / prt); }
    else if (j == ny) { amu1 = C2(o->amu, i, j - 1) + 0.; amk1 = o->cp * (C2(o->amu, i, j - 1) / o->pr + 0. / prt); }
    else {
      amu1 = (C2(o->amu, i, j) + C2(o->amu, i, j - 1) + 0. + 0.) * 0.5;
      amk1 = o->cp * ((C2(o->amu, i, j - 1) + C2(o->amu, i, j)) / o->pr + (0. + 0.) / prt) * 0.5;
    }
    dix = (C2(o->x1, i + 1, j - 1) + C2(o->x1, i + 1, j) - C2(o->x1, i - 1, j - 1) - C2(o->x1, i - 1, j)) * 0.25;
    diy = (C2(o->y1, i + 1, j - 1) + C2(o->y1, i + 1, j) - C2(o->y1, i - 1, j - 1) - C2(o->y1, i - 1, j)) * 0.25;
    djx = C2(o->x1, i, j) - C2(o->x1, i, j - 1);
    djy = C2(o->y1, i, j) - C2(o->y1, i, j - 1);
    diu = (C2(o->uu, i + 1, j - 1) + C2(o->uu, i + 1, j) - C2(o->uu, i - 1, j - 1) - C2(o->uu, i - 1, j)) * 0.25;
    div = (C2(o->vv, i + 1, j - 1) + C2(o->vv, i + 1, j) - C2(o->vv, i - 1, j - 1) - C2(o->vv, i - 1, j)) * 0.25;
    dit = (C2(o->T, i + 1, j - 1) + C2(o->T, i + 1, j) - C2(o->T, i - 1, j - 1) - C2(o->T, i - 1, j)) * 0.25;
    dju = C2(o->uu, i, j) - C2(o->uu, i, j - 1);
    djv = C2(o->vv, i, j) - C2(o->vv, i, j - 1);
    djt = C2(o->T, i, j) - C2(o->T, i, j - 1);
    u1 = (C2(o->uu, i, j) + C2(o->uu, i, j - 1)) * 0.5;
    v1 = (C2(o->vv, i, j) + C2(o->vv, i, j - 1)) * 0.5;
  }
  double ds = 1. / (dix * djy - djx * diy);
  double ux = (diu * djy - dju * diy) * ds;
  double vx = (div * djy - djv * diy) * ds;
  double tx = (dit * djy - djt * diy) * ds;
  double uy = (-diu * djx + dju * dix) * ds;
  double vy = (-div * djx + djv * dix) * ds;
  double ty = (-dit * djx + djt * dix) * ds;
  *t11 = ((4. / 3.) * ux - (2. / 3.) * vy) * amu1;
  *t22 = ((4. / 3.) * vy - (2. / 3.) * ux) * amu1;
  *t12 = (uy + vx) * amu1;
  *t33 = amu1 * (-(2. / 3.) * ux - (2. / 3.) * vy);
  *qx = amk1 * tx;
  *qy = amk1 * ty;
  if (u1o) *u1o = u1;
  if (v1o) *v1o = v1;
}

/* Convert_NS: S -> NS conservative U, face NS stress and the high-order terms
 * (HoT) of the GSIS constitutive correction, src/Hybrid_Sub.f90:191-461. */
int orc_convert_ns(orc_t *o) {
  const int nx = o->nx, ny = o->ny;
  const double g = o->gamma, ma = o->ma;
  /* bilinear corner extrapolation (:221-224) */
  for (int k = 1; k <= 11; k++) {
    S_(k, 0, 0) = S_(k, 1, 0) + S_(k, 0, 1) - S_(k, 1, 1);
    S_(k, nx, 0) = S_(k, nx, 1) + S_(k, nx - 1, 0) - S_(k, nx - 1, 1);
    S_(k, 0, ny) = S_(k, 1, ny) + S_(k, 0, ny - 1) - S_(k, 1, ny - 1);
    S_(k, nx, ny) = S_(k, nx - 1, ny) + S_(k, nx, ny - 1) - S_(k, nx - 1, ny - 1);
  }
  const double sq = sqrt(2.0 / g);
  for (int i = 0; i <= nx; i++)
    for (int j = 0; j <= ny; j++) {
      U4(o->U, 1, i, j) = S_(1, i, j);
      double tu = S_(2, i, j) / ma * sq;
      U4(o->U, 2, i, j) = U4(o->U, 1, i, j) * tu;
      double tv = S_(3, i, j) / ma * sq;
      U4(o->U, 3, i, j) = U4(o->U, 1, i, j) * tv;
      double tt = S_(4, i, j);
      U4(o->U, 4, i, j) = U4(o->U, 1, i, j) * o->cv * tt + 0.5 * U4(o->U, 1, i, j) * (tu * tu + tv * tv);
    }
  /* linear ghost extrapolation, whole-array statements in order (:236-239) */
  for (int j = -1; j <= ny + 1; j++)
    for (int m = 1; m <= 4; m++) U4(o->U, m, 0, j) = 2.0 * U4(o->U, m, 1, j) - U4(o->U, m, 2, j);
  for (int j = -1; j <= ny + 1; j++)
    for (int m = 1; m <= 4; m++) U4(o->U, m, nx, j) = 2.0 * U4(o->U, m, nx - 1, j) - U4(o->U, m, nx - 2, j);
  for (int i = -1; i <= nx + 1; i++)
    for (int m = 1; m <= 4; m++) U4(o->U, m, i, 0) = 2.0 * U4(o->U, m, i, 1) - U4(o->U, m, i, 2);
  for (int i = -1; i <= nx + 1; i++)
    for (int m = 1; m <= 4; m++) U4(o->U, m, i, ny) = 2.0 * U4(o->U, m, i, ny - 1) - U4(o->U, m, i, ny - 2);
  int bad = orc_comput_prims(o);
  if (bad) return GSIS_ETEMP;
  orc_get_viscous(o);
  if (!o->if_viscous) return 0;
  const double gm2 = g * (ma * ma);
  for (int j = 1; j <= ny - 1; j++)
    for (int i = 1; i <= nx; i++) {
      double t11, t22, t12, t33, qx, qy;
      orc_ns_face_stress(o, 0, i, j, &t11, &t22, &t12, &t33, &qx, &qy, NULL, NULL);
      U4(o->taonsi_in, 1, i, j) = t11; U4(o->taonsi_in, 2, i, j) = t22;
      U4(o->taonsi_in, 3, i, j) = t12; U4(o->taonsi_in, 4, i, j) = t33;
      Q2(o->qi_in, 1, i, j) = qx; Q2(o->qi_in, 2, i, j) = qy;
      if (i != 1 && i != nx) {
        double tmp = (S_(5, i, j) + S_(6, i, j) + S_(10, i, j) + S_(5, i - 1, j) + S_(6, i - 1, j) + S_(10, i - 1, j)) / 6.0;
        U4(o->taonsi_g, 1, i, j) = -((S_(5, i, j) + S_(5, i - 1, j)) / 2.0 - tmp) / gm2 - t11;
        U4(o->taonsi_g, 2, i, j) = -((S_(6, i, j) + S_(6, i - 1, j)) / 2.0 - tmp) / gm2 - t22;
        U4(o->taonsi_g, 3, i, j) = -(S_(7, i, j) + S_(7, i - 1, j)) / 2.0 / gm2 - t12;
        U4(o->taonsi_g, 4, i, j) = -((S_(10, i, j) + S_(10, i - 1, j)) / 2.0 - tmp) / gm2 - t33;
        Q2(o->qi_g, 1, i, j) = -(S_(8, i - 1, j) + S_(8, i, j)) / 2.0 * ((double)sqrtf(2.0f)) / ((ma * ma * ma) * pow(g, 1.5)) - qx;
        Q2(o->qi_g, 2, i, j) = -(S_(9, i - 1, j) + S_(9, i, j)) / 2.0 * ((double)sqrtf(2.0f)) / ((ma * ma * ma) * pow(g, 1.5)) - qy;
      } else {
        int ic = (i == 1) ? i : i - 1;
        double tmp = (S_(5, ic, j) + S_(6, ic, j) + S_(10, ic, j)) / 3.0;
        U4(o->taonsi_g, 1, i, j) = -(S_(5, ic, j) - tmp) / gm2 - t11;
        U4(o->taonsi_g, 2, i, j) = -(S_(6, ic, j) - tmp) / gm2 - t22;
        U4(o->taonsi_g, 3, i, j) = -S_(7, ic, j) / gm2 - t12;
        U4(o->taonsi_g, 4, i, j) = -(S_(10, ic, j) - tmp) / gm2 - t33;
        Q2(o->qi_g, 1, i, j) = -S_(8, ic, j) * ((double)sqrtf(2.0f)) / ((ma * ma * ma) * pow(g, 1.5)) - qx;
        Q2(o->qi_g, 2, i, j) = -S_(9, ic, j) * ((double)sqrtf(2.0f)) / ((ma * ma * ma) * pow(g, 1.5)) - qy;
      }
    }
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx - 1; i++) {
      double t11, t22, t12, t33, qx, qy;
      orc_ns_face_stress(o, 1, i, j, &t11, &t22, &t12, &t33, &qx, &qy, NULL, NULL);
      U4(o->taonsj_in, 1, i, j) = t11; U4(o->taonsj_in, 2, i, j) = t22;
      U4(o->taonsj_in, 3, i, j) = t12; U4(o->taonsj_in, 4, i, j) = t33;
      Q2(o->qj_in, 1, i, j) = qx; Q2(o->qj_in, 2, i, j) = qy;
      if (j != 1 && j != ny) {
        double tmp = (S_(5, i, j) + S_(6, i, j) + S_(10, i, j) + S_(5, i, j - 1) + S_(6, i, j - 1) + S_(10, i, j - 1)) / 6.0;
        U4(o->taonsj_g, 1, i, j) = -((S_(5, i, j) + S_(5, i, j - 1)) / 2.0 - tmp) / gm2 - t11;
        U4(o->taonsj_g, 2, i, j) = -((S_(6, i, j) + S_(6, i, j - 1)) / 2.0 - tmp) / gm2 - t22;
        U4(o->taonsj_g, 3, i, j) = -(S_(7, i, j) + S_(7, i, j - 1)) / 2.0 / gm2 - t12;
        U4(o->taonsj_g, 4, i, j) = -((S_(10, i, j) + S_(10, i, j - 1)) / 2.0 - tmp) / gm2 - t33;
        Q2(o->qj_g, 1, i, j) = -(S_(8, i, j - 1) + S_(8, i, j)) / 2.0 * ((double)sqrtf(2.0f)) / ((ma * ma * ma) * pow(g, 1.5)) - qx;
        Q2(o->qj_g, 2, i, j) = -(S_(9, i, j - 1) + S_(9, i, j)) / 2.0 * ((double)sqrtf(2.0f)) / ((ma * ma * ma) * pow(g, 1.5)) - qy;
      } else {
        int jc = (j == 1) ? j : j - 1;
        double tmp = (S_(5, i, jc) + S_(6, i, jc) + S_(10, i, jc)) / 3.0;
        U4(o->taonsj_g, 1, i, j) = -(S_(5, i, jc) - tmp) / gm2 - t11;
        U4(o->taonsj_g, 2, i, j) = -(S_(6, i, jc) - tmp) / gm2 - t22;
        U4(o->taonsj_g, 3, i, j) = -S_(7, i, jc) / gm2 - t12;
        U4(o->taonsj_g, 4, i, j) = -(S_(10, i, jc) - tmp) / gm2 - t33;
        Q2(o->qj_g, 1, i, j) = -S_(8, i, jc) * ((double)sqrtf(2.0f)) / ((ma * ma * ma) * pow(g, 1.5)) - qx;
        Q2(o->qj_g, 2, i, j) = -S_(9, i, jc) * ((double)sqrtf(2.0f)) / ((ma * ma * ma) * pow(g, 1.5)) - qy;
      }
    }
  return 0;
}

int orc_synthetic_ns(orc_t *o); /* defined in gsis_oracle_ns.c */

/* out_mass of DETFP: the sum of S_GKUA(1)*vol over the interior cells, j
   outer and i inner (src/GKUA_DETFP.f90:134-138) */
double orc_out_mass(const orc_t *o) {
  double m = 0.0;
  for (int j = 1; j <= o->ny - 1; j++)
    for (int i = 1; i <= o->nx - 1; i++) m = m + S_(1, i, j) * G2(o->vol, i, j);
  return m;
}

/* MassCorrect on: Init_mass is the out_mass of the current state, as after a
   restart (src/Backup_write_read.f90:102-104) */
void orc_set_mass_correct(orc_t *o, int on) {
  o->mass_correct = on;
  if (on) o->init_mass = orc_out_mass(o);
}

/* MassCorrect (src/GKUA_DETFP.f90:153-172): S_GKUA(1) of every interior cell
   scaled by Init_mass/out_mass */
static void mass_correct(orc_t *o) {
  const double out_mass = orc_out_mass(o);
  for (int j = 1; j <= o->ny - 1; j++)
    for (int i = 1; i <= o->nx - 1; i++) S_(1, i, j) = S_(1, i, j) * o->init_mass / out_mass;
}

/* Hybrid_Time_advance_LU_SGS, src/Hybrid_Sub.f90:4-160 */
int orc_step(orc_t *o, int nsteps) {
  for (int s = 0; s < nsteps; s++) {
    for (int i = 1; i <= o->nx - 1; i++)
      for (int j = 1; j <= o->ny - 1; j++)
        for (int k = 1; k <= 10; k++) S0_(k, i, j) = S_(k, i, j);
    orc_phy_boundary(o);
    kinetic_solve(o);
    orc_detfp(o);
    /* the place of the reference's (commented-out) call MassCorrect, src/Hybrid_Sub.f90:42 */
    if (o->mass_correct) mass_correct(o);
    int rc = orc_convert_ns(o);
    if (rc) return rc;
    if (o->mode == GSIS_MODE_M1) {
      rc = orc_synthetic_ns(o);
      if (rc) return rc;
    }
    o->tt += o->dt;
    o->kstep += 1;
  }
  return 0;
}

/* DETFP_res + GaussLegendre_rs, src/GKUA_DETFP_RES.f90:4-136 */
int orc_residual(orc_t *o, double res_rms[4], double res_max[4]) {
  const int nx = o->nx, ny = o->ny;
  memset(o->S00, 0, sizeof(double) * 11 * NI7 * NJ7);
  for (int j = 1; j <= ny - 1; j++)
    for (int i = 1; i <= nx - 1; i++) {
      double acc[4] = {0};
      for (int jv = 1; jv <= o->nvj; jv++) {
        const double vy = o->vj[jv - 1];
        for (int iv = 1; iv <= o->nvi; iv++) {
          const double vx = o->vi[iv - 1];
          const double wgt = o->wi[iv - 1] * o->wj[jv - 1];
          const double *Dv = fvel(o, o->F00, iv, jv);
          double g = FV(Dv, 1, i, j), h = FV(Dv, 2, i, j);
          double vx1 = vx * g, vy1 = vy * g;
          acc[0] += wgt * g;
          acc[1] += wgt * vx1;
          acc[2] += wgt * vy1;
          acc[3] += wgt * (h + vx * vx1 + vy * vy1);
        }
      }
      for (int k = 0; k < 4; k++) S00_(k + 1, i, j) = acc[k];
    }
  double rms[4] = {0}, mx[4] = {0};
  for (int j = 1; j <= ny - 1; j++)
    for (int i = 1; i <= nx - 1; i++)
      for (int k = 1; k <= 4; k++) {
        double r = S00_(k, i, j);
        if (fabs(r) > mx[k - 1]) mx[k - 1] = fabs(r);
        rms[k - 1] += r * r;
      }
  const double ncell = (double)(nx - 1) * (ny - 1);
  for (int k = 0; k < 4; k++) {
    res_rms[k] = sqrt(rms[k] / (1.0 * ncell));
    if (res_max) res_max[k] = mx[k];
  }
  return isnan(res_rms[0]) ? GSIS_ENAN : 0;
}
