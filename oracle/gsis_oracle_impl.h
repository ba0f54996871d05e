/*
 * gsis_oracle_impl.h -- private layout of the CPU oracle.  TEST INFRASTRUCTURE ONLY.
 * Arrays mirror the reference's Fortran layouts (column-major, same lower bounds).
 */
#ifndef GSIS_ORACLE_IMPL_H
#define GSIS_ORACLE_IMPL_H
#include "gsis_oracle.h"

#define PI_D 3.1415926535897932 /* const_var PI, src/COMMONDATA.f90:8 */

/* ---- layout helpers (Fortran column-major with lower bounds) ---- */
/* FRD(2,-3:nx+3,-3:ny+3,nvi,nvj) */
#define NI7 (o->nx + 7)
#define NJ7 (o->ny + 7)
#define FPLANE ((size_t)2 * NI7 * NJ7)
#define FV(base, ird, i, j) (base)[((ird)-1) + 2 * (((i) + 3) + (size_t)NI7 * ((j) + 3))]
/* S_GKUA(11,-3:nx+3,-3:ny+3) */
#define S_(k, i, j) o->S[((k)-1) + 11 * (((i) + 3) + (size_t)NI7 * ((j) + 3))]
#define S0_(k, i, j) o->S0[((k)-1) + 11 * (((i) + 3) + (size_t)NI7 * ((j) + 3))]
#define S00_(k, i, j) o->S00[((k)-1) + 11 * (((i) + 3) + (size_t)NI7 * ((j) + 3))]
/* (-1:nx+1,-1:ny+1) */
#define NI3 (o->nx + 3)
#define C2(a, i, j) (a)[((i) + 1) + (size_t)NI3 * ((j) + 1)]
/* (4,-1:nx+1,-1:ny+1) */
#define U4(a, m, i, j) (a)[((m)-1) + 4 * (((i) + 1) + (size_t)NI3 * ((j) + 1))]
#define Q2(a, m, i, j) (a)[((m)-1) + 2 * (((i) + 1) + (size_t)NI3 * ((j) + 1))]
/* (nx,ny) 1-based */
#define G2(a, i, j) (a)[((i)-1) + (size_t)o->nx * ((j)-1)]
/* node coordinates (-1:nx+2,-1:ny+2) */
#define NI4 (o->nx + 4)
#define XN(a, i, j) (a)[((i) + 1) + (size_t)NI4 * ((j) + 1)]

struct orc {
  int nx, ny, nvi, nvj;
  double *vi, *vj, *wi, *wj;
  double *x, *y, *x1, *y1;
  double *vol, *si, *sj, *ni1, *ni2, *nj1, *nj2;
  int nbc;
  gsis_bc *bc;
  double cfc, x1s, cpr, twall, dt, ma, gamma, re, pr, cv, cp, aoa;
  int scheme, mode, if_viscous, local_dt, ns_maxit;
  double cfl, dtmax, dtmin, ns_tol;
  /* state */
  double *F, *F0, *F00;
  double *S, *S0, *S00;
  double *CF, *U, *U0, *U2;
  double *taonsi_g, *taonsj_g, *qi_g, *qj_g;
  double *taonsi_in, *taonsj_in, *qi_in, *qj_in;
  double *swall; /* per wall/farfield BC: (10, nk) packed in BC order */
  int *swall_off;
  /* primitives of comput_duvtpckw (src/sub_NS_singlegrid.f90:121-157) */
  double *d, *uu, *vv, *T, *p, *amu;
  /* synthetic NS work */
  double *res, *du, *dtl, *lci, *lcj, *lvi, *lvj;
  int kstep, ns_iter;
  double tt;
  /* MassCorrect (src/GKUA_DETFP.f90:153-172): on/off and Init_mass */
  int mass_correct;
  double init_mass;
};

/* shared internals (gsis_oracle.c) */
int orc_comput_prims(orc_t *o);
void orc_get_viscous(orc_t *o);
void orc_ns_face_stress(const orc_t *o, int dir, int i, int j, double *t11, double *t22, double *t12,
                        double *t33, double *qx, double *qy, double *u1o, double *v1o);
#endif
