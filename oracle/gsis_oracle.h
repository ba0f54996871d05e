/*
 * gsis_oracle.h -- CPU oracle of the GSIS kinetic loop.  TEST INFRASTRUCTURE ONLY:
 * loaded by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg as the
 * reference values, never by the product path.  Arrays use the reference's Fortran
 * layouts, documented in include/gsis_capi.h.
 */
#ifndef GSIS_ORACLE_H
#define GSIS_ORACLE_H
#include <stddef.h>
#include "../include/gsis_capi.h"

#ifdef __cplusplus
extern "C" {
#endif
typedef struct orc orc_t;
orc_t *orc_create(const gsis_params *p);
void orc_destroy(orc_t *o);
size_t orc_frd_len(const orc_t *o);
void orc_set_frd(orc_t *o, const double *frd);
void orc_get_frd(const orc_t *o, double *frd);
void orc_get_frd00(const orc_t *o, double *frd);
void orc_get_frd0(const orc_t *o, double *frd);
void orc_get_prims(const orc_t *o, double *out);
void orc_feq(orc_t *o);
void orc_set_macro(orc_t *o, const double *s, const double *cf, const double *u, const double *u0);
void orc_get_macro(const orc_t *o, double *s, double *s0, double *cf, double *u, double *taonsi_g,
                   double *taonsj_g, double *qi_g, double *qj_g, double *swall);
void orc_counters(const orc_t *o, int *kstep, double *tt, int *ns_iter);
void orc_set_counters(orc_t *o, int kstep, double tt);
void orc_phy_boundary(orc_t *o);
void orc_detfp(orc_t *o);
int orc_convert_ns(orc_t *o);
int orc_step(orc_t *o, int nsteps);
int orc_residual(orc_t *o, double res_rms[4], double res_max[4]);
double orc_out_mass(const orc_t *o);
void orc_set_mass_correct(orc_t *o, int on);
double orc_ns_advance(orc_t *o, int n);
void orc_ns_work(const orc_t *o, double *res, double *du, double *lc);
#ifdef __cplusplus
}
#endif
#endif
