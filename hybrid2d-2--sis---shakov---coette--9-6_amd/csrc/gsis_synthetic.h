// gsis_synthetic.h -- synthetic NS solve of mode M1 (src/Hybrid_Sub.f90:45-153).
#pragma once
#include "../../include/gsis_capi.h"
#include "gsis_device.h"

namespace gsis {

struct SyntheticState {
  bool on = false;
  double tol = 0.0;
  int maxit = 100000;
  int local_dt = 1, if_viscous = 1;
  double cfl = 20.0, dtmax = 10.0, dtmin = 1e-7;
  double *s00 = nullptr;               // [4][ncell]  S00_GKUA(1:4)
  double *res = nullptr, *du = nullptr; // [4][ncell]  B%Res, B%dU
  double *dtl = nullptr, *lci = nullptr, *lcj = nullptr, *lvi = nullptr, *lvj = nullptr;  // [ncell]
  double *fi = nullptr, *fj = nullptr; // [4][ncell]  face fluxes of Residual_FVM
  int *iters = nullptr;                // [0]: inner iterations of the last step, [1]: T<0.1 flag
  void *block = nullptr;               // single allocation holding all of the above
};

int synthetic_init(Dev &d, const gsis_params &p, SyntheticState &st);
int synthetic_step(const Dev &d, SyntheticState &st, hipStream_t s, int *iters);
int synthetic_read_iters(SyntheticState &st, hipStream_t s, int *iters, int *bad_t);
void synthetic_free(SyntheticState &st);
const char *synthetic_error();

}  // namespace gsis
