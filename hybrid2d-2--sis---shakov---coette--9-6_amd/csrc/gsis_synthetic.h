// gsis_synthetic.h -- synthetic NS solve of mode M1 (src/Hybrid_Sub.f90:45-153).
#pragma once
#include "../../include/gsis_capi.h"
#include "gsis_device.h"

namespace gsis {

struct SyntheticState {
  bool on = false;
  void *mem = nullptr;
};

int synthetic_init(Dev &d, const gsis_params &p, SyntheticState &st);
int synthetic_step(const Dev &d, SyntheticState &st, hipStream_t s, int *iters);
void synthetic_free(SyntheticState &st);
const char *synthetic_error();

}  // namespace gsis
