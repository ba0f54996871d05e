// gsis_synthetic.hip -- synthetic NS solve of mode M1 (src/Hybrid_Sub.f90:45-153).
#include "gsis_synthetic.h"

namespace gsis {

int synthetic_init(Dev &, const gsis_params &, SyntheticState &st) {
  st.on = false;
  return GSIS_EUNSUPPORTED;
}
int synthetic_step(const Dev &, SyntheticState &, hipStream_t, int *) { return GSIS_EUNSUPPORTED; }
void synthetic_free(SyntheticState &) {}
const char *synthetic_error() { return "mode M1 (synthetic NS solve) is not built in this library"; }

}  // namespace gsis
