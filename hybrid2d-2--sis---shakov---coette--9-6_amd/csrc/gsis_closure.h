// gsis_closure.h -- the learned HoT closure in the synthetic NS solve
// (gsis_closure.hip): closure faces, network inputs and targets, acceptance
// test and HoT updates.
#pragma once
#include "gsis_device.h"
#include "gsis_mlp.h"

namespace gsis {

// Device buffers of the closure of one handle.
struct ClosureBufs {
  int nfi = 0, nf = 0;       // closure i-faces, all closure faces
  double *saved = nullptr;   // [12][ncell]: taoi, taoj, qi, qj of Convert_NS
  float *x = nullptr;        // [nf][8] network inputs
  float *ybase = nullptr;    // [nf][5] network at the Convert_NS state
  float *y = nullptr;        // [nf][5] network at the current state
  float *t = nullptr;        // [nf][5] DVM HoT of Convert_NS
  int *accept = nullptr;     // [nf] 1 where the closure is used
  int *count = nullptr;      // accepted faces of the last step
  double *usave = nullptr;   // [4][ncell]: U at the start of the inner loop
  int refreshes = 0;         // HoT updates of the last step
  int rejected = 0;          // the last step fell back to the DVM HoT
  bool replace = false;      // the network's HoT replace the DVM HoT (warm-up) instead of updating them
  void *mem = nullptr;
  size_t bytes = 0;
};

// number of closure faces (nfi of them i-faces)
int closure_faces(const Dev &d, int *nfi);
float closure_log10_kn(const Dev &d);
int closure_alloc(const Dev &d, ClosureBufs &b);
void closure_free(ClosureBufs &b);
// network inputs of the closure faces from the Convert_NS state (prim) or the
// current synthetic state (U)
void launch_cl_features(const Dev &d, int nfi, int nf, bool from_prim, float *x, hipStream_t s);
// the five DVM HoT of the closure faces
void launch_cl_targets(const Dev &d, int nfi, int nf, float *y, hipStream_t s);
// start of a synthetic solve: keep the DVM HoT and the starting state U,
// evaluate the network at the Convert_NS state, decide which faces use the
// closure (with replace, every face with a finite prediction, whose HoT are
// set to the prediction at once)
int closure_begin(const Dev &d, ClosureBufs &b, MlpDev &net, float gate, bool replace, hipStream_t s);
// between inner-loop chunks: HoT of the accepted faces from the current state
int closure_refresh(const Dev &d, ClosureBufs &b, MlpDev &net, hipStream_t s);
// end of the synthetic solve: the DVM HoT back in place
int closure_end(const Dev &d, ClosureBufs &b, hipStream_t s);
// the closure rejected: the DVM HoT and the inner loop's starting state U back
int closure_reject(const Dev &d, ClosureBufs &b, hipStream_t s);

}  // namespace gsis
