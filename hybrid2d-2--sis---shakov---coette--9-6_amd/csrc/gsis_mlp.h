// gsis_mlp.h -- the learned HoT closure: an MLP of 8 inputs, 4 hidden
// layers of width 128 with SiLU and 5 outputs, BF16 weights and activations
// with FP32 accumulation on the MFMA units (gsis_mlp.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/gsis_capi.h"

namespace gsis {

constexpr int kMlpIn = GSIS_MLP_NIN, kMlpOut = GSIS_MLP_NOUT, kMlpW = GSIS_MLP_WIDTH;
constexpr int kMlpLayers = GSIS_MLP_HIDDEN + 1;  // weight matrices
constexpr int kMlpOutPad = 32;                   // output rows padded to one MFMA block
constexpr int kMlpBatchMax = 8192;               // samples per pass

// rows (outputs) and columns (inputs) of layer l, and its padded row count
__host__ __device__ inline int mlp_rows(int l) { return l == kMlpLayers - 1 ? kMlpOut : kMlpW; }
__host__ __device__ inline int mlp_cols(int l) { return l == 0 ? kMlpIn : kMlpW; }
__host__ __device__ inline int mlp_rows_pad(int l) { return l == kMlpLayers - 1 ? kMlpOutPad : kMlpW; }
constexpr int kMlpInPad = 32;  // input columns padded to one MFMA block
__host__ __device__ inline int mlp_cols_pad(int l) { return l == 0 ? kMlpInPad : kMlpW; }

// Device state of one network: FP32 master parameters [W1 b1 W2 b2 ... W5 b5]
// (W row-major, outputs by inputs), their BF16 copies for the MFMA kernels in
// both orientations (W and its transpose, rows and columns padded with zeros
// to whole 32 x 32 blocks), Adam moments, and the work buffers of one batch.
struct MlpDev {
  float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;  // [nparams]
  unsigned short *w[kMlpLayers] = {}, *wt[kMlpLayers] = {};
  float scale[2 * kMlpIn + 2 * kMlpOut] = {};  // in mean, in std, out mean, out std (host copy)
  float *dscale = nullptr;                     // the same on the device
  // batch buffers: inputs X0 (+ transpose), per layer pre-activations Z (fp32),
  // activations H (bf16, + transpose), the GEMM output C (fp32), gradients
  unsigned short *x0 = nullptr, *x0t = nullptr;
  float *z[kMlpLayers] = {};
  unsigned short *h[kMlpLayers] = {}, *ht[kMlpLayers] = {};
  float *c = nullptr;
  unsigned short *dz = nullptr, *dzt = nullptr;
  float *dh = nullptr;
  float *xin = nullptr, *yin = nullptr;  // staged samples (raw)
  int *idx = nullptr;                    // batch sample indices
  double *loss = nullptr;
  size_t cap = 0;  // samples the staged buffers hold
  hipStream_t s = nullptr;
  int step = 0;  // Adam steps taken
  std::vector<void *> allocs;
};

int mlp_nparams();
// offset of W_l and b_l in the parameter vector
int mlp_woff(int l);
int mlp_boff(int l);

// forward pass of n device-resident raw inputs x [n x 8] into raw outputs
// y [n x 5], enqueued on stream s (no synchronisation)
int mlp_forward_dev(MlpDev &d, const float *x, int n, float *y, hipStream_t s);

}  // namespace gsis

struct gsis_mlp {
  gsis::MlpDev d;
};
