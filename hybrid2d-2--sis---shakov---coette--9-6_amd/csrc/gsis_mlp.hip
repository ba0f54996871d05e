// gsis_mlp.hip -- hand-written gfx950 kernels of the learned HoT closure:
// forward, backward and Adam for an 8-128-128-128-128-5 SiLU network with BF16
// weights and activations and FP32 accumulation on the matrix cores.
//
// Every product of the network is one shape of GEMM, C[M x N] = A[M x K] B[K x N]
// with both operands stored K-contiguous (A row-major, B as its transpose Bt
// row-major), so that each lane's MFMA operand is one 8-byte load:
//   forward   Z_l  = H_{l-1} W_l^T    A = H_{l-1} [batch x in],  Bt = W_l   [out x in]
//   backward  dH_l = dZ_{l+1} W_{l+1} A = dZ_{l+1} [batch x out], Bt = W^T [in x out]
//             dW_l = dZ_l^T H_{l-1}   A = dZ_l^T [out x batch],  Bt = H^T   [in x batch]
// so the forward pass stores each activation in both orientations, and the
// backward pass each gradient.  One wavefront computes one 32 x 32 block of C
// with v_mfma_f32_32x32x8_bf16 (A: lane l holds row l%32, k = 4(l/32)..+3;
// B likewise for column l%32; C: lane l holds column l%32 and rows
// 8(r/4) + 4(l/32) + r%4 of its 16 accumulators r).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "gsis_mlp.h"

namespace gsis {

void set_last_error(const std::string &msg);

namespace {
int mfail(int code, const std::string &msg) {
  set_last_error(msg);
  return code;
}
}  // namespace

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

int mlp_woff(int l) {
  int o = 0;
  for (int k = 0; k < l; k++) o += mlp_rows(k) * mlp_cols(k) + mlp_rows(k);
  return o;
}
int mlp_boff(int l) { return mlp_woff(l) + mlp_rows(l) * mlp_cols(l); }
int mlp_nparams() { return mlp_woff(kMlpLayers); }

// FP32 -> BF16, round to nearest even
__host__ __device__ inline unsigned short to_bf16(float f) {
  unsigned u;
  memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

__device__ inline float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ inline float silu_grad(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// C[M x N] (fp32, leading dimension ldc) = A[M x K] Bt[N x K]^T; M, N multiples
// of 32, K of 8.  One wavefront per 32 x 32 block, four per workgroup.
__global__ void __launch_bounds__(256) k_gemm_nt(const unsigned short *A, int lda, const unsigned short *Bt, int ldb,
                                                 float *C, int ldc, int M, int N, int K) {
  const int lane = threadIdx.x & 63;
  const int blk = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int nbn = N / 32;
  if (blk >= (M / 32) * nbn) return;
  const int m0 = (blk / nbn) * 32, n0 = (blk % nbn) * 32;
  const int r = lane & 31, kq = (lane >> 5) * 4;
  const bf16x4 *a = reinterpret_cast<const bf16x4 *>(A + (size_t)(m0 + r) * lda + kq);
  const bf16x4 *b = reinterpret_cast<const bf16x4 *>(Bt + (size_t)(n0 + r) * ldb + kq);
  f32x16 acc = {0};
#pragma unroll 8
  for (int k = 0; k < K; k += 8) acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a[k / 4], b[k / 4], acc, 0, 0, 0);
  for (int q = 0; q < 16; q++) {
    const int row = m0 + (q >> 2) * 8 + (lane >> 5) * 4 + (q & 3);
    C[(size_t)row * ldc + n0 + r] = acc[q];
  }
}

// the batch's inputs, standardised, as BF16 X0 [Mp x 32] and X0^T [32 x Mp];
// rows past n and columns past 8 are zero
__global__ void k_mlp_input(const float *x, const int *idx, int n, int Mp, const float *sc, unsigned short *x0,
                            unsigned short *x0t) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= Mp * kMlpInPad) return;
  const int row = t / kMlpInPad, f = t % kMlpInPad;
  float v = 0.0f;
  if (row < n && f < kMlpIn) {
    const int sidx = idx ? idx[row] : row;
    v = (x[(size_t)sidx * kMlpIn + f] - sc[f]) / sc[kMlpIn + f];
  }
  const unsigned short b = to_bf16(v);
  x0[(size_t)row * kMlpInPad + f] = b;
  x0t[(size_t)f * Mp + row] = b;
}

// Z = C + bias; H = SiLU(Z) in BF16, both orientations (hidden layers)
__global__ void k_mlp_bias_silu(const float *c, const float *bias, int Mp, int N, float *z, unsigned short *h,
                                unsigned short *ht) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= Mp * N) return;
  const int row = t / N, col = t % N;
  const float v = c[t] + bias[col];
  z[t] = v;
  const unsigned short b = to_bf16(silu(v));
  h[t] = b;
  ht[(size_t)col * Mp + row] = b;
}

// the output layer: Y = C + bias on the 5 outputs (standardised units)
__global__ void k_mlp_out(const float *c, const float *bias, int n, float *y) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * kMlpOut) return;
  const int row = t / kMlpOut, col = t % kMlpOut;
  y[t] = c[(size_t)row * kMlpOutPad + col] + bias[col];
}

// de-standardised outputs of a forward pass
__global__ void k_mlp_unscale(float *y, int n, const float *sc) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * kMlpOut) return;
  const int col = t % kMlpOut;
  y[t] = y[t] * sc[2 * kMlpIn + kMlpOut + col] + sc[2 * kMlpIn + col];
}

// loss = mean (Y - T)^2 over the n x 5 standardised outputs; dY = 2 (Y - T)/(5n)
// as BF16 dY [Mp x 32] and dY^T [32 x Mp] (zero outside the n x 5 block), and
// the output bias gradient.  One workgroup of 256 threads.
__global__ void k_mlp_loss(const float *yhat, const float *t, const int *idx, int n, int Mp, const float *sc,
                           unsigned short *dy, unsigned short *dyt, float *gb, double *loss) {
  __shared__ double red[256];
  __shared__ float bsum[256][kMlpOut];
  double acc = 0.0;
  float bs[kMlpOut] = {0, 0, 0, 0, 0};
  for (int e = threadIdx.x; e < Mp * kMlpOutPad; e += blockDim.x) {
    const int row = e / kMlpOutPad, col = e % kMlpOutPad;
    float g = 0.0f;
    if (row < n && col < kMlpOut) {
      const int sidx = idx ? idx[row] : row;
      const float tt = (t[(size_t)sidx * kMlpOut + col] - sc[2 * kMlpIn + col]) / sc[2 * kMlpIn + kMlpOut + col];
      const float r = yhat[(size_t)row * kMlpOut + col] - tt;
      acc += (double)r * r;
      g = 2.0f * r / (float)(kMlpOut * n);
      bs[col] += g;
    }
    const unsigned short b = to_bf16(g);
    dy[e] = b;
    dyt[(size_t)col * Mp + row] = b;
  }
  red[threadIdx.x] = acc;
  for (int k = 0; k < kMlpOut; k++) bsum[threadIdx.x][k] = bs[k];
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      red[threadIdx.x] += red[threadIdx.x + s];
      for (int k = 0; k < kMlpOut; k++) bsum[threadIdx.x][k] += bsum[threadIdx.x + s][k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    *loss = red[0] / (double)(kMlpOut * n);
    for (int k = 0; k < kMlpOut; k++) gb[k] = bsum[0][k];
  }
}

// dZ = dH * SiLU'(Z) as BF16, both orientations; the rows of samples past n
// are zero.  The bias gradient (column sums of dZ) goes to gb, one column per
// thread of a second pass (k_mlp_colsum).
__global__ void k_mlp_silu_back(const float *dh, const float *z, int n, int Mp, int N, unsigned short *dz,
                                unsigned short *dzt, float *dzf) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= Mp * N) return;
  const int row = t / N, col = t % N;
  const float g = row < n ? dh[t] * silu_grad(z[t]) : 0.0f;
  dzf[t] = g;
  const unsigned short b = to_bf16(g);
  dz[t] = b;
  dzt[(size_t)col * Mp + row] = b;
}

// column sums of a [Mp x N]: one workgroup of 256 per column
__global__ void __launch_bounds__(256) k_mlp_colsum(const float *a, int Mp, int N, float *out) {
  __shared__ float red[256];
  const int col = blockIdx.x;
  float s = 0.0f;
  for (int row = threadIdx.x; row < Mp; row += 256) s += a[(size_t)row * N + col];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[col] = red[0];
}

// gradient of W_l [rows x cols] from the GEMM output [rows_pad x ldc]
__global__ void k_mlp_take(const float *c, int ldc, int rows, int cols, float *gw) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < rows * cols) gw[t] = c[(size_t)(t / cols) * ldc + t % cols];
}

// BF16 copies of W_l: W [rows_pad x cols_pad] and W^T [cols_pad x rows_pad],
// the padding zero
__global__ void k_mlp_pack(const float *w, int rows, int rows_pad, int cols, int cols_pad, unsigned short *wb,
                           unsigned short *wtb) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows_pad * cols_pad) return;
  const int r = t / cols_pad, c = t % cols_pad;
  const unsigned short b = r < rows && c < cols ? to_bf16(w[(size_t)r * cols + c]) : (unsigned short)0;
  wb[t] = b;
  wtb[(size_t)c * rows_pad + r] = b;
}

__global__ void k_mlp_adam(float *p, const float *g, float *m1, float *m2, int n, float lr, float b1, float b2,
                           float eps, float c1, float c2) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float gt = g[t];
  const float a = b1 * m1[t] + (1.0f - b1) * gt;
  const float b = b2 * m2[t] + (1.0f - b2) * gt * gt;
  m1[t] = a;
  m2[t] = b;
  p[t] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

namespace {

unsigned nblk(size_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

void gemm(const unsigned short *A, int lda, const unsigned short *Bt, int ldb, float *C, int ldc, int M, int N, int K,
          hipStream_t s) {
  const int blocks = (M / 32) * (N / 32);
  hipLaunchKernelGGL(k_gemm_nt, dim3((blocks + 3) / 4), dim3(256), 0, s, A, lda, Bt, ldb, C, ldc, M, N, K);
}

template <typename P>
int malloc_d(MlpDev &d, P **p, size_t bytes) {
  void *q = nullptr;
  if (hipMalloc(&q, std::max(bytes, (size_t)8)) != hipSuccess) return mfail(GSIS_EHIP, "mlp: hipMalloc failed");
  if (hipMemset(q, 0, std::max(bytes, (size_t)8)) != hipSuccess) return mfail(GSIS_EHIP, "mlp: hipMemset failed");
  d.allocs.push_back(q);
  *p = reinterpret_cast<P *>(q);
  return 0;
}

void pack(MlpDev &d) {
  for (int l = 0; l < kMlpLayers; l++) {
    const int rp = mlp_rows_pad(l), cp = mlp_cols_pad(l);
    hipLaunchKernelGGL(k_mlp_pack, dim3(nblk((size_t)rp * cp)), dim3(256), 0, d.s, d.p + mlp_woff(l), mlp_rows(l), rp,
                       mlp_cols(l), cp, d.w[l], d.wt[l]);
  }
}

// forward pass of Mp rows (inputs already in x0/x0t): standardised outputs of
// the first n rows into yhat [n x 5]
void forward(MlpDev &d, int n, int Mp, float *yhat) {
  const unsigned short *a = d.x0;
  int ka = kMlpInPad;
  for (int l = 0; l < kMlpLayers; l++) {
    const int rp = mlp_rows_pad(l);
    gemm(a, ka, d.w[l], ka, d.c, rp, Mp, rp, ka, d.s);
    if (l < kMlpLayers - 1) {
      hipLaunchKernelGGL(k_mlp_bias_silu, dim3(nblk((size_t)Mp * rp)), dim3(256), 0, d.s, d.c, d.p + mlp_boff(l), Mp,
                         rp, d.z[l], d.h[l], d.ht[l]);
      a = d.h[l];
      ka = rp;
    } else {
      hipLaunchKernelGGL(k_mlp_out, dim3(nblk((size_t)n * kMlpOut)), dim3(256), 0, d.s, d.c, d.p + mlp_boff(l), n,
                         yhat);
    }
  }
}

// loss and gradient of one batch (x0/x0t staged, n rows of Mp), into d.g
void backward(MlpDev &d, const float *t, const int *idx, int n, int Mp, float *yhat) {
  forward(d, n, Mp, yhat);
  const int L = kMlpLayers - 1;
  // the output layer: dY, its bias gradient, dW5 = dY^T H4
  hipLaunchKernelGGL(k_mlp_loss, dim3(1), dim3(256), 0, d.s, yhat, t, idx, n, Mp, d.dscale, d.dz, d.dzt,
                     d.g + mlp_boff(L), d.loss);
  for (int l = L; l >= 0; l--) {
    const int rp = mlp_rows_pad(l), c = mlp_cols(l), cp = mlp_cols_pad(l);
    // dW_l = dZ_l^T [rp x Mp] . H_{l-1}^T [cp x Mp]^T
    const unsigned short *hprev_t = l > 0 ? d.ht[l - 1] : d.x0t;
    gemm(d.dzt, Mp, hprev_t, Mp, d.c, cp, rp, cp, Mp, d.s);
    hipLaunchKernelGGL(k_mlp_take, dim3(nblk((size_t)mlp_rows(l) * c)), dim3(256), 0, d.s, d.c, cp, mlp_rows(l), c,
                       d.g + mlp_woff(l));
    if (l == 0) break;
    // dH_{l-1} = dZ_l [Mp x rp] . (W_l^T [c x rp])^T, then dZ_{l-1}
    gemm(d.dz, rp, d.wt[l], rp, d.dh, c, Mp, c, rp, d.s);
    float *dzf = d.c;  // the fp32 dZ for the bias gradient (the GEMM output is consumed)
    hipLaunchKernelGGL(k_mlp_silu_back, dim3(nblk((size_t)Mp * c)), dim3(256), 0, d.s, d.dh, d.z[l - 1], n, Mp, c,
                       d.dz, d.dzt, dzf);
    hipLaunchKernelGGL(k_mlp_colsum, dim3(c), dim3(256), 0, d.s, dzf, Mp, c, d.g + mlp_boff(l - 1));
  }
}

int ensure_cap(MlpDev &d, size_t n) {
  if (n <= d.cap) return 0;
  for (void *q : {(void *)d.xin, (void *)d.yin, (void *)d.idx}) {
    if (q) {
      (void)hipFree(q);
      d.allocs.erase(std::remove(d.allocs.begin(), d.allocs.end(), q), d.allocs.end());
    }
  }
  int rc;
  if ((rc = malloc_d(d, &d.xin, n * kMlpIn * 4)) || (rc = malloc_d(d, &d.yin, n * kMlpOut * 4)) ||
      (rc = malloc_d(d, &d.idx, n * 4)))
    return rc;
  d.cap = n;
  return 0;
}

int sync(MlpDev &d) {
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(d.s);
  if (e != hipSuccess) return mfail(GSIS_EHIP, std::string("mlp: ") + hipGetErrorString(e));
  return 0;
}

}  // namespace

// Standardised forward pass of n device-resident raw inputs (x: n x 8) into
// y (n x 5, raw units), on the network's stream.  Used by the closure in the
// synthetic solve.
int mlp_forward_dev(MlpDev &d, const float *x, int n, float *y, hipStream_t s) {
  const hipStream_t keep = d.s;
  d.s = s;
  for (int b0 = 0; b0 < n; b0 += kMlpBatchMax) {
    const int nb = std::min(kMlpBatchMax, n - b0), Mp = (nb + 63) / 64 * 64;
    hipLaunchKernelGGL(k_mlp_input, dim3(nblk((size_t)Mp * kMlpInPad)), dim3(256), 0, s, x + (size_t)b0 * kMlpIn,
                       (const int *)nullptr, nb, Mp, d.dscale, d.x0, d.x0t);
    forward(d, nb, Mp, y + (size_t)b0 * kMlpOut);
    hipLaunchKernelGGL(k_mlp_unscale, dim3(nblk((size_t)nb * kMlpOut)), dim3(256), 0, s, y + (size_t)b0 * kMlpOut,
                       nb, d.dscale);
  }
  d.s = keep;
  return 0;
}

}  // namespace gsis

using namespace gsis;

extern "C" {

int gsis_mlp_nparams(void) { return mlp_nparams(); }

int gsis_mlp_destroy(gsis_mlp *m) {
  if (!m) return 0;
  if (m->d.s) (void)hipStreamSynchronize(m->d.s);
  for (void *q : m->d.allocs) (void)hipFree(q);
  if (m->d.s) (void)hipStreamDestroy(m->d.s);
  delete m;
  return 0;
}

int gsis_mlp_create(gsis_mlp **out) {
  if (!out) return mfail(GSIS_EARG, "out is NULL");
  *out = nullptr;
  gsis_mlp *m = new gsis_mlp();
  MlpDev &d = m->d;
  auto bail = [&](int rc) {
    gsis_mlp_destroy(m);
    return rc;
  };
  if (hipStreamCreateWithFlags(&d.s, hipStreamNonBlocking) != hipSuccess)
    return bail(mfail(GSIS_EHIP, "mlp: hipStreamCreate failed"));
  const size_t np = mlp_nparams();
  int rc;
  if ((rc = malloc_d(d, &d.p, np * 4)) || (rc = malloc_d(d, &d.g, np * 4)) || (rc = malloc_d(d, &d.m1, np * 4)) ||
      (rc = malloc_d(d, &d.m2, np * 4)) || (rc = malloc_d(d, &d.dscale, sizeof(d.scale))))
    return bail(rc);
  for (int l = 0; l < kMlpLayers; l++) {
    const size_t e = (size_t)mlp_rows_pad(l) * mlp_cols_pad(l);
    if ((rc = malloc_d(d, &d.w[l], e * 2)) || (rc = malloc_d(d, &d.wt[l], e * 2))) return bail(rc);
  }
  const size_t B = kMlpBatchMax;
  if ((rc = malloc_d(d, &d.x0, B * kMlpInPad * 2)) || (rc = malloc_d(d, &d.x0t, B * kMlpInPad * 2)) ||
      (rc = malloc_d(d, &d.c, B * kMlpW * 4)) || (rc = malloc_d(d, &d.dz, B * kMlpW * 2)) ||
      (rc = malloc_d(d, &d.dzt, B * kMlpW * 2)) || (rc = malloc_d(d, &d.dh, B * kMlpW * 4)) ||
      (rc = malloc_d(d, &d.loss, 8)))
    return bail(rc);
  for (int l = 0; l < kMlpLayers - 1; l++)
    if ((rc = malloc_d(d, &d.z[l], B * kMlpW * 4)) || (rc = malloc_d(d, &d.h[l], B * kMlpW * 2)) ||
        (rc = malloc_d(d, &d.ht[l], B * kMlpW * 2)))
      return bail(rc);
  for (int k = 0; k < kMlpIn; k++) d.scale[kMlpIn + k] = 1.0f;
  for (int k = 0; k < kMlpOut; k++) d.scale[2 * kMlpIn + kMlpOut + k] = 1.0f;
  if (hipMemcpy(d.dscale, d.scale, sizeof(d.scale), hipMemcpyHostToDevice) != hipSuccess)
    return bail(mfail(GSIS_EHIP, "mlp: scaling upload failed"));
  pack(d);
  if ((rc = sync(d))) return bail(rc);
  *out = m;
  return 0;
}

int gsis_mlp_set_params(gsis_mlp *m, const float *params) {
  if (!m || !params) return mfail(GSIS_EARG, "NULL argument");
  MlpDev &d = m->d;
  if (hipMemcpyAsync(d.p, params, (size_t)mlp_nparams() * 4, hipMemcpyHostToDevice, d.s) != hipSuccess)
    return mfail(GSIS_EHIP, "mlp: parameter upload failed");
  (void)hipMemsetAsync(d.m1, 0, (size_t)mlp_nparams() * 4, d.s);
  (void)hipMemsetAsync(d.m2, 0, (size_t)mlp_nparams() * 4, d.s);
  d.step = 0;
  pack(d);
  return sync(d);
}

int gsis_mlp_get_params(gsis_mlp *m, float *params) {
  if (!m || !params) return mfail(GSIS_EARG, "NULL argument");
  MlpDev &d = m->d;
  if (hipMemcpyAsync(params, d.p, (size_t)mlp_nparams() * 4, hipMemcpyDeviceToHost, d.s) != hipSuccess)
    return mfail(GSIS_EHIP, "mlp: parameter download failed");
  return sync(d);
}

int gsis_mlp_set_scaling(gsis_mlp *m, const float *scaling) {
  if (!m || !scaling) return mfail(GSIS_EARG, "NULL argument");
  MlpDev &d = m->d;
  for (int k = 0; k < 2 * kMlpIn + 2 * kMlpOut; k++) {
    const bool is_std = (k >= kMlpIn && k < 2 * kMlpIn) || k >= 2 * kMlpIn + kMlpOut;
    if (is_std && !(scaling[k] > 0.0f)) return mfail(GSIS_EARG, "mlp: standard deviations must be positive");
    d.scale[k] = scaling[k];
  }
  if (hipMemcpyAsync(d.dscale, d.scale, sizeof(d.scale), hipMemcpyHostToDevice, d.s) != hipSuccess)
    return mfail(GSIS_EHIP, "mlp: scaling upload failed");
  return sync(d);
}

int gsis_mlp_forward(gsis_mlp *m, const float *x, int32_t n, float *y) {
  if (!m || !x || !y || n < 0) return mfail(GSIS_EARG, "bad argument");
  if (n == 0) return 0;
  MlpDev &d = m->d;
  int rc = ensure_cap(d, (size_t)n);
  if (rc) return rc;
  if (hipMemcpyAsync(d.xin, x, (size_t)n * kMlpIn * 4, hipMemcpyHostToDevice, d.s) != hipSuccess)
    return mfail(GSIS_EHIP, "mlp: input upload failed");
  mlp_forward_dev(d, d.xin, n, d.yin, d.s);
  if (hipMemcpyAsync(y, d.yin, (size_t)n * kMlpOut * 4, hipMemcpyDeviceToHost, d.s) != hipSuccess)
    return mfail(GSIS_EHIP, "mlp: output download failed");
  return sync(d);
}

int gsis_mlp_loss_grad(gsis_mlp *m, const float *x, const float *y, int32_t n, double *loss, float *grad) {
  if (!m || !x || !y || n < 1 || n > kMlpBatchMax) return mfail(GSIS_EARG, "bad argument (1 <= n <= 8192)");
  MlpDev &d = m->d;
  int rc = ensure_cap(d, (size_t)n);
  if (rc) return rc;
  if (hipMemcpyAsync(d.xin, x, (size_t)n * kMlpIn * 4, hipMemcpyHostToDevice, d.s) != hipSuccess ||
      hipMemcpyAsync(d.yin, y, (size_t)n * kMlpOut * 4, hipMemcpyHostToDevice, d.s) != hipSuccess)
    return mfail(GSIS_EHIP, "mlp: sample upload failed");
  const int Mp = (n + 63) / 64 * 64;
  float *yhat = d.dh;  // n x 5 standardised outputs, before dh is needed
  hipLaunchKernelGGL(k_mlp_input, dim3(nblk((size_t)Mp * kMlpInPad)), dim3(256), 0, d.s, d.xin, (const int *)nullptr, n,
                     Mp, d.dscale, d.x0, d.x0t);
  // the forward pass writes yhat into the first n*5 floats of dh, which the
  // backward pass overwrites only after the loss kernel has read it
  backward(d, d.yin, nullptr, n, Mp, yhat);
  if (loss && hipMemcpyAsync(loss, d.loss, 8, hipMemcpyDeviceToHost, d.s) != hipSuccess)
    return mfail(GSIS_EHIP, "mlp: loss download failed");
  if (grad && hipMemcpyAsync(grad, d.g, (size_t)mlp_nparams() * 4, hipMemcpyDeviceToHost, d.s) != hipSuccess)
    return mfail(GSIS_EHIP, "mlp: gradient download failed");
  return sync(d);
}

int gsis_mlp_train(gsis_mlp *m, const float *x, const float *y, int32_t n, int32_t epochs, int32_t batch, float lr,
                   uint64_t seed, double *loss_per_epoch) {
  if (!m || !x || !y || n < 1 || epochs < 0 || batch < 64 || batch > kMlpBatchMax || batch % 64)
    return mfail(GSIS_EARG, "bad argument (batch a multiple of 64 in 64..8192)");
  MlpDev &d = m->d;
  int rc = ensure_cap(d, (size_t)n);
  if (rc) return rc;
  if (hipMemcpyAsync(d.xin, x, (size_t)n * kMlpIn * 4, hipMemcpyHostToDevice, d.s) != hipSuccess ||
      hipMemcpyAsync(d.yin, y, (size_t)n * kMlpOut * 4, hipMemcpyHostToDevice, d.s) != hipSuccess)
    return mfail(GSIS_EHIP, "mlp: sample upload failed");
  std::vector<int> perm(n);
  for (int k = 0; k < n; k++) perm[k] = k;
  std::mt19937_64 rng(seed);
  const int np = mlp_nparams();
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  std::vector<double> lossb;
  float *yhat = nullptr;
  if ((rc = malloc_d(d, &yhat, (size_t)batch * kMlpOut * 4))) return rc;
  for (int ep = 0; ep < epochs; ep++) {
    std::shuffle(perm.begin(), perm.end(), rng);
    if (hipMemcpyAsync(d.idx, perm.data(), (size_t)n * 4, hipMemcpyHostToDevice, d.s) != hipSuccess)
      return mfail(GSIS_EHIP, "mlp: index upload failed");
    const int nb = n / batch;  // full batches only
    double sum = 0.0;
    for (int b = 0; b < std::max(nb, 1); b++) {
      const int cnt = nb > 0 ? batch : n, Mp = (cnt + 63) / 64 * 64;
      const int *idx = d.idx + (size_t)b * batch;
      hipLaunchKernelGGL(k_mlp_input, dim3(nblk((size_t)Mp * kMlpInPad)), dim3(256), 0, d.s, d.xin, idx, cnt, Mp,
                         d.dscale, d.x0, d.x0t);
      backward(d, d.yin, idx, cnt, Mp, yhat);
      d.step += 1;
      const float c1 = 1.0f - std::pow(b1, (float)d.step), c2 = 1.0f - std::pow(b2, (float)d.step);
      hipLaunchKernelGGL(k_mlp_adam, dim3(nblk(np)), dim3(256), 0, d.s, d.p, d.g, d.m1, d.m2, np, lr, b1, b2, eps, c1,
                         c2);
      pack(d);
      double l = 0.0;
      if (hipMemcpyAsync(&l, d.loss, 8, hipMemcpyDeviceToHost, d.s) != hipSuccess ||
          hipStreamSynchronize(d.s) != hipSuccess)
        return mfail(GSIS_EHIP, "mlp: loss download failed");
      sum += l;
    }
    lossb.push_back(sum / std::max(nb, 1));
  }
  if ((rc = sync(d))) return rc;
  if (loss_per_epoch) std::copy(lossb.begin(), lossb.end(), loss_per_epoch);
  (void)hipFree(yhat);
  d.allocs.erase(std::remove(d.allocs.begin(), d.allocs.end(), (void *)yhat), d.allocs.end());
  return 0;
}

}  // extern "C"
