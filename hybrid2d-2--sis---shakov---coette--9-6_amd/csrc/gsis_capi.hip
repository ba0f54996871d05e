// gsis_capi.hip -- host side of the C-ABI declared in include/gsis_capi.h.
//
// Owns the device state of one GSIS case and orders the kernels of one outer
// iteration exactly as Hybrid_Time_advance_LU_SGS does (src/Hybrid_Sub.f90:4-160):
//   S0<-S, kinetic BCs in input order, kinetic RHS + LU-SGS sweep, moments +
//   DETFP + DETFPwall, Convert_NS (U, HoT), [synthetic NS], counters.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gsis_capi.h"
#include "gsis_closure.h"
#include "gsis_kernels.h"
#include "gsis_synthetic.h"

using gsis::BcDev;
using gsis::Dev;

namespace {
thread_local std::string g_err;
int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}
}  // namespace

namespace gsis {
// the message gsis_last_error reports, for the entry points of other files
void set_last_error(const std::string &msg) { g_err = msg; }
}  // namespace gsis

#define HIPCHK(expr)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return fail(GSIS_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));              \
  } while (0)

struct gsis_handle {
  gsis_params p{};
  std::vector<gsis_bc> bc;
  std::vector<BcDev> bcd;
  Dev d{};
  hipStream_t s = nullptr;
  int dev = 0;  // device current at gsis_create
  int prec = 0;
  size_t tsize = 8;
  void *FG = nullptr, *FH = nullptr, *slot = nullptr, *RG = nullptr, *RH = nullptr;
  std::vector<void *> allocs;
  double *cpar = nullptr;  // per-cell constants of the fused kinetic kernel
  gsis::StripBufs strip;   // seam buffers of the column-wavefront kernel
  int *wall_list = nullptr;
  int nwall = 0;
  double *swall = nullptr;
  int swall_len = 0;
  double *resbuf = nullptr;
  int kstep = 0;
  int res_kstep = -1;  // step count after the last step whose kinetic RHS was kept
  double tt = 0.0;
  int ns_iter = 0;
  gsis::SyntheticState syn{};
  // learned HoT closure of the synthetic solve (gsis_set_closure)
  gsis_mlp *net = nullptr;
  float cl_gate = 0.0f;
  int cl_every = 0;
  gsis_mlp *warm_net = nullptr;  // gsis_set_closure_warmup: network and steps left
  int warm_steps = 0, warm_every = 0;
  gsis::ClosureBufs cl{};
  // MassCorrect (gsis_set_mass_correct): on/off, Init_mass, out_mass on the device
  bool mass_on = false;
  double init_mass = 0.0;
  double *mwork = nullptr;
  // optional kernel timing (gsis_timing): events around the kinetic kernels
  bool timing = false;
  hipEvent_t ev[8] = {};
  double t_ms[4] = {0, 0, 0, 0};
  int t_n = 0;
  bool t_pending = false;
};

namespace {

template <typename P>
int dalloc(gsis_handle *h, P **ptr, size_t bytes) {
  void *q = nullptr;
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) return fail(GSIS_EHIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  e = hipMemset(q, 0, bytes);
  if (e != hipSuccess) return fail(GSIS_EHIP, std::string("hipMemset: ") + hipGetErrorString(e));
  h->allocs.push_back(q);
  *ptr = reinterpret_cast<P *>(q);
  return 0;
}

int bc_nk(const gsis_bc &b) { return (b.face == 1 || b.face == 3) ? (b.jend - b.jst + 1) : (b.iend - b.ist + 1); }

int validate(const gsis_params *p) {
  if (!p) return fail(GSIS_EARG, "params is NULL");
  if (p->nx < 5 || p->ny < 5) return fail(GSIS_EARG, "nx and ny must be >= 5 nodes");
  if (p->nvi < 1 || p->nvj < 1 || p->nvi > gsis::kMaxV1 || p->nvj > gsis::kMaxV1)
    return fail(GSIS_EARG, "nvi/nvj out of range (1..512)");
  if (!p->vi || !p->vj || !p->wi || !p->wj || !p->x || !p->y || !p->vol || !p->si || !p->sj || !p->ni1 ||
      !p->ni2 || !p->nj1 || !p->nj2)
    return fail(GSIS_EARG, "a required array pointer is NULL");
  if (p->nbc < 0 || (p->nbc > 0 && !p->bc)) return fail(GSIS_EARG, "bad BC list");
  if (p->scheme != GSIS_SCHEME_UP1 && p->scheme != GSIS_SCHEME_NND && p->scheme != GSIS_SCHEME_WENO3)
    return fail(GSIS_EUNSUPPORTED, "kinetic scheme must be GSIS_SCHEME_UP1, GSIS_SCHEME_NND or GSIS_SCHEME_WENO3");
  if (p->precision != GSIS_PREC_FP64 && p->precision != GSIS_PREC_FP32)
    return fail(GSIS_EARG, "precision must be GSIS_PREC_FP64 or GSIS_PREC_FP32");
  if (p->mode != GSIS_MODE_M0 && p->mode != GSIS_MODE_M1) return fail(GSIS_EARG, "unknown mode");
  if (p->dt <= 0.0) return fail(GSIS_EARG, "dt must be positive");
  if (p->walk_tiles < 0 || p->kstep_show < 0) return fail(GSIS_EARG, "walk_tiles and kstep_show must be >= 0");
  if (p->split_kernels < GSIS_KERNEL_AUTO || p->split_kernels > GSIS_KERNEL_STRIP16)
    return fail(GSIS_EARG, "split_kernels must be 0 (automatic), 1 (split), 2 (tile walker) or 3 (16-column strips)");
  for (int b = 0; b < p->nbc; b++) {
    const gsis_bc &c = p->bc[b];
    bool ok = false;
    if (c.face == 1) ok = c.ist == 1 && c.iend == 1;
    if (c.face == 3) ok = c.ist == p->nx && c.iend == p->nx;
    if (c.face == 2) ok = c.jst == 1 && c.jend == 1;
    if (c.face == 4) ok = c.jst == p->ny && c.jend == p->ny;
    if (!ok) return fail(GSIS_EARG, "BC " + std::to_string(b) + ": face does not match its index range");
    if (c.ist < 1 || c.iend > p->nx || c.jst < 1 || c.jend > p->ny || c.ist > c.iend || c.jst > c.jend)
      return fail(GSIS_EARG, "BC " + std::to_string(b) + ": index range outside the block");
    // symmetry and inner faces are passive in the kinetic step: no ghost fill
    // (src/GKUA_boundary.f90:25-35) and no face override (src/GKUA_Residual.f90:189-192)
    if (c.code != GSIS_BC_WALL && c.code != GSIS_BC_FARFIELD && c.code != GSIS_BC_PERIODIC &&
        c.code != GSIS_BC_SYMMETRY && c.code < 1)
      return fail(GSIS_EUNSUPPORTED, "BC " + std::to_string(b) +
                                         ": only wall, farfield, periodic, symmetry and inner faces are implemented");
  }
  return 0;
}

template <typename T>
void phy_boundary(gsis_handle *h) {
  // GKUA_Phy_boundary (src/GKUA_boundary.f90:2-46) applies the BCs in input
  // order.  Diffuse walls read interior cells and write only their own face
  // slot, so they commute with every other BC and go in one launch; far-field
  // and periodic fills (which may share corner ghost cells) keep their order.
  T *FG = (T *)h->FG, *FH = (T *)h->FH, *slot = (T *)h->slot;
  gsis::WallSet ws{};
  for (const BcDev &b : h->bcd) {
    if (b.code != GSIS_BC_WALL) continue;
    const int nn = (b.face == 1 || b.face == 3) ? (b.jend - b.jst) : (b.iend - b.ist);
    if (nn <= 0) continue;
    if (ws.n == gsis::kMaxWalls) {
      gsis::launch_walls<T>(h->d, ws, FG, FH, slot, h->s);
      ws = gsis::WallSet{};
    }
    ws.b[ws.n] = b;
    ws.start[ws.n + 1] = ws.start[ws.n] + nn;
    ws.n++;
  }
  gsis::launch_walls<T>(h->d, ws, FG, FH, slot, h->s);
  for (const BcDev &b : h->bcd) {
    if (b.code == GSIS_BC_FARFIELD) gsis::launch_farfield<T>(h->d, b, FG, FH, h->s);
    else if (b.code == GSIS_BC_PERIODIC) gsis::launch_periodic<T>(h->d, b, FG, FH, h->s);
  }
}

template <typename T>
void moments(gsis_handle *h, bool timed = false) {
  if (timed && h->timing) (void)hipEventRecord(h->ev[4], h->s);
  gsis::launch_moments<T>(h->d, (T *)h->FG, (T *)h->FH, h->s);
  if (timed && h->timing) { (void)hipEventRecord(h->ev[5], h->s); h->t_pending = true; }
  gsis::launch_det_wall(h->d, h->wall_list, h->nwall, h->swall, h->s);
}

// the last timing event of a step: the end of the synthetic solve in M1,
// else the end of the moments kernel
hipEvent_t last_event(gsis_handle *h) { return h->ev[h->p.mode == GSIS_MODE_M1 ? 7 : 5]; }

void harvest_timing(gsis_handle *h) {
  if (!h->t_pending) return;
  float ms = 0.f;
  const int nt = h->p.mode == GSIS_MODE_M1 ? 4 : 3;
  for (int k = 0; k < nt; k++)
    if (hipEventElapsedTime(&ms, h->ev[2 * k], h->ev[2 * k + 1]) == hipSuccess) h->t_ms[k] += ms;
  h->t_n += 1;
  h->t_pending = false;
}

template <typename T>
void kinetic(gsis_handle *h) {
  if (h->timing) {
    if (h->t_pending) { (void)hipEventSynchronize(last_event(h)); harvest_timing(h); }
    (void)hipEventRecord(h->ev[0], h->s);
  }
  const int show = h->p.kstep_show;
  const int store = (show <= 0 || (h->kstep + 1) % show == 0) ? 1 : 0;
  if (store) h->res_kstep = h->kstep + 1;
  T *FG = (T *)h->FG, *FH = (T *)h->FH, *RG = (T *)h->RG, *RH = (T *)h->RH;
  const T *slot = (const T *)h->slot;
  const int sel = h->p.split_kernels;
  if (sel != GSIS_KERNEL_SPLIT) {
    // timing: ev[2]..ev[3] bracket the one-pass kernel itself
    if (h->timing) (void)hipEventRecord(h->ev[1], h->s);
    bool ok = false;
    if (sel == GSIS_KERNEL_AUTO || sel == GSIS_KERNEL_STRIP16) {
      h->strip.ev_before = h->timing ? h->ev[2] : nullptr;
      h->strip.ev_after = h->timing ? h->ev[3] : nullptr;
      ok = gsis::launch_strip<T>(h->d, h->cpar, FG, slot, RG, store, sel == GSIS_KERNEL_STRIP16 ? 16 : 8, h->strip,
                                 h->s);
    }
    if (!ok) {
      if (h->timing) (void)hipEventRecord(h->ev[2], h->s);
      ok = gsis::launch_walk<T>(h->d, h->cpar, FG, slot, RG, store, h->p.walk_tiles, h->s);
      if (ok && h->timing) (void)hipEventRecord(h->ev[3], h->s);
    }
    if (ok) {
      if (store) gsis::launch_res4<T>(h->d, RG, RH, h->s);
      return;
    }
    if (h->timing) (void)hipEventRecord(h->ev[0], h->s);
  }
  gsis::launch_rhs<T>(h->d, FG, FH, slot, RG, RH, h->s);
  if (h->timing) { (void)hipEventRecord(h->ev[1], h->s); (void)hipEventRecord(h->ev[2], h->s); }
  gsis::launch_sweep<T>(h->d, FG, FH, RG, RH, h->s);
  if (h->timing) (void)hipEventRecord(h->ev[3], h->s);
}

// The synthetic solve with the learned closure: the inner loop in launches of
// cl_every iterations, the HoT of the accepted closure faces updated from the
// network between them.  The residual check: when an updated launch ends
// with a T<0.1 stop or with max |Res_4| grown by more than kClosureGrow over
// the launch before, the step falls back to the DVM HoT and continues the
// loop from its starting state without the network.
constexpr double kClosureGrow = 1.5;

int synthetic_with_closure(gsis_handle *h, gsis_mlp *m, float gate, int every, bool replace) {
  gsis::MlpDev &net = m->d;
  int rc = gsis::synthetic_begin(h->d, h->syn, h->s);
  if (rc) return fail(rc, gsis::synthetic_error());
  if (gsis::closure_begin(h->d, h->cl, net, gate, replace, h->s)) return fail(GSIS_EHIP, "closure: start failed");
  bool on = true;
  double res_prev = 0.0;
  for (int k0 = 1;;) {
    rc = gsis::synthetic_run(h->d, h->syn, k0, on ? k0 + every - 1 : 0x7fffffff, h->s);
    if (rc) return fail(rc, gsis::synthetic_error());
    bool more = false, bad = false;
    int it = 0;
    double res = 0.0;
    rc = gsis::synthetic_state(h->syn, h->s, &more, &it, &bad, &res);
    if (rc) return fail(rc, gsis::synthetic_error());
    // the residual check, once the network's HoT are in use (replace: from
    // the first launch on); growth is judged against the launch before
    const bool net_used = replace || h->cl.refreshes > 0;
    const bool grew = k0 > 1 && !(res <= kClosureGrow * res_prev);
    if (on && net_used && (bad || grew || res != res)) {
      if (gsis::closure_reject(h->d, h->cl, h->s)) return fail(GSIS_EHIP, "closure: fallback failed");
      on = false;
      if (it >= h->syn.maxit) break;
      k0 = it + 1;
      continue;
    }
    if (!more) break;
    if (on && gsis::closure_refresh(h->d, h->cl, net, h->s)) return fail(GSIS_EHIP, "closure: HoT update failed");
    res_prev = res;
    k0 = it + 1;
  }
  if (gsis::closure_end(h->d, h->cl, h->s)) return fail(GSIS_EHIP, "closure: restoring the DVM HoT failed");
  gsis::synthetic_finish(h->d, h->syn, h->s);
  return 0;
}

// the synthetic solve of one step: with the network's HoT in place of the
// DVM's during a warm-up, with the closure's updates, or plain
int synthetic_any(gsis_handle *h) {
  if (h->warm_net && h->warm_steps > 0) {
    h->warm_steps -= 1;
    return synthetic_with_closure(h, h->warm_net, 0.0f, h->warm_every, true);
  }
  if (h->net) return synthetic_with_closure(h, h->net, h->cl_gate, h->cl_every, false);
  const int rc = gsis::synthetic_step(h->d, h->syn, h->s);
  return rc ? fail(rc, gsis::synthetic_error()) : 0;
}

template <typename T>
int one_step(gsis_handle *h) {
  gsis::launch_save_s0(h->d, h->s);
  phy_boundary<T>(h);
  kinetic<T>(h);
  moments<T>(h, true);
  // the place of the reference's commented-out call MassCorrect (src/Hybrid_Sub.f90:42)
  if (h->mass_on) gsis::launch_mass_correct(h->d, h->mwork, h->init_mass, h->s);
  gsis::launch_convert(h->d, h->s);
  if (h->p.if_viscous) gsis::launch_hot(h->d, h->s);
  if (h->p.mode == GSIS_MODE_M1) {
    if (h->timing) (void)hipEventRecord(h->ev[6], h->s);
    int rc = synthetic_any(h);
    if (rc) return rc;
    if (h->timing) (void)hipEventRecord(h->ev[7], h->s);
  }
  HIPCHK(hipGetLastError());
  h->tt += h->p.dt;
  h->kstep += 1;
  return 0;
}

int check_err_flag(gsis_handle *h) {
  int e[2] = {0, 0};
  HIPCHK(hipMemcpyAsync(e, h->d.err, sizeof(e), hipMemcpyDeviceToHost, h->s));
  HIPCHK(hipStreamSynchronize(h->s));
  if (e[0] & 1) return fail(GSIS_ETEMP, "T < 0.1 in an interior cell (comput_duvtpckw stop)");
  if (e[1]) return fail(GSIS_EHIP, "kinetic kernel: a wait for the neighbouring strip timed out");
  if (h->p.mode == GSIS_MODE_M1) {
    int it = 0, bad = 0;
    int rc = gsis::synthetic_read_iters(h->syn, h->s, &it, &bad);
    if (rc) return fail(rc, gsis::synthetic_error());
    h->ns_iter = it;
    if (bad) return fail(GSIS_ETEMP, "T < 0.1 in an interior cell during the synthetic NS solve");
  }
  return 0;
}

// FRD between the host's Fortran array and the device layout, a chunk of
// consecutive velocities at a time through a device stage (each velocity's
// 2(nx+7)(ny+7) doubles are contiguous in the Fortran array).  With stage ==
// nullptr a stage of at most kStageBytes (at least one velocity) is allocated
// for the call.
constexpr size_t kStageBytes = (size_t)64 << 20;

int frd_transfer(gsis_handle *h, double *frd, bool up, double *stage, size_t stage_bytes) {
  const Dev &d = h->d;
  const size_t blk = (size_t)2 * (d.nx + 7) * (d.ny + 7);
  bool own = false;
  if (!stage) {
    stage_bytes = std::max(blk * 8, std::min(blk * 8 * (size_t)d.nv, kStageBytes));
    HIPCHK(hipMalloc(&stage, stage_bytes));
    own = true;
  }
  const int nvc = (int)std::min((size_t)d.nv, stage_bytes / (blk * 8));
  hipError_t e = hipSuccess;
  for (int v0 = 0; v0 < d.nv && e == hipSuccess; v0 += nvc) {
    const int n = std::min(nvc, d.nv - v0);
    double *host = frd + blk * (size_t)v0;
    if (up) {
      e = hipMemcpyAsync(stage, host, blk * n * 8, hipMemcpyHostToDevice, h->s);
      if (e != hipSuccess) break;
      if (h->prec == GSIS_PREC_FP32)
        gsis::launch_frd_in_chunk<float>(d, stage, v0, n, (float *)h->FG, (float *)h->FH, (float *)h->slot, h->s);
      else
        gsis::launch_frd_in_chunk<double>(d, stage, v0, n, (double *)h->FG, (double *)h->FH, (double *)h->slot,
                                          h->s);
    } else {
      if (h->prec == GSIS_PREC_FP32)
        gsis::launch_frd_out_chunk<float>(d, stage, v0, n, (const float *)h->FG, (const float *)h->FH,
                                          (const float *)h->slot, h->s);
      else
        gsis::launch_frd_out_chunk<double>(d, stage, v0, n, (const double *)h->FG, (const double *)h->FH,
                                           (const double *)h->slot, h->s);
      e = hipGetLastError();
      if (e != hipSuccess) break;
      e = hipMemcpyAsync(host, stage, blk * n * 8, hipMemcpyDeviceToHost, h->s);
    }
    if (e == hipSuccess) e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->s);
  if (own) (void)hipFree(stage);
  if (e != hipSuccess)
    return fail(GSIS_EHIP, std::string(up ? "upload_frd: " : "download_frd: ") + hipGetErrorString(e));
  return 0;
}

// Stage of gsis_download_frd_noalloc: a file-scope device array, one per
// GPU, held until the process ends.  Host code never uses its name as an
// address (that is a host-side stand-in); scratch_on() looks up each device's
// copy once with hipGetSymbolAddress and keeps the pointer.
constexpr size_t kScratchBytes = (size_t)32 << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

int scratch_on(int dev, double **out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (dev < 0 || dev >= kMaxDevices) return fail(GSIS_EARG, "device ordinal out of range for the scratch table");
  if (!g_scratch_ptr[dev]) {
    int cur = -1;
    HIPCHK(hipGetDevice(&cur));
    if (cur != dev) HIPCHK(hipSetDevice(dev));
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess) return fail(GSIS_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double *>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace

// ---- batches of instances (gsis_batch_*) ----------------------------------
struct gsis_batch {
  std::vector<gsis_handle *> hs;
  int prec = 0, scheme = 0;
  int nw = 8, order = 0;  // strip width and job order of the captured step
  hipStream_t s = nullptr;
  void *djobs = nullptr;  // device table of StripJob<T>
  int *epoch = nullptr;   // [0] launch epoch, [1] a seam wait timed out
  std::vector<int *> flags;
  int nstrip = 0, nseam = 0, ncellb = 0;
  gsis::NsJob *nsjobs = nullptr;  // device table of the inner loops run in the graph
  std::vector<bool> in_graph;     // M1 instance whose synthetic solve is captured
  // synthetic solves inside the graph (one workgroup each): -1 automatic (when
  // at least kNsGraphMin instances qualify), 0 never, 1 for every qualifying one
  int ns_graph = -1;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<void *> allocs;
};

namespace {

// One workgroup per instance pays off with many instances: measured 3.0x
// faster than one cooperative launch per instance for 64 instances of side
// 64-160, 0.6x for 16 of side 48-96, 0.17x for one of side 256.
constexpr int kNsGraphMin = 32;

void batch_drop_graph(gsis_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// job order: as given (0) or the largest kinetic work first (1), so that a
// persistent launch starts its longest pipelines first
std::vector<int> batch_order(const gsis_batch *b, int order) {
  std::vector<int> ix(b->hs.size());
  for (size_t k = 0; k < ix.size(); k++) ix[k] = (int)k;
  if (order == 1) {
    auto work = [&](int k) { const Dev &d = b->hs[k]->d; return (double)d.nx * d.ny * d.nv; };
    std::stable_sort(ix.begin(), ix.end(), [&](int x, int y) { return work(x) > work(y); });
  }
  return ix;
}

template <typename T>
int batch_jobs(gsis_batch *b) {
  std::vector<gsis::StripJob<T>> jobs;
  int a = 0, c = 0, e = 0;
  for (int k : batch_order(b, b->order)) {
    gsis_handle *h = b->hs[k];
    gsis::StripJob<T> j{};
    j.d = h->d;
    j.cp = reinterpret_cast<const T *>(h->cpar);
    j.F = (T *)h->FG;
    j.R = (T *)h->RG;
    j.slot = (const T *)h->slot;
    j.msg = (T *)h->strip.msg;
    j.dright = (T *)h->strip.dright;
    j.flag = b->flags[k];
    j.oflags = h->strip.oflags;
    j.ntiles = (h->d.nv + 63) / 64;
    j.strip0 = a;
    j.seam0 = c;
    j.cell0 = e;
    a += gsis::strip_blocks(h->d, b->nw);
    c += gsis::seam_blocks(h->d, b->nw);
    e += gsis::cell_blocks(h->d);
    jobs.push_back(j);
  }
  b->nstrip = a;
  b->nseam = c;
  b->ncellb = e;
  HIPCHK(hipMemcpy(b->djobs, jobs.data(), jobs.size() * sizeof(jobs[0]), hipMemcpyHostToDevice));
  return 0;
}

// the M0 step of every instance as one graph: per instance S0 <- S and the
// kinetic BCs, the batched kinetic step, per instance moments + DETFPwall,
// Convert_NS and the HoT
template <typename T>
int batch_capture(gsis_batch *b) {
  batch_drop_graph(b);
  int rc = batch_jobs<T>(b);
  if (rc) return rc;
  {  // which M1 instances run their inner loop inside the graph
    std::vector<gsis::NsJob> jobs;
    b->in_graph.assign(b->hs.size(), false);
    for (size_t k = 0; k < b->hs.size(); k++) {
      gsis_handle *h = b->hs[k];
      if (h->p.mode != GSIS_MODE_M1 || h->net || (h->warm_net && h->warm_steps > 0) ||
          !gsis::synthetic_batchable(h->syn))
        continue;
      b->in_graph[k] = true;
      jobs.push_back(gsis::NsJob{h->d, h->syn});
    }
    if (b->ns_graph == 0 || (b->ns_graph < 0 && (int)jobs.size() < kNsGraphMin)) {
      b->in_graph.assign(b->hs.size(), false);
      jobs.clear();
    }
    if (!jobs.empty()) {
      if (!b->nsjobs) {
        void *q = nullptr;
        HIPCHK(hipMalloc(&q, b->hs.size() * sizeof(gsis::NsJob)));
        b->allocs.push_back(q);
        b->nsjobs = reinterpret_cast<gsis::NsJob *>(q);
      }
      HIPCHK(hipMemcpy(b->nsjobs, jobs.data(), jobs.size() * sizeof(gsis::NsJob), hipMemcpyHostToDevice));
    }
  }
  std::vector<hipStream_t> own(b->hs.size());
  for (size_t k = 0; k < b->hs.size(); k++) {
    own[k] = b->hs[k]->s;
    b->hs[k]->s = b->s;
  }
  hipError_t e = hipStreamBeginCapture(b->s, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    for (gsis_handle *h : b->hs) {
      gsis::launch_save_s0(h->d, b->s);
      phy_boundary<T>(h);
    }
    gsis::launch_strip_batch<T>((const gsis::StripJob<T> *)b->djobs, (int)b->hs.size(), b->nw, b->scheme, b->nstrip,
                                b->nseam, b->ncellb, b->epoch, b->epoch + 1, b->s);
    for (gsis_handle *h : b->hs) {
      moments<T>(h);
      if (h->mass_on) gsis::launch_mass_correct(h->d, h->mwork, h->init_mass, b->s);
      gsis::launch_convert(h->d, b->s);
      if (h->p.if_viscous) gsis::launch_hot(h->d, b->s);
    }
    // the synthetic solves of the M1 instances whose inner loop fits one
    // workgroup and that use no closure: all in one launch, one workgroup each
    int nj = 0;
    for (size_t k = 0; k < b->hs.size(); k++)
      if (b->in_graph[k]) gsis::synthetic_begin(b->hs[k]->d, b->hs[k]->syn, b->s);
    for (size_t k = 0; k < b->hs.size(); k++) nj += b->in_graph[k] ? 1 : 0;
    if (nj > 0) gsis::launch_ns_batch(b->nsjobs, nj, b->s);
    for (size_t k = 0; k < b->hs.size(); k++)
      if (b->in_graph[k]) gsis::synthetic_finish(b->hs[k]->d, b->hs[k]->syn, b->s);
    const hipError_t e2 = hipStreamEndCapture(b->s, &b->graph);
    e = e2 != hipSuccess ? e2 : hipGetLastError();
  }
  for (size_t k = 0; k < b->hs.size(); k++) b->hs[k]->s = own[k];
  if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    batch_drop_graph(b);
    return fail(GSIS_EHIP, std::string("batch: graph capture failed: ") + hipGetErrorString(e));
  }
  return 0;
}

int batch_capture_any(gsis_batch *b) {
  return b->prec == GSIS_PREC_FP32 ? batch_capture<float>(b) : batch_capture<double>(b);
}

int batch_replay(gsis_batch *b, int nsteps) {
  bool m1 = false;  // an M1 instance whose synthetic solve runs outside the graph
  for (size_t q = 0; q < b->hs.size(); q++) {
    HIPCHK(hipStreamSynchronize(b->hs[q]->s));
    m1 = m1 || (b->hs[q]->p.mode == GSIS_MODE_M1 && !b->in_graph[q]);
  }
  for (int k = 0; k < nsteps; k++) {
    HIPCHK(hipGraphLaunch(b->exec, b->s));
    if (!m1) continue;
    // the synthetic solve of each other M1 instance after the captured part
    // of its step (a cooperative launch, or the closure's host-side loop)
    HIPCHK(hipStreamSynchronize(b->s));
    for (size_t q = 0; q < b->hs.size(); q++) {
      gsis_handle *h = b->hs[q];
      if (h->p.mode != GSIS_MODE_M1 || b->in_graph[q]) continue;
      int rc = synthetic_any(h);
      if (rc) return rc;
    }
    for (gsis_handle *h : b->hs) HIPCHK(hipStreamSynchronize(h->s));
  }
  HIPCHK(hipStreamSynchronize(b->s));
  for (gsis_handle *h : b->hs) {
    h->tt += nsteps * h->p.dt;
    h->kstep += nsteps;
  }
  int w[2] = {0, 0};
  HIPCHK(hipMemcpy(w, b->epoch, sizeof(w), hipMemcpyDeviceToHost));
  if (w[1]) return fail(GSIS_EHIP, "batch: a wait for a neighbouring strip timed out");
  for (gsis_handle *h : b->hs) {
    const int rc = check_err_flag(h);
    if (rc) return rc;
  }
  return 0;
}

}  // namespace

extern "C" {

const char *gsis_last_error(void) { return g_err.c_str(); }

int gsis_create(const gsis_params *p, gsis_handle **out) {
  if (!out) return fail(GSIS_EARG, "out is NULL");
  *out = nullptr;
  int rc = validate(p);
  if (rc) return rc;
  gsis_handle *h = new gsis_handle();
  h->p = *p;
  h->bc.assign(p->bc, p->bc + p->nbc);
  h->prec = p->precision;
  h->tsize = (p->precision == GSIS_PREC_FP32) ? sizeof(float) : sizeof(double);
  const int nx = p->nx, ny = p->ny;
  Dev &d = h->d;
  d.nx = nx; d.ny = ny; d.nvi = p->nvi; d.nvj = p->nvj; d.nv = p->nvi * p->nvj;
  d.NI = nx + 3; d.NJ = ny + 3; d.ncell = d.NI * d.NJ; d.GI = nx + 2;
  d.L = (nx > ny ? nx : ny) + 3;
  d.cfc = p->cfc; d.x1s = p->x1s; d.cpr = p->cpr; d.twall = p->twall; d.dt = p->dt;
  d.ma = p->ma; d.gamma = p->gamma; d.re = p->re; d.pr = p->pr; d.cv = p->cv; d.cp = p->cp; d.aoa = p->aoa;
  d.scheme = p->scheme;
  auto cleanup = [&](int code) {
    gsis_destroy(h);
    return code;
  };
  if (hipGetDevice(&h->dev) != hipSuccess) return cleanup(fail(GSIS_EHIP, "hipGetDevice failed"));
  if (hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking) != hipSuccess)
    return cleanup(fail(GSIS_EHIP, "hipStreamCreate failed"));
  // velocity grid
  double *vi, *vj, *wi, *wj;
  if ((rc = dalloc(h, &vi, 8 * p->nvi)) || (rc = dalloc(h, &wi, 8 * p->nvi)) || (rc = dalloc(h, &vj, 8 * p->nvj)) ||
      (rc = dalloc(h, &wj, 8 * p->nvj)))
    return cleanup(rc);
  if (hipMemcpy(vi, p->vi, 8 * p->nvi, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(wi, p->wi, 8 * p->nvi, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(vj, p->vj, 8 * p->nvj, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(wj, p->wj, 8 * p->nvj, hipMemcpyHostToDevice) != hipSuccess)
    return cleanup(fail(GSIS_EHIP, "velocity grid upload failed"));
  d.vi = vi; d.vj = vj; d.wi = wi; d.wj = wj;
  {  // row factors of the moments, in the order the moments kernel would form them
    std::vector<double> rf(4 * (size_t)p->nvj);
    for (int jv = 0; jv < p->nvj; jv++) {
      const double wy = p->wj[jv], vy = p->vj[jv];
      const double w1 = wy * vy, w2 = w1 * vy, w3 = w2 * vy;
      rf[4 * jv] = wy; rf[4 * jv + 1] = w1; rf[4 * jv + 2] = w2; rf[4 * jv + 3] = w3;
    }
    double *rfd;
    if ((rc = dalloc(h, &rfd, 8 * rf.size()))) return cleanup(rc);
    if (hipMemcpy(rfd, rf.data(), 8 * rf.size(), hipMemcpyHostToDevice) != hipSuccess)
      return cleanup(fail(GSIS_EHIP, "velocity grid upload failed"));
    d.rowf = rfd;
  }
  // geometry (nx,ny) Fortran -> zero-padded (nx+2)x(ny+2)
  const size_t ng = (size_t)(nx + 2) * (ny + 2);
  const double *src[7] = {p->vol, p->si, p->sj, p->ni1, p->ni2, p->nj1, p->nj2};
  const double *dst[7];
  std::vector<double> tmp(ng);
  for (int a = 0; a < 7; a++) {
    std::fill(tmp.begin(), tmp.end(), 0.0);
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) tmp[(size_t)j * (nx + 2) + i] = src[a][(size_t)(i - 1) + (size_t)nx * (j - 1)];
    double *g;
    if ((rc = dalloc(h, &g, 8 * ng))) return cleanup(rc);
    if (hipMemcpy(g, tmp.data(), 8 * ng, hipMemcpyHostToDevice) != hipSuccess)
      return cleanup(fail(GSIS_EHIP, "geometry upload failed"));
    dst[a] = g;
  }
  d.vol = dst[0]; d.si = dst[1]; d.sj = dst[2]; d.ni1 = dst[3]; d.ni2 = dst[4]; d.nj1 = dst[5]; d.nj2 = dst[6];
  // node coordinates: Fortran x(-1:nx+2,-1:ny+2) has exactly the device layout
  const size_t nn = (size_t)(nx + 4) * (ny + 4);
  double *xn, *yn;
  if ((rc = dalloc(h, &xn, 8 * nn)) || (rc = dalloc(h, &yn, 8 * nn))) return cleanup(rc);
  if (hipMemcpy(xn, p->x, 8 * nn, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(yn, p->y, 8 * nn, hipMemcpyHostToDevice) != hipSuccess)
    return cleanup(fail(GSIS_EHIP, "coordinate upload failed"));
  d.xn = xn; d.yn = yn;
  // cell centres (src/sub_init.f90:814-819)
  std::vector<double> cx(d.ncell), cy(d.ncell);
  for (int j = -1; j <= ny + 1; j++)
    for (int i = -1; i <= nx + 1; i++) {
      auto X = [&](const double *a, int ii, int jj) { return a[(size_t)(ii + 1) + (size_t)(nx + 4) * (jj + 1)]; };
      const int c = (j + 1) * d.NI + (i + 1);
      cx[c] = (X(p->x, i, j) + X(p->x, i + 1, j) + X(p->x, i, j + 1) + X(p->x, i + 1, j + 1)) * 0.25;
      cy[c] = (X(p->y, i, j) + X(p->y, i + 1, j) + X(p->y, i, j + 1) + X(p->y, i + 1, j + 1)) * 0.25;
    }
  double *x1, *y1;
  if ((rc = dalloc(h, &x1, 8 * d.ncell)) || (rc = dalloc(h, &y1, 8 * d.ncell))) return cleanup(rc);
  if (hipMemcpy(x1, cx.data(), 8 * d.ncell, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(y1, cy.data(), 8 * d.ncell, hipMemcpyHostToDevice) != hipSuccess)
    return cleanup(fail(GSIS_EHIP, "cell centre upload failed"));
  d.x1 = x1; d.y1 = y1;
  // macroscopic state
  const size_t N = d.ncell;
  if ((rc = dalloc(h, &d.S, 8 * 11 * N)) || (rc = dalloc(h, &d.S0, 8 * 10 * N)) || (rc = dalloc(h, &d.CF, 8 * N)) ||
      (rc = dalloc(h, &d.U, 8 * 4 * N)) || (rc = dalloc(h, &d.U0, 8 * 4 * N)) || (rc = dalloc(h, &d.U2, 8 * 4 * N)) ||
      (rc = dalloc(h, &d.taoi, 8 * 4 * N)) || (rc = dalloc(h, &d.taoj, 8 * 4 * N)) ||
      (rc = dalloc(h, &d.qi, 8 * 2 * N)) || (rc = dalloc(h, &d.qj, 8 * 2 * N)) ||
      (rc = dalloc(h, &d.prim, 8 * 5 * N)) || (rc = dalloc(h, &d.res4, 8 * 4 * N)) || (rc = dalloc(h, &d.err, 64)) ||
      // ny+3 planes of <= 3*1024 constants (k_walk), or kNCellPar per interior cell (k_strip)
      (rc = dalloc(h, &h->cpar, 8 * std::max((size_t)(ny + 3) * 3 * 1024,
                                             (size_t)(nx - 1) * (ny - 1) * gsis::kNCellPar))) ||
      (rc = dalloc(h, &h->resbuf, 8 * 8)))
    return cleanup(rc);
  {  // Creat_Mesh sets U, U0, U2 to (1,0,0,1) (src/sub_init.f90:159-162)
    std::vector<double> u(4 * N, 0.0);
    for (size_t c = 0; c < N; c++) { u[c] = 1.0; u[3 * N + c] = 1.0; }
    if (hipMemcpy(d.U, u.data(), 8 * 4 * N, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d.U0, u.data(), 8 * 4 * N, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d.U2, u.data(), 8 * 4 * N, hipMemcpyHostToDevice) != hipSuccess)
      return cleanup(fail(GSIS_EHIP, "U upload failed"));
  }
  // face override flags and the DETFPwall node list
  std::vector<uint8_t> oi1(ny + 2, 0), oi3(ny + 2, 0), oj2(nx + 2, 0), oj4(nx + 2, 0);
  std::vector<int> wl;
  int soff = 0;
  for (int b = 0; b < p->nbc; b++) {
    const gsis_bc &c = p->bc[b];
    BcDev bd{c.face, c.ist, c.iend, c.jst, c.jend, c.code, c.uw, c.vw, -1};
    if (c.code == GSIS_BC_WALL || c.code == GSIS_BC_FARFIELD) {
      bd.swall_off = soff;
      if (c.face == 1) for (int j = c.jst; j <= c.jend - 1; j++) oi1[j] = 1;
      if (c.face == 3) for (int j = c.jst; j <= c.jend - 1; j++) oi3[j] = 1;
      if (c.face == 2) for (int i = c.ist; i <= c.iend - 1; i++) oj2[i] = 1;
      if (c.face == 4) for (int i = c.ist; i <= c.iend - 1; i++) oj4[i] = 1;
      const int nk = bc_nk(c);
      for (int kk = 0; kk < nk - 1; kk++) {  // DETFPwall skips the last node
        int g1, g2, a1, a2, b1, b2;
        if (c.face == 1) { int j = c.jst + kk; g1 = c.ist - 1; g2 = j; a1 = c.ist; a2 = j; b1 = c.ist + 1; b2 = j; }
        else if (c.face == 2) { int i = c.ist + kk; g1 = i; g2 = c.jst - 1; a1 = i; a2 = c.jst; b1 = i; b2 = c.jst + 1; }
        else if (c.face == 3) { int j = c.jst + kk; g1 = c.iend; g2 = j; a1 = c.iend - 1; a2 = j; b1 = c.iend - 2; b2 = j; }
        else { int i = c.ist + kk; g1 = i; g2 = c.jend; a1 = i; a2 = c.jend - 1; b1 = i; b2 = c.jend - 2; }
        wl.push_back(gsis::cidx(d, g1, g2));
        wl.push_back(gsis::cidx(d, a1, a2));
        wl.push_back(gsis::cidx(d, b1, b2));
        wl.push_back(soff + 10 * kk);
      }
      soff += 10 * nk;
    }
    h->bcd.push_back(bd);
  }
  h->swall_len = soff;
  h->nwall = (int)wl.size() / 4;
  uint8_t *o1, *o3, *o2, *o4;
  if ((rc = dalloc(h, &o1, ny + 2)) || (rc = dalloc(h, &o3, ny + 2)) || (rc = dalloc(h, &o2, nx + 2)) ||
      (rc = dalloc(h, &o4, nx + 2)) || (rc = dalloc(h, &h->wall_list, 4 * wl.size())) ||
      (rc = dalloc(h, &h->swall, 8 * (size_t)soff)))
    return cleanup(rc);
  if (hipMemcpy(o1, oi1.data(), ny + 2, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(o3, oi3.data(), ny + 2, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(o2, oj2.data(), nx + 2, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(o4, oj4.data(), nx + 2, hipMemcpyHostToDevice) != hipSuccess ||
      (!wl.empty() && hipMemcpy(h->wall_list, wl.data(), 4 * wl.size(), hipMemcpyHostToDevice) != hipSuccess))
    return cleanup(fail(GSIS_EHIP, "BC table upload failed"));
  d.ovrI1 = o1; d.ovrI3 = o3; d.ovrJ2 = o2; d.ovrJ4 = o4;
  {  // the same flags as 32-bit words for the column-wavefront kernel's scalar loads
    std::vector<int> of(ny + 2 + nx + 2, 0);
    for (int j = 0; j < ny + 2; j++) of[j] = (oi1[j] ? 1 : 0) | (oi3[j] ? 2 : 0);
    for (int i = 0; i < nx + 2; i++) of[ny + 2 + i] = (oj2[i] ? 1 : 0) | (oj4[i] ? 2 : 0);
    int *ofd;
    if ((rc = dalloc(h, &ofd, sizeof(int) * of.size()))) return cleanup(rc);
    if (hipMemcpy(ofd, of.data(), sizeof(int) * of.size(), hipMemcpyHostToDevice) != hipSuccess)
      return cleanup(fail(GSIS_EHIP, "BC table upload failed"));
    h->strip.oflags = ofd;
  }
  // distributions
  const size_t nf = (size_t)d.ncell * d.nv * h->tsize;
  const size_t ns = (size_t)4 * d.L * 2 * d.nv * h->tsize;
  // G plane, H plane and the wall slots in one allocation (the fused kernel
  // addresses all three with 32-bit offsets from one base)
  if ((rc = dalloc(h, &h->FG, 2 * nf + ns)) || (rc = dalloc(h, &h->RG, 2 * nf))) return cleanup(rc);
  h->FH = (char *)h->FG + nf;
  h->slot = (char *)h->FG + 2 * nf;
  h->RH = (char *)h->RG + nf;
  {  // seam buffers of the column-wavefront kernel, sized for 8 columns per workgroup
    const size_t rows = gsis::strip_rows(d, 8);
    gsis::StripBufs &sb = h->strip;
    if (rows > 0) {
      if ((rc = dalloc(h, &sb.msg, rows * 256 * h->tsize)) || (rc = dalloc(h, &sb.dright, rows * 128 * h->tsize)) ||
          (rc = dalloc(h, &sb.flag, rows * sizeof(int))))
        return cleanup(rc);
      sb.msg_elems = rows * 256;
      sb.dright_elems = rows * 128;
      sb.flag_elems = rows;
    }
    sb.err = d.err + 1;
  }
  if (p->mode == GSIS_MODE_M1) {
    rc = gsis::synthetic_init(h->d, *p, h->syn);
    if (rc) return cleanup(fail(rc, gsis::synthetic_error()));
  }
  *out = h;
  return 0;
}

int gsis_closure_samples(gsis_handle *h, float *x, float *y, int64_t cap, int64_t *n) {
  if (!h || !n) return fail(GSIS_EARG, "NULL argument");
  int nfi = 0;
  const int nf = gsis::closure_faces(h->d, &nfi);
  *n = nf;
  if (!x && !y) return 0;
  if (cap < nf) return fail(GSIS_EARG, "closure samples: cap is below the number of closure faces");
  if (nf == 0) return 0;
  float *dx = nullptr, *dy = nullptr;
  HIPCHK(hipMalloc(&dx, (size_t)nf * (GSIS_MLP_NIN + GSIS_MLP_NOUT) * sizeof(float)));
  dy = dx + (size_t)nf * GSIS_MLP_NIN;
  gsis::launch_cl_features(h->d, nfi, nf, true, dx, h->s);
  gsis::launch_cl_targets(h->d, nfi, nf, dy, h->s);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && x) e = hipMemcpyAsync(x, dx, (size_t)nf * GSIS_MLP_NIN * 4, hipMemcpyDeviceToHost, h->s);
  if (e == hipSuccess && y) e = hipMemcpyAsync(y, dy, (size_t)nf * GSIS_MLP_NOUT * 4, hipMemcpyDeviceToHost, h->s);
  if (e == hipSuccess) e = hipStreamSynchronize(h->s);
  (void)hipFree(dx);
  if (e != hipSuccess) return fail(GSIS_EHIP, std::string("closure samples: ") + hipGetErrorString(e));
  return 0;
}

int gsis_set_mass_correct(gsis_handle *h, int32_t on) {
  if (!h) return fail(GSIS_EARG, "NULL handle");
  if (!on) {
    h->mass_on = false;
    return 0;
  }
  if (!h->mwork) {
    int rc = dalloc(h, &h->mwork, 64);
    if (rc) return rc;
  }
  gsis::launch_out_mass(h->d, h->mwork, h->s);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(&h->init_mass, h->mwork, sizeof(double), hipMemcpyDeviceToHost, h->s));
  HIPCHK(hipStreamSynchronize(h->s));
  h->mass_on = true;
  return 0;
}

int gsis_set_closure(gsis_handle *h, gsis_mlp *net, double gate, int32_t every) {
  if (!h) return fail(GSIS_EARG, "NULL handle");
  if (!net) {
    h->net = nullptr;
    return 0;
  }
  if (h->p.mode != GSIS_MODE_M1) return fail(GSIS_EARG, "the closure needs the synthetic solve (mode M1)");
  if (every < 1 || !(gate >= 0.0)) return fail(GSIS_EARG, "closure: every >= 1 and gate >= 0 required");
  if (gsis::closure_alloc(h->d, h->cl)) return fail(GSIS_EHIP, "closure: device allocation failed");
  h->net = net;
  h->cl_gate = (float)gate;
  h->cl_every = every;
  return 0;
}

int gsis_set_closure_warmup(gsis_handle *h, gsis_mlp *net, int32_t nsteps, int32_t every) {
  if (!h) return fail(GSIS_EARG, "NULL handle");
  if (!net || nsteps <= 0) {
    h->warm_net = nullptr;
    h->warm_steps = 0;
    return 0;
  }
  if (h->p.mode != GSIS_MODE_M1) return fail(GSIS_EARG, "the closure needs the synthetic solve (mode M1)");
  if (every < 1) return fail(GSIS_EARG, "closure: every >= 1 required");
  if (gsis::closure_alloc(h->d, h->cl)) return fail(GSIS_EHIP, "closure: device allocation failed");
  h->warm_net = net;
  h->warm_steps = nsteps;
  h->warm_every = every;
  return 0;
}

int gsis_closure_stats(gsis_handle *h, int64_t *faces, int64_t *accepted, int32_t *refreshes, int32_t *rejected) {
  if (!h) return fail(GSIS_EARG, "NULL handle");
  int acc = 0;
  if (h->cl.count) {
    HIPCHK(hipMemcpyAsync(&acc, h->cl.count, sizeof(int), hipMemcpyDeviceToHost, h->s));
    HIPCHK(hipStreamSynchronize(h->s));
  }
  if (faces) *faces = h->cl.mem ? h->cl.nf : gsis::closure_faces(h->d, nullptr);
  if (accepted) *accepted = acc;
  if (refreshes) *refreshes = h->cl.refreshes;
  if (rejected) *rejected = h->cl.rejected;
  return 0;
}

int gsis_timing(gsis_handle *h, int32_t enable, double *ms, int32_t *nsteps) {
  if (!h) return fail(GSIS_EARG, "NULL handle");
  if (enable && !h->ev[0])
    for (int k = 0; k < 8; k++) HIPCHK(hipEventCreate(&h->ev[k]));
  HIPCHK(hipStreamSynchronize(h->s));
  harvest_timing(h);
  if (ms) for (int k = 0; k < 4; k++) ms[k] = h->t_ms[k];
  if (nsteps) *nsteps = h->t_n;
  if (enable >= 0) {
    h->timing = enable != 0;
    for (int k = 0; k < 4; k++) h->t_ms[k] = 0.0;
    h->t_n = 0;
  }
  return 0;
}

int gsis_destroy(gsis_handle *h) {
  if (!h) return 0;
  if (h->s) (void)hipStreamSynchronize(h->s);
  for (int k = 0; k < 8; k++)
    if (h->ev[k]) (void)hipEventDestroy(h->ev[k]);
  gsis::synthetic_free(h->syn);
  gsis::closure_free(h->cl);
  for (void *q : h->allocs) (void)hipFree(q);
  if (h->s) (void)hipStreamDestroy(h->s);
  delete h;
  return 0;
}

int gsis_upload_frd(gsis_handle *h, const double *frd) {
  if (!h || !frd) return fail(GSIS_EARG, "NULL argument");
  return frd_transfer(h, const_cast<double *>(frd), true, nullptr, 0);
}

int gsis_download_frd(gsis_handle *h, double *frd) {
  if (!h || !frd) return fail(GSIS_EARG, "NULL argument");
  return frd_transfer(h, frd, false, nullptr, 0);
}

int gsis_download_frd_noalloc(gsis_handle *h, double *frd) {
  if (!h || !frd) return fail(GSIS_EARG, "NULL argument");
  const size_t blk_bytes = (size_t)2 * (h->d.nx + 7) * (h->d.ny + 7) * sizeof(double);
  if (blk_bytes > kScratchBytes)
    return fail(GSIS_EARG, "download_frd_noalloc: one velocity of FRD takes " + std::to_string(blk_bytes) +
                               " bytes, more than the " + std::to_string(kScratchBytes) +
                               "-byte device stage; use gsis_download_frd");
  double *stage = nullptr;
  const int rc = scratch_on(h->dev, &stage);
  if (rc) return rc;
  return frd_transfer(h, frd, false, stage, kScratchBytes);
}

int gsis_upload_macro(gsis_handle *h, const double *s, const double *cf, const double *u, const double *u0) {
  if (!h) return fail(GSIS_EARG, "NULL handle");
  const Dev &d = h->d;
  const size_t N = d.ncell;
  HIPCHK(hipStreamSynchronize(h->s));
  if (s) {
    std::vector<double> t(11 * N, 0.0);
    for (int j = -1; j <= d.ny + 1; j++)
      for (int i = -1; i <= d.nx + 1; i++) {
        const int c = gsis::cidx(d, i, j);
        for (int k = 0; k < 11; k++) t[k * N + c] = s[k + 11 * ((size_t)(i + 3) + (size_t)(d.nx + 7) * (j + 3))];
      }
    HIPCHK(hipMemcpy(d.S, t.data(), 8 * 11 * N, hipMemcpyHostToDevice));
  }
  if (cf) HIPCHK(hipMemcpy(d.CF, cf, 8 * N, hipMemcpyHostToDevice));
  double *targets[2] = {d.U, d.U0};
  const double *srcs[2] = {u, u0};
  for (int a = 0; a < 2; a++) {
    if (!srcs[a]) continue;
    std::vector<double> t(4 * N);
    for (size_t c = 0; c < N; c++)
      for (int m = 0; m < 4; m++) t[m * N + c] = srcs[a][m + 4 * c];
    HIPCHK(hipMemcpy(targets[a], t.data(), 8 * 4 * N, hipMemcpyHostToDevice));
  }
  return 0;
}

int gsis_download_macro(gsis_handle *h, double *s, double *s0, double *cf, double *u, double *taonsi_g,
                        double *taonsj_g, double *qi_g, double *qj_g, double *swall) {
  if (!h) return fail(GSIS_EARG, "NULL handle");
  const Dev &d = h->d;
  const size_t N = d.ncell;
  HIPCHK(hipStreamSynchronize(h->s));
  int rc = check_err_flag(h);
  auto get_s = [&](const double *dev, int ncomp, double *out) -> int {
    std::vector<double> t(ncomp * N);
    HIPCHK(hipMemcpy(t.data(), dev, 8 * ncomp * N, hipMemcpyDeviceToHost));
    const size_t tot = (size_t)11 * (d.nx + 7) * (d.ny + 7);
    std::memset(out, 0, 8 * tot);
    for (int j = -1; j <= d.ny + 1; j++)
      for (int i = -1; i <= d.nx + 1; i++) {
        const int c = gsis::cidx(d, i, j);
        for (int k = 0; k < ncomp; k++) out[k + 11 * ((size_t)(i + 3) + (size_t)(d.nx + 7) * (j + 3))] = t[k * N + c];
      }
    return 0;
  };
  auto get_v = [&](const double *dev, int ncomp, double *out) -> int {
    std::vector<double> t(ncomp * N);
    HIPCHK(hipMemcpy(t.data(), dev, 8 * ncomp * N, hipMemcpyDeviceToHost));
    for (size_t c = 0; c < N; c++)
      for (int m = 0; m < ncomp; m++) out[m + ncomp * c] = t[m * N + c];
    return 0;
  };
  int r2 = 0;
  if (s && (r2 = get_s(d.S, 11, s))) return r2;
  if (s0 && (r2 = get_s(d.S0, 10, s0))) return r2;
  if (cf) HIPCHK(hipMemcpy(cf, d.CF, 8 * N, hipMemcpyDeviceToHost));
  if (u && (r2 = get_v(d.U, 4, u))) return r2;
  if (taonsi_g && (r2 = get_v(d.taoi, 4, taonsi_g))) return r2;
  if (taonsj_g && (r2 = get_v(d.taoj, 4, taonsj_g))) return r2;
  if (qi_g && (r2 = get_v(d.qi, 2, qi_g))) return r2;
  if (qj_g && (r2 = get_v(d.qj, 2, qj_g))) return r2;
  if (swall && h->swall_len > 0) HIPCHK(hipMemcpy(swall, h->swall, 8 * (size_t)h->swall_len, hipMemcpyDeviceToHost));
  return rc;
}

int gsis_refresh_moments(gsis_handle *h) {
  if (!h) return fail(GSIS_EARG, "NULL handle");
  if (h->prec == GSIS_PREC_FP32) { phy_boundary<float>(h); moments<float>(h); }
  else { phy_boundary<double>(h); moments<double>(h); }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(h->s));
  return 0;
}

int gsis_step(gsis_handle *h, int32_t nsteps) {
  if (!h) return fail(GSIS_EARG, "NULL handle");
  if (nsteps < 0) return fail(GSIS_EARG, "nsteps < 0");
  for (int k = 0; k < nsteps; k++) {
    int rc = (h->prec == GSIS_PREC_FP32) ? one_step<float>(h) : one_step<double>(h);
    if (rc) return rc;
  }
  return 0;
}

int gsis_residual(gsis_handle *h, double res_rms[4]) {
  if (!h || !res_rms) return fail(GSIS_EARG, "NULL argument");
  if (h->res_kstep < 0)
    return fail(GSIS_EARG, "no kinetic RHS kept yet: the first is kept after a step count that is a multiple of "
                           "kstep_show (" + std::to_string(h->p.kstep_show) + ")");
  gsis::launch_residual(h->d, h->resbuf, h->s);
  double r[8];
  HIPCHK(hipMemcpyAsync(r, h->resbuf, sizeof(r), hipMemcpyDeviceToHost, h->s));
  HIPCHK(hipStreamSynchronize(h->s));
  for (int k = 0; k < 4; k++) res_rms[k] = r[k];
  int rc = check_err_flag(h);
  if (rc) return rc;
  if (std::isnan(res_rms[0])) return fail(GSIS_ENAN, "NaN detected in the residual");
  return 0;
}

int gsis_counters(gsis_handle *h, int32_t *kstep, double *tt, int32_t *ns_iter) {
  if (!h) return fail(GSIS_EARG, "NULL handle");
  HIPCHK(hipStreamSynchronize(h->s));
  int rc = check_err_flag(h);
  if (rc) return rc;
  if (kstep) *kstep = h->kstep;
  if (tt) *tt = h->tt;
  if (ns_iter) *ns_iter = h->ns_iter;
  return 0;
}

int gsis_inner_exchange(gsis_handle *dst, int32_t ksub, gsis_handle *src, int32_t msub, int32_t orient) {
  if (!dst || !src) return fail(GSIS_EARG, "NULL handle");
  if (ksub < 1 || ksub > (int32_t)dst->bcd.size() || msub < 1 || msub > (int32_t)src->bcd.size())
    return fail(GSIS_EARG, "subface number out of range");
  const BcDev &bd = dst->bcd[ksub - 1], &bs = src->bcd[msub - 1];
  if (bd.code < 1) return fail(GSIS_EARG, "subface ksub of dst is not an inner face (code >= 1)");
  if (dst->prec != src->prec) return fail(GSIS_EARG, "the two blocks have different precisions");
  if (dst->d.nvi != src->d.nvi || dst->d.nvj != src->d.nvj)
    return fail(GSIS_EARG, "the two blocks have different velocity grids");
  if (dst->dev != src->dev) return fail(GSIS_EARG, "the two blocks live on different devices");
  // n1 cells along each face (src/GKUA_boundary.f90:427); the reference's
  // buffer is sized by this block's face, so the neighbour's must match
  const int n1 = bd.iend - bd.ist + bd.jend - bd.jst, n1s = bs.iend - bs.ist + bs.jend - bs.jst;
  if (n1 < 1 || n1 != n1s) return fail(GSIS_EARG, "the two subfaces have different cell counts");
  auto launch = [&]() {
    if (dst->prec == GSIS_PREC_FP32)
      gsis::launch_inner<float>(dst->d, bd, (float *)dst->FG, (float *)dst->FH, src->d, bs, (const float *)src->FG,
                                (const float *)src->FH, n1, orient, dst->s);
    else
      gsis::launch_inner<double>(dst->d, bd, (double *)dst->FG, (double *)dst->FH, src->d, bs,
                                 (const double *)src->FG, (const double *)src->FH, n1, orient, dst->s);
    return hipGetLastError();
  };
  if (src == dst) {
    HIPCHK(launch());
    return 0;
  }
  // two streams: the copy follows src's queued work, and src's later work follows the copy
  hipEvent_t e = nullptr;
  HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  hipError_t r = hipEventRecord(e, src->s);
  if (r == hipSuccess) r = hipStreamWaitEvent(dst->s, e, 0);
  if (r == hipSuccess) r = launch();
  if (r == hipSuccess) r = hipEventRecord(e, dst->s);
  if (r == hipSuccess) r = hipStreamWaitEvent(src->s, e, 0);
  (void)hipEventDestroy(e);
  if (r != hipSuccess) return fail(GSIS_EHIP, std::string("inner_exchange: ") + hipGetErrorString(r));
  return 0;
}

int gsis_set_counters(gsis_handle *h, int32_t kstep, double tt) {
  if (!h) return fail(GSIS_EARG, "NULL handle");
  if (kstep < 0) return fail(GSIS_EARG, "kstep must be >= 0");
  // the counters are read when a step is queued; no wait on the stream
  h->kstep = kstep;
  h->tt = tt;
  return 0;
}

int gsis_fill_frd_uniform(gsis_handle *h, const double *g, const double *hv) {
  if (!h || !g || !hv) return fail(GSIS_EARG, "NULL argument");
  const size_t nv = h->d.nv;
  double *dv = nullptr;
  HIPCHK(hipMalloc(&dv, 2 * nv * sizeof(double)));
  hipError_t e = hipMemcpyAsync(dv, g, nv * 8, hipMemcpyHostToDevice, h->s);
  if (e == hipSuccess) e = hipMemcpyAsync(dv + nv, hv, nv * 8, hipMemcpyHostToDevice, h->s);
  if (e == hipSuccess) {
    if (h->prec == GSIS_PREC_FP32)
      gsis::launch_fill_uniform<float>(h->d, dv, dv + nv, (float *)h->FG, (float *)h->FH, (float *)h->slot, h->s);
    else
      gsis::launch_fill_uniform<double>(h->d, dv, dv + nv, (double *)h->FG, (double *)h->FH, (double *)h->slot, h->s);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->s);
  (void)hipFree(dv);
  if (e != hipSuccess) return fail(GSIS_EHIP, std::string("fill_frd_uniform: ") + hipGetErrorString(e));
  return 0;
}

int gsis_batch_create(gsis_handle *const *hs, int32_t n, gsis_batch **out) {
  if (!out || !hs || n < 1) return fail(GSIS_EARG, "batch: NULL argument or no instance");
  *out = nullptr;
  for (int k = 0; k < n; k++) {
    const gsis_handle *h = hs[k];
    if (!h) return fail(GSIS_EARG, "batch: NULL handle");
    if (h->prec != hs[0]->prec || h->d.scheme != hs[0]->d.scheme || h->dev != hs[0]->dev)
      return fail(GSIS_EARG, "batch: instances must share precision, flux scheme and device");
    if (h->d.nx - 1 < 2 || h->d.ny - 1 < 2 || gsis::strip_rows(h->d, 8) * 256 > h->strip.msg_elems)
      return fail(GSIS_EUNSUPPORTED, "batch: instance " + std::to_string(k) + " does not fit the column-wavefront kernel");
    for (int m = 0; m < k; m++)
      if (hs[m] == h) return fail(GSIS_EARG, "batch: a handle is listed twice");
  }
  gsis_batch *b = new gsis_batch();
  b->hs.assign(hs, hs + n);
  b->prec = hs[0]->prec;
  b->scheme = hs[0]->d.scheme;
  auto bail = [&](int rc) {
    gsis_batch_destroy(b);
    return rc;
  };
  auto alloc = [&](void **p, size_t bytes) {
    if (hipMalloc(p, bytes) != hipSuccess || hipMemset(*p, 0, bytes) != hipSuccess) return false;
    b->allocs.push_back(*p);
    return true;
  };
  if (hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(GSIS_EHIP, "batch: hipStreamCreate failed"));
  const size_t jb = (size_t)n * std::max(sizeof(gsis::StripJob<double>), sizeof(gsis::StripJob<float>));
  if (!alloc(&b->djobs, jb) || !alloc((void **)&b->epoch, 64)) return bail(fail(GSIS_EHIP, "batch: allocation failed"));
  for (int k = 0; k < n; k++) {
    void *f = nullptr;
    if (!alloc(&f, std::max<size_t>(gsis::strip_rows(hs[k]->d, 8), 1) * sizeof(int)))
      return bail(fail(GSIS_EHIP, "batch: allocation failed"));
    b->flags.push_back((int *)f);
  }
  const int rc = batch_capture_any(b);
  if (rc) return bail(rc);
  *out = b;
  return 0;
}

int gsis_batch_step(gsis_batch *b, int32_t nsteps) {
  if (!b || nsteps < 0) return fail(GSIS_EARG, "batch: bad argument");
  return batch_replay(b, nsteps);
}

int gsis_batch_tune(gsis_batch *b, int32_t steps_per_trial, int32_t *nw, int32_t *order, double *ms_per_step) {
  if (!b || steps_per_trial < 1) return fail(GSIS_EARG, "batch: bad argument");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  double best = 1e300;
  int bnw = b->nw, bord = b->order;
  int rc = 0;
  for (int cw : {8, 16}) {
    for (int co : {0, 1}) {
      b->nw = cw;
      b->order = co;
      if ((rc = batch_capture_any(b))) break;
      (void)hipEventRecord(e0, b->s);
      if ((rc = batch_replay(b, steps_per_trial))) break;
      (void)hipEventRecord(e1, b->s);
      (void)hipEventSynchronize(e1);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (ms / steps_per_trial < best) {
        best = ms / steps_per_trial;
        bnw = cw;
        bord = co;
      }
    }
    if (rc) break;
  }
  // the synthetic solves: the other placement than the current one, with the
  // best strip width and order (one more trial, when any solve qualifies)
  int bns = b->ns_graph;
  if (!rc) {
    b->nw = bnw;
    b->order = bord;
    rc = batch_capture_any(b);
    int qualify = 0, now = 0;
    for (size_t k = 0; k < b->hs.size(); k++) {
      const gsis_handle *h = b->hs[k];
      qualify += h->p.mode == GSIS_MODE_M1 && !h->net && !(h->warm_net && h->warm_steps > 0) &&
                 gsis::synthetic_batchable(h->syn);
      now += b->in_graph[k] ? 1 : 0;
    }
    if (!rc && qualify > 0) {
      const int cur = now > 0 ? 1 : 0;
      b->ns_graph = 1 - cur;
      if (!(rc = batch_capture_any(b))) {
        (void)hipEventRecord(e0, b->s);
        if (!(rc = batch_replay(b, steps_per_trial))) {
          (void)hipEventRecord(e1, b->s);
          (void)hipEventSynchronize(e1);
          float ms = 0.f;
          (void)hipEventElapsedTime(&ms, e0, e1);
          bns = ms / steps_per_trial < best ? 1 - cur : cur;
          best = std::min(best, (double)ms / steps_per_trial);
        }
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  b->nw = bnw;
  b->order = bord;
  b->ns_graph = bns;
  if ((rc = batch_capture_any(b))) return rc;
  if (nw) *nw = bnw;
  if (order) *order = bord;
  if (ms_per_step) *ms_per_step = best;
  return 0;
}

int gsis_batch_set_ns_graph(gsis_batch *b, int32_t mode) {
  if (!b || mode < -1 || mode > 1) return fail(GSIS_EARG, "batch: bad argument");
  b->ns_graph = mode;
  return batch_capture_any(b);
}

int gsis_batch_ns_graph(const gsis_batch *b, int32_t *count) {
  if (!b || !count) return fail(GSIS_EARG, "batch: NULL argument");
  int n = 0;
  for (bool x : b->in_graph) n += x ? 1 : 0;
  *count = n;
  return 0;
}

int gsis_batch_destroy(gsis_batch *b) {
  if (!b) return 0;
  if (b->s) (void)hipStreamSynchronize(b->s);
  batch_drop_graph(b);
  for (void *q : b->allocs) (void)hipFree(q);
  if (b->s) (void)hipStreamDestroy(b->s);
  delete b;
  return 0;
}

}  // extern "C"
