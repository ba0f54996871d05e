// gsis_kernels.h -- launch wrappers of the gfx950 kernels (gsis_kernels.hip).
#pragma once
#include "gsis_device.h"

namespace gsis {

constexpr int kMaxV1 = 512;  // max discrete velocities per direction (LDS tables)

template <typename T> void launch_wall(const Dev &d, const BcDev &b, T *FG, T *FH, T *slot, hipStream_t s);
// up to kMaxWalls diffuse-wall subfaces launched together; start[k] is the
// first block of wall k (start[n] = total face nodes)
constexpr int kMaxWalls = 16;
struct WallSet {
  int n;
  int start[kMaxWalls + 1];
  BcDev b[kMaxWalls];
};
template <typename T> void launch_walls(const Dev &d, const WallSet &ws, T *FG, T *FH, T *slot, hipStream_t s);
template <typename T> void launch_farfield(const Dev &d, const BcDev &b, T *FG, T *FH, hipStream_t s);
template <typename T> void launch_periodic(const Dev &d, const BcDev &b, T *FG, T *FH, hipStream_t s);
// inner-face exchange of LAP = 2 layers, FRD(1:2) and S_GKUA(1:10); n1 cells along the face
template <typename T>
void launch_inner(const Dev &dd, const BcDev &bd, T *FGd, T *FHd, const Dev &ds, const BcDev &bs, const T *FGs,
                  const T *FHs, int n1, int orient, hipStream_t s);
template <typename T> void launch_rhs(const Dev &d, const T *FG, const T *FH, const T *slot, T *RG, T *RH, hipStream_t s);
template <typename T> void launch_sweep(const Dev &d, T *FG, T *FH, const T *RG, const T *RH, hipStream_t s);
template <typename T> void launch_moments(const Dev &d, T *FG, T *FH, hipStream_t s);
constexpr int kNCellPar = 20;  // per-cell constants of the fused kinetic kernel
// Fused kinetic step over F = [G plane | H plane] (and R likewise); false when
// the mesh does not fit the fused kernel.
template <typename T> bool launch_walk(const Dev &d, double *cpar, T *F, const T *slot, T *R, int store_rhs, int tiles, hipStream_t s);
// Buffers of the column-wavefront kernel (k_strip): seam messages
// [seam][tile][row][4][64] and the next strip's Delta* [seam][tile][row][2][64]
// (elements of the working precision), per-row ready flags holding the
// launch epoch, and a word set when a wait for a flag timed out.
struct StripBufs {
  void *msg = nullptr, *dright = nullptr;
  int *flag = nullptr, *err = nullptr;
  // face-override words: ny+2 rows (bit 0 i-face 1, bit 1 i-face nx), then
  // nx+2 columns (bit 0 j-face 1, bit 1 j-face ny)
  const int *oflags = nullptr;
  size_t msg_elems = 0, dright_elems = 0, flag_elems = 0;
  int epoch = 0;
  // when set, recorded on the stream just before and after k_strip itself
  hipEvent_t ev_before = nullptr, ev_after = nullptr;
};
// seam rows (seam x tile x row) of a mesh with nw columns per workgroup
size_t strip_rows(const Dev &d, int nw);
// Column-wavefront kinetic step with nw (8 or 16) columns per workgroup, plus
// the seam update; false when the mesh or the buffers do not fit.
template <typename T> bool launch_strip(const Dev &d, double *cpar, T *F, const T *slot, T *R, int store_rhs, int nw, StripBufs &b, hipStream_t s);
template <typename T> void launch_res4(const Dev &d, const T *RG, const T *RH, hipStream_t s);
// One instance of a batched kinetic step: its device state and the first
// workgroup of its share of each batched launch.  The seam flags are the
// batch's own (their values are the batch's launch epochs).
template <typename T>
struct StripJob {
  Dev d;
  const T *cp;  // cell constants in T (k_cellpar), by cell
  T *F, *R;
  const T *slot;
  T *msg, *dright;
  int *flag;
  const int *oflags;
  int ntiles;
  int strip0, seam0, cell0;  // first workgroup in the strip, seam and cell-constant launches
};
// workgroups of one job in the batched launches with nw columns per strip workgroup
int strip_blocks(const Dev &d, int nw);
int seam_blocks(const Dev &d, int nw);
int cell_blocks(const Dev &d);
// the kinetic step of every job (djobs: device table, host copies of the
// counts): epoch += 1, cell constants, one column-wavefront launch over all
// jobs' strips (no RHS kept), the seam update; scheme and T shared by all jobs
template <typename T>
void launch_strip_batch(const StripJob<T> *djobs, int njobs, int nw, int scheme, int nstrip, int nseam, int ncellb,
                        int *epoch, int *werr, hipStream_t s);
// every cell's distribution set to g[v] (G) and h[v] (H), wall slots zero
template <typename T>
void launch_fill_uniform(const Dev &d, const double *g, const double *h, T *FG, T *FH, T *slot, hipStream_t s);
// FRD (Fortran layout) to and from the device layout for velocities v0..v0+nvc-1 through a stage holding that
// contiguous slice of the Fortran FRD array (2(nx+7)(ny+7) doubles per velocity)
template <typename T>
void launch_frd_out_chunk(const Dev &d, double *stage, int v0, int nvc, const T *FG, const T *FH, const T *slot,
                          hipStream_t s);
template <typename T>
void launch_frd_in_chunk(const Dev &d, const double *stage, int v0, int nvc, T *FG, T *FH, T *slot, hipStream_t s);
void launch_save_s0(const Dev &d, hipStream_t s);
// out_mass (sum of S_GKUA(1)*vol over interior cells) into out[0]
void launch_out_mass(const Dev &d, double *out, hipStream_t s);
// MassCorrect: out_mass into work[0], then S_GKUA(1) scaled by init/out_mass
void launch_mass_correct(const Dev &d, double *work, double init, hipStream_t s);
void launch_det_wall(const Dev &d, const int *list, int n, double *swall, hipStream_t s);
void launch_convert(const Dev &d, hipStream_t s);
void launch_hot(const Dev &d, hipStream_t s);
void launch_residual(const Dev &d, double *out, hipStream_t s);

}  // namespace gsis
