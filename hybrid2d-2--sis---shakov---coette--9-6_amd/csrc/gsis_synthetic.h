// gsis_synthetic.h -- synthetic NS solve of mode M1 (src/Hybrid_Sub.f90:45-153).
#pragma once
#include "../../include/gsis_capi.h"
#include "gsis_device.h"

namespace gsis {

// Per-cell inputs of the LU-SGS sweeps of the cells i, j in [3, n-3], written
// by the flux workgroups of the inner loop: field k of each cell.  Stored by
// anti-diagonal plane p = i+j (the sweep order) and by the 64 rows of a sweep
// wavefront, two fields to a 16-byte pair: block ((p-6)*nw + (j-3)/64) holds
// field k of row j at (k/2)*128 + 2*((j-3)%64) + k%2, so a wavefront reads two
// fields of a plane with one contiguous 16-byte-per-lane load.
enum NsRec {
  kRecRes = 0,   // Res(1:4)
  kRecU = 4,     // U(1:4) before this iteration's update
  kRecRa = 8,    // 1/alfa, the inverse diagonal
  kRecCbi = 9,   // Lci/2 + Lvi: the weight of dU in the i-direction off-diagonal terms
  kRecCbj = 10,  // Lcj/2 + Lvj
  kRecHsi = 11,  // si(i,j)/2
  kRecHsj = 12,  // sj(i,j)/2
  kRecRr = 13,   // 1/U(1), the reciprocal of the record's density (as the sweep forms it)
  kRecNi = 14,   // ni1(i,j), ni2(i,j)
  kRecNj = 16,   // nj1(i,j), nj2(i,j)
  kNRec = 18     // nine pairs
};

struct SyntheticState {
  bool on = false;
  double tol = 0.0;
  int maxit = 100000;
  int local_dt = 1, if_viscous = 1;
  double cfl = 20.0, dtmax = 10.0, dtmin = 1e-7;
  int rs = 0, np = 0;        // sweep rows (ny-5) and planes (nx+ny-11); 0 when no cell is swept
  int nw = 0;                // sweep wavefronts holding rows, ceil(rs/64)
  int ntx = 0, nty = 0;      // flux tiles per direction
  int nflux = 0;             // flux workgroups of the inner-loop launch
  int block = 256;           // threads per workgroup of the inner-loop launch
  bool axis = false;         // every swept cell's face normals are exactly (1,0) and (0,1)
  double *s00 = nullptr;     // [4][ncell]  S00_GKUA(1:4)
  double *rec = nullptr;     // [np][nw][kNRec/2][64][2] sweep inputs, see NsRec
  double *dstar = nullptr;   // [np][nw][2][64][2] dU after the forward sweep (components in pairs)
  double *part = nullptr;    // [nflux][8]: per flux workgroup {iteration done (int), T<0.1 (int), max |Res(4)|}
  int *ctl = nullptr;        // [1] go (next iteration or -1), [3] a wait timed out
  int *iters = nullptr;      // [0]: inner iterations of the last step, [1]: T<0.1 flag,
                             // [2]: the launch stopped at kstop with the loop unfinished,
                             // [4:6]: max |Res_4| of the launch's last iteration (double)
  void *mem = nullptr;       // single allocation holding all of the above
  int k0 = 1;                // first inner iteration of a launch
  int kstop = 0x7fffffff;    // last inner iteration a launch may run
};

int synthetic_init(Dev &d, const gsis_params &p, SyntheticState &st);
// one synthetic solve: synthetic_begin, synthetic_run(1, no limit), synthetic_finish
int synthetic_step(const Dev &d, SyntheticState &st, hipStream_t s);
// the parts of a synthetic solve: the kinetic increment into the synthetic
// state (src/Hybrid_Sub.f90:45-94); inner iterations k0..at most kstop of the
// loop (src/Hybrid_Sub.f90:99-104), k0 > 1 continuing a launch that stopped
// at kstop = k0 - 1; the correction of S (src/Hybrid_Sub.f90:108-153)
int synthetic_begin(const Dev &d, SyntheticState &st, hipStream_t s);
int synthetic_run(const Dev &d, SyntheticState &st, int k0, int kstop, hipStream_t s);
void synthetic_finish(const Dev &d, SyntheticState &st, hipStream_t s);
// after synthetic_run: whether the loop stopped at kstop unfinished, the
// inner iterations run so far, the T<0.1 stop and max |Res_4| of the last
// iteration (synchronises the stream)
int synthetic_state(SyntheticState &st, hipStream_t s, bool *unfinished, int *iters, bool *bad_t, double *res);
int synthetic_read_iters(SyntheticState &st, hipStream_t s, int *iters, int *bad_t);
// One instance's inner loop in a batched launch (k_ns_batch)
struct NsJob {
  Dev d;
  SyntheticState st;
};
// whether an instance's inner loop fits one 256-thread workgroup (at most
// 256 swept rows)
bool synthetic_batchable(const SyntheticState &st);
// the inner loops of n instances (device table djobs), one workgroup each,
// between their synthetic_begin and synthetic_finish
void launch_ns_batch(const NsJob *djobs, int n, hipStream_t s);
void synthetic_free(SyntheticState &st);
const char *synthetic_error();

}  // namespace gsis
