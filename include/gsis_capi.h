/*
 * gsis_capi.h -- C-ABI of the MI355X-native GSIS kinetic loop.
 *
 * The library replaces the body of one outer GSIS iteration of the Fortran
 * code Hybrid2D (subroutine Hybrid_Time_advance_LU_SGS, src/Hybrid_Sub.f90:4-160,
 * called from src/opencfd_ec2d.f90:46) and the residual monitor DETFP_res
 * (src/GKUA_DETFP_RES.f90:4-60, called from src/opencfd_ec2d.f90:47).
 *
 * The reference keeps all state in Fortran module globals (src/COMMONDATA.f90:44-170);
 * here every array crosses the boundary explicitly, as plain pointers in the
 * reference's own Fortran layout (column-major, ghost lower bounds as noted),
 * so that a Fortran host passes its allocatable arrays unchanged through
 * ISO_C_BINDING.  The library copies in during create/upload and copies out
 * during download; it never keeps a host pointer past a call.
 *
 * Every entry point returns 0 on success or one of the GSIS_E* codes; the
 * message of the last failure is returned by gsis_last_error().  No C++
 * exception crosses this boundary.  A handle is single-threaded and
 * non-re-entrant.
 */
#ifndef GSIS_CAPI_H
#define GSIS_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* return codes */
#define GSIS_OK 0
#define GSIS_EARG 1         /* bad argument / unsupported size */
#define GSIS_EHIP 2         /* HIP runtime error */
#define GSIS_ENAN 3         /* NaN in the residual (reference: src/IO.f90:236-244) */
#define GSIS_EUNSUPPORTED 4 /* option not implemented */
#define GSIS_ETEMP 5        /* T < 0.1 in an interior cell (src/sub_NS_singlegrid.f90:148-151) */

/* boundary codes, as stored after negation in read_bcin (src/sub_init.f90:278,
 * constants src/COMMONDATA.f90:14) */
#define GSIS_BC_WALL (-2)
#define GSIS_BC_FARFIELD (-4)
#define GSIS_BC_PERIODIC (-5) /* "inlet": periodic copy, src/GKUA_boundary.f90:219-273 */
#define GSIS_BC_SYMMETRY (-3) /* symmetry or slide wall.  The kinetic step neither fills its
                                 ghost cells (src/GKUA_boundary.f90:25-35) nor overrides its
                                 fluxes (src/GKUA_Residual.f90:189-192): the ghost cells keep
                                 what was uploaded.  Its only other effect in the reference is
                                 velocity-space bookkeeping (src/Symetric_X.f90:3-46,
                                 src/sub_init.f90:279-281,480-533). */
/* A code >= 1 is an inner face: the number of the neighbour block, as set from
 * the neighbour line of generic.inp (src/sub_init.f90:282-284).  The kinetic step
 * treats it like GSIS_BC_SYMMETRY; its ghost cells change only through
 * gsis_inner_exchange. */

/* kinetic flux schemes (src/COMMONDATA.f90:18) */
#define GSIS_SCHEME_UP1 0
#define GSIS_SCHEME_NND 1
#define GSIS_SCHEME_WENO3 3 /* third-order WENO, src/GKUA_Residual.f90:167-185 */

/* step modes */
#define GSIS_MODE_M0 0 /* the step as shipped: DVM implicit step + HoT, synthetic solve off */
#define GSIS_MODE_M1 1 /* synthetic NS solve enabled (src/Hybrid_Sub.f90:45-153) */

/* arithmetic of the distribution sweeps */
#define GSIS_PREC_FP64 0
#define GSIS_PREC_FP32 1 /* FP32 storage + sweeps, FP64 moments/macro/synthetic */

/* kinetic kernels (gsis_params.split_kernels); every choice gives the same results */
#define GSIS_KERNEL_AUTO 0    /* one-pass column-wavefront kernel, 8 columns per workgroup, when the mesh fits */
#define GSIS_KERNEL_SPLIT 1   /* separate RHS and LU-SGS sweep kernels */
#define GSIS_KERNEL_TILES 2   /* one-pass velocity-tile walker (at most 128 cell columns) */
#define GSIS_KERNEL_STRIP16 3 /* column-wavefront kernel with 16 columns per workgroup */

/* One boundary subface, mirrors BC_MSG_TYPE (src/COMMONDATA.f90:34-41) as
 * filled by read_bcin (src/sub_init.f90:266-278).  Node indices are 1-based. */
typedef struct gsis_bc {
  int32_t face;  /* 1 i-min, 2 j-min, 3 i-max, 4 j-max (src/sub_init.f90:269-277) */
  int32_t ist, iend, jst, jend;
  int32_t code;  /* GSIS_BC_* */
  double uw, vw; /* diffuse-wall velocity in sqrt(2RT) units; the reference
                    hard-codes 0 (src/GKUA_boundary.f90:66-67) */
} gsis_bc;

/* Everything the reference reads from module globals during one step. */
typedef struct gsis_params {
  int32_t nx, ny;   /* node counts; cells i=1..nx-1, j=1..ny-1 */
  int32_t nvi, nvj; /* discrete velocities per direction (NVIt, NVJt) */
  const double *vi, *vj; /* nodes B%vi(1:nvi), B%vj(1:nvj) */
  /* effective 1-D weights: the 2-D quadrature weight of (iv,jv) is wi[iv]*wj[jv].
     Gauss-Legendre: wi = FBAVX(ivp)*wi_GL (src/GKUA_DETFP.f90:314);
     Gauss-Hermite: wi = wi_GH*exp(vi^2) (src/GKUA_DETFP.f90:200-202). */
  const double *wi, *wj;
  /* node coordinates with LAP=2 ghost layers, Fortran x(1-2:nx+2,1-2:ny+2),
     as produced by Update_coordinate_buffer_onemesh (src/sub_init.f90:719-823) */
  const double *x, *y;
  /* geometry, Fortran (nx,ny) arrays from Creat_Mesh (src/sub_init.f90:183-208) */
  const double *vol, *si, *sj, *ni1, *ni2, *nj1, *nj2;
  int32_t nbc;
  const gsis_bc *bc; /* in the order of generic.inp (the order the BCs are applied) */
  /* kinetic constants (src/sub_init.f90:14,65-66; src/sub_init.f90:402) */
  double cfc, x1s, cpr, twall, dt;
  /* NS-side constants (src/sub_init.f90:16,32-35; src/opencfd_ec2d.f90:134) */
  double ma, gamma, re, pr, cv, cp, aoa;
  int32_t scheme;     /* GSIS_SCHEME_* (Iflag_Scheme, src/COMMONDATA.f90:18) */
  int32_t mode;       /* GSIS_MODE_* */
  int32_t precision;  /* GSIS_PREC_* */
  int32_t if_viscous; /* If_viscous (src/opencfd_ec2d.f90:137) */
  /* synthetic NS solve (mode M1): local time step and inner-loop control
     (src/sub_NS_singlegrid.f90:160-183, src/Hybrid_Sub.f90:99-104) */
  int32_t local_dt;
  double cfl, dtmax, dtmin;
  double ns_tol;    /* inner loop runs while max |Res_4| > ns_tol (reference 1e-6) */
  int32_t ns_maxit; /* safety cap on inner iterations (0 = 100000) */
  /* residual cadence (Kstep_show, src/opencfd_ec2d.f90:47): the kinetic RHS
     of a step is kept for gsis_residual only when the step count after it is a
     multiple of kstep_show; 0 keeps it after every step */
  int32_t kstep_show;
  int32_t split_kernels; /* GSIS_KERNEL_*: which kinetic kernels run (a one-pass kernel that does not
                            fit the mesh falls back to the next: strips, tile walker, split) */
  int32_t walk_tiles;    /* velocity tiles per workgroup of the tile walker; 0 = automatic */
} gsis_params;

typedef struct gsis_handle gsis_handle;

/* Create the device state.  Replaces the allocation part of Alloc_GKUA_FRD
 * (src/sub_init.f90:329-386); the state starts zeroed like the reference's. */
int gsis_create(const gsis_params *p, gsis_handle **out);

/* Upload B%FRD, Fortran FRD(2,-3:nx+3,-3:ny+3,nvi,nvj).  Used after the
 * initial Maxwellian fill (src/GKUA_sudufenbu.f90:22-46) and on restart
 * (src/Backup_write_read.f90:93-109). */
int gsis_upload_frd(gsis_handle *h, const double *frd);

/* Upload S_GKUA(11,-3:nx+3,-3:ny+3), CF(-1:nx+1,-1:ny+1), U(4,-1:nx+1,-1:ny+1).
 * u0 (the synthetic solver's state B%U0, same shape as U) may be NULL; the
 * reference initialises it to (1,0,0,1) (src/sub_init.f90:159). */
int gsis_upload_macro(gsis_handle *h, const double *s, const double *cf, const double *u,
                      const double *u0);

/* Kinetic BCs + moments of the current distribution, the pair that
 * GKUA_sudufenbu runs after the Maxwellian fill (src/GKUA_sudufenbu.f90:43,53):
 * GKUA_Phy_boundary then DETFP (with DETFPwall). */
int gsis_refresh_moments(gsis_handle *h);

/* nsteps outer iterations of Hybrid_Time_advance_LU_SGS (src/Hybrid_Sub.f90:4-160).
 * Asynchronous: returns once the work is queued on the handle's stream. */
int gsis_step(gsis_handle *h, int32_t nsteps);

/* RMS of the four conserved moments of the kinetic RHS of the last step
 * whose RHS was kept (see kstep_show), B%Res_rms(1:4) of DETFP_res
 * (src/GKUA_DETFP_RES.f90:45-54).
 * Returns GSIS_ENAN when res_rms[0] is NaN (src/IO.f90:236-244), and
 * GSIS_EARG when no step of this handle has kept its RHS yet. */
int gsis_residual(gsis_handle *h, double res_rms[4]);

/* Download macro state.  Any pointer may be NULL.  Shapes (Fortran):
 *   s, s0   : (11,-3:nx+3,-3:ny+3)     S_GKUA, S0_GKUA
 *   cf      : (-1:nx+1,-1:ny+1)
 *   u       : (4,-1:nx+1,-1:ny+1)
 *   taonsi_g, taonsj_g : (4,-1:nx+1,-1:ny+1)   qi_g, qj_g : (2,-1:nx+1,-1:ny+1)
 *   swall   : for each wall/farfield BC in list order, (10, nk) with nk the
 *             BC's node count along the face (BC%Swall(1:10,...)). */
int gsis_download_macro(gsis_handle *h, double *s, double *s0, double *cf, double *u,
                        double *taonsi_g, double *taonsj_g, double *qi_g, double *qj_g,
                        double *swall);

/* Download B%FRD (same layout as gsis_upload_frd), for backup_Mesco_write
 * (src/Backup_write_read.f90:150-209). */
int gsis_download_frd(gsis_handle *h, double *frd);

/* gsis_download_frd without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The transpose goes through a 32 MiB device stage that the library holds in
 * a file-scope array on each GPU until the process ends, a chunk of whole
 * velocities at a time.  Returns GSIS_EARG when one velocity of FRD,
 * 2(nx+7)(ny+7) doubles, does not fit the stage.  Calls on handles of the same
 * GPU share that stage: calls that overlap in time on one GPU race. */
int gsis_download_frd_noalloc(gsis_handle *h, double *frd);

/* Step counter and accumulated time (Mesh(1)%Kstep, Mesh(1)%tt,
 * src/Hybrid_Sub.f90:156-157) and, in mode M1, the last inner iteration count
 * (MP%iter). */
int gsis_counters(gsis_handle *h, int32_t *kstep, double *tt, int32_t *ns_iter);

/* Restore the step counter and accumulated time that init_flow_read takes
 * from restart{1,2}/flag.dat (src/Backup_write_read.f90:20-23,29-32), after
 * gsis_upload_frd and gsis_refresh_moments have restored the distribution
 * (backup_Mesco_read, :93-109).  The residual cadence (kstep_show) follows the
 * restored count.  kstep must be >= 0. */
int gsis_set_counters(gsis_handle *h, int32_t kstep, double tt);

/* GKUA_inner_mesco_boundary + GKUA_inner_macro_boundary (src/GKUA_boundary.f90:400-633)
 * for one inner subface: the LAP = 2 cell layers next to subface msub of block
 * src are copied into the ghost layers of subface ksub of block dst, FRD(1:2)
 * for every velocity and S_GKUA(1:10).  ksub and msub are 1-based positions in
 * each handle's BC list; ksub's code must be >= 1.  orient == 2 keeps the
 * direction along the face, any other value reverses it (Bc%orient from
 * Get_neighb_info, src/sub_init.f90:636-693).  Both faces must have the same
 * cell count, and both handles the same precision, velocity grid and device;
 * src may equal dst.  Asynchronous, on dst's stream, ordered after the work
 * already queued on src's stream and before src's later work.  The reference
 * runs it when a restart is read (src/Backup_write_read.f90:106-107), after the
 * distribution of every block is in place. */
int gsis_inner_exchange(gsis_handle *dst, int32_t ksub, gsis_handle *src, int32_t msub, int32_t orient);

/* Kernel timing with HIP events on the handle's stream.  No counterpart in
 * the reference: it exists for benchmarks.  enable=1 clears and starts the
 * accumulators, 0 clears and stops them, -1 only reads.  ms[0..3] receive the
 * summed durations (ms) of the kinetic RHS kernel, the LU-SGS sweep kernel
 * (for the one-pass kernels: the whole kinetic step), the moments kernel and,
 * in mode M1, the synthetic NS solve over *nsteps timed steps. */
int gsis_timing(gsis_handle *h, int32_t enable, double *ms, int32_t *nsteps);

int gsis_destroy(gsis_handle *h);

const char *gsis_last_error(void);

/* MassCorrect (src/GKUA_DETFP.f90:153-172), whose call the reference's step
 * keeps commented out (src/Hybrid_Sub.f90:42): with on != 0, after the
 * moments of every later step, S_GKUA(1) of every interior cell is scaled by
 * Init_mass/out_mass, out_mass being the sum of S_GKUA(1)*vol over the
 * interior cells (src/GKUA_DETFP.f90:134-138).  Init_mass is the out_mass of
 * the current state, as the reference sets it after a restart
 * (src/Backup_write_read.f90:102-104).  Off by default (the step as shipped);
 * set it before creating a batch that holds the handle. */
int gsis_set_mass_correct(gsis_handle *h, int32_t on);

/* ---- Learned closure of the high-order terms (HoT) ------------------------
 * No counterpart in the reference.  Convert_NS computes the HoT that the
 * synthetic NS equations add to their stress and heat flux from the DVM
 * moments (src/Hybrid_Sub.f90:317-349, :419-448); this network predicts the
 * five that the synthetic solve uses, HoT tau_11, tau_22, tau_12 and q_x, q_y
 * on an NS face, from 8 local inputs: the strain rates du/dx, dv/dy,
 * du/dy + dv/dx and the temperature gradient dT/dx, dT/dy of the face (the
 * gradients of Convert_NS), density and temperature averaged over the two
 * cells, and log10 Kn.  4 hidden layers of 128 SiLU units; BF16 weights and
 * activations, FP32 accumulation (MFMA), FP32 master weights and Adam.
 * Inputs and outputs are standardised with per-feature mean and standard
 * deviation (gsis_mlp_set_scaling).  Arrays are row-major, one sample per row. */
#define GSIS_MLP_NIN 8
#define GSIS_MLP_NOUT 5
#define GSIS_MLP_WIDTH 128
#define GSIS_MLP_HIDDEN 4
typedef struct gsis_mlp gsis_mlp;

/* A network with all parameters zero, on the current device. */
int gsis_mlp_create(gsis_mlp **out);
/* Number of parameters: W1 (128x8), b1, W2, b2, W3, b3, W4 (128x128 each),
 * b4, W5 (5x128), b5 in this order, each W row-major (output by input). */
int gsis_mlp_nparams(void);
int gsis_mlp_set_params(gsis_mlp *m, const float *params);
int gsis_mlp_get_params(gsis_mlp *m, float *params);
/* scaling[0:8] input means, [8:16] input standard deviations, [16:21] output
 * means, [21:26] output standard deviations */
int gsis_mlp_set_scaling(gsis_mlp *m, const float *scaling);
/* y (n x 5) = network(x (n x 8)), in the features' and HoT's own units. */
int gsis_mlp_forward(gsis_mlp *m, const float *x, int32_t n, float *y);
/* Mean squared error of the standardised outputs against y (n x 5) and its
 * gradient with respect to the parameters (nparams floats), n <= 8192. */
int gsis_mlp_loss_grad(gsis_mlp *m, const float *x, const float *y, int32_t n, double *loss, float *grad);
/* Adam on the mean squared error of the standardised outputs: epochs over
 * the n samples in shuffled batches of `batch` (<= 8192, a multiple of 64),
 * learning rate lr; loss_per_epoch (epochs doubles, may be NULL) receives each
 * epoch's mean batch loss. */
int gsis_mlp_train(gsis_mlp *m, const float *x, const float *y, int32_t n, int32_t epochs, int32_t batch, float lr,
                   uint64_t seed, double *loss_per_epoch);
int gsis_mlp_destroy(gsis_mlp *m);

/* Closure faces of a handle's mesh: the NS faces whose gradient stencil holds
 * only physical cells, i-faces i = 2..nx-1 on rows j = 2..ny-2 (row-major, i
 * fastest), then j-faces i = 2..nx-2, j = 2..ny-1.  Their 8 network inputs,
 * computed from the macroscopic state of the last Convert_NS, and the 5 HoT
 * of that Convert_NS (the network's targets): x (cap x 8), y (cap x 5); *n
 * receives the number of closure faces; x and y may be NULL to query it.
 * Fails with GSIS_EARG when cap < *n. */
int gsis_closure_samples(gsis_handle *h, float *x, float *y, int64_t cap, int64_t *n);
/* Use the network in the synthetic NS solve (mode M1) of later steps:
 * every `every` inner iterations the HoT of each accepted closure face is set
 * to HoT_DVM + N(state now) - N(state of Convert_NS), N the network, so the
 * synthetic equations follow how the HoT change with the macroscopic state;
 * the fixed point is the one without the network.  A face is accepted when
 * the network reproduces its DVM HoT at the step start within `gate`
 * standard deviations of the training targets in every output; the others
 * keep the DVM HoT (the fallback).  Residual check: when a launch after an
 * update ends with T < 0.1 or with max |Res_4| more than 1.5 times that of
 * the launch before, the step falls back to the DVM HoT on every face and
 * continues the inner loop from its starting state.  net NULL switches the
 * closure off.  The network must outlive its use and is not used
 * concurrently by two handles. */
int gsis_set_closure(gsis_handle *h, gsis_mlp *net, double gate, int32_t every);
/* Warm start: in the synthetic solve of the next nsteps steps (mode M1) the
 * HoT of every closure face with a finite prediction are the network's,
 * N(state now), refreshed every `every` inner iterations, in place of the DVM
 * HoT (which at the start of a run, from a Maxwellian, only cancel the NS
 * stresses).  The same residual check as gsis_set_closure falls back to the
 * DVM HoT.  Later steps are plain (or use gsis_set_closure when set), so the
 * fixed point is the plain one.  net NULL or nsteps <= 0 cancels. */
int gsis_set_closure_warmup(gsis_handle *h, gsis_mlp *net, int32_t nsteps, int32_t every);
/* Closure faces, accepted faces and HoT refreshes of the last step, and
 * whether its residual check fell back to the DVM HoT (1) or not (0). */
int gsis_closure_stats(gsis_handle *h, int64_t *faces, int64_t *accepted, int32_t *refreshes, int32_t *rejected);

/* ---- Initial distribution and batches of instances ----------------------
 * No counterparts in the reference. */

/* The Maxwellian fill of GKUA_sudufenbu (src/GKUA_sudufenbu.f90:22-46) for a
 * uniform initial state, on the device: every cell's G and H set to g[v] and
 * h[v] (nvi*nvj values each, velocity iv fastest), the wall slots zero; the
 * same device state as gsis_upload_frd of the FRD array that fill makes,
 * without the host array. */
int gsis_fill_frd_uniform(gsis_handle *h, const double *g, const double *hv);

/* A batch steps n handles together: the kinetic step of all instances is one
 * launch of the column-wavefront kernel over every instance's strips, and
 * the step of all instances up to the synthetic solve (S0, kinetic BCs, that
 * launch, moments, DETFPwall, Convert_NS, HoT) is captured once as a hipGraph
 * and replayed per step.  The synthetic solves of instances in mode M1 run
 * either inside the graph, all in one launch with one workgroup per
 * instance, or after the graph, one cooperative launch per instance (the
 * only way for instances with a closure network or more than 256 swept rows).
 * Inside is the default when at least 32 instances qualify; gsis_batch_tune
 * times the other placement too, gsis_batch_set_ns_graph chooses it.  Both
 * give the same results.  Instances may differ in mesh, velocity grid, gas, walls and mode; they
 * must share precision, flux scheme and device, and fit the
 * column-wavefront kernel.  A batch step keeps no kinetic RHS (gsis_residual
 * reports the last one a gsis_step kept).  The handles must not be stepped
 * on their own while a gsis_batch_step runs; they may be between calls. */
typedef struct gsis_batch gsis_batch;
int gsis_batch_create(gsis_handle *const *hs, int32_t n, gsis_batch **out);
int gsis_batch_step(gsis_batch *b, int32_t nsteps);
/* Autotuner: runs steps_per_trial real batch steps with each strip width
 * (8 or 16 columns per workgroup) and job order (as given, or the largest
 * kinetic work first), then with the best of those and the other placement
 * of the synthetic solves (when any qualifies), keeps the fastest (every
 * choice computes the same step) and reports it with its time per step. */
int gsis_batch_tune(gsis_batch *b, int32_t steps_per_trial, int32_t *nw, int32_t *order, double *ms_per_step);
/* Placement of the M1 instances' synthetic solves: -1 automatic (inside the
 * graph when at least 32 instances qualify), 0 after the graph, 1 inside the
 * graph for every instance that qualifies; the step is captured again. */
int gsis_batch_set_ns_graph(gsis_batch *b, int32_t mode);
/* The number of instances whose synthetic solve runs inside the graph. */
int gsis_batch_ns_graph(const gsis_batch *b, int32_t *count);
int gsis_batch_destroy(gsis_batch *b);

#ifdef __cplusplus
}
#endif

#endif /* GSIS_CAPI_H */
