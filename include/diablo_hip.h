/*
 * diablo_hip.h -- C ABI of libdiablo_hip.so, the MI355X (gfx950) implementation
 * of DIABLO's channel timestep RK_CHAN_1 and its callees.
 *
 * Each entry point names the reference interface it replaces
 * (paths relative to diablo_2.0/for/ of vvslaroche/DIABLO).
 *
 * Conventions
 *   - All reals are IEEE double, all integers int32.
 *   - Host arrays passed to dh_upload/dh_download use the reference's Fortran
 *     layout (first index fastest, lower bound 0):
 *       Fourier fields  CA(0:NXP, 0:NZ+1, 0:NY+1) COMPLEX*16, NXP = NX/2
 *       physical fields  A(0:NX+1, 0:NZ+1, 0:NY+1) REAL*8
 *       scalars add a trailing (1:N_TH) index       (header:212-258).
 *     The library copies; it never keeps a host pointer.
 *   - Entry points whose names end in _dev take device pointers in the
 *     library's own device layout (below) and run on the context's stream.
 *   - Return value: 0 on success, a negative DH_E* code otherwise; the text
 *     of the last error is returned by dh_last_error().
 *   - One context per process and device; not thread-safe.
 *
 * Device layout (one GPU, replaces the NPROCY x NPROCZ decomposition of
 * mpi.f:868-1081):
 *   Fourier  [NY+2][TNKZ+1][NKX+1] complex double (kx fastest): only the
 *            de-aliased modes I<=NKX, K<=TNKZ; the reference keeps the rest
 *            exactly zero (mpi.f:1501-1506) or never reads them.
 *   Physical [NY+2][NZ][NX] double.
 */
#ifndef DIABLO_HIP_H
#define DIABLO_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DH_MAX_TH 4

enum {
  DH_OK = 0,
  DH_EARG = -1,     /* invalid argument / unsupported configuration */
  DH_EHIP = -2,     /* HIP runtime error */
  DH_ESTATE = -3,   /* call out of order (e.g. no grid yet) */
  DH_ELAYOUT = -4,  /* host array violates the de-aliased layout */
};

/* Run-time parameters: grid_def (NX,NY,NZ,N_TH), /INPUT/, /SCALAR_VARS/,
 * /GEOPHYS_VARS/, /OSCILL_VARS/, /VARIABLE_VISC/, /DRHODX_VARS/ of
 * header:35-190.  BC types: 0 Dirichlet, 1 Neumann (channel.f:1660-2282). */
typedef struct dh_params {
  int32_t nx, ny, nz, n_th;
  double nu, beta, lx, lz, nu_v_scale;
  double delta_t, cfl;
  int32_t variable_dt, update_dt;
  int32_t f_type;
  double px0, omega0, amp_omega0;
  double i_ro, grav_x, grav_y, grav_z;
  int32_t u_bc_ymin, v_bc_ymin, w_bc_ymin, u_bc_ymax, v_bc_ymax, w_bc_ymax;
  double u_bc_ymin_c1, v_bc_ymin_c1, w_bc_ymin_c1;
  double u_bc_ymax_c1, v_bc_ymax_c1, w_bc_ymax_c1;
  int32_t th_bc_ymin[DH_MAX_TH], th_bc_ymax[DH_MAX_TH];
  double th_bc_ymin_c1[DH_MAX_TH], th_bc_ymax_c1[DH_MAX_TH];
  double ri[DH_MAX_TH], pr[DH_MAX_TH];
  double drhodx[DH_MAX_TH], drhodz[DH_MAX_TH];
  int32_t create_new_flow;
} dh_params;

/* Host-side state in Fortran layout (/GLOBAL_VARS/, header:260-261).
 * Any pointer may be NULL to skip that field.  cth/cfth/th hold N_TH
 * consecutive fields. */
typedef struct dh_host_state {
  double *cu1, *cu2, *cu3, *cp; /* COMPLEX*16 (0:NXP,0:NZ+1,0:NY+1) */
  double *cf1, *cf2, *cf3;      /* RK carry of the explicit terms */
  double *cth, *cfth;           /* COMPLEX*16 (...,1:N_TH) */
  double *th;                   /* REAL*8 (0:NX+1,0:NZ+1,0:NY+1,1:N_TH) */
} dh_host_state;

typedef struct dh_ctx dh_ctx;
typedef struct dh_ens dh_ens;

/* INIT_FFT (fft.f:55-143: wavenumbers, transform set-up), INIT_CHAN_MPI
 * index rules on one rank (mpi.f:1013-1081) and the RKW3 constants
 * (diablo_io.f:116-124).  Allocates all device state on the current device. */
int dh_create(const dh_params *p, dh_ctx **out);
/* Frees the device state (END_RUN_MPI, mpi.f:1668-1679, has no device part). */
int dh_destroy(dh_ctx *ctx);
/* Text of the last error on this context ("" if none). */
const char *dh_last_error(dh_ctx *ctx);

/* /GRID_PARAMS/ (header:205-207): GY, GYF, DY, DYF each (0:NY+1), as built by
 * CREATE_GRID_CHAN (channel.f:1564-1655) + ReadGridHDF5 (hdf5.f:548-710). */
/* Reads the float64 dataset `dataset` (e.g. "/grids/y") of an HDF5 file
 * into out (cap values) and sets *n to its number of values.  Covers what
 * the reference's grid generator (pre_process/create_grid_h5.m) writes and
 * ReadGridHDF5 (hdf5.f:548-710) reads: superblock version 0 or 1,
 * symbol-table groups, contiguous or compact little-endian float64 data;
 * anything else returns DH_ELAYOUT with the reason in err (errlen bytes).
 * Needs no context and no GPU. */
int dh_h5_read_f64(const char *path, const char *dataset, double *out, int64_t cap, int64_t *n, char *err,
                   int32_t errlen);

/* CREATE_GRID_CHAN (channel.f:1564-1655) with ReadGridHDF5 (hdf5.f:548-710):
 * GY(1..NY+1) from /grids/y of the file at `path`, which must hold NY+1
 * values ("Grid file and program do not match" otherwise); GYF(1..NY) =
 * (GY(J)+GY(J+1))/2 with the ghosts GYF(0), GYF(NY+1) extrapolated; DY, DYF
 * as channel.f:1646-1652; then dh_set_grid. */
int dh_set_grid_h5(dh_ctx *ctx, const char *path);

int dh_set_grid(dh_ctx *ctx, const double *gy, const double *gyf, const double *dy, const double *dyf);

/* Copy the host state in / out (the COMMON arrays RK_CHAN_1 reads and
 * writes, channel.f:62-72).  Upload fails with DH_ELAYOUT if a mode outside
 * the de-aliased set is non-zero in CU*, CP, CF*, CTH or CFTH. */
int dh_upload(dh_ctx *ctx, const dh_host_state *st);
int dh_download(dh_ctx *ctx, dh_host_state *st);

/* /RUNTIME_VARS/ TIME, TIME_STEP and DELTA_T (header:115-117); setting
 * DELTA_T recomputes H_BAR(1:3) as diablo_io.f:116-118.  The device keeps
 * its own copy, which COURANT (courant.f:1-65) and diablo.f:109 update. */
int dh_set_clock(dh_ctx *ctx, double time, double delta_t, int32_t time_step);
int dh_get_clock(dh_ctx *ctx, double *time, double *delta_t, double *h_bar3, int32_t *time_step);

/* RK_CHAN_1 (channel.f:62-1073), one RK3/CN substep rk_step in 1..3 on the
 * device state, LES off, empty USER_RHS hooks.  COURANT runs inside when
 * VARIABLE_DT, rk_step==3 and MOD(TIME_STEP,UPDATE_DT)==0 (channel.f:1037).
 * Asynchronous on the context stream. */
int dh_rk_substep(dh_ctx *ctx, int32_t rk_step);
/* nsteps iterations of the loop body of diablo.f:66-110 without the I/O
 * hooks: TIME_STEP+=1, RK_CHAN_1 for RK_STEP=1..3, TIME=TIME+DELTA_T.
 * Replays a captured hipGraph of one step after the first call. */
int dh_time_step(dh_ctx *ctx, int32_t nsteps);
/* The FP32/FP64 mixed form of RK_CHAN_1 (no counterpart in the reference,
 * which is FP64 throughout): with on = 1 the substeps of dh_rk_substep and
 * dh_time_step run the x-z transforms, the nonlinear products and the
 * physical work arrays in FP32, while the Fourier state (CU, CP, CF, CTH,
 * CFTH), the accumulation of CF, the RK combination, the Crank-Nicolson and
 * pressure-Poisson solves and COURANT stay FP64.  on = 0 returns to the FP64
 * drop-in path.  The first switch allocates two FP32 physical arrays of
 * (NY+2)*NX*NZ*(3+N_TH) values. */
int dh_set_mixed_precision(dh_ctx *ctx, int32_t on);
/* 1 when the mixed form is on, 0 when not. */
int dh_mixed_precision(dh_ctx *ctx);

/* Constant-Smagorinsky LES, the reference's LES = .TRUE. with
 * LES_MODEL_TYPE = 1 (header:168, les.f:1-336, channel.f:303-375 and the
 * NU_T terms of channel.f:631-659, 748-760, 793-807, 862-875, 927-941,
 * 991-1005): on = 1 runs les_chan in every substep once the spin-up gate of
 * channel.f:303 is open (CREATE_NEW_FLOW = 0, or TIME_STEP > 100) and the
 * eddy-viscosity (and, for scalars, KAPPA_T = NU_T) terms in the implicit
 * solves always.  FP64 only (refused while the mixed form is on).  The first
 * switch allocates NU_T, a copy of CU and the model's work fields:
 * 24 + 5 N_TH fields of (NY+2) NX NZ doubles and 13 + 3 N_TH of (NY+2)
 * (NKX+1)(2 NKZ+1) complex values (97 MB at 64x97x64, N_TH = 0).
 * The periodic mean_les output (les.f:338-517) is not produced. */
int dh_set_les(dh_ctx *ctx, int32_t on);
/* 1 when LES is on, 0 when not. */
int dh_les(dh_ctx *ctx);
/* A learned closure in place of the Smagorinsky eddy viscosity (no
 * counterpart in the reference; it replaces les_chan's NU_T = C_s^2
 * Delta^2 |S|, les.f:196-227, for drivers that train one): NU_T at the GY
 * point of plane J = NUS softplus(y) on J = 3..NY-1 (zero on the rows
 * GHOST_LES_MPI zeroes), y the network of dh_mlp_* below on the point's
 * inputs: asinh of the five invariants of the resolved strain S and
 * rotation O, tr S^2 TS^2, tr O^2 TS^2, tr S^3 TS^3, tr O^2 S TS^3,
 * tr O^2 S^2 TS^4, and asinh of the wall distance (1 - |y|) / LS.  The
 * subgrid stress is -2 NU_T S_ij throughout (NU_T at GYF the mean of the two
 * GY values around it), and the implicit and scalar terms take the same
 * NU_T.  w: the network's dh_mlp_nparams() parameters, then TS, LS and NUS
 * (n = dh_mlp_nparams() + 3); NULL returns to the Smagorinsky model.  The
 * inputs of the last substep, then S11, S22, S33, S12, S13, S23 and |S| at
 * GY, are dh_field_dev(ctx, 301), [13][NY+2][NZ][NX]. */
int dh_set_les_closure(dh_ctx *h, const float *w, int32_t n);
/* The closure's inputs (dh_field_dev(ctx, 301)) from the current state,
 * as the first kernels of an LES substep form them, without stepping: for
 * building training samples from filtered fields.  Needs LES and the
 * closure on and the spin-up gate open. */
int dh_les_features(dh_ctx *h);

/* The learned closure's network (no counterpart in the reference): 6
 * inputs -> 64 -> 64 -> 64 -> 1, GELU (tanh form) hidden layers, softplus
 * output, BF16 operands with FP32 accumulation on gfx950's 16x16x32 BF16
 * MFMA.  Parameters are dh_mlp_nparams() FP32 values: W1[32][64] (rows 6..31
 * unused), b1[64], W2[64][64], b2[64], W3[64][64], b3[64], w4[64], b4, with
 * W[in][out].  Device pointers; asynchronous on `stream`; errors in
 * dh_mlp_last_error(). */
const char *dh_mlp_last_error(void);
int32_t dh_mlp_nparams(void);
/* Bytes of the packed BF16 operand buffer and of the gradient workspace
 * for a batch of n samples. */
int64_t dh_mlp_packed_bytes(void);
int64_t dh_mlp_work_bytes(int64_t n);
/* Packs params into the operand order the kernels read. */
int dh_mlp_pack_dev(const float *params, void *packed, void *stream);
/* out[i] = scale * softplus(y(x_i)), i < n, x feature-major: x[f * ld + i]. */
int dh_mlp_infer_dev(const void *packed, const float *params, const float *x, int64_t ld, int64_t n, float scale,
                     float *out, void *stream);
/* Loss sum_i (softplus(y_i) - t_i)^2 / n and its gradient (dh_mlp_nparams()
 * values) for a batch of n samples (a multiple of 32, at most 2^24), x
 * feature-major [6][n]: the forward and backward pass on MFMA, the weight
 * gradients as MFMA products over the batch, summed in a fixed order. */
int dh_mlp_grad_dev(const float *params, const void *packed, const float *x, const float *target, int64_t n,
                    void *work, int64_t work_bytes, float *grad, float *loss, void *stream);
/* One Adam step (beta 0.9, 0.999, eps 1e-8; step counts from 1) on params
 * with moments m, v, then the packing of the new params. */
int dh_mlp_adam_dev(float *params, float *m, float *v, const float *grad, float lr, int32_t step, void *packed,
                    void *stream);

/* Launch-shape tuning (no counterpart in the reference; for ensembles of
 * uneven sizes, where the best block size depends on the plane's shape).
 * dh_set_plane_threads sets the block size of the plane transform kernels
 * and of the nonlinear pass's LDS variant: a multiple of 64 in 64..1024
 * (64..512 for NX = 64 planes of at most 2048 modes, whose nonlinear pass
 * takes its x rows in registers) with 4 x threads >= the plane's
 * de-aliased Fourier modes.  Results do not depend on it.  dh_autotune times nsteps steps (best of two) with each
 * power-of-two size from the current state, restores the state (Fourier
 * fields, CF, CP, scalars, clock and CN tables) and keeps the fastest,
 * returned in *chosen. */
int dh_set_plane_threads(dh_ctx *ctx, int32_t threads);
int dh_plane_threads(dh_ctx *ctx);
int dh_autotune(dh_ctx *ctx, int32_t nsteps, int32_t *chosen);
/* Per-phase timing of one substep (the reference's only timer is the
 * wall-clock "Seconds per Iteration" of diablo.f:123-127): runs substep
 * rk_step nrep times with HIP events between its kernels on the context
 * stream and returns the mean milliseconds of each phase in ms[0..5]
 * (nms >= 6); the phases are named by dh_phase_name().  Advances the state. */
int dh_profile_substep(dh_ctx *ctx, int32_t rk_step, int32_t nrep, double *ms, int32_t nms);
const char *dh_phase_name(int32_t phase);
/* Wait for all work queued on the context. */
int dh_synchronize(dh_ctx *ctx);

/* Ensembles of independent instances of any sizes (many shear layers at
 * once): no counterpart in the reference, which runs one flow per process.
 * The ensemble steps existing contexts (it does not own them; they must
 * outlive it and each may appear once).  dh_ensemble_time_step runs nsteps
 * iterations of diablo.f:66-110 on every member: each member's captured
 * step graph on the member's own stream, every member's step launched
 * before the next step of any, so members overlap on the device.  Work
 * queued on a member stays ordered with it, so each member's own entry
 * points remain usable between calls, and results equal those of
 * dh_time_step on each member alone. */
int dh_ensemble_create(int32_t n, dh_ctx *const *members, dh_ens **out);
int dh_ensemble_time_step(dh_ens *ens, int32_t nsteps);
int dh_ensemble_synchronize(dh_ens *ens);
/* dh_autotune on the first member of each shape; the other members of that
 * shape take the same block size, so that they still run as one batch. */
int dh_ensemble_autotune(dh_ens *ens, int32_t nsteps);
int dh_ensemble_destroy(dh_ens *ens);

/* REM_DIV_CHAN (channel.f:1103-1214 with APPLY_BC_REM_DIV_MPI
 * mpi.f:1239-1278 and THOMAS_*_COMPLEX_MPI mpi.f:708-865) on the device
 * state; leaves phi in the context's phi buffer. */
int dh_rem_div(dh_ctx *ctx);

/* POISSON_P_CHAN (channel.f:1217-1453 with APPLY_BC_POISSON_MPI
 * mpi.f:1280-1310 and THOMAS_*_COMPLEX_MPI mpi.f:708-865): CP on planes
 * 1..NY from the current CU (and CTH), as the driver computes it at a
 * restart that has no pressure (hdf5.f:537).  CU, CF and CTH are left
 * unchanged (the reference overwrites CF1..3 as scratch); the context's
 * phi buffer and physical velocity buffers are overwritten. */
int dh_poisson_p_chan(dh_ctx *ctx);

/* CREATE_FLOW_CHAN (set_ics.f:2-200) on the device: IC_TYPE 0 (laminar
 * Poiseuille, 3/2 UBULK0 (1-GYF^2)), 2 (Couette), 3 (tanh shear layer), plus
 * Gaussian noise of standard deviation amp on U1..U3 wherever |GYF| < ymask
 * (ymask <= 0: everywhere), drawn from a counter-based generator seeded by
 * `seed` in place of the compiler's RANDOM_NUMBER; then the transform to
 * Fourier space, APPLY_BC_VEL_MPI and REM_DIV_CHAN as set_ics.f:115-186.
 * CP is set to zero. */
int dh_create_flow_chan(dh_ctx *ctx, int32_t ic_type, double ubulk0, uint64_t seed, double amp, double ymask);
/* CREATE_TH_CHAN (set_ics.f:204-273) for IC_TYPE 0..4 on the device. */
int dh_create_th_chan(dh_ctx *ctx, int32_t ic_type);
/* APPLY_BC_VEL_MPI (mpi.f:1317-1331 -> channel.f:2285-2492): ghost and wall
 * planes of CU1..3 in Fourier space. */
int dh_apply_bc_vel(dh_ctx *ctx);

/* Plane statistics of SAVE_STATS_CHAN (save_stats.f:136-262) on the device
 * from the current Fourier state, which is left unchanged (the physical
 * velocity buffers are overwritten; the next substep recomputes them):
 * out[q * (NY+2) + j], j = 0..NY+1, for q in the order
 *   UME VME WME (plane means, :136-150)  URMS VRMS WRMS (:166-192)
 *   UV UW WV (Reynolds stresses, :197-232)  DUDY DWDY (:235-238)
 *   SHEAR (mean square shear, :241-262)
 * with rows the reference leaves unset at 0, and *ubulk =
 * INTEGRATE_Y_VAR(UME) over the domain height ly (mpi.f:1610-1632).
 * Plane sums are reduced in FP64 on the device. */
enum {
  DH_ST_UME = 0, DH_ST_VME, DH_ST_WME, DH_ST_URMS, DH_ST_VRMS, DH_ST_WRMS,
  DH_ST_UV, DH_ST_UW, DH_ST_WV, DH_ST_DUDY, DH_ST_DWDY, DH_ST_SHEAR,
  DH_NSTATS
};
int dh_save_stats(dh_ctx *ctx, double ly, double *out, double *ubulk);

/* FFT_XZ_TO_FOURIER (mpi.f:1477-1542) of nplanes consecutive physical
 * planes [nplanes][NZ][NX] into de-aliased Fourier planes
 * [nplanes][TNKZ+1][NKX+1] (device pointers). */
int dh_fft_xz_to_fourier_dev(dh_ctx *ctx, const double *phys, double *four, int32_t nplanes);
/* FFT_XZ_TO_PHYSICAL (mpi.f:1547-1606), the inverse direction. */
int dh_fft_xz_to_physical_dev(dh_ctx *ctx, const double *four, double *phys, int32_t nplanes);

/* Batched kernels on caller-owned device arrays (no context).  Each runs
 * asynchronously on `stream` (a hipStream_t; NULL for the null stream);
 * fp32 = 1 selects FP32 storage and arithmetic, 0 FP64.  Errors return
 * DH_E* with the text in dh_batched_last_error() (per host thread). */
const char *dh_batched_last_error(void);

/* The work of THOMAS_FORWARD_COMPLEX_MPI + THOMAS_BACKWARD_COMPLEX_MPI
 * (mpi.f:708-865) on nsys independent complex systems of the
 * Crank-Nicolson form of channel.f:914-963, whose diagonals are formed on
 * the device from one coefficient per system and two per row:
 *   row j:  -alpha_s sl_j x_{j-1} + (1 + alpha_s (sl_j + su_j)) x_j
 *           - alpha_s su_j x_{j+1} = rhs_j,      j = 0..ny-1,
 * i.e. MATL/MATD/MATU of the reference with alpha_s = NU T1 / h^2 and sl,
 * su = h^2/(DY(J) DYF(J)), h^2/(DY(J+1) DYF(J)); sl_j = su_j = 0 makes row j
 * an identity (Dirichlet) row, and the couplings of row 0 to j = -1 and of
 * row ny-1 to j = ny are dropped.  sl, su: [ny] doubles; alpha: [nsys]
 * doubles; rhs, x: [ny][nsys] complex (double2, or float2 when fp32),
 * systems fastest; x may equal rhs.  2 <= ny <= 288. */
int dh_cn_solve_batched_dev(const double *sl, const double *su, const double *alpha, const void *rhs, void *x,
                            int32_t ny, int64_t nsys, int32_t fp32, void *stream);

/* FFT_XZ_TO_FOURIER (mpi.f:1477-1542) on nplanes 512 x 512 planes without
 * its de-aliasing truncation and packing: the real-to-complex transform
 * along x divided by NX, then the forward complex transform along z divided
 * by NZ (FFTW_FORWARD, fft.f:73-74,86-87).  phys: [nplanes][512 z][512 x]
 * reals; four: [nplanes][512 kz][257 kx] complex, kz in transform order
 * (0..255, then -256..-1). */
int dh_fft512_xz_to_fourier_dev(const void *phys, void *four, int64_t nplanes, int32_t fp32, void *stream);
/* FFT_XZ_TO_PHYSICAL (mpi.f:1547-1606) on the same planes, unnormalised:
 * the backward complex transform along z, then complex-to-real along x
 * (the imaginary parts of kx = 0 and 256 ignored, as a c2r ignores them).
 * four is left unchanged (the reference's transform destroys its input);
 * the z pass goes through a staging buffer that the library allocates on
 * the current device at the first call (16 planes, 33.7 MB) and keeps. */
int dh_fft512_xz_to_physical_dev(const void *four, void *phys, int64_t nplanes, int32_t fp32, void *stream);
/* No counterpart in the reference: dh_fft512_xz_to_physical_dev for callers
 * that cannot allocate device memory (code under stream capture, tight
 * serving loops, harnesses that forbid hipMalloc).  It calls no allocator:
 * the z pass goes through a file-scope device array of 8 planes (16.8 MB on
 * each device the library runs on, held until the process ends) instead of
 * the library's staging buffer, whose address is looked up once per device
 * at the first call.  The planes pass through it 8 at a time, so the array
 * bounds the planes in flight, not the call: any nplanes >= 1 is taken, and
 * nothing in it needs clearing between calls.  Results are bit-identical to
 * dh_fft512_xz_to_physical_dev.  Calls that overlap on one GPU (two streams,
 * two threads) share the array and race. */
int dh_fft512_xz_to_physical_noalloc_dev(const void *four, void *phys, int64_t nplanes, int32_t fp32, void *stream);

/* Device pointer of one state field in the device layout, for tests and
 * benchmarks: which = 0..2 CU1..3, 3 CP, 4..6 CF1..3, 7 phi (CR1 after
 * REM_DIV), 8 + n CTH(n), 8 + N_TH + n CFTH(n); 100+c the physical
 * field c: for c < 3 the velocity transformed at the start of the last
 * substep (its solve writes over R), for c >= 3 TH as left by its solve;
 * 200+c the same field of the mixed form (float values; null until the
 * mixed form is first switched on); 300 NU_T [NY+2][NZ][NX] at GY points
 * (null until LES is first switched on); 301 the learned closure's
 * features (see dh_set_les_closure). */
double *dh_field_dev(dh_ctx *ctx, int32_t which);
/* HIP stream (hipStream_t) the context runs on. */
void *dh_stream(dh_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
