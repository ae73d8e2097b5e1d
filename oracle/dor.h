/*
 * dor.h -- CPU oracle for DIABLO's channel timestep (test infrastructure).
 *
 * THIS IS TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only tests/, the smoke()
 * check of __graft_entry__.py and the cpu_baseline leg of bench.py may load
 * it, and only as the checker.  The product (libdiablo_hip.so) never links
 * or calls it.
 *
 * It is a plain-C FP64 restatement of the reference's one-rank MPI build
 * (USE_MPI=.TRUE., NPROCY=NPROCZ=1) of RK_CHAN_1 and its callees.  Every
 * array keeps the reference's Fortran layout (first index fastest, lower
 * bound 0) and the physical/Fourier EQUIVALENCE aliasing of
 * diablo_2.0/for/header:245-255, so intermediate states can be compared
 * byte for byte with what the reference would hold.
 *
 * The reference itself cannot be built in this image: it needs MPI
 * (header_mpi:3 includes mpif.h), FFTW3 (fft.f:69) and, for its grid reader,
 * HDF5; none is installed, and stand-ins for them are not allowed.  The
 * FFTW transforms are therefore restated from their published definition
 * (unnormalised DFT, forward sign e^{-i}) with a textbook mixed-radix FFT.
 */
#ifndef DOR_H
#define DOR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DOR_MAX_TH 4

typedef struct dor_params {
  int nx, ny, nz, n_th;
  double nu, beta, lx, lz, nu_v_scale;
  double delta_t, cfl;
  int variable_dt, update_dt;
  int f_type;
  double px0, omega0, amp_omega0;
  double i_ro, grav_x, grav_y, grav_z;
  int u_bc_ymin, v_bc_ymin, w_bc_ymin, u_bc_ymax, v_bc_ymax, w_bc_ymax;
  double u_bc_ymin_c1, v_bc_ymin_c1, w_bc_ymin_c1;
  double u_bc_ymax_c1, v_bc_ymax_c1, w_bc_ymax_c1;
  int th_bc_ymin[DOR_MAX_TH], th_bc_ymax[DOR_MAX_TH];
  double th_bc_ymin_c1[DOR_MAX_TH], th_bc_ymax_c1[DOR_MAX_TH];
  double ri[DOR_MAX_TH], pr[DOR_MAX_TH];
  double drhodx[DOR_MAX_TH], drhodz[DOR_MAX_TH];
  int create_new_flow;
  /* LES (header:168): constant Smagorinsky, LES_MODEL_TYPE=1 (les.f:75-267) */
  int les;
} dor_params;

typedef struct dor_state dor_state;

/* Array identifiers for dor_array(). */
enum {
  DOR_CU1 = 0, DOR_CU2, DOR_CU3, DOR_CP, DOR_CR1, DOR_CR2, DOR_CR3,
  DOR_CF1, DOR_CF2, DOR_CF3, DOR_CS1, /* complex (0:NXP,0:NZ+1,0:NY+1) */
  DOR_CTH, DOR_CFTH, DOR_CRTH,        /* complex, x N_TH */
  DOR_TH, DOR_RTH,                    /* real (0:NX+1,0:NZ+1,0:NY+1), x N_TH */
  DOR_GY, DOR_GYF, DOR_DY, DOR_DYF,   /* real (0:NY+1) */
  DOR_KX, DOR_KX2, DOR_KZ, DOR_KZ2,   /* real (0:NXP), (0:TNKZ) */
  DOR_HBAR, DOR_BETABAR, DOR_ZETABAR, /* real (3) */
  DOR_NUT,                            /* real (0:NX+1,0:NZ+1,0:NY+1), LES only */
  DOR_NARRAYS
};

dor_state *dor_create(const dor_params *p);
void dor_destroy(dor_state *s);
/* Pointer to an array of the state (complex arrays as interleaved doubles);
 * *n receives the number of doubles. */
double *dor_array(dor_state *s, int which, int64_t *n);

/* Grid: GY(1..NY+1) given; computes GYF, ghosts, DY, DYF as
 * ReadGridHDF5 (hdf5.f:690-705) + CREATE_GRID_CHAN (channel.f:1646-1652). */
void dor_set_grid_gy(dor_state *s, const double *gy_1_to_nyp1);
/* Uniform/stretched y-grid generator of pre_process/create_grid_h5.m:28-79
 * (grid_type 1: tanh both ends, 2: cubic centre-refined). */
void dor_make_gy(int ny, double L, double cs, int grid_type, double *gy_out);

/* INIT_FFT wavenumbers (fft.f:77-100) and RKW3 constants (diablo_io.f:116-124). */
void dor_init(dor_state *s);

void dor_get_clock(dor_state *s, double *time, double *delta_t, int *time_step);
void dor_set_clock(dor_state *s, double time, double delta_t, int time_step);

/* Hot-path routines (one rank). */
void dor_fft_xz_to_fourier(dor_state *s, double *v, int jmin, int jmax);
void dor_fft_xz_to_physical(dor_state *s, double *vv, int jmin, int jmax);
void dor_rk_chan_1(dor_state *s, int rk_step);
void dor_rem_div_chan(dor_state *s);
/* POISSON_P_CHAN (channel.f:1217-1453): CP(J), J=1..NY, from CU and CTH;
 * CF1..3 and CS1 are left holding scratch, as in the reference. */
void dor_poisson_p_chan(dor_state *s);
void dor_courant(dor_state *s);
/* One full time step as diablo.f:66-110 (3 substeps, then TIME+=DELTA_T). */
void dor_time_step(dor_state *s);

/* Thomas sweeps of mpi.f:630-703,767-809 (real) and 708-764,812-865
 * (complex) on one rank, on caller-provided Fortran arrays:
 * real: A,B,C,G are (0:inx-1, 0:iny+1); complex: A,B,C real (0:inx,0:iny+1),
 * G complex interleaved. */
void dor_thomas_real(double *a, double *b, double *c, double *g, int iny, int inx);
void dor_thomas_complex(double *a, double *b, double *c, double *g, int iny, int inx);

/* Plain complex DFT helper exposed for pinning tests: n-point, sign -1/+1,
 * unnormalised, interleaved complex in/out (out of place). */
void dor_fft_c2c(int n, int sign, const double *in, double *out);

/* Initial conditions (restatement of set_ics.f:2-200 with a seeded
 * counter-based RNG in place of the compiler's RANDOM_NUMBER).
 * ic_type as IC_TYPE; noise: Gaussian, std `amp`, applied where
 * |GYF(J)| < ymask (ymask <= 0: everywhere).  Then FFT, APPLY_BC_VEL_*,
 * REM_DIV_CHAN as set_ics.f:115-186. */
void dor_create_flow_chan(dor_state *s, int ic_type, double ubulk0,
                          uint64_t seed, double amp, double ymask);
/* CREATE_TH_CHAN (set_ics.f:204-273) for ic_type 1..4. */
void dor_create_th_chan(dor_state *s, int ic_type);
void dor_apply_bc_vel(dor_state *s);
/* les_chan (les.f:1-336, one rank, LES_MODEL_TYPE=1): with velocity (and
 * CF1..3) in Fourier space, sets NU_T and adds the SGS stress divergence to
 * CF1..3; on return U1..3 are in physical space.  The periodic mean_les
 * output (les.f:338-517) is left out: it writes a file and changes no state. */
void dor_les_chan(dor_state *s);
/* LES filter widths (les.f:196-217), -2 C_SMAG^2 absorbed: dyw at GY,
 * dyfw at GYF, (0:NY+1), set for J=1..NY+1. */
void dor_les_widths(dor_state *s, double *dyw, double *dyfw);

/* Plane statistics of SAVE_STATS_CHAN (save_stats.f:136-262) from the
 * current Fourier state, without touching it: out[q*(NY+2) + j] for
 * q = DOR_ST_* and j = 0..NY+1 (rows the reference leaves unset are 0),
 * *ubulk = INTEGRATE_Y_VAR of UME over the height ly (mpi.f:1610-1632). */
enum {
  DOR_ST_UME = 0, DOR_ST_VME, DOR_ST_WME, DOR_ST_URMS, DOR_ST_VRMS, DOR_ST_WRMS,
  DOR_ST_UV, DOR_ST_UW, DOR_ST_WV, DOR_ST_DUDY, DOR_ST_DWDY, DOR_ST_SHEAR,
  DOR_NSTATS
};
void dor_save_stats(dor_state *s, double ly, double *out, double *ubulk);
/* Deterministic generator used for ICs (SplitMix64 + Box-Muller). */
double dor_gauss(uint64_t seed, uint64_t counter);

#ifdef __cplusplus
}
#endif
#endif
