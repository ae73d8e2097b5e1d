/*
 * dor.c -- CPU oracle (test infrastructure only) for DIABLO's channel
 * timestep.  Plain C99, FP64.  See dor.h for the contract and for why the
 * reference itself is not built here.
 *
 * Every routine follows one reference routine statement by statement; the
 * citation above each function names the file:line range it restates
 * (paths relative to diablo_2.0/for/).  Arrays keep the Fortran layout and
 * the physical <-> Fourier EQUIVALENCE aliasing of header:245-255.
 *
 * Loops over independent y-planes or (x,z) columns are OpenMP-parallel; each
 * iteration touches only its own plane/column, so results do not depend on
 * the thread count.
 */
#include "dor.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

/* ------------------------------------------------------------------------ */
/* FFT: textbook recursive mixed-radix Cooley-Tukey (decimation in time),    */
/* unnormalised, out of place.  Stands in for FFTW's dft/r2c/c2r plans      */
/* (fft.f:73-76,86-89), which compute the same unnormalised DFTs.           */
/* ------------------------------------------------------------------------ */

typedef struct {
  int n, sign;
  cplx *tw; /* tw[j] = exp(sign*2*pi*i*j/n) */
} fft_plan;

static void plan_init(fft_plan *p, int n, int sign) {
  p->n = n;
  p->sign = sign;
  p->tw = (cplx *)malloc(sizeof(cplx) * (size_t)n);
  const long double PI_L = 3.141592653589793238462643383279502884L;
  for (int j = 0; j < n; ++j) {
    long double a = 2.0L * PI_L * (long double)j / (long double)n;
    p->tw[j] = (double)cosl(a) + (double)(sign * sinl(a)) * I;
  }
}

static void plan_free(fft_plan *p) {
  free(p->tw);
  p->tw = NULL;
}

static int smallest_factor(int n) {
  if (n % 4 == 0) return 4;
  if (n % 2 == 0) return 2;
  for (int f = 3; f * f <= n; f += 2)
    if (n % f == 0) return f;
  return n;
}

/* y[0..n) = DFT_n of x[0], x[s], x[2s], ...  (n divides plan->n). */
static void fft_rec(const fft_plan *P, int n, const cplx *x, int s, cplx *y) {
  if (n == 1) {
    y[0] = x[0];
    return;
  }
  int p = smallest_factor(n);
  int m = n / p;
  for (int r = 0; r < p; ++r) fft_rec(P, m, x + (size_t)r * s, s * p, y + (size_t)r * m);
  int N = P->n;
  int step = N / n; /* W_n^j = tw[j*step] */
  cplx t[64];
  cplx *tt = (p <= 64) ? t : (cplx *)malloc(sizeof(cplx) * (size_t)p);
  for (int k = 0; k < m; ++k) {
    for (int r = 0; r < p; ++r) {
      cplx v = y[(size_t)r * m + k];
      int e = (int)(((long)r * k) % n);
      tt[r] = (e == 0) ? v : v * P->tw[(size_t)e * step];
    }
    for (int q = 0; q < p; ++q) {
      cplx acc = 0.0;
      for (int r = 0; r < p; ++r) {
        /* tt[r] already carries W_n^{r k}; the rest is W_n^{r q m} = W_p^{r q} */
        int e2 = (int)(((long)r * q) % p) * m;
        acc += (e2 == 0) ? tt[r] : tt[r] * P->tw[(size_t)e2 * step];
      }
      y[(size_t)k + (size_t)q * m] = acc;
    }
  }
  if (tt != t) free(tt);
}

static void fft_exec(const fft_plan *P, const cplx *in, int istride, cplx *out) {
  fft_rec(P, P->n, in, istride, out);
}

/* exposed for pinning tests */
void dor_fft_c2c(int n, int sign, const double *in, double *out) {
  fft_plan p;
  plan_init(&p, n, sign);
  fft_exec(&p, (const cplx *)in, 1, (cplx *)out);
  plan_free(&p);
}

/* ------------------------------------------------------------------------ */
/* State                                                                    */
/* ------------------------------------------------------------------------ */

struct dor_state {
  dor_params p;
  int NX, NY, NZ, NTH, NKX, NKZ, TNKZ, NXP, NZP, NXM;
  int JSTART, JEND, JSTART_TH[DOR_MAX_TH], JEND_TH[DOR_MAX_TH];
  double TIME, DELTA_T;
  int TIME_STEP, RK_STEP;
  double H_BAR[3], BETA_BAR[3], ZETA_BAR[3];
  double PI;
  /* grid, 0:NY+1 */
  double *GY, *GYF, *DY, *DYF;
  double DX1, DZ1; /* DX(I)=LX/NX, DZ(K)=LZ/NZ (channel.f:1572-1583) */
  /* wavenumbers */
  double *KX, *KX2, *KZ, *KZ2; /* 0:NXP, 0:TNKZ */
  /* fields: Fourier complex (0:NXP,0:NZ+1,0:NY+1); physical views alias */
  cplx *CU1, *CU2, *CU3, *CP, *CR1, *CR2, *CR3, *CF1, *CF2, *CF3, *CS1;
  cplx *CTH, *CFTH, *CRTH; /* x NTH, not aliased (header:256-258) */
  double *TH, *RTH;        /* real (0:NX+1,0:NZ+1,0:NY+1) x NTH */
  size_t cplane, cfield;   /* complex elements per plane / per field */
  size_t rplane, rfield;   /* real elements per plane / per field */
  fft_plan px_f, pz_f, pz_b, px_b;
  /* REM_DIV work arrays (header:226-228), persistent like COMMON */
  double *MATL_C, *MATD_C, *MATU_C;
  cplx *VEC_C;
  /* LES (header:164, header_les:5-33): NU_T at GY points (KAPPA_T, set to
   * 1*NU_T for every scalar at channel.f:313-321, is NU_T itself), the
   * strain fields CSij1..6 and CTEMP, each EQUIVALENCEd with its physical
   * view Sij / TEMP */
  double *NU_T;
  cplx *CSij[6], *CTEMP;
};

/* Fortran-layout index helpers */
#define CIDX(s, i, k, j) ((size_t)(i) + (size_t)((s)->NXP + 1) * ((size_t)(k) + (size_t)((s)->NZ + 2) * (size_t)(j)))
#define RIDX(s, i, k, j) ((size_t)(i) + (size_t)((s)->NX + 2) * ((size_t)(k) + (size_t)((s)->NZ + 2) * (size_t)(j)))
#define PHYS(ptr) ((double *)(ptr))

static void *zalloc(size_t bytes) {
  void *p = calloc(1, bytes);
  return p;
}

dor_state *dor_create(const dor_params *p) {
  if (p->nx % 2 || p->nx < 4 || p->nz < 1 || p->ny < 4 || p->n_th < 0 || p->n_th > DOR_MAX_TH)
    return NULL;
  dor_state *s = (dor_state *)zalloc(sizeof(dor_state));
  s->p = *p;
  s->NX = p->nx;
  s->NY = p->ny;
  s->NZ = p->nz;
  s->NTH = p->n_th;
  s->NKX = s->NX / 3;
  s->NKZ = s->NZ / 3;
  s->TNKZ = 2 * s->NKZ;
  s->NXP = s->NX / 2; /* header_mpi:13, NPROCZ=1 */
  s->NZP = s->NZ;
  s->NXM = s->NX - 1;
  s->DELTA_T = p->delta_t;
  s->cplane = (size_t)(s->NXP + 1) * (size_t)(s->NZ + 2);
  s->cfield = s->cplane * (size_t)(s->NY + 2);
  s->rplane = (size_t)(s->NX + 2) * (size_t)(s->NZ + 2);
  s->rfield = s->rplane * (size_t)(s->NY + 2);
  size_t nyy = (size_t)s->NY + 2;
  s->GY = zalloc(8 * nyy);
  s->GYF = zalloc(8 * nyy);
  s->DY = zalloc(8 * nyy);
  s->DYF = zalloc(8 * nyy);
  s->KX = zalloc(8 * (size_t)(s->NXP + 1));
  s->KX2 = zalloc(8 * (size_t)(s->NXP + 1));
  s->KZ = zalloc(8 * (size_t)(s->TNKZ + 1));
  s->KZ2 = zalloc(8 * (size_t)(s->TNKZ + 1));
  cplx **f[] = {&s->CU1, &s->CU2, &s->CU3, &s->CP, &s->CR1, &s->CR2,
                &s->CR3, &s->CF1, &s->CF2, &s->CF3, &s->CS1};
  for (size_t i = 0; i < sizeof(f) / sizeof(f[0]); ++i) *f[i] = zalloc(sizeof(cplx) * s->cfield);
  int nth = s->NTH > 0 ? s->NTH : 1;
  s->CTH = zalloc(sizeof(cplx) * s->cfield * nth);
  s->CFTH = zalloc(sizeof(cplx) * s->cfield * nth);
  s->CRTH = zalloc(sizeof(cplx) * s->cfield * nth);
  s->TH = zalloc(8 * s->rfield * nth);
  s->RTH = zalloc(8 * s->rfield * nth);
  s->MATL_C = zalloc(8 * (size_t)(s->NXP + 1) * nyy);
  s->MATD_C = zalloc(8 * (size_t)(s->NXP + 1) * nyy);
  s->MATU_C = zalloc(8 * (size_t)(s->NXP + 1) * nyy);
  s->VEC_C = zalloc(sizeof(cplx) * (size_t)(s->NXP + 1) * nyy);
  if (p->les) {
    s->NU_T = zalloc(8 * s->rfield);
    for (int q = 0; q < 6; ++q) s->CSij[q] = zalloc(sizeof(cplx) * s->cfield);
    s->CTEMP = zalloc(sizeof(cplx) * s->cfield);
  }
  plan_init(&s->px_f, s->NX, -1);
  plan_init(&s->px_b, s->NX, +1);
  plan_init(&s->pz_f, s->NZ, -1);
  plan_init(&s->pz_b, s->NZ, +1);
  dor_init(s);
  return s;
}

void dor_destroy(dor_state *s) {
  if (!s) return;
  free(s->GY); free(s->GYF); free(s->DY); free(s->DYF);
  free(s->KX); free(s->KX2); free(s->KZ); free(s->KZ2);
  free(s->CU1); free(s->CU2); free(s->CU3); free(s->CP); free(s->CR1);
  free(s->CR2); free(s->CR3); free(s->CF1); free(s->CF2); free(s->CF3);
  free(s->CS1); free(s->CTH); free(s->CFTH); free(s->CRTH); free(s->TH);
  free(s->RTH); free(s->MATL_C); free(s->MATD_C); free(s->MATU_C); free(s->VEC_C);
  free(s->NU_T); free(s->CTEMP);
  for (int q = 0; q < 6; ++q) free(s->CSij[q]);
  plan_free(&s->px_f); plan_free(&s->px_b); plan_free(&s->pz_f); plan_free(&s->pz_b);
  free(s);
}

double *dor_array(dor_state *s, int which, int64_t *n) {
  int nth = s->NTH > 0 ? s->NTH : 1;
  int64_t cf = (int64_t)s->cfield * 2, rf = (int64_t)s->rfield;
  switch (which) {
    case DOR_CU1: *n = cf; return (double *)s->CU1;
    case DOR_CU2: *n = cf; return (double *)s->CU2;
    case DOR_CU3: *n = cf; return (double *)s->CU3;
    case DOR_CP: *n = cf; return (double *)s->CP;
    case DOR_CR1: *n = cf; return (double *)s->CR1;
    case DOR_CR2: *n = cf; return (double *)s->CR2;
    case DOR_CR3: *n = cf; return (double *)s->CR3;
    case DOR_CF1: *n = cf; return (double *)s->CF1;
    case DOR_CF2: *n = cf; return (double *)s->CF2;
    case DOR_CF3: *n = cf; return (double *)s->CF3;
    case DOR_CS1: *n = cf; return (double *)s->CS1;
    case DOR_CTH: *n = cf * nth; return (double *)s->CTH;
    case DOR_CFTH: *n = cf * nth; return (double *)s->CFTH;
    case DOR_CRTH: *n = cf * nth; return (double *)s->CRTH;
    case DOR_TH: *n = rf * nth; return s->TH;
    case DOR_RTH: *n = rf * nth; return s->RTH;
    case DOR_GY: *n = s->NY + 2; return s->GY;
    case DOR_GYF: *n = s->NY + 2; return s->GYF;
    case DOR_DY: *n = s->NY + 2; return s->DY;
    case DOR_DYF: *n = s->NY + 2; return s->DYF;
    case DOR_KX: *n = s->NXP + 1; return s->KX;
    case DOR_KX2: *n = s->NXP + 1; return s->KX2;
    case DOR_KZ: *n = s->TNKZ + 1; return s->KZ;
    case DOR_KZ2: *n = s->TNKZ + 1; return s->KZ2;
    case DOR_HBAR: *n = 3; return s->H_BAR;
    case DOR_BETABAR: *n = 3; return s->BETA_BAR;
    case DOR_ZETABAR: *n = 3; return s->ZETA_BAR;
    case DOR_NUT: *n = s->NU_T ? rf : 0; return s->NU_T;
  }
  *n = 0;
  return NULL;
}

/* INIT_FFT wavenumbers fft.f:77-100; INIT_CHAN_MPI index rules
 * mpi.f:1022-1046 (RANKY=0 branch on one rank); RKW3 constants
 * diablo_io.f:116-124; grid spacing in x,z channel.f:1572-1583. */
void dor_init(dor_state *s) {
  s->PI = 4.0 * atan(1.0);
  /* fft.f:79-82: KX(I)=(I+NXPP*RANKZ)*(2.*PI)/LX for I=0..NXPP-1; KX(NXP) unset */
  for (int i = 0; i < s->NXP; ++i) {
    s->KX[i] = (double)i * (2. * s->PI) / s->p.lx;
    s->KX2[i] = s->KX[i] * s->KX[i];
  }
  s->KX[s->NXP] = 0.0;
  s->KX2[s->NXP] = 0.0;
  /* fft.f:91-100 */
  for (int k = 0; k <= s->NKZ; ++k) s->KZ[k] = (double)k * (2. * s->PI) / s->p.lz;
  for (int k = 1; k <= s->NKZ; ++k) s->KZ[s->TNKZ + 1 - k] = -(double)k * (2. * s->PI) / s->p.lz;
  for (int k = 0; k <= s->TNKZ; ++k) s->KZ2[k] = s->KZ[k] * s->KZ[k];
  /* mpi.f:1022-1046 */
  s->JEND = s->NY;
  s->JSTART = 2;
  for (int n = 0; n < s->NTH; ++n) {
    s->JEND_TH[n] = s->NY;
    s->JSTART_TH[n] = (s->p.th_bc_ymin[n] == 1) ? 1 : 2;
  }
  s->DX1 = s->p.lx / s->NX;
  s->DZ1 = s->p.lz / s->NZ;
  s->H_BAR[0] = s->DELTA_T * (8.0 / 15.0);
  s->H_BAR[1] = s->DELTA_T * (2.0 / 15.0);
  s->H_BAR[2] = s->DELTA_T * (5.0 / 15.0);
  s->BETA_BAR[0] = 1.0;
  s->BETA_BAR[1] = 25.0 / 8.0;
  s->BETA_BAR[2] = 9.0 / 4.0;
  s->ZETA_BAR[0] = 0.0;
  s->ZETA_BAR[1] = -17.0 / 8.0;
  s->ZETA_BAR[2] = -5.0 / 4.0;
}

void dor_get_clock(dor_state *s, double *time, double *delta_t, int *time_step) {
  *time = s->TIME;
  *delta_t = s->DELTA_T;
  *time_step = s->TIME_STEP;
}

void dor_set_clock(dor_state *s, double time, double delta_t, int time_step) {
  s->TIME = time;
  s->TIME_STEP = time_step;
  if (delta_t != s->DELTA_T) {
    s->DELTA_T = delta_t;
    s->H_BAR[0] = delta_t * (8.0 / 15.0);
    s->H_BAR[1] = delta_t * (2.0 / 15.0);
    s->H_BAR[2] = delta_t * (5.0 / 15.0);
  }
}

/* ReadGridHDF5 hdf5.f:690-705 (GYF interior and ghosts, RANKY=0=NPROCY-1)
 * then CREATE_GRID_CHAN channel.f:1646-1652 (DY, DYF). */
void dor_set_grid_gy(dor_state *s, const double *gy) {
  int NY = s->NY;
  memset(s->GY, 0, 8 * (size_t)(NY + 2));
  memset(s->GYF, 0, 8 * (size_t)(NY + 2));
  memset(s->DY, 0, 8 * (size_t)(NY + 2));
  memset(s->DYF, 0, 8 * (size_t)(NY + 2));
  for (int j = 1; j <= NY + 1; ++j) s->GY[j] = gy[j - 1];
  for (int j = 1; j <= NY; ++j) s->GYF[j] = 0.5 * (s->GY[j] + s->GY[j + 1]);
  s->GYF[0] = 2.0 * s->GYF[1] - s->GYF[2];
  s->GYF[NY + 1] = 2.0 * s->GYF[NY] - s->GYF[NY - 1];
  for (int j = 1; j <= NY + 1; ++j) s->DY[j] = s->GYF[j] - s->GYF[j - 1];
  for (int j = 1; j <= NY; ++j) s->DYF[j] = s->GY[j + 1] - s->GY[j];
  s->DYF[NY + 1] = s->DYF[NY];
}

/* pre_process/create_grid_h5.m:28-79.  Output GY(1..N+1) in gy_out[0..N]. */
void dor_make_gy(int N, double L, double CS, int grid_type, double *out) {
  double *G = (double *)calloc((size_t)N + 3, 8);
  double *GF = (double *)calloc((size_t)N + 3, 8);
  for (int J = 1; J <= N + 1; ++J) {
    double t = (2.0 * (J - 1)) / N - 1.0;
    if (grid_type == 1)
      G[J + 1] = (L / 2.0) * tanh(CS * t) / tanh(CS);
    else
      G[J + 1] = (L / 2.0) * (CS * t * t * t + t) / (CS + 1);
  }
  for (int J = 1; J <= N; ++J) GF[J + 1] = (G[J + 1] + G[J + 2]) / 2.0;
  double gf_lower = GF[2], gf_upper = GF[N + 1], g_lower = G[2], g_upper = G[N + 2];
  for (int J = 1; J <= N; ++J) GF[J + 1] = GF[J + 1] * (g_upper - g_lower) / (gf_upper - gf_lower);
  for (int J = 1; J <= N + 1; ++J) G[J + 1] = G[J + 1] * (g_upper - g_lower) / (gf_upper - gf_lower);
  gf_lower = GF[2];
  double shift = g_lower - gf_lower;
  for (int J = 1; J <= N + 1; ++J) G[J + 1] = G[J + 1] + shift;
  for (int J = 1; J <= N + 1; ++J) out[J - 1] = G[J + 1];
  free(G);
  free(GF);
}

/* ------------------------------------------------------------------------ */
/* FFT_XZ_TO_FOURIER  mpi.f:1477-1542 (one rank: MPI_ALLTOALL with XY2ZY_1 */
/* -> XY2ZY_2 is the copy TMP2(I,K)=TMP_1A(I,K), I<NXP, K<NZ, mpi.f:928-954)*/
/* ------------------------------------------------------------------------ */
void dor_fft_xz_to_fourier(dor_state *s, double *V, int JMIN, int JMAX) {
  int NX = s->NX, NZ = s->NZ, NXP = s->NXP, NKX = s->NKX, NKZ = s->NKZ;
  cplx *VV = (cplx *)V;
#pragma omp parallel
  {
    cplx *xin = (cplx *)malloc(sizeof(cplx) * (size_t)(NX > NZ ? NX : NZ));
    cplx *xout = (cplx *)malloc(sizeof(cplx) * (size_t)(NX > NZ ? NX : NZ));
    cplx *TMP = (cplx *)malloc(sizeof(cplx) * (size_t)(NX / 2 + 1) * (size_t)NZ);
    cplx *TMP1A = (cplx *)malloc(sizeof(cplx) * (size_t)(NX / 2) * (size_t)NZ);
#pragma omp for schedule(static)
    for (int J = JMIN; J <= JMAX; ++J) {
      /* x R2C (plan fft.f:73-74: n=NX, NZP rows, in dist NX+2, out NX/2+1) */
      for (int K = 0; K < s->NZP; ++K) {
        for (int i = 0; i < NX; ++i) xin[i] = V[RIDX(s, i, K, J)];
        fft_exec(&s->px_f, xin, 1, xout);
        for (int ix = 0; ix <= NX / 2; ++ix) TMP[ix + (size_t)(NX / 2 + 1) * K] = xout[ix];
        for (int ix = 0; ix <= NKX; ++ix) {
          cplx t = TMP[ix + (size_t)(NX / 2 + 1) * K] / NX;
          TMP[ix + (size_t)(NX / 2 + 1) * K] = t;
          TMP1A[ix + (size_t)(NX / 2) * K] = t;
        }
        for (int ix = NKX + 1; ix <= NX / 2; ++ix) TMP[ix + (size_t)(NX / 2 + 1) * K] = 0.0;
        for (int ix = NKX + 1; ix <= NX / 2 - 1; ++ix) TMP1A[ix + (size_t)(NX / 2) * K] = TMP[ix + (size_t)(NX / 2 + 1) * K];
      }
      /* transpose (one rank) and zero padding, mpi.f:1510-1523 */
      for (int K = 0; K <= NZ - 1; ++K) {
        for (int ix = 0; ix <= NXP - 1; ++ix) VV[CIDX(s, ix, K, J)] = TMP1A[ix + (size_t)(NX / 2) * K];
        VV[CIDX(s, NXP, K, J)] = 0.0;
      }
      for (int K = NZ; K <= NZ + 1; ++K)
        for (int ix = 0; ix <= NXP; ++ix) VV[CIDX(s, ix, K, J)] = 0.0;
      /* forward c2c along z, in place, I=0..NXP-1 (plan fft.f:86-87,
       * executed through the r2c entry point at mpi.f:1528) */
      for (int ix = 0; ix <= NXP - 1; ++ix) {
        for (int k = 0; k < NZ; ++k) xin[k] = VV[CIDX(s, ix, k, J)];
        fft_exec(&s->pz_f, xin, 1, xout);
        for (int k = 0; k < NZ; ++k) VV[CIDX(s, ix, k, J)] = xout[k];
      }
      for (int K = 0; K <= NKZ; ++K)
        for (int ix = 0; ix <= NXP; ++ix) VV[CIDX(s, ix, K, J)] = VV[CIDX(s, ix, K, J)] / NZ;
      for (int K = 1; K <= NKZ; ++K)
        for (int ix = 0; ix <= NXP; ++ix)
          VV[CIDX(s, ix, NKZ + K, J)] = VV[CIDX(s, ix, NZ - 1 + K - NKZ, J)] / NZ;
    }
    free(xin); free(xout); free(TMP); free(TMP1A);
  }
}

/* FFT_XZ_TO_PHYSICAL mpi.f:1547-1606 */
void dor_fft_xz_to_physical(dor_state *s, double *Vd, int JMIN, int JMAX) {
  int NX = s->NX, NZ = s->NZ, NXP = s->NXP, NKZ = s->NKZ;
  cplx *VV = (cplx *)Vd;
  double *V = Vd;
#pragma omp parallel
  {
    cplx *xin = (cplx *)malloc(sizeof(cplx) * (size_t)(NX > NZ ? NX : NZ));
    cplx *xout = (cplx *)malloc(sizeof(cplx) * (size_t)(NX > NZ ? NX : NZ));
    cplx *TMP = (cplx *)malloc(sizeof(cplx) * (size_t)(NX / 2 + 1) * (size_t)NZ);
#pragma omp for schedule(static)
    for (int J = JMIN; J <= JMAX; ++J) {
      /* unpack, mpi.f:1564-1574 */
      for (int K = NKZ; K >= 1; --K)
        for (int ix = 0; ix <= NXP; ++ix)
          VV[CIDX(s, ix, NZ - 1 + K - NKZ, J)] = VV[CIDX(s, ix, NKZ + K, J)];
      for (int K = NKZ + 1; K <= NZ - NKZ - 1; ++K)
        for (int ix = 0; ix <= NXP; ++ix) VV[CIDX(s, ix, K, J)] = 0.0;
      /* inverse c2c along z (FFTW_BACKWARD, unnormalised), I=0..NXP-1 */
      for (int ix = 0; ix <= NXP - 1; ++ix) {
        for (int k = 0; k < NZ; ++k) xin[k] = VV[CIDX(s, ix, k, J)];
        fft_exec(&s->pz_b, xin, 1, xout);
        for (int k = 0; k < NZ; ++k) VV[CIDX(s, ix, k, J)] = xout[k];
      }
      /* transpose (one rank), mpi.f:1579-1596: TMP(I<NX/2)=VV, Nyquist 0 */
      for (int K = 0; K <= s->NZP - 1; ++K) {
        for (int ix = 0; ix <= NX / 2 - 1; ++ix) TMP[ix + (size_t)(NX / 2 + 1) * K] = VV[CIDX(s, ix, K, J)];
        TMP[NX / 2 + (size_t)(NX / 2 + 1) * K] = 0.0;
      }
      /* C2R along x (plan fft.f:75-76, unnormalised).  The input is taken as
       * the half spectrum of a Hermitian sequence; the imaginary parts of the
       * DC and Nyquist terms do not enter the result. */
      for (int K = 0; K <= s->NZP - 1; ++K) {
        const cplx *X = TMP + (size_t)(NX / 2 + 1) * K;
        xin[0] = creal(X[0]);
        for (int k = 1; k < NX / 2; ++k) {
          xin[k] = X[k];
          xin[NX - k] = conj(X[k]);
        }
        xin[NX / 2] = creal(X[NX / 2]);
        fft_exec(&s->px_b, xin, 1, xout);
        for (int i = 0; i < NX; ++i) V[RIDX(s, i, K, J)] = creal(xout[i]);
      }
    }
    free(xin); free(xout); free(TMP);
  }
}

/* ------------------------------------------------------------------------ */
/* Thomas sweeps, one rank (RANKY=0=NPROCY-1)                               */
/* ------------------------------------------------------------------------ */

/* THOMAS_FORWARD_REAL_MPI mpi.f:630-703 then THOMAS_BACKWARD_REAL_MPI
 * mpi.f:767-809.  A,B,C,G are (0:INX-1, 0:INY+1). */
void dor_thomas_real(double *A, double *B, double *C, double *G, int INY, int INX) {
#define M2(a, i, j) a[(size_t)(i) + (size_t)INX * (size_t)(j)]
  for (int J = 1; J <= INY; ++J)
    for (int ix = 0; ix <= INX - 1; ++ix) {
      M2(A, ix, J) = -M2(A, ix, J) / M2(B, ix, J - 1);
      M2(B, ix, J) = M2(B, ix, J) + M2(A, ix, J) * M2(C, ix, J - 1);
      M2(G, ix, J) = M2(G, ix, J) + M2(A, ix, J) * M2(G, ix, J - 1);
    }
  {
    int J = INY + 1;
    for (int ix = 0; ix <= INX - 1; ++ix) {
      M2(A, ix, J) = -M2(A, ix, J) / M2(B, ix, J - 1);
      M2(B, ix, J) = M2(B, ix, J) + M2(A, ix, J) * M2(C, ix, J - 1);
      M2(G, ix, J) = M2(G, ix, J) + M2(A, ix, J) * M2(G, ix, J - 1);
    }
  }
  for (int ix = 0; ix <= INX - 1; ++ix) {
    M2(G, ix, INY + 1) = M2(G, ix, INY + 1) / M2(B, ix, INY + 1);
    for (int J = INY; J >= 0; --J) {
      if (M2(B, ix, J) != 0.0)
        M2(G, ix, J) = (M2(G, ix, J) - M2(C, ix, J) * M2(G, ix, J + 1)) / M2(B, ix, J);
      else
        M2(G, ix, J) = 0.0;
    }
  }
#undef M2
}

/* THOMAS_FORWARD_COMPLEX_MPI mpi.f:708-764 then THOMAS_BACKWARD_COMPLEX_MPI
 * mpi.f:812-865.  A,B,C real (0:INX, 0:INY+1); G complex (0:INX, 0:INY+1). */
void dor_thomas_complex(double *A, double *B, double *C, double *Gd, int INY, int INX) {
  cplx *G = (cplx *)Gd;
#define M2(a, i, j) a[(size_t)(i) + (size_t)(INX + 1) * (size_t)(j)]
  for (int ix = 0; ix <= INX; ++ix) {
    for (int J = 2; J <= INY; ++J) {
      if (M2(B, ix, J - 1) != 0.0)
        M2(A, ix, J) = -M2(A, ix, J) / M2(B, ix, J - 1);
      else
        M2(A, ix, J) = 0.0;
      M2(B, ix, J) = M2(B, ix, J) + M2(A, ix, J) * M2(C, ix, J - 1);
      M2(G, ix, J) = M2(G, ix, J) + M2(A, ix, J) * M2(G, ix, J - 1);
    }
  }
  for (int ix = 0; ix <= INX; ++ix) {
    if (M2(B, ix, INY) != 0.0)
      M2(G, ix, INY) = M2(G, ix, INY) / M2(B, ix, INY);
    else
      M2(G, ix, INY) = 0.0;
    for (int J = INY - 1; J >= 1; --J) {
      if (M2(B, ix, J) != 0.0)
        M2(G, ix, J) = (M2(G, ix, J) - M2(C, ix, J) * M2(G, ix, J + 1)) / M2(B, ix, J);
      else
        M2(G, ix, J) = 0.0;
    }
  }
#undef M2
}

/* ------------------------------------------------------------------------ */
/* Boundary rows: APPLY_BC_{1,2,3,TH}_{LOWER,UPPER} channel.f:1660-2282,     */
/* dispatched by APPLY_BC_{U1,U2,U3,TH}_MPI mpi.f:1083-1236 (both ends on   */
/* one rank).  MAT arrays are (0:NX-1, 0:NY+1).                             */
/* ------------------------------------------------------------------------ */
#define MI(i, j) ((size_t)(i) + (size_t)NX * (size_t)(j))

/* U1/U3/TH share the same row pattern: channel.f:1660-1778 (U1),
 * 1962-2084 (U3), 2088-2241 (TH). */
/* The eddy viscosity (or diffusivity) part of the implicit rows J=j0..j1 of
 * column K (channel.f:794-807, 927-941, 991-1005): NU_T at GY points. */
static void les_rows_gy(dor_state *s, int K, int j0, int j1, double TEMP2, double *L, double *D, double *Uu) {
  const int NX = s->NX, NXM = s->NXM;
  const double *DY = s->DY, *DYF = s->DYF, *NU_T = s->NU_T;
  for (int J = j0; J <= j1; ++J)
    for (int i = 0; i <= NXM; ++i) {
      const double nj = NU_T[RIDX(s, i, K, J)], np = NU_T[RIDX(s, i, K, J + 1)];
      L[i + (size_t)NX * J] = L[i + (size_t)NX * J] - TEMP2 * nj / (DY[J] * DYF[J]);
      D[i + (size_t)NX * J] = D[i + (size_t)NX * J] + TEMP2 * np / (DY[J + 1] * DYF[J]) + TEMP2 * nj / (DY[J] * DYF[J]);
      Uu[i + (size_t)NX * J] = Uu[i + (size_t)NX * J] - TEMP2 * np / (DY[J + 1] * DYF[J]);
    }
}

static void bc_u13th(int NX, int NY, const double *DY, int bc_min, double c1_min, int bc_max,
                     double c1_max, double *L, double *D, double *U, double *VEC) {
  if (bc_min == 0) {
    for (int i = 0; i < NX; ++i) {
      L[MI(i, 0)] = 0.; D[MI(i, 0)] = 1.; U[MI(i, 0)] = 0.; VEC[MI(i, 0)] = 0.;
      L[MI(i, 1)] = 0.; D[MI(i, 1)] = 1.; U[MI(i, 1)] = 0.; VEC[MI(i, 1)] = c1_min;
    }
  } else {
    for (int i = 0; i < NX; ++i) {
      L[MI(i, 0)] = 0.; D[MI(i, 0)] = 1.; U[MI(i, 0)] = 0.; VEC[MI(i, 0)] = 0.;
    }
    for (int i = 0; i < NX; ++i) {
      L[MI(i, 1)] = 0.; D[MI(i, 1)] = -1.; U[MI(i, 1)] = 1.; VEC[MI(i, 1)] = DY[2] * c1_min;
    }
  }
  if (bc_max == 0) {
    for (int i = 0; i < NX; ++i) {
      L[MI(i, NY + 1)] = 0.; D[MI(i, NY + 1)] = 1.; U[MI(i, NY + 1)] = 0.; VEC[MI(i, NY + 1)] = 0.;
      L[MI(i, NY)] = 0.; D[MI(i, NY)] = 1.; U[MI(i, NY)] = 0.; VEC[MI(i, NY)] = c1_max;
    }
  } else {
    for (int i = 0; i < NX; ++i) {
      L[MI(i, NY)] = -1.; D[MI(i, NY)] = 1.; U[MI(i, NY)] = 0.; VEC[MI(i, NY)] = DY[NY] * c1_max;
    }
    for (int i = 0; i < NX; ++i) {
      L[MI(i, NY + 1)] = 0.; D[MI(i, NY + 1)] = 1.; U[MI(i, NY + 1)] = 0.; VEC[MI(i, NY + 1)] = 0.;
    }
  }
}

/* APPLY_BC_2_LOWER channel.f:1782-1822, APPLY_BC_2_UPPER :1864-1892 */
static void bc_u2(int NX, int NY, const double *DYF, int bc_min, double c1_min, int bc_max,
                  double c1_max, double *L, double *D, double *U, double *VEC) {
  if (bc_min == 0) {
    for (int i = 0; i < NX; ++i) {
      L[MI(i, 1)] = 0.; D[MI(i, 1)] = 1.; U[MI(i, 1)] = 0.; VEC[MI(i, 1)] = c1_min;
      L[MI(i, 2)] = 0.; D[MI(i, 2)] = 1.; U[MI(i, 2)] = 0.; VEC[MI(i, 2)] = c1_min;
    }
  } else if (bc_min == 1) {
    for (int i = 0; i < NX; ++i) {
      D[MI(i, 1)] = -1.; U[MI(i, 1)] = 1.; L[MI(i, 1)] = 0.; VEC[MI(i, 1)] = DYF[1] * c1_min;
    }
  }
  for (int i = 0; i < NX; ++i) {
    L[MI(i, 0)] = 0.; D[MI(i, 0)] = 1.; U[MI(i, 0)] = 0.; VEC[MI(i, 0)] = 0.;
  }
  if (bc_max == 0) {
    for (int i = 0; i < NX; ++i) {
      L[MI(i, NY + 1)] = 0.; D[MI(i, NY + 1)] = 1.; U[MI(i, NY + 1)] = 0.; VEC[MI(i, NY + 1)] = c1_max;
      L[MI(i, NY)] = 0.; D[MI(i, NY)] = 1.; U[MI(i, NY)] = 0.; VEC[MI(i, NY)] = c1_max;
    }
  } else if (bc_max == 1) {
    for (int i = 0; i < NX; ++i) {
      L[MI(i, NY + 1)] = -1.; D[MI(i, NY + 1)] = 1.; U[MI(i, NY + 1)] = 0.; VEC[MI(i, NY + 1)] = DYF[NY] * c1_max;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* RK_CHAN_1  channel.f:62-1073, LES=.FALSE., USE_MPI=.TRUE., one rank       */
/* ------------------------------------------------------------------------ */
/* Filter widths of les.f:196-217 with -2 C_SMAG^2 absorbed, C_SMAG=0.13,
 * beta_sgs=1: DYFW(J) = -2 C^2 (DX(1) DYF(J) 2 DZ(1))^(2/3) at GYF and
 * DYW(J) likewise with DY(J) at GY, J=1..NY+1 (others 0). */
void dor_les_widths(dor_state *s, double *dyw, double *dyfw) {
  const double C_SMAG = 0.13, beta_sgs = 1.0;
  for (int J = 0; J <= s->NY + 1; ++J) dyw[J] = dyfw[J] = 0.0;
  for (int J = 1; J <= s->NY + 1; ++J) {
    dyfw[J] = -2.0 * pow(C_SMAG, 2.0) * pow(s->DX1 * beta_sgs * s->DYF[J] * 2.0 * s->DZ1 * beta_sgs, 2.0 / 3.0);
    dyw[J] = -2.0 * pow(C_SMAG, 2.0) * pow(s->DX1 * beta_sgs * s->DY[J] * 2.0 * s->DZ1 * beta_sgs, 2.0 / 3.0);
  }
}

/* APPLY_BC_LES les.f:805-862: with Neumann U or W, the ghost planes copy
 * the first interior plane so the model sees no wall gradient (Fourier). */
static void apply_bc_les(dor_state *s) {
  const int NY = s->NY, NXP = s->NXP, TNKZ = s->TNKZ;
#define C3(a, i, k, j) a[CIDX(s, i, k, j)]
  for (int K = 0; K <= TNKZ; ++K)
    for (int i = 0; i <= NXP - 1; ++i) {
      if (s->p.u_bc_ymax == 1) { C3(s->CU1, i, K, NY) = C3(s->CU1, i, K, NY - 1); C3(s->CU1, i, K, NY + 1) = C3(s->CU1, i, K, NY); }
      if (s->p.w_bc_ymax == 1) { C3(s->CU3, i, K, NY) = C3(s->CU3, i, K, NY - 1); C3(s->CU3, i, K, NY + 1) = C3(s->CU3, i, K, NY); }
      if (s->p.u_bc_ymin == 1) { C3(s->CU1, i, K, 1) = C3(s->CU1, i, K, 2); C3(s->CU1, i, K, 0) = C3(s->CU1, i, K, 1); }
      if (s->p.w_bc_ymin == 1) { C3(s->CU3, i, K, 1) = C3(s->CU3, i, K, 2); C3(s->CU3, i, K, 0) = C3(s->CU3, i, K, 1); }
    }
#undef C3
}

/* APPLY_BC_VEL_PHYS_MPI mpi.f:1366-1378 -> APPLY_BC_VEL_PHYS_LOWER/UPPER
 * channel.f:2495-2650 (one rank: both walls), on physical U1..3. */
static void apply_bc_vel_phys(dor_state *s) {
  const int NY = s->NY, NZP = s->NZP, NXM = s->NXM;
  const double *DY = s->DY, *DYF = s->DYF;
  double *U1 = PHYS(s->CU1), *U2 = PHYS(s->CU2), *U3 = PHYS(s->CU3);
  const dor_params *p = &s->p;
#define R3D(a, i, k, j) a[RIDX(s, i, k, j)]
  for (int K = 0; K <= NZP - 1; ++K)
    for (int i = 0; i <= NXM; ++i) {
      if (p->u_bc_ymin == 0) { R3D(U1, i, K, 0) = p->u_bc_ymin_c1; R3D(U1, i, K, 1) = p->u_bc_ymin_c1; }
      else { R3D(U1, i, K, 1) = R3D(U1, i, K, 2) - DY[2] * p->u_bc_ymin_c1; R3D(U1, i, K, 0) = R3D(U1, i, K, 1) - DY[1] * p->u_bc_ymin_c1; }
      if (p->w_bc_ymin == 0) { R3D(U3, i, K, 0) = p->w_bc_ymin_c1; R3D(U3, i, K, 1) = p->w_bc_ymin_c1; }
      else { R3D(U3, i, K, 1) = R3D(U3, i, K, 2) - DY[2] * p->w_bc_ymin_c1; R3D(U3, i, K, 0) = R3D(U3, i, K, 1) - DY[1] * p->w_bc_ymin_c1; }
      if (p->v_bc_ymin == 0) { R3D(U2, i, K, 1) = 2.0 * p->v_bc_ymin_c1 - R3D(U2, i, K, 2); R3D(U2, i, K, 0) = R3D(U2, i, K, 1); }
      else { R3D(U2, i, K, 1) = R3D(U2, i, K, 2) - DYF[1] * p->v_bc_ymin_c1; R3D(U2, i, K, 0) = R3D(U2, i, K, 1) - DYF[1] * p->v_bc_ymin_c1; }
    }
  for (int K = 0; K <= NZP - 1; ++K)
    for (int i = 0; i <= NXM; ++i) {
      if (p->u_bc_ymax == 0) { R3D(U1, i, K, NY) = p->u_bc_ymax_c1; R3D(U1, i, K, NY + 1) = p->u_bc_ymax_c1; }
      else { R3D(U1, i, K, NY) = R3D(U1, i, K, NY - 1) + DY[NY] * p->u_bc_ymax_c1; R3D(U1, i, K, NY + 1) = R3D(U1, i, K, NY) + DY[NY] * p->u_bc_ymax_c1; }
      if (p->w_bc_ymax == 0) { R3D(U3, i, K, NY) = p->w_bc_ymax_c1; R3D(U3, i, K, NY + 1) = p->w_bc_ymax_c1; }
      else { R3D(U3, i, K, NY) = R3D(U3, i, K, NY - 1) + DY[NY] * p->w_bc_ymax_c1; R3D(U3, i, K, NY + 1) = R3D(U3, i, K, NY) + DY[NY] * p->w_bc_ymax_c1; }
      if (p->v_bc_ymax == 0) R3D(U2, i, K, NY + 1) = 2.0 * p->v_bc_ymax_c1 - R3D(U2, i, K, NY);
      else R3D(U2, i, K, NY + 1) = R3D(U2, i, K, NY) + DY[NY] * p->v_bc_ymax_c1;
    }
#undef R3D
}

/* les_chan les.f:1-336 (LES_MODEL_TYPE=1) with compute_strain_chan
 * les.f:526-569 and GHOST_LES_MPI's NPROCY=1 branch mpi.f:493-509. */
void dor_les_chan(dor_state *s) {
  const int NY = s->NY, NXP = s->NXP, TNKZ = s->TNKZ, NZP = s->NZP, NXM = s->NXM;
  const int JSTART = s->JSTART, JEND = s->JEND;
  const double *DY = s->DY, *DYF = s->DYF, *KX = s->KX, *KZ = s->KZ;
  cplx *CU1 = s->CU1, *CU2 = s->CU2, *CU3 = s->CU3;
  cplx *CF1 = s->CF1, *CF2 = s->CF2, *CF3 = s->CF3;
  cplx *CS11 = s->CSij[0], *CS22 = s->CSij[1], *CS33 = s->CSij[2];
  cplx *CS4 = s->CSij[3], *CS5 = s->CSij[4], *CS6 = s->CSij[5], *CTEMP = s->CTEMP;
  double *S11 = PHYS(CS11), *S22 = PHYS(CS22), *S33 = PHYS(CS33), *S4 = PHYS(CS4), *S5 = PHYS(CS5), *S6 = PHYS(CS6);
  double *TEMP = PHYS(CTEMP), *S1 = PHYS(s->CS1), *NU_T = s->NU_T;
  double *U1 = PHYS(CU1), *U3 = PHYS(CU3);
#define C3(a, i, k, j) a[CIDX(s, i, k, j)]
#define R3D(a, i, k, j) a[RIDX(s, i, k, j)]
#define CIKX(i) (I * KX[i])
#define CIKZ(k) (I * KZ[k])
  /* les.f:45-57, RANKY = 0 = NPROCY-1: the upper-wall branch */
  const int J1 = JSTART, J2 = NY - 1;
  /* les.f:62-73 (GHOST_CHAN_MPI is a no-op with NPROCY=1, mpi.f:25) */
  dor_apply_bc_vel(s);
  apply_bc_les(s);
  /* compute_strain_chan les.f:536-564 */
  for (int J = 1; J <= NY; ++J)
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) {
        C3(CS11, i, K, J) = CIKX(i) * C3(CU1, i, K, J);
        C3(CS22, i, K, J) = (C3(CU2, i, K, J + 1) - C3(CU2, i, K, J)) / DYF[J];
        C3(CS33, i, K, J) = CIKZ(K) * C3(CU3, i, K, J);
        C3(CS5, i, K, J) = 0.5 * (CIKZ(K) * C3(CU1, i, K, J) + CIKX(i) * C3(CU3, i, K, J));
      }
  for (int J = 1; J <= NY + 1; ++J)
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) {
        C3(CS4, i, K, J) = 0.5 * ((C3(CU1, i, K, J) - C3(CU1, i, K, J - 1)) / DY[J] + CIKX(i) * C3(CU2, i, K, J));
        C3(CS6, i, K, J) = 0.5 * (CIKZ(K) * C3(CU2, i, K, J) + (C3(CU3, i, K, J) - C3(CU3, i, K, J - 1)) / DY[J]);
      }
  for (int q = 0; q < 6; ++q) dor_fft_xz_to_physical(s, (double *)s->CSij[q], 0, NY + 1);
  /* les.f:86-91: the mean velocity profiles */
  double *U1b = malloc(8 * (size_t)(NY + 2)), *U3b = malloc(8 * (size_t)(NY + 2));
  for (int J = 0; J <= NY + 1; ++J) {
    U1b[J] = creal(C3(CU1, 0, 0, J));
    U3b[J] = creal(C3(CU3, 0, 0, J));
  }
  /* les.f:103-105 */
  dor_fft_xz_to_physical(s, (double *)CU1, 0, NY + 1);
  dor_fft_xz_to_physical(s, (double *)CU2, 0, NY + 1);
  dor_fft_xz_to_physical(s, (double *)CU3, 0, NY + 1);
  /* |S| at GYF, les.f:113-134 */
  for (int J = JSTART; J <= JEND; ++J)
    for (int K = 0; K <= NZP - 1; ++K)
      for (int i = 0; i <= NXM; ++i) {
        double a4 = 0.5 * (R3D(S4, i, K, J + 1) + R3D(S4, i, K, J) - 0.5 * (U1b[J] - U1b[J - 1]) / DY[J] -
                           0.5 * (U1b[J + 1] - U1b[J]) / DY[J + 1]);
        double a6 = 0.5 * (R3D(S6, i, K, J + 1) + R3D(S6, i, K, J) - 0.5 * (U3b[J] - U3b[J - 1]) / DY[J] -
                           0.5 * (U3b[J + 1] - U3b[J]) / DY[J + 1]);
        double s11 = R3D(S11, i, K, J), s22 = R3D(S22, i, K, J), s33 = R3D(S33, i, K, J), s5 = R3D(S5, i, K, J);
        R3D(S1, i, K, J) = sqrt(2.0 * (s11 * s11) + 4.0 * (a4 * a4) + 4.0 * (s5 * s5) + 2.0 * (s22 * s22) +
                                4.0 * (a6 * a6) + 2.0 * (s33 * s33));
      }
  /* |S| at GY, les.f:137-160 */
  for (int J = 2; J <= NY + 1; ++J)
    for (int K = 0; K <= NZP - 1; ++K)
      for (int i = 0; i <= NXM; ++i) {
#define GYI(A) ((R3D(A, i, K, J) * DYF[J - 1] + R3D(A, i, K, J - 1) * DYF[J]) / (2.0 * DY[J]))
        double i11 = GYI(S11), i5 = GYI(S5), i22 = GYI(S22), i33 = GYI(S33);
#undef GYI
        double a4 = R3D(S4, i, K, J) - 0.5 * (U1b[J] - U1b[J - 1]) / DY[J];
        double a6 = R3D(S6, i, K, J) - 0.5 * (U3b[J] - U3b[J - 1]) / DY[J];
        R3D(TEMP, i, K, J) = sqrt(2.0 * (i11 * i11) + 4.0 * (a4 * a4) + 4.0 * (i5 * i5) + 2.0 * (i22 * i22) +
                                  4.0 * (a6 * a6) + 2.0 * (i33 * i33));
      }
  /* |S| S_ij, les.f:164-190 (the dU/dy parts of S4 and S6 are implicit) */
  for (int J = JSTART; J <= JEND; ++J)
    for (int K = 0; K <= NZP - 1; ++K)
      for (int i = 0; i <= NXM; ++i) {
        R3D(S11, i, K, J) = R3D(S1, i, K, J) * R3D(S11, i, K, J);
        R3D(S5, i, K, J) = R3D(S1, i, K, J) * R3D(S5, i, K, J);
        R3D(S22, i, K, J) = 0.0;
        R3D(S33, i, K, J) = R3D(S1, i, K, J) * R3D(S33, i, K, J);
      }
  for (int J = 2; J <= NY + 1; ++J)
    for (int K = 0; K <= NZP - 1; ++K)
      for (int i = 0; i <= NXM; ++i) {
        R3D(S4, i, K, J) = R3D(TEMP, i, K, J) * (R3D(S4, i, K, J) - 0.5 * (R3D(U1, i, K, J) - R3D(U1, i, K, J - 1)) / DY[J]);
        R3D(S6, i, K, J) = R3D(TEMP, i, K, J) * (R3D(S6, i, K, J) - 0.5 * (R3D(U3, i, K, J) - R3D(U3, i, K, J - 1)) / DY[J]);
      }
  /* filter widths with -2 C_SMAG^2 absorbed, les.f:196-217 */
  double *DYFW = calloc((size_t)(NY + 2), 8), *DYW = calloc((size_t)(NY + 2), 8);
  dor_les_widths(s, DYW, DYFW);
  /* NU_T at GY, les.f:221-227, then GHOST_LES_MPI (NPROCY=1) */
  for (int J = 2; J <= NY; ++J)
    for (int K = 0; K <= NZP - 1; ++K)
      for (int i = 0; i <= NXM; ++i) R3D(NU_T, i, K, J) = -0.5 * DYW[J] * R3D(TEMP, i, K, J);
  for (int K = 0; K <= NZP - 1; ++K)
    for (int i = 0; i <= NXM; ++i) {
      R3D(NU_T, i, K, 1) = 0.0;
      R3D(NU_T, i, K, 2) = 0.0;
      R3D(NU_T, i, K, NY) = 0.0;
      R3D(NU_T, i, K, NY + 1) = 0.0;
    }
  /* les.f:237-264 */
  dor_fft_xz_to_fourier(s, S11, 0, NY + 1);
  dor_fft_xz_to_fourier(s, S33, 0, NY + 1);
  dor_fft_xz_to_fourier(s, S4, 0, NY + 1);
  dor_fft_xz_to_fourier(s, S5, 0, NY + 1);
  dor_fft_xz_to_fourier(s, S6, 0, NY + 1);
  for (int J = 1; J <= NY + 1; ++J)
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) {
        C3(CS11, i, K, J) = DYFW[J] * C3(CS11, i, K, J);
        C3(CS5, i, K, J) = DYFW[J] * C3(CS5, i, K, J);
        C3(CS33, i, K, J) = DYFW[J] * C3(CS33, i, K, J);
      }
  for (int J = 2; J <= NY + 1; ++J)
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) {
        C3(CS4, i, K, J) = DYW[J] * C3(CS4, i, K, J);
        C3(CS6, i, K, J) = DYW[J] * C3(CS6, i, K, J);
      }
  /* SGS divergence into CF, les.f:280-304 */
  for (int J = J1; J <= J2; ++J)
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) {
        C3(CF1, i, K, J) = C3(CF1, i, K, J) - CIKX(i) * C3(CS11, i, K, J) -
                           (C3(CS4, i, K, J + 1) - C3(CS4, i, K, J)) / DYF[J] - CIKZ(K) * C3(CS5, i, K, J);
        C3(CF3, i, K, J) = C3(CF3, i, K, J) - CIKX(i) * C3(CS5, i, K, J) -
                           (C3(CS6, i, K, J + 1) - C3(CS6, i, K, J)) / DYF[J] - CIKZ(K) * C3(CS33, i, K, J);
      }
  for (int J = 2; J <= NY; ++J)
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i)
        C3(CF2, i, K, J) = C3(CF2, i, K, J) - CIKX(i) * C3(CS4, i, K, J) - CIKZ(K) * C3(CS6, i, K, J);
  /* explicit cross terms, les.f:307-336 */
  for (int c = 0; c < 2; ++c) {
    const double *Uc = c ? U3 : U1;
    for (int K = 0; K <= NZP - 1; ++K)
      for (int i = 0; i <= NXM; ++i)
        for (int J = 2; J <= NY; ++J)
          R3D(TEMP, i, K, J) = R3D(NU_T, i, K, J) * (R3D(Uc, i, K, J) - R3D(Uc, i, K, J - 1)) / DY[J];
    dor_fft_xz_to_fourier(s, TEMP, 2, NY);
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i)
        for (int J = 2; J <= NY; ++J)
          C3(CF2, i, K, J) = C3(CF2, i, K, J) + (c ? CIKZ(K) : CIKX(i)) * C3(CTEMP, i, K, J);
  }
  free(U1b); free(U3b); free(DYFW); free(DYW);
#undef C3
#undef R3D
#undef CIKX
#undef CIKZ
}

void dor_rk_chan_1(dor_state *s, int RK_STEP) {
  const int NX = s->NX, NY = s->NY, NXP = s->NXP, TNKZ = s->TNKZ, NZP = s->NZP, NXM = s->NXM;
  const int JSTART = s->JSTART, JEND = s->JEND;
  const double NU = s->p.nu, BETA = s->p.beta, NU_V_SCALE = s->p.nu_v_scale;
  const double *DY = s->DY, *DYF = s->DYF, *KX = s->KX, *KZ = s->KZ, *KX2 = s->KX2, *KZ2 = s->KZ2;
  cplx *CU1 = s->CU1, *CU2 = s->CU2, *CU3 = s->CU3, *CP = s->CP;
  cplx *CR1 = s->CR1, *CR2 = s->CR2, *CR3 = s->CR3;
  cplx *CF1 = s->CF1, *CF2 = s->CF2, *CF3 = s->CF3, *CS1 = s->CS1;
  double *U1 = PHYS(CU1), *U2 = PHYS(CU2), *U3 = PHYS(CU3);
  double *R1 = PHYS(CR1), *R2 = PHYS(CR2), *R3 = PHYS(CR3), *S1 = PHYS(CS1);
  s->RK_STEP = RK_STEP;
  const int rk = RK_STEP - 1;
  /* channel.f:77-81 */
  const double TEMP1 = NU * s->H_BAR[rk] / 2.0;
  const double TEMP2 = s->H_BAR[rk] / 2.0;
  const double TEMP3 = s->ZETA_BAR[rk] * s->H_BAR[rk];
  const double TEMP4 = s->H_BAR[rk];
  const double TEMP5 = s->BETA_BAR[rk] * s->H_BAR[rk];
#define C3(a, i, k, j) a[CIDX(s, i, k, j)]
#define R3D(a, i, k, j) a[RIDX(s, i, k, j)]
#define CIKX(i) (I * KX[i])
#define CIKZ(k) (I * KZ[k])

  /* channel.f:88-102 */
  for (int J = JSTART; J <= JEND; ++J)
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) {
        C3(CR1, i, K, J) = C3(CU1, i, K, J);
        C3(CR3, i, K, J) = C3(CU3, i, K, J);
      }
  for (int J = 2; J <= NY; ++J)
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) C3(CR2, i, K, J) = C3(CU2, i, K, J);
  /* channel.f:105-121 */
  if (RK_STEP > 1) {
    for (int J = JSTART; J <= JEND; ++J)
      for (int K = 0; K <= TNKZ; ++K)
        for (int i = 0; i <= NXP - 1; ++i) {
          C3(CR1, i, K, J) = C3(CR1, i, K, J) + TEMP3 * C3(CF1, i, K, J);
          C3(CR3, i, K, J) = C3(CR3, i, K, J) + TEMP3 * C3(CF3, i, K, J);
        }
    for (int J = 2; J <= NY; ++J)
      for (int K = 0; K <= TNKZ; ++K)
        for (int i = 0; i <= NXP - 1; ++i) C3(CR2, i, K, J) = C3(CR2, i, K, J) + TEMP3 * C3(CF2, i, K, J);
  }
  /* channel.f:124-130 */
  for (int J = 2; J <= NY; ++J)
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) C3(CS1, i, K, J) = (C3(CP, i, K, J) - C3(CP, i, K, J - 1)) / DY[J];
  /* channel.f:133-148 */
  for (int J = JSTART; J <= JEND; ++J)
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) {
        C3(CR1, i, K, J) = C3(CR1, i, K, J) - TEMP4 * (CIKX(i) * C3(CP, i, K, J));
        C3(CR3, i, K, J) = C3(CR3, i, K, J) - TEMP4 * (CIKZ(K) * C3(CP, i, K, J));
      }
  for (int J = 2; J <= NY; ++J)
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) C3(CR2, i, K, J) = C3(CR2, i, K, J) - TEMP4 * C3(CS1, i, K, J);
  /* channel.f:155-167 (RANKZ=0) */
  if (s->p.f_type == 1) {
    for (int J = JSTART; J <= JEND; ++J) C3(CR1, 0, 0, J) = C3(CR1, 0, 0, J) - TEMP4 * s->p.px0;
  } else if (s->p.f_type == 2) {
    for (int J = JSTART; J <= JEND; ++J)
      C3(CR1, 0, 0, J) = C3(CR1, 0, 0, J) - TEMP4 * (s->p.px0 + s->p.amp_omega0 * cos(s->p.omega0 * s->TIME));
  }
  /* channel.f:175-189 */
  for (int J = JSTART; J <= JEND; ++J)
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) {
        double f = NU * pow(KX2[i] + KZ2[K], BETA);
        C3(CF1, i, K, J) = -(f * C3(CU1, i, K, J));
        C3(CF3, i, K, J) = -(f * C3(CU3, i, K, J));
      }
  for (int J = 2; J <= NY; ++J)
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) {
        double f = NU * pow(KX2[i] + KZ2[K], BETA);
        C3(CF2, i, K, J) = -(f * C3(CU2, i, K, J));
      }
  /* channel.f:193-200 */
  for (int K = 0; K <= TNKZ; ++K)
    for (int i = 0; i <= NXP - 1; ++i)
      for (int J = JSTART; J <= JEND; ++J) {
        C3(CF1, i, K, J) = C3(CF1, i, K, J) + s->p.i_ro * C3(CU3, i, K, J);
        C3(CF3, i, K, J) = C3(CF3, i, K, J) - s->p.i_ro * C3(CU1, i, K, J);
      }
  /* channel.f:203-293, per scalar */
  for (int N = 0; N < s->NTH; ++N) {
    cplx *CTH = s->CTH + s->cfield * N, *CFTH = s->CFTH + s->cfield * N, *CRTH = s->CRTH + s->cfield * N;
    const double RI = s->p.ri[N], PR = s->p.pr[N];
    for (int J = 2; J <= NY; ++J) {
      for (int K = 1; K <= TNKZ; ++K)
        for (int i = 0; i <= NXP - 1; ++i)
          C3(CF2, i, K, J) = C3(CF2, i, K, J) +
                             RI * s->p.grav_y * (C3(CTH, i, K, J) * DYF[J - 1] + C3(CTH, i, K, J - 1) * DYF[J]) / (2.0 * DY[J]);
      int K = 0;
      for (int i = 1; i <= NXP - 1; ++i) /* ISTART=1 on RANKZ=0 */
        C3(CF2, i, K, J) = C3(CF2, i, K, J) +
                           RI * s->p.grav_y * (C3(CTH, i, K, J) * DYF[J - 1] + C3(CTH, i, K, J - 1) * DYF[J]) / (2.0 * DY[J]);
    }
    for (int J = JSTART; J <= JEND; ++J)
      for (int K = 0; K <= TNKZ; ++K)
        for (int i = 0; i <= NXP - 1; ++i) C3(CF1, i, K, J) = C3(CF1, i, K, J) + RI * s->p.grav_x * C3(CTH, i, K, J);
    for (int J = JSTART; J <= JEND; ++J)
      for (int K = 0; K <= TNKZ; ++K)
        for (int i = 0; i <= NXP - 1; ++i) C3(CF3, i, K, J) = C3(CF3, i, K, J) + RI * s->p.grav_z * C3(CTH, i, K, J);
    int JS = s->JSTART_TH[N], JE = s->JEND_TH[N];
    for (int J = JS; J <= JE; ++J)
      for (int K = 0; K <= TNKZ; ++K)
        for (int i = 0; i <= NXP - 1; ++i) C3(CRTH, i, K, J) = C3(CTH, i, K, J);
    if (RK_STEP > 1)
      for (int J = JS; J <= JE; ++J)
        for (int K = 0; K <= TNKZ; ++K)
          for (int i = 0; i <= NXP - 1; ++i) C3(CRTH, i, K, J) = C3(CRTH, i, K, J) + TEMP3 * C3(CFTH, i, K, J);
    for (int J = JS; J <= JE; ++J)
      for (int K = 0; K <= TNKZ; ++K)
        for (int i = 0; i <= NXP - 1; ++i)
          C3(CFTH, i, K, J) = -(((NU / PR) * pow(KX2[i] + KZ2[K], BETA)) * C3(CTH, i, K, J));
    for (int J = JS; J <= JE; ++J)
      for (int K = 0; K <= TNKZ; ++K)
        for (int i = 0; i <= NXP - 1; ++i)
          C3(CFTH, i, K, J) = C3(CFTH, i, K, J) - C3(CU1, i, K, J) * s->p.drhodx[N] - C3(CU3, i, K, J) * s->p.drhodz[N];
  }
  /* USER_RHS_CHAN_FOURIER (channel.f:297): empty in the reference. */

  /* channel.f:303-391: the subgrid model after a spin-up of 100 steps of a
   * new flow (it leaves U1..3 physical), else the plain transforms */
  const int les_now = s->p.les && (!s->p.create_new_flow || s->TIME_STEP > 100);
  if (les_now) {
    dor_les_chan(s);
    apply_bc_vel_phys(s);
    /* channel.f:313-375: eddy diffusivity KAPPA_T = NU_T of the scalars'
     * horizontal fluxes, explicit */
    for (int N = 0; N < s->NTH; ++N) {
      cplx *CTH = s->CTH + s->cfield * N, *CFTH = s->CFTH + s->cfield * N;
      for (int d = 0; d < 2; ++d) {
        for (int J = 1; J <= NY + 1; ++J)
          for (int K = 0; K <= TNKZ; ++K)
            for (int i = 0; i <= NXP - 1; ++i) C3(CS1, i, K, J) = (d ? CIKZ(K) : CIKX(i)) * C3(CTH, i, K, J);
        dor_fft_xz_to_physical(s, (double *)CS1, 0, NY + 1);
        for (int J = 1; J <= NY + 1; ++J)
          for (int K = 0; K <= NZP - 1; ++K)
            for (int i = 0; i <= NXM; ++i) R3D(S1, i, K, J) = R3D(s->NU_T, i, K, J) * R3D(S1, i, K, J);
        dor_fft_xz_to_fourier(s, S1, 0, NY + 1);
        for (int J = s->JSTART_TH[N]; J <= s->JEND_TH[N]; ++J)
          for (int K = 0; K <= TNKZ; ++K)
            for (int i = 0; i <= NXP - 1; ++i) C3(CFTH, i, K, J) = C3(CFTH, i, K, J) + (d ? CIKZ(K) : CIKX(i)) * C3(CS1, i, K, J);
      }
    }
  } else {
    dor_fft_xz_to_physical(s, (double *)CU1, 0, NY + 1);
    dor_fft_xz_to_physical(s, (double *)CU2, 0, NY + 1);
    dor_fft_xz_to_physical(s, (double *)CU3, 0, NY + 1);
  }
  for (int N = 0; N < s->NTH; ++N) {
    memcpy(CS1, s->CTH + s->cfield * N, sizeof(cplx) * s->cfield);
    dor_fft_xz_to_physical(s, (double *)CS1, 0, NY + 1);
    memcpy(s->TH + s->rfield * N, S1, 8 * s->rfield);
  }

  /* Nonlinear terms, channel.f:410-571 */
  /* U1*U3 */
  for (int J = JSTART; J <= JEND; ++J)
    for (int K = 0; K <= NZP - 1; ++K)
      for (int i = 0; i <= NXM; ++i) R3D(S1, i, K, J) = R3D(U3, i, K, J) * R3D(U1, i, K, J);
  dor_fft_xz_to_fourier(s, S1, 0, NY + 1);
  for (int J = JSTART; J <= JEND; ++J)
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) {
        C3(CF1, i, K, J) = C3(CF1, i, K, J) - CIKZ(K) * C3(CS1, i, K, J);
        C3(CF3, i, K, J) = C3(CF3, i, K, J) - CIKX(i) * C3(CS1, i, K, J);
      }
  /* U1*U1 */
  for (int J = JSTART; J <= JEND; ++J)
    for (int K = 0; K <= NZP - 1; ++K)
      for (int i = 0; i <= NXM; ++i) R3D(S1, i, K, J) = R3D(U1, i, K, J) * R3D(U1, i, K, J);
  dor_fft_xz_to_fourier(s, S1, 0, NY + 1);
  for (int J = JSTART; J <= JEND; ++J)
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) C3(CF1, i, K, J) = C3(CF1, i, K, J) - CIKX(i) * C3(CS1, i, K, J);
  /* U3*U3 */
  for (int J = JSTART; J <= JEND; ++J)
    for (int K = 0; K <= NZP - 1; ++K)
      for (int i = 0; i <= NXM; ++i) R3D(S1, i, K, J) = R3D(U3, i, K, J) * R3D(U3, i, K, J);
  dor_fft_xz_to_fourier(s, S1, 0, NY + 1);
  for (int J = JSTART; J <= JEND; ++J)
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) C3(CF3, i, K, J) = C3(CF3, i, K, J) - CIKZ(K) * C3(CS1, i, K, J);
  /* U1*U2 */
  for (int J = 2; J <= NY; ++J)
    for (int K = 0; K <= NZP - 1; ++K)
      for (int i = 0; i <= NXM; ++i)
        R3D(S1, i, K, J) = ((DYF[J] * R3D(U1, i, K, J) + DYF[J - 1] * R3D(U1, i, K, J - 1)) / (2. * DY[J])) * R3D(U2, i, K, J);
  dor_fft_xz_to_fourier(s, S1, 0, NY + 1);
  for (int J = 2; J <= NY; ++J)
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) C3(CF2, i, K, J) = C3(CF2, i, K, J) - CIKX(i) * C3(CS1, i, K, J);
  /* U3*U2 */
  for (int J = 2; J <= NY; ++J)
    for (int K = 0; K <= NZP - 1; ++K)
      for (int i = 0; i <= NXM; ++i)
        R3D(S1, i, K, J) = ((DYF[J] * R3D(U3, i, K, J) + DYF[J - 1] * R3D(U3, i, K, J - 1)) / (2. * DY[J])) * R3D(U2, i, K, J);
  dor_fft_xz_to_fourier(s, S1, 0, NY + 1);
  for (int J = 2; J <= NY; ++J)
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) C3(CF2, i, K, J) = C3(CF2, i, K, J) - CIKZ(K) * C3(CS1, i, K, J);
  /* d/dy(U1*U2) */
  for (int J = JSTART; J <= JEND; ++J)
    for (int K = 0; K <= NZP - 1; ++K)
      for (int i = 0; i <= NXM; ++i)
        R3D(S1, i, K, J) = (R3D(U1, i, K, J + 1) * R3D(U2, i, K, J + 1) + R3D(U1, i, K, J) * R3D(U2, i, K, J + 1) -
                            R3D(U1, i, K, J) * R3D(U2, i, K, J) - R3D(U1, i, K, J - 1) * R3D(U2, i, K, J)) /
                           (2.0 * DYF[J]);
  dor_fft_xz_to_fourier(s, S1, 0, NY + 1);
  for (int J = JSTART; J <= JEND; ++J)
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) C3(CF1, i, K, J) = C3(CF1, i, K, J) - C3(CS1, i, K, J);
  /* d/dy(U3*U2) */
  for (int J = JSTART; J <= JEND; ++J)
    for (int K = 0; K <= NZP - 1; ++K)
      for (int i = 0; i <= NXM; ++i)
        R3D(S1, i, K, J) = (R3D(U3, i, K, J + 1) * R3D(U2, i, K, J + 1) + R3D(U3, i, K, J) * R3D(U2, i, K, J + 1) -
                            R3D(U3, i, K, J) * R3D(U2, i, K, J) - R3D(U3, i, K, J - 1) * R3D(U2, i, K, J)) /
                           (2.0 * DYF[J]);
  dor_fft_xz_to_fourier(s, S1, 0, NY + 1);
  for (int J = JSTART; J <= JEND; ++J)
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) C3(CF3, i, K, J) = C3(CF3, i, K, J) - C3(CS1, i, K, J);
  /* d/dy(U2*U2) */
  for (int J = 2; J <= NY; ++J)
    for (int K = 0; K <= NZP - 1; ++K)
      for (int i = 0; i <= NXM; ++i) {
        double a = R3D(U2, i, K, J) + R3D(U2, i, K, J + 1);
        double b = R3D(U2, i, K, J) + R3D(U2, i, K, J - 1);
        R3D(S1, i, K, J) = (0.25 * (a * a) - 0.25 * (b * b)) / DY[J];
      }
  dor_fft_xz_to_fourier(s, S1, 0, NY + 1);
  for (int J = 2; J <= NY; ++J)
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) C3(CF2, i, K, J) = C3(CF2, i, K, J) - C3(CS1, i, K, J);
  /* USER_RHS_CHAN_PHYSICAL (channel.f:578): empty in the reference. */

  /* channel.f:581-595 */
  for (int J = JSTART; J <= JEND; ++J)
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) {
        C3(CR1, i, K, J) = C3(CR1, i, K, J) + TEMP5 * C3(CF1, i, K, J);
        C3(CR3, i, K, J) = C3(CR3, i, K, J) + TEMP5 * C3(CF3, i, K, J);
      }
  for (int J = 2; J <= NY; ++J)
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) C3(CR2, i, K, J) = C3(CR2, i, K, J) + TEMP5 * C3(CF2, i, K, J);
  /* channel.f:598-600 */
  dor_fft_xz_to_physical(s, (double *)CR1, 0, NY + 1);
  dor_fft_xz_to_physical(s, (double *)CR2, 2, NY);
  dor_fft_xz_to_physical(s, (double *)CR3, 0, NY + 1);
  /* channel.f:606-626 explicit Crank-Nicolson part */
  for (int J = JSTART; J <= JEND; ++J)
    for (int K = 0; K <= NZP - 1; ++K)
      for (int i = 0; i <= NXM; ++i) {
        R3D(R1, i, K, J) = R3D(R1, i, K, J) + TEMP1 * NU_V_SCALE *
                               (((R3D(U1, i, K, J + 1) - R3D(U1, i, K, J)) / DY[J + 1] -
                                 (R3D(U1, i, K, J) - R3D(U1, i, K, J - 1)) / DY[J]) / DYF[J]);
        R3D(R3, i, K, J) = R3D(R3, i, K, J) + TEMP1 * NU_V_SCALE *
                               (((R3D(U3, i, K, J + 1) - R3D(U3, i, K, J)) / DY[J + 1] -
                                 (R3D(U3, i, K, J) - R3D(U3, i, K, J - 1)) / DY[J]) / DYF[J]);
      }
  for (int J = 2; J <= NY; ++J)
    for (int K = 0; K <= NZP - 1; ++K)
      for (int i = 0; i <= NXM; ++i)
        R3D(R2, i, K, J) = R3D(R2, i, K, J) + TEMP1 * NU_V_SCALE *
                               (((R3D(U2, i, K, J + 1) - R3D(U2, i, K, J)) / DYF[J] -
                                 (R3D(U2, i, K, J) - R3D(U2, i, K, J - 1)) / DYF[J - 1]) / DY[J]);
  /* channel.f:631-659: the eddy viscosity's explicit Crank-Nicolson half
   * (NU_T at GY; interpolated to GYF for U2) */
  const double *NU_T = s->NU_T;
  if (s->p.les) {
    for (int J = JSTART; J <= JEND; ++J)
      for (int K = 0; K <= NZP - 1; ++K)
        for (int i = 0; i <= NXM; ++i) {
          R3D(R1, i, K, J) = R3D(R1, i, K, J) + TEMP2 * ((R3D(NU_T, i, K, J + 1) * (R3D(U1, i, K, J + 1) - R3D(U1, i, K, J)) / DY[J + 1] -
                                                          R3D(NU_T, i, K, J) * (R3D(U1, i, K, J) - R3D(U1, i, K, J - 1)) / DY[J]) / DYF[J]);
          R3D(R3, i, K, J) = R3D(R3, i, K, J) + TEMP2 * ((R3D(NU_T, i, K, J + 1) * (R3D(U3, i, K, J + 1) - R3D(U3, i, K, J)) / DY[J + 1] -
                                                          R3D(NU_T, i, K, J) * (R3D(U3, i, K, J) - R3D(U3, i, K, J - 1)) / DY[J]) / DYF[J]);
        }
    for (int J = 2; J <= NY; ++J)
      for (int K = 0; K <= NZP - 1; ++K)
        for (int i = 0; i <= NXM; ++i)
          R3D(R2, i, K, J) = R3D(R2, i, K, J) +
                             TEMP2 * ((0.5 * (R3D(NU_T, i, K, J) + R3D(NU_T, i, K, J + 1)) * (R3D(U2, i, K, J + 1) - R3D(U2, i, K, J)) / DYF[J] -
                                       0.5 * (R3D(NU_T, i, K, J) + R3D(NU_T, i, K, J - 1)) * (R3D(U2, i, K, J) - R3D(U2, i, K, J - 1)) / DYF[J - 1]) /
                                      DY[J]);
  }

  /* Scalars, channel.f:663-838 */
  for (int N = 0; N < s->NTH; ++N) {
    cplx *CFTH = s->CFTH + s->cfield * N, *CRTH = s->CRTH + s->cfield * N;
    double *TH = s->TH + s->rfield * N, *RTH = s->RTH + s->rfield * N;
    const double PR = s->p.pr[N];
    int JS = s->JSTART_TH[N], JE = s->JEND_TH[N];
    for (int J = JS; J <= JE; ++J)
      for (int K = 0; K <= NZP - 1; ++K)
        for (int i = 0; i <= NXM; ++i) R3D(S1, i, K, J) = R3D(TH, i, K, J) * R3D(U1, i, K, J);
    dor_fft_xz_to_fourier(s, S1, 0, NY + 1);
    for (int J = JS; J <= JE; ++J)
      for (int K = 0; K <= TNKZ; ++K)
        for (int i = 0; i <= NXP - 1; ++i) C3(CFTH, i, K, J) = C3(CFTH, i, K, J) - CIKX(i) * C3(CS1, i, K, J);
    for (int J = JS; J <= JE; ++J)
      for (int K = 0; K <= NZP - 1; ++K)
        for (int i = 0; i <= NXM; ++i) R3D(S1, i, K, J) = R3D(TH, i, K, J) * R3D(U3, i, K, J);
    dor_fft_xz_to_fourier(s, S1, 0, NY + 1);
    for (int J = JS; J <= JE; ++J)
      for (int K = 0; K <= TNKZ; ++K)
        for (int i = 0; i <= NXP - 1; ++i) C3(CFTH, i, K, J) = C3(CFTH, i, K, J) - CIKZ(K) * C3(CS1, i, K, J);
    for (int J = JS; J <= JE; ++J)
      for (int K = 0; K <= NZP - 1; ++K)
        for (int i = 0; i <= NXM; ++i)
          R3D(S1, i, K, J) = (R3D(TH, i, K, J + 1) * R3D(U2, i, K, J + 1) + R3D(TH, i, K, J) * R3D(U2, i, K, J + 1) -
                              R3D(TH, i, K, J) * R3D(U2, i, K, J) - R3D(TH, i, K, J - 1) * R3D(U2, i, K, J)) /
                             (2.0 * DYF[J]);
    dor_fft_xz_to_fourier(s, S1, 0, NY + 1);
    for (int J = JS; J <= JE; ++J)
      for (int K = 0; K <= TNKZ; ++K)
        for (int i = 0; i <= NXP - 1; ++i) C3(CFTH, i, K, J) = C3(CFTH, i, K, J) - C3(CS1, i, K, J);
    for (int J = JS; J <= JE; ++J)
      for (int K = 0; K <= TNKZ; ++K)
        for (int i = 0; i <= NXP - 1; ++i) C3(CRTH, i, K, J) = C3(CRTH, i, K, J) + TEMP5 * C3(CFTH, i, K, J);
    memcpy(CS1, CRTH, sizeof(cplx) * s->cfield);
    dor_fft_xz_to_physical(s, (double *)CS1, 0, NY + 1);
    memcpy(RTH, S1, 8 * s->rfield);
    for (int J = JS; J <= JE; ++J)
      for (int K = 0; K <= NZP - 1; ++K)
        for (int i = 0; i <= NXM; ++i)
          R3D(RTH, i, K, J) = R3D(RTH, i, K, J) + (TEMP1 / PR) * NU_V_SCALE *
                                  (((R3D(TH, i, K, J + 1) - R3D(TH, i, K, J)) / DY[J + 1] -
                                    (R3D(TH, i, K, J) - R3D(TH, i, K, J - 1)) / DY[J]) / DYF[J]);
    /* channel.f:748-760: eddy diffusivity, explicit half */
    if (s->p.les)
      for (int J = JS; J <= JE; ++J)
        for (int K = 0; K <= NZP - 1; ++K)
          for (int i = 0; i <= NXM; ++i)
            R3D(RTH, i, K, J) = R3D(RTH, i, K, J) + TEMP2 * ((R3D(NU_T, i, K, J + 1) * (R3D(TH, i, K, J + 1) - R3D(TH, i, K, J)) / DY[J + 1] -
                                                              R3D(NU_T, i, K, J) * (R3D(TH, i, K, J) - R3D(TH, i, K, J - 1)) / DY[J]) / DYF[J]);
    /* implicit solve, channel.f:771-833 */
#pragma omp parallel
    {
      size_t msz = (size_t)NX * (size_t)(NY + 2);
      double *L = malloc(8 * msz), *D = malloc(8 * msz), *Uu = malloc(8 * msz), *VEC = malloc(8 * msz);
      for (size_t q = 0; q < msz; ++q) { L[q] = 0.; D[q] = 1.; Uu[q] = 0.; VEC[q] = 0.; }
#pragma omp for schedule(static)
      for (int K = 0; K <= NZP - 1; ++K) {
        for (int J = JS; J <= JE; ++J)
          for (int i = 0; i <= NXM; ++i) {
            L[MI(i, J)] = -(TEMP1 / PR * NU_V_SCALE) / (DY[J] * DYF[J]);
            D[MI(i, J)] = 1. + (TEMP1 / PR * NU_V_SCALE) / (DY[J + 1] * DYF[J]) + (TEMP1 / PR * NU_V_SCALE) / (DY[J] * DYF[J]);
            Uu[MI(i, J)] = -(TEMP1 / PR * NU_V_SCALE) / (DY[J + 1] * DYF[J]);
            VEC[MI(i, J)] = R3D(RTH, i, K, J);
          }
        if (s->p.les) les_rows_gy(s, K, JS, JE, TEMP2, L, D, Uu); /* channel.f:793-807 */
        bc_u13th(NX, NY, DY, s->p.th_bc_ymin[N], s->p.th_bc_ymin_c1[N], s->p.th_bc_ymax[N], s->p.th_bc_ymax_c1[N], L, D, Uu, VEC);
        dor_thomas_real(L, D, Uu, VEC, NY, NX);
        for (int J = JS - 1; J <= JE + 1; ++J)
          for (int i = 0; i <= NXM; ++i) R3D(TH, i, K, J) = VEC[MI(i, J)];
      }
      free(L); free(D); free(Uu); free(VEC);
    }
  }

  /* Velocity implicit solves, channel.f:842-1029 */
  for (int comp = 2; comp <= 4; ++comp) { /* order: U2 (2), U1 (3), U3 (4) */
    double *Uc = (comp == 2) ? U2 : (comp == 3) ? U1 : U3;
    double *Rc = (comp == 2) ? R2 : (comp == 3) ? R1 : R3;
#pragma omp parallel
    {
      size_t msz = (size_t)NX * (size_t)(NY + 2);
      double *L = malloc(8 * msz), *D = malloc(8 * msz), *Uu = malloc(8 * msz), *VEC = malloc(8 * msz);
      for (size_t q = 0; q < msz; ++q) { L[q] = 0.; D[q] = 1.; Uu[q] = 0.; VEC[q] = 0.; }
#pragma omp for schedule(static)
      for (int K = 0; K <= NZP - 1; ++K) {
        if (comp == 2) {
          for (int J = 2; J <= NY; ++J)
            for (int i = 0; i <= NXM; ++i) {
              L[MI(i, J)] = -TEMP1 * NU_V_SCALE / (DYF[J - 1] * DY[J]);
              D[MI(i, J)] = 1. + TEMP1 * NU_V_SCALE / (DYF[J] * DY[J]) + TEMP1 * NU_V_SCALE / (DYF[J - 1] * DY[J]);
              Uu[MI(i, J)] = -TEMP1 * NU_V_SCALE / (DYF[J] * DY[J]);
              VEC[MI(i, J)] = R3D(Rc, i, K, J);
            }
          if (s->p.les) /* channel.f:862-875 */
            for (int J = 2; J <= NY; ++J)
              for (int i = 0; i <= NXM; ++i) {
                L[MI(i, J)] = L[MI(i, J)] - TEMP2 * 0.5 * (R3D(NU_T, i, K, J) + R3D(NU_T, i, K, J - 1)) / (DYF[J - 1] * DY[J]);
                D[MI(i, J)] = D[MI(i, J)] + TEMP2 * 0.5 * (R3D(NU_T, i, K, J) + R3D(NU_T, i, K, J + 1)) / (DYF[J] * DY[J]) +
                              TEMP2 * 0.5 * (R3D(NU_T, i, K, J) + R3D(NU_T, i, K, J - 1)) / (DYF[J - 1] * DY[J]);
                Uu[MI(i, J)] = Uu[MI(i, J)] - TEMP2 * 0.5 * (R3D(NU_T, i, K, J) + R3D(NU_T, i, K, J + 1)) / (DYF[J] * DY[J]);
              }
          bc_u2(NX, NY, DYF, s->p.v_bc_ymin, s->p.v_bc_ymin_c1, s->p.v_bc_ymax, s->p.v_bc_ymax_c1, L, D, Uu, VEC);
        } else {
          for (int J = JSTART; J <= JEND; ++J)
            for (int i = 0; i <= NXM; ++i) {
              L[MI(i, J)] = -TEMP1 * NU_V_SCALE / (DY[J] * DYF[J]);
              D[MI(i, J)] = 1. - TEMP1 * NU_V_SCALE * (-1. / (DY[J + 1] * DYF[J]) - 1. / (DY[J] * DYF[J]));
              Uu[MI(i, J)] = -TEMP1 * NU_V_SCALE / (DY[J + 1] * DYF[J]);
              VEC[MI(i, J)] = R3D(Rc, i, K, J);
            }
          if (s->p.les) les_rows_gy(s, K, JSTART, JEND, TEMP2, L, D, Uu); /* channel.f:927-941, 991-1005 */
          if (comp == 3)
            bc_u13th(NX, NY, DY, s->p.u_bc_ymin, s->p.u_bc_ymin_c1, s->p.u_bc_ymax, s->p.u_bc_ymax_c1, L, D, Uu, VEC);
          else
            bc_u13th(NX, NY, DY, s->p.w_bc_ymin, s->p.w_bc_ymin_c1, s->p.w_bc_ymax, s->p.w_bc_ymax_c1, L, D, Uu, VEC);
        }
        dor_thomas_real(L, D, Uu, VEC, NY, NX);
        int j0 = (comp == 2) ? 1 : JSTART - 1, j1 = (comp == 2) ? NY + 1 : JEND + 1;
        for (int J = j0; J <= j1; ++J)
          for (int i = 0; i <= NXM; ++i) R3D(Uc, i, K, J) = VEC[MI(i, J)];
      }
      free(L); free(D); free(Uu); free(VEC);
    }
  }

  /* channel.f:1037-1040 */
  if (s->p.variable_dt && RK_STEP == 3 && (s->TIME_STEP % s->p.update_dt) == 0) dor_courant(s);

  /* channel.f:1043-1050 */
  dor_fft_xz_to_fourier(s, U1, 0, NY + 1);
  dor_fft_xz_to_fourier(s, U2, 0, NY + 1);
  dor_fft_xz_to_fourier(s, U3, 0, NY + 1);
  for (int N = 0; N < s->NTH; ++N) {
    memcpy(S1, s->TH + s->rfield * N, 8 * s->rfield);
    dor_fft_xz_to_fourier(s, S1, 0, NY + 1);
    memcpy(s->CTH + s->cfield * N, CS1, sizeof(cplx) * s->cfield);
  }

  dor_rem_div_chan(s);

  /* channel.f:1059-1065 */
  for (int J = JSTART; J <= JEND; ++J)
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) C3(CP, i, K, J) = C3(CP, i, K, J) + C3(CR1, i, K, J) / TEMP4;
  /* GHOST_CHAN_MPI: no-op when NPROCY=1 (mpi.f:25) */
  (void)KZ;
#undef C3
#undef R3D
#undef CIKX
#undef CIKZ
}

/* REM_DIV_CHAN channel.f:1103-1214 (USE_MPI branch) with
 * APPLY_BC_REM_DIV_MPI mpi.f:1239-1278 */
void dor_rem_div_chan(dor_state *s) {
  const int NY = s->NY, NXP = s->NXP, TNKZ = s->TNKZ;
  double *MATL_C = s->MATL_C, *MATD_C = s->MATD_C, *MATU_C = s->MATU_C;
  cplx *VEC_C = s->VEC_C;
  const double *DY = s->DY, *DYF = s->DYF, *KX = s->KX, *KZ = s->KZ, *KX2 = s->KX2, *KZ2 = s->KZ2;
#define MC(i, j) ((size_t)(i) + (size_t)(NXP + 1) * (size_t)(j))
#define C3(a, i, k, j) a[CIDX(s, i, k, j)]
  for (int J = 0; J <= NY + 1; ++J)
    for (int i = 0; i <= NXP - 1; ++i) {
      MATL_C[MC(i, J)] = 0.;
      MATD_C[MC(i, J)] = 1.;
      MATU_C[MC(i, J)] = 0.;
      VEC_C[MC(i, J)] = 0.;
    }
  for (int K = 0; K <= TNKZ; ++K) {
    for (int J = 1; J <= NY; ++J)
      for (int i = 0; i <= NXP - 1; ++i) {
        MATL_C[MC(i, J)] = 1. / (DY[J] * DYF[J]);
        MATD_C[MC(i, J)] = -KX2[i] - KZ2[K] - 1. / (DY[J + 1] * DYF[J]) - 1. / (DY[J] * DYF[J]);
        MATU_C[MC(i, J)] = 1. / (DY[J + 1] * DYF[J]);
      }
    for (int J = 1; J <= NY; ++J)
      for (int i = 0; i <= NXP - 1; ++i)
        VEC_C[MC(i, J)] = ((I * KX[i]) * C3(s->CU1, i, K, J) + (C3(s->CU2, i, K, J + 1) - C3(s->CU2, i, K, J)) / DYF[J] +
                           (I * KZ[K]) * C3(s->CU3, i, K, J));
    /* APPLY_BC_REM_DIV_MPI */
    for (int i = 0; i <= NXP - 1; ++i) {
      if (K == 0 && i == 0) {
        MATL_C[MC(i, 1)] = 0.; MATD_C[MC(i, 1)] = 1.; MATU_C[MC(i, 1)] = 0.; VEC_C[MC(i, 1)] = 0.;
      } else {
        MATL_C[MC(i, 1)] = 0.; MATD_C[MC(i, 1)] = 1.; MATU_C[MC(i, 1)] = -1.; VEC_C[MC(i, 1)] = 0.;
      }
    }
    for (int i = 0; i <= NXP - 1; ++i) {
      MATL_C[MC(i, NY)] = 1.; MATD_C[MC(i, NY)] = -1.; MATU_C[MC(i, NY)] = 0.; VEC_C[MC(i, NY)] = 0.;
    }
    dor_thomas_complex(MATL_C, MATD_C, MATU_C, (double *)VEC_C, NY, NXP);
    for (int J = 1; J <= NY; ++J)
      for (int i = 0; i <= NXP - 1; ++i) C3(s->CR1, i, K, J) = VEC_C[MC(i, J)];
  }
  for (int J = 1; J <= NY; ++J)
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) {
        C3(s->CU1, i, K, J) = C3(s->CU1, i, K, J) - (I * KX[i]) * C3(s->CR1, i, K, J);
        C3(s->CU3, i, K, J) = C3(s->CU3, i, K, J) - (I * KZ[K]) * C3(s->CR1, i, K, J);
      }
  for (int J = 2; J <= NY; ++J)
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i)
        C3(s->CU2, i, K, J) = C3(s->CU2, i, K, J) - (C3(s->CR1, i, K, J) - C3(s->CR1, i, K, J - 1)) / DY[J];
#undef MC
#undef C3
}

/* POISSON_P_CHAN channel.f:1217-1453, USE_MPI branch: boundary rows of
 * APPLY_BC_POISSON_MPI mpi.f:1280-1310 and THOMAS_*_COMPLEX_MPI
 * mpi.f:708-865.  CP(J), J=1..NY, from CU (and CTH): the Poisson equation
 * for the pressure with right-hand side (dui/dxj)(duj/dxi) plus buoyancy.
 * As in the reference, CF1..CF3 (and their physical views) and CS1 are
 * used as scratch and hold junk on return. */
void dor_poisson_p_chan(dor_state *s) {
  const int NY = s->NY, NXP = s->NXP, TNKZ = s->TNKZ, NZP = s->NZP, NXM = s->NXM;
  const double *DY = s->DY, *DYF = s->DYF, *GYF = s->GYF, *KX = s->KX, *KZ = s->KZ, *KX2 = s->KX2, *KZ2 = s->KZ2;
  double *MATL_C = s->MATL_C, *MATD_C = s->MATD_C, *MATU_C = s->MATU_C;
  cplx *VEC_C = s->VEC_C;
  cplx *CU1 = s->CU1, *CU2 = s->CU2, *CU3 = s->CU3, *CF1 = s->CF1, *CF2 = s->CF2, *CF3 = s->CF3, *CS1 = s->CS1;
  double *F1 = PHYS(CF1), *F2 = PHYS(CF2), *F3 = PHYS(CF3);
#define MC(i, j) ((size_t)(i) + (size_t)(NXP + 1) * (size_t)(j))
#define C3(a, i, k, j) a[CIDX(s, i, k, j)]
#define R3D(a, i, k, j) a[RIDX(s, i, k, j)]
  /* diagonal terms: (dU1/dx)^2 + (dU2/dy)^2 + (dU3/dz)^2, :1230-1270 */
  for (int J = 2; J <= NY; ++J)
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) {
        C3(CF1, i, K, J) = (I * KX[i]) * C3(CU1, i, K, J);
        C3(CF2, i, K, J) = (C3(CU2, i, K, J + 1) - C3(CU2, i, K, J)) / DYF[J];
        C3(CF3, i, K, J) = (I * KZ[K]) * C3(CU3, i, K, J);
      }
  dor_fft_xz_to_physical(s, F1, 0, NY + 1);
  dor_fft_xz_to_physical(s, F2, 0, NY + 1);
  dor_fft_xz_to_physical(s, F3, 0, NY + 1);
  for (int J = 2; J <= NY; ++J)
    for (int K = 0; K <= NZP - 1; ++K)
      for (int i = 0; i <= NXM; ++i) {
        R3D(F1, i, K, J) = R3D(F1, i, K, J) * R3D(F1, i, K, J);
        R3D(F2, i, K, J) = R3D(F2, i, K, J) * R3D(F2, i, K, J);
        R3D(F3, i, K, J) = R3D(F3, i, K, J) * R3D(F3, i, K, J);
      }
  dor_fft_xz_to_fourier(s, F1, 0, NY + 1);
  dor_fft_xz_to_fourier(s, F2, 0, NY + 1);
  dor_fft_xz_to_fourier(s, F3, 0, NY + 1);
  for (int J = 2; J <= NY; ++J)
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) C3(CS1, i, K, J) = C3(CF1, i, K, J) + C3(CF2, i, K, J) + C3(CF3, i, K, J);
  /* off-diagonal terms 2 (a)(b), :1280-1390 */
  for (int term = 0; term < 3; ++term) {
    for (int J = 2; J <= NY; ++J)
      for (int K = 0; K <= TNKZ; ++K)
        for (int i = 0; i <= NXP - 1; ++i) {
          if (term == 0) {
            C3(CF1, i, K, J) = (C3(CU1, i, K, J + 1) - C3(CU1, i, K, J - 1)) / (2. * DYF[J]);
            C3(CF2, i, K, J) = (I * KX[i]) * 0.5 * (C3(CU2, i, K, J) + C3(CU2, i, K, J + 1));
          } else if (term == 1) {
            C3(CF1, i, K, J) = (C3(CU3, i, K, J + 1) - C3(CU3, i, K, J - 1)) / (2. * DYF[J]);
            C3(CF2, i, K, J) = (I * KZ[K]) * 0.5 * (C3(CU2, i, K, J) + C3(CU2, i, K, J + 1));
          } else {
            C3(CF1, i, K, J) = (I * KZ[K]) * C3(CU1, i, K, J);
            C3(CF2, i, K, J) = (I * KX[i]) * C3(CU3, i, K, J);
          }
        }
    dor_fft_xz_to_physical(s, F1, 0, NY + 1);
    dor_fft_xz_to_physical(s, F2, 0, NY + 1);
    for (int J = 2; J <= NY; ++J)
      for (int K = 0; K <= NZP - 1; ++K)
        for (int i = 0; i <= NXM; ++i) R3D(F1, i, K, J) = 2. * R3D(F1, i, K, J) * R3D(F2, i, K, J);
    dor_fft_xz_to_fourier(s, F1, 0, NY + 1);
    for (int J = 2; J <= NY; ++J)
      for (int K = 0; K <= TNKZ; ++K)
        for (int i = 0; i <= NXP - 1; ++i) C3(CS1, i, K, J) = C3(CS1, i, K, J) + C3(CF1, i, K, J);
  }
  /* buoyancy, skipping the (0,0) mode, :1394-1406 */
  for (int N = 0; N < s->NTH; ++N) {
    const cplx *CTH = s->CTH + s->cfield * N;
    const double RI = s->p.ri[N];
    for (int J = 2; J <= NY; ++J)
      for (int K = 0; K <= TNKZ; ++K)
        for (int i = 0; i <= NXP - 1; ++i)
          if (i != 0 || K != 0)
            C3(CS1, i, K, J) =
                C3(CS1, i, K, J) + RI * (C3(CTH, i, K, J + 1) - C3(CTH, i, K, J - 1)) / (GYF[J + 1] - GYF[J - 1]);
  }
  /* the tridiagonal systems, :1410-1444 */
  for (int J = 0; J <= NY + 1; ++J)
    for (int i = 0; i <= NXP - 1; ++i) {
      MATL_C[MC(i, J)] = 0.;
      MATD_C[MC(i, J)] = 1.;
      MATU_C[MC(i, J)] = 0.;
      VEC_C[MC(i, J)] = 0.;
    }
  for (int K = 0; K <= TNKZ; ++K) {
    for (int J = 2; J <= NY; ++J)
      for (int i = 0; i <= NXP - 1; ++i) {
        MATL_C[MC(i, J)] = 1. / (DY[J] * DYF[J]);
        MATD_C[MC(i, J)] = -KX2[i] - KZ2[K] - 1. / (DY[J + 1] * DYF[J]) - 1. / (DY[J] * DYF[J]);
        MATU_C[MC(i, J)] = 1. / (DY[J + 1] * DYF[J]);
        VEC_C[MC(i, J)] = -1. * C3(CS1, i, K, J);
      }
    /* APPLY_BC_POISSON_MPI */
    for (int i = 0; i <= NXP - 1; ++i) {
      if (i == 0 && K == 0) {
        MATD_C[MC(i, 1)] = 1.; MATU_C[MC(i, 1)] = 0.; VEC_C[MC(i, 1)] = 0.;
      } else {
        MATD_C[MC(i, 1)] = 1.; MATU_C[MC(i, 1)] = -1.; VEC_C[MC(i, 1)] = 0.;
      }
    }
    for (int i = 0; i <= NXP - 1; ++i) {
      MATD_C[MC(i, NY)] = -1.; MATL_C[MC(i, NY)] = 1.; VEC_C[MC(i, NY)] = 0.;
    }
    dor_thomas_complex(MATL_C, MATD_C, MATU_C, (double *)VEC_C, NY, NXP);
    for (int J = 1; J <= NY; ++J)
      for (int i = 0; i <= NXP - 1; ++i) C3(s->CP, i, K, J) = VEC_C[MC(i, J)];
  }
#undef MC
#undef C3
#undef R3D
}

/* COURANT courant.f:1-65 (get_minimum_mpi is the identity on one rank) */
void dor_courant(dor_state *s) {
  const int NY = s->NY, NZP = s->NZP, NXM = s->NXM;
  double *U1 = PHYS(s->CU1), *U2 = PHYS(s->CU2), *U3 = PHYS(s->CU3);
  double dt = 999.0;
  double dx1 = s->DX1, dy1 = s->DY[1];
  double m = dx1 < dy1 ? dx1 : dy1;
  double c = 0.5 * (m * m) / s->p.nu;
  dt = dt < c ? dt : c;
  for (int n = 0; n < s->NTH; ++n) {
    double v = dt * s->p.nu / (s->p.nu / s->p.pr[n]);
    dt = dt < v ? dt : v;
  }
  if (s->p.i_ro != 0.0) {
    double v = 2.0 * s->PI / fabs(s->p.i_ro) / 20.0;
    dt = dt < v ? dt : v;
  }
  for (int n = 0; n < s->NTH; ++n) {
    if (s->p.ri[n] != 0.0) {
      double Nmax = sqrt(fabs(s->p.ri[n] * s->p.th_bc_ymin_c1[n]));
      double v = (double)0.1f * 2.0 * s->PI / Nmax; /* 0.1 is a default-real literal */
      dt = dt < v ? dt : v;
    }
  }
  for (int J = 1; J <= NY; ++J)
    for (int K = 0; K <= NZP - 1; ++K)
      for (int i = 0; i <= NXM; ++i) {
        double dt_x = s->p.cfl * dx1 / fabs(U1[RIDX(s, i, K, J)]);
        double dt_y = s->p.cfl * s->DY[J] / fabs(U2[RIDX(s, i, K, J)]);
        double dt_z = s->p.cfl * s->DZ1 / fabs(U3[RIDX(s, i, K, J)]);
        double v = dt;
        if (dt_x < v) v = dt_x;
        if (dt_y < v) v = dt_y;
        if (dt_z < v) v = dt_z;
        dt = v;
      }
  if (dt <= 0)
    s->DELTA_T = 0.0001;
  else if (dt >= 999.)
    s->DELTA_T = 999.0;
  else
    s->DELTA_T = dt;
  s->H_BAR[0] = s->DELTA_T * (8.0 / 15.0);
  s->H_BAR[1] = s->DELTA_T * (2.0 / 15.0);
  s->H_BAR[2] = s->DELTA_T * (5.0 / 15.0);
}

/* diablo.f:66-110 for one step, stats/flow saving hooks omitted. */
void dor_time_step(dor_state *s) {
  s->TIME_STEP += 1;
  for (int rk = 1; rk <= 3; ++rk) dor_rk_chan_1(s, rk);
  s->TIME = s->TIME + s->DELTA_T;
}

/* ------------------------------------------------------------------------ */
/* Initial conditions                                                       */
/* ------------------------------------------------------------------------ */

static uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

double dor_gauss(uint64_t seed, uint64_t counter) {
  uint64_t a = splitmix64(seed * 0x100000001B3ull ^ splitmix64(2 * counter));
  uint64_t b = splitmix64(seed * 0x100000001B3ull ^ splitmix64(2 * counter + 1));
  double u1 = ((double)(a >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  double u2 = ((double)(b >> 11)) * (1.0 / 9007199254740992.0);
  return sqrt(-2.0 * log(u1)) * cos(2.0 * 3.141592653589793 * u2);
}

/* APPLY_BC_VEL_LOWER channel.f:2285-2400 and APPLY_BC_VEL_UPPER
 * :2402-2492 (RANKZ=0), via APPLY_BC_VEL_MPI mpi.f:1317-1331 */
void dor_apply_bc_vel(dor_state *s) {
  const int NY = s->NY, NXP = s->NXP, TNKZ = s->TNKZ;
  const double *DY = s->DY, *DYF = s->DYF;
  cplx *CU1 = s->CU1, *CU2 = s->CU2, *CU3 = s->CU3;
#define C3(a, i, k, j) a[CIDX(s, i, k, j)]
  const dor_params *p = &s->p;
  /* lower */
  if (p->u_bc_ymin == 0) {
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) { C3(CU1, i, K, 0) = 0.; C3(CU1, i, K, 1) = 0.; }
    C3(CU1, 0, 0, 1) = p->u_bc_ymin_c1;
    C3(CU1, 0, 0, 0) = p->u_bc_ymin_c1;
  } else {
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) { C3(CU1, i, K, 1) = C3(CU1, i, K, 2); C3(CU1, i, K, 0) = C3(CU1, i, K, 1); }
    C3(CU1, 0, 0, 1) = C3(CU1, 0, 0, 2) - DY[2] * p->u_bc_ymin_c1;
    C3(CU1, 0, 0, 0) = C3(CU1, 0, 0, 1) - DY[1] * p->u_bc_ymin_c1;
  }
  if (p->w_bc_ymin == 0) {
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) { C3(CU3, i, K, 0) = 0.; C3(CU3, i, K, 1) = 0.; }
    C3(CU3, 0, 0, 1) = p->w_bc_ymin_c1;
    C3(CU3, 0, 0, 0) = p->w_bc_ymin_c1;
  } else {
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) { C3(CU3, i, K, 1) = C3(CU3, i, K, 2); C3(CU3, i, K, 0) = C3(CU3, i, K, 1); }
    C3(CU3, 0, 0, 1) = C3(CU3, 0, 0, 2) - DY[2] * p->w_bc_ymin_c1;
    C3(CU3, 0, 0, 0) = C3(CU3, 0, 0, 1) - DY[1] * p->w_bc_ymin_c1;
  }
  if (p->v_bc_ymin == 0) {
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) {
        C3(CU2, i, K, 1) = 2.0 * p->v_bc_ymin_c1 - C3(CU2, i, K, 2);
        C3(CU2, i, K, 0) = C3(CU2, i, K, 1);
      }
  } else {
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) { C3(CU2, i, K, 1) = C3(CU2, i, K, 2); C3(CU2, i, K, 0) = C3(CU2, i, K, 1); }
    C3(CU2, 0, 0, 1) = C3(CU2, 0, 0, 2) - DYF[1] * p->v_bc_ymin_c1;
    C3(CU2, 0, 0, 0) = C3(CU2, 0, 0, 1) - DYF[1] * p->v_bc_ymin_c1;
  }
  /* upper */
  if (p->u_bc_ymax == 0) {
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) { C3(CU1, i, K, NY) = 0.; C3(CU1, i, K, NY + 1) = 0.; }
    C3(CU1, 0, 0, NY) = p->u_bc_ymax_c1;
    C3(CU1, 0, 0, NY + 1) = p->u_bc_ymax_c1;
  } else {
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) { C3(CU1, i, K, NY) = C3(CU1, i, K, NY - 1); C3(CU1, i, K, NY + 1) = C3(CU1, i, K, NY); }
    C3(CU1, 0, 0, NY) = C3(CU1, 0, 0, NY - 1) + DY[NY] * p->u_bc_ymax_c1;
    C3(CU1, 0, 0, NY + 1) = C3(CU1, 0, 0, NY) + DY[NY] * p->u_bc_ymax_c1;
  }
  if (p->w_bc_ymax == 0) {
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) { C3(CU3, i, K, NY) = 0.; C3(CU3, i, K, NY + 1) = 0.; }
    C3(CU3, 0, 0, NY) = p->w_bc_ymax_c1;
    C3(CU3, 0, 0, NY + 1) = p->w_bc_ymax_c1;
    C3(CU3, 0, 0, NY + 1) = 0.0;
  } else {
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) { C3(CU3, i, K, NY) = C3(CU3, i, K, NY - 1); C3(CU3, i, K, NY + 1) = C3(CU3, i, K, NY); }
    C3(CU3, 0, 0, NY) = C3(CU3, 0, 0, NY - 1) + DY[NY] * p->w_bc_ymax_c1;
    C3(CU3, 0, 0, NY + 1) = C3(CU3, 0, 0, NY) + DY[NY] * p->w_bc_ymax_c1;
  }
  if (p->v_bc_ymax == 0) {
    C3(CU2, 0, 0, NY + 1) = 2.0 * p->v_bc_ymax_c1 - C3(CU2, 0, 0, NY);
  } else {
    for (int K = 0; K <= TNKZ; ++K)
      for (int i = 0; i <= NXP - 1; ++i) { C3(CU2, i, K, NY) = C3(CU2, i, K, NY - 1); C3(CU2, i, K, NY + 1) = C3(CU2, i, K, NY); }
    C3(CU2, 0, 0, NY + 1) = C3(CU2, 0, 0, NY) + DY[NY] * p->v_bc_ymax_c1;
  }
#undef C3
}

/* CREATE_FLOW_CHAN set_ics.f:2-200 (one rank, USE_MPI=.TRUE.) with the
 * compiler RNG replaced by dor_gauss and an optional |y| mask. */
void dor_create_flow_chan(dor_state *s, int ic_type, double ubulk0, uint64_t seed, double amp, double ymask) {
  const int NY = s->NY, NZP = s->NZP, NXM = s->NXM;
  double *U1 = PHYS(s->CU1), *U2 = PHYS(s->CU2), *U3 = PHYS(s->CU3), *P = PHYS(s->CP);
  memset(s->CU1, 0, sizeof(cplx) * s->cfield);
  memset(s->CU2, 0, sizeof(cplx) * s->cfield);
  memset(s->CU3, 0, sizeof(cplx) * s->cfield);
  memset(s->CP, 0, sizeof(cplx) * s->cfield);
  for (int J = 0; J <= NY; ++J)
    for (int K = 0; K <= NZP - 1; ++K)
      for (int i = 0; i <= NXM; ++i) {
        double g = s->GYF[J];
        double u1 = 0.0;
        if (ic_type == 0) u1 = (3. / 2.) * ubulk0 * (1.0 - g * g);
        else if (ic_type == 2) u1 = g;
        else if (ic_type == 3) u1 = tanh(g);
        U1[RIDX(s, i, K, J)] = u1;
        U2[RIDX(s, i, K, J)] = 0.;
        U3[RIDX(s, i, K, J)] = 0.;
      }
  for (int J = 0; J <= NY + 1; ++J) {
    int on = (ymask <= 0.0) || (fabs(s->GYF[J]) < ymask);
    if (!on) continue;
    for (int K = 0; K <= NZP - 1; ++K)
      for (int i = 0; i <= NXM; ++i) {
        uint64_t c = 3ull * ((uint64_t)i + (uint64_t)s->NX * ((uint64_t)K + (uint64_t)s->NZ * (uint64_t)J));
        U1[RIDX(s, i, K, J)] += amp * dor_gauss(seed, c);
        U2[RIDX(s, i, K, J)] += amp * dor_gauss(seed, c + 1);
        U3[RIDX(s, i, K, J)] += amp * dor_gauss(seed, c + 2);
      }
  }
  dor_fft_xz_to_fourier(s, U1, 0, NY + 1);
  dor_fft_xz_to_fourier(s, U2, 0, NY + 1);
  dor_fft_xz_to_fourier(s, U3, 0, NY + 1);
  dor_fft_xz_to_fourier(s, P, 0, NY + 1);
  dor_apply_bc_vel(s);
  dor_rem_div_chan(s);
}

/* CREATE_TH_CHAN set_ics.f:204-273 */
void dor_create_th_chan(dor_state *s, int ic_type) {
  const int NY = s->NY, NZP = s->NZP, NXM = s->NXM;
  for (int N = 0; N < s->NTH; ++N) {
    double *TH = s->TH + s->rfield * N;
    memset(TH, 0, 8 * s->rfield);
    for (int K = 0; K <= NZP - 1; ++K)
      for (int i = 0; i <= NXM; ++i)
        for (int J = 1; J <= NY; ++J) {
          double g = s->GYF[J], v = 0.0;
          if (ic_type == 0) {
            double gx = (i * s->p.lx) / s->NX;
            v = sin(2.0 * s->PI * gx / s->p.lx) / (4.0 * s->PI * s->PI);
          } else if (ic_type == 1 || ic_type == 2) {
            if (s->p.th_bc_ymin[N] == 0 && s->p.th_bc_ymax[N] == 0)
              v = (g <= 2.0) ? (s->p.th_bc_ymax_c1[N] - s->p.th_bc_ymin_c1[N]) * (g + 1.) / 2.0 + s->p.th_bc_ymin_c1[N]
                             : s->p.th_bc_ymax_c1[N];
            else if (s->p.th_bc_ymin[N] == 1 && s->p.th_bc_ymax[N] == 1)
              v = s->p.th_bc_ymax_c1[N] * g;
          } else {
            v = tanh(g);
          }
          TH[RIDX(s, i, K, J)] = v;
        }
    double *S1 = PHYS(s->CS1);
    memcpy(S1, TH, 8 * s->rfield);
    dor_fft_xz_to_fourier(s, S1, 0, NY + 1);
    memcpy(s->CTH + s->cfield * N, s->CS1, sizeof(cplx) * s->cfield);
  }
}

/* SAVE_STATS_CHAN subset, save_stats.f:136-262 (one rank, RANKZ=0): mean
 * profiles from the (0,0) modes (:136-150), the velocities transformed to
 * physical space from a copy of CU (the reference saves CU in CR and
 * transforms CU in place, :130-160), rms (:166-192), Reynolds stresses on
 * the GY/GYF grids (:197-232), mean gradients (:235-238) and the mean
 * square shear (:241-262); UBULK by INTEGRATE_Y_VAR (mpi.f:1610-1632). */
void dor_save_stats(dor_state *s, double ly, double *out, double *ubulk) {
#define R3D(a, i, k, j) a[RIDX(s, i, k, j)]
  const int NY = s->NY, NZP = s->NZP, NXM = s->NXM, NX = s->NX, NZ = s->NZ;
  const double *DY = s->DY, *DYF = s->DYF, *GYF = s->GYF;
  const size_t ny2 = (size_t)NY + 2;
  memset(out, 0, sizeof(double) * ny2 * DOR_NSTATS);
  double *ume = out + DOR_ST_UME * ny2, *vme = out + DOR_ST_VME * ny2, *wme = out + DOR_ST_WME * ny2;
  double *urms = out + DOR_ST_URMS * ny2, *vrms = out + DOR_ST_VRMS * ny2, *wrms = out + DOR_ST_WRMS * ny2;
  double *uv = out + DOR_ST_UV * ny2, *uw = out + DOR_ST_UW * ny2, *wv = out + DOR_ST_WV * ny2;
  double *dudy = out + DOR_ST_DUDY * ny2, *dwdy = out + DOR_ST_DWDY * ny2, *shear = out + DOR_ST_SHEAR * ny2;
  for (size_t j = 0; j < ny2; ++j) {
    ume[j] = creal(s->CU1[CIDX(s, 0, 0, j)]);
    vme[j] = creal(s->CU2[CIDX(s, 0, 0, j)]);
    wme[j] = creal(s->CU3[CIDX(s, 0, 0, j)]);
  }
  /* INTEGRATE_Y_VAR(UME): RES = sum_{j=2..NY} 0.5 (VAR(j)+VAR(j-1)) DY(j), / LY */
  {
    double res = 0.;
    for (int j = 2; j <= NY; ++j) res = res + 0.5 * (ume[j] + ume[j - 1]) * DY[j];
    *ubulk = res / ly;
  }
  cplx *C[3];
  const cplx *src[3] = {s->CU1, s->CU2, s->CU3};
  for (int c = 0; c < 3; ++c) {
    C[c] = malloc(sizeof(cplx) * s->cfield);
    memcpy(C[c], src[c], sizeof(cplx) * s->cfield);
    dor_fft_xz_to_physical(s, (double *)C[c], 0, NY + 1);
  }
  const double *U1 = PHYS(C[0]), *U2 = PHYS(C[1]), *U3 = PHYS(C[2]);
  for (int j = 1; j <= NY; ++j) {
    double su = 0., sv = 0., sw = 0., suv = 0., suw = 0., swv = 0., ssh = 0.;
    for (int k = 0; k <= NZP - 1; ++k)
      for (int i = 0; i <= NXM; ++i) {
        const double u = R3D(U1, i, k, j), w = R3D(U3, i, k, j);
        su = su + (u - ume[j]) * (u - ume[j]);
        sv = sv + 0.5 * ((R3D(U2, i, k, j) - vme[j]) * (R3D(U2, i, k, j) - vme[j]) +
                         (R3D(U2, i, k, j + 1) - vme[j + 1]) * (R3D(U2, i, k, j + 1) - vme[j + 1]));
        sw = sw + (w - wme[j]) * (w - wme[j]);
        suv = suv + (DYF[j - 1] * (u - ume[j]) + DYF[j] * (R3D(U1, i, k, j - 1) - ume[j - 1])) / (2.0 * DY[j]) *
                        (R3D(U2, i, k, j) - vme[j]);
        swv = swv + (DYF[j - 1] * (w - wme[j]) + DYF[j] * (R3D(U3, i, k, j - 1) - wme[j - 1])) / (2.0 * DY[j]) *
                        (R3D(U2, i, k, j) - vme[j]);
        suw = suw + (u - ume[j]) * (w - wme[j]);
        const double a = (R3D(U1, i, k, j + 1) - R3D(U1, i, k, j - 1)) / (2.0 * DYF[j]);
        const double b = (R3D(U3, i, k, j + 1) - R3D(U3, i, k, j - 1)) / (2.0 * DYF[j]);
        ssh = ssh + a * a + b * b;
      }
    const double n = (double)NZ * (double)NX;
    urms[j] = sqrt(su / n);
    vrms[j] = sqrt(sv / n);
    wrms[j] = sqrt(sw / n);
    uv[j] = suv / n;
    uw[j] = suw / n;
    wv[j] = swv / n;
    dudy[j] = (ume[j] - ume[j - 1]) / (GYF[j] - GYF[j - 1]);
    dwdy[j] = (wme[j] - wme[j - 1]) / (GYF[j] - GYF[j - 1]);
    shear[j] = ssh / (double)(NX * NZ);
  }
  for (int c = 0; c < 3; ++c) free(C[c]);
#undef R3D
}
