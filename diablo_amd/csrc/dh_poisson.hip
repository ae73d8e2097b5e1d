// POISSON_P_CHAN (channel.f:1217-1453, USE_MPI branch with
// APPLY_BC_POISSON_MPI, mpi.f:1280-1310): the pressure CP of the current
// velocity, as at a restart without a pressure field (hdf5.f:537).
//
// Right-hand side CS1 = (dU1/dx)^2 + (dU2/dy)^2 + (dU3/dz)^2
//                     + 2 (dU1/dy)(dU2/dx) + 2 (dU3/dy)(dU2/dz) + 2 (dU1/dz)(dU3/dx)
//                     + sum_n RI(n) dTH(n)/dy   (not on the (0,0) mode),
// each product formed in physical space, in four stages of derivative
// fields (Fourier) -> transform -> product -> transform back -> accumulate.
// The Poisson matrix of each (kx,kz) mode is the one REM_DIV_CHAN solves
// (rows 2..NY-1 equal, boundary rows 1 and NY equal wherever the sweeps
// read them), so its tabulated multipliers and pivots are reused.
//
// Scratch: the context's CUN buffers (Fourier terms), the physical U
// buffers of U1..U3, and phi (CS1).  CU, CF and CTH are left unchanged (the
// reference uses CF1..3 as its scratch); CP changes on planes 1..NY.
#include "dh_fft_dev.h"

namespace dh {

struct PpArgs {
  Geo g;
  const double2 *cu[3];
  const double2 *cth;  // [n_th][NY+2][pf]
  long long cfs;       // Fourier field stride (NY+2)*pf
  int n_th;
  double ri[DH_MAX_TH];
  const double *gyf;
  double2 *t[3];   // Fourier terms of the current stage
  double2 *cs;     // CS1, planes 2..NY
  double *ph[3];   // physical views of t[0..2]
  long long npts;  // physical points of one field, (NY+2)*NX*NZ
};

// Fourier fields of stage `stage` on planes 2..NY (zero elsewhere; the
// products of the other planes never reach CS1), channel.f:1230-1238,
// 1271-1278, 1305-1312, 1339-1346.
__global__ void k_pp_terms(PpArgs a, int stage) {
#pragma clang fp contract(off)
  const Geo &g = a.g;
  const int J = blockIdx.y;
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= g.pf) return;
  const int kzp = (int)fdiv((unsigned)m, g.f_nkx1), kx = m - kzp * g.nkx1;
  const long long pf = g.pf, o = (long long)J * pf + m;
  double2 r0 = make_double2(0.0, 0.0), r1 = r0, r2 = r0;
  if (J >= 2 && J <= g.ny) {
    const double kxv = g.kx[kx], kzv = g.kz[kzp];
    if (stage == 0) {
      r0 = cmul_ik(kxv, a.cu[0][o]);
      const double2 d = csub(a.cu[1][o + pf], a.cu[1][o]);
      r1 = make_double2(d.x / g.dyf[J], d.y / g.dyf[J]);
      r2 = cmul_ik(kzv, a.cu[2][o]);
    } else if (stage == 1 || stage == 2) {
      const double2 *C = a.cu[stage == 1 ? 0 : 2];
      const double2 d = csub(C[o + pf], C[o - pf]);
      const double den = 2. * g.dyf[J];
      r0 = make_double2(d.x / den, d.y / den);
      // CIKX(I)*0.5*(CU2(J)+CU2(J+1)) (CIKZ for stage 2): (i k 0.5) times the sum
      r1 = cmul_ik((stage == 1 ? kxv : kzv) * 0.5, cadd(a.cu[1][o], a.cu[1][o + pf]));
    } else {
      r0 = cmul_ik(kzv, a.cu[0][o]);
      r1 = cmul_ik(kxv, a.cu[2][o]);
    }
  }
  a.t[0][o] = r0;
  a.t[1][o] = r1;
  if (stage == 0) a.t[2][o] = r2;
}

// Products in physical space: F_i = F_i**2 (stage 0), F1 = 2 F1 F2 (others).
__global__ void k_pp_products(PpArgs a, int stage) {
#pragma clang fp contract(off)
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.npts) return;
  if (stage == 0) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double v = a.ph[i][e];
      a.ph[i][e] = v * v;
    }
  } else {
    a.ph[0][e] = 2. * a.ph[0][e] * a.ph[1][e];
  }
}

// CS1 on planes 2..NY: the sum of stage 0, then each cross term in turn
// (channel.f:1263-1268, 1296-1302, 1330-1336, 1364-1370); stage 4 adds
// buoyancy, skipping the (0,0) mode (:1394-1406).
__global__ void k_pp_accum(PpArgs a, int stage) {
#pragma clang fp contract(off)
  const Geo &g = a.g;
  const int J = blockIdx.y + 2;
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= g.pf || J > g.ny) return;
  const long long pf = g.pf, o = (long long)J * pf + m;
  if (stage == 0) {
    a.cs[o] = cadd(cadd(a.t[0][o], a.t[1][o]), a.t[2][o]);
  } else if (stage < 4) {
    a.cs[o] = cadd(a.cs[o], a.t[0][o]);
  } else {
    const int kzp = (int)fdiv((unsigned)m, g.f_nkx1), kx = m - kzp * g.nkx1;
    if (kx == 0 && kzp == 0) return;
    double2 c = a.cs[o];
    const double dg = a.gyf[J + 1] - a.gyf[J - 1];
    for (int n = 0; n < a.n_th; ++n) {
      const double2 *T = a.cth + n * a.cfs;
      const double2 d = csub(T[o + pf], T[o - pf]);
      c = make_double2(c.x + (a.ri[n] * d.x) / dg, c.y + (a.ri[n] * d.y) / dg);
    }
    a.cs[o] = c;
  }
}

// One thread per mode: THOMAS_FORWARD/BACKWARD_COMPLEX_MPI (mpi.f:708-865)
// over rows 1..NY with right-hand side -CS1 on rows 2..NY-1 and 0 on the
// boundary rows, using the grid's multipliers ta, pivots tb and reciprocal
// pivots tr.  The forward values G_J are kept in CP, which the backward
// sweep overwrites with the solution.
__global__ void k_pp_solve(Geo g, const double2 *cs, const double *ta, const double *tb, const double *tr,
                           const double *ru, double2 *cp) {
#pragma clang fp contract(off)
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= g.pf) return;
  const int NY = g.ny;
  const long long pf = g.pf;
  const int kzp = (int)fdiv((unsigned)m, g.f_nkx1), kx = m - kzp * g.nkx1;
  const bool mean = (kx == 0 && kzp == 0);
  double2 gp = make_double2(0.0, 0.0);  // G_1 = 0
  for (int J = 2; J <= NY; ++J) {
    const long long o = (long long)J * pf + m;
    double2 v = make_double2(0.0, 0.0);
    if (J < NY) v = make_double2(-1. * cs[o].x, -1. * cs[o].y);
    const double A = ta[o];
    gp = make_double2(v.x + A * gp.x, v.y + A * gp.y);
    cp[o] = gp;
  }
  const double bny = tb[(long long)NY * pf + m];
  double2 x;
  x.x = (bny != 0.0) ? gp.x / bny : 0.0;
  x.y = (bny != 0.0) ? gp.y / bny : 0.0;
  cp[(long long)NY * pf + m] = x;
  for (int J = NY - 1; J >= 2; --J) {
    const long long o = (long long)J * pf + m;
    const double B = tb[o], R = tr[o], U = ru[J];
    const double2 G = cp[o];
    const double xr = div_by_recip(G.x - U * x.x, B, R);
    const double xi = div_by_recip(G.y - U * x.y, B, R);
    x.x = (B != 0.0) ? xr : 0.0;
    x.y = (B != 0.0) ? xi : 0.0;
    cp[o] = x;
  }
  // row 1: B = 1, G = 0, U = 0 (mean mode) or -1
  const double U1 = mean ? 0.0 : -1.0;
  cp[pf + m] = make_double2((0.0 - U1 * x.x) / 1.0, (0.0 - U1 * x.y) / 1.0);
}

hipError_t launch_poisson_p(Ctx &c) {
  const Geo &g = c.g;
  PpArgs a{};
  a.g = g;
  for (int i = 0; i < 3; ++i) a.cu[i] = c.cu[i];
  a.cth = c.cth;
  a.cfs = (long long)(g.ny + 2) * g.pf;
  a.n_th = c.p.n_th;
  for (int n = 0; n < DH_MAX_TH; ++n) a.ri[n] = c.p.ri[n];
  a.gyf = c.d_gyf;
  const long long fs = (long long)(g.ny + 2) * g.pp;
  for (int i = 0; i < 3; ++i) {
    a.t[i] = c.cun[i];
    a.ph[i] = c.uphys + i * fs;
  }
  a.cs = c.phi;
  a.npts = fs;
  const dim3 gm((g.pf + 255) / 256, g.ny + 2), ga((g.pf + 255) / 256, g.ny - 1);
  const int gp = (int)((fs + 255) / 256);
  hipError_t e;
  for (int stage = 0; stage < 4; ++stage) {
    const int nf = (stage == 0) ? 3 : 2, nb = (stage == 0) ? 3 : 1;
    hipLaunchKernelGGL(k_pp_terms, gm, dim3(256), 0, c.stream, a, stage);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    for (int i = 0; i < nf; ++i)
      if ((e = launch_fft_to_phys(c, a.t[i], a.ph[i], g.ny + 2)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_pp_products, dim3(gp), dim3(256), 0, c.stream, a, stage);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    for (int i = 0; i < nb; ++i)
      if ((e = launch_fft_to_four(c, a.ph[i], a.t[i], g.ny + 2)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_pp_accum, ga, dim3(256), 0, c.stream, a, stage);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  if (a.n_th > 0) {
    hipLaunchKernelGGL(k_pp_accum, ga, dim3(256), 0, c.stream, a, 4);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_pp_solve, dim3((g.pf + 255) / 256), dim3(256), 0, c.stream, g, (const double2 *)c.phi,
                     (const double *)c.rd_ta, (const double *)c.rd_tb, (const double *)c.rd_tr,
                     (const double *)c.rd_u, c.cp);
  return hipGetLastError();
}

}  // namespace dh
