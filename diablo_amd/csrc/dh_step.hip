// Kernels of one RK3/Crank-Nicolson substep (RK_CHAN_1, channel.f:62-1073)
// other than the plain plane transforms:
//   k_nonlinear   explicit Fourier right-hand side, nonlinear products and
//                 their forward transforms, RK combination and the inverse
//                 transform of the right-hand side (channel.f:77-600);
//   k_solve       explicit half of the y-viscous term plus the real
//                 tridiagonal Crank-Nicolson solve of every (x,z) column, and
//                 the CFL minimum (channel.f:606-1029, courant.f:34-43);
//   k_rem_div     divergence, pressure-Poisson solve per (kx,kz), projection
//                 and pressure update (channel.f:1103-1214, 1059-1065);
//   k_courant_fin the rest of COURANT (courant.f:45-62) on the device.
#include "dh_fft_dev.h"

namespace dh {

constexpr int kMaxQ = 8;  // Fourier modes owned per thread in k_nonlinear

struct StepArgs {
  Geo g;
  Clock *clk;
  int rk;  // 0..2
  const double2 *cu[3];
  const double2 *cp;
  double2 *cf[3];
  const double2 *cth;
  double2 *cfth;
  const double *uphys;
  double *rphys;
  long long fs;   // physical field stride (NY+2)*NX*NZ
  long long cfs;  // Fourier field stride (NY+2)*pf
  double nu, beta, i_ro, grav_x, grav_y, grav_z, px0, omega0, amp_omega0;
  int f_type, n_th;
  double ri[DH_MAX_TH], pr[DH_MAX_TH], drhodx[DH_MAX_TH], drhodz[DH_MAX_TH];
  int jstart_th[DH_MAX_TH], jend_th[DH_MAX_TH];
};

__device__ __forceinline__ double kpow(double x, double beta) { return (beta == 1.0) ? x : pow(x, beta); }

// Physical product p (0..2) of component c at plane J for the two samples
// stored in one complex LDS slot (x = 2m, 2m+1).  channel.f:410-571 (c<3)
// and 669-720 (scalars).
__device__ __forceinline__ double2 product_pair(const StepArgs &a, int c, int p, int J, long long off) {
  const long long pp = a.g.pp;
  const double *U1 = a.uphys, *U2 = a.uphys + a.fs, *U3 = a.uphys + 2 * a.fs;
  auto ld = [&](const double *F, int j) -> double2 {
    return *reinterpret_cast<const double2 *>(F + (long long)j * pp + off);
  };
  const double *dy = a.g.dy, *dyf = a.g.dyf;
  if (c == 0 || c == 2) {
    const double *Uc = (c == 0) ? U1 : U3;
    if (p == 0) {  // U3*U1
      double2 u1 = ld(U1, J), u3 = ld(U3, J);
      return make_double2(u3.x * u1.x, u3.y * u1.y);
    } else if (p == 1) {  // U1*U1 or U3*U3
      double2 u = ld(Uc, J);
      return make_double2(u.x * u.x, u.y * u.y);
    } else {  // d/dy (Uc U2), staggered
      double2 up = ld(Uc, J + 1), u0 = ld(Uc, J), um = ld(Uc, J - 1);
      double2 vp = ld(U2, J + 1), v0 = ld(U2, J);
      const double den = 2.0 * dyf[J];
      return make_double2((up.x * vp.x + u0.x * vp.x - u0.x * v0.x - um.x * v0.x) / den,
                          (up.y * vp.y + u0.y * vp.y - u0.y * v0.y - um.y * v0.y) / den);
    }
  } else if (c == 1) {
    if (p <= 1) {  // interp_y(U1 or U3) * U2
      const double *Uc = (p == 0) ? U1 : U3;
      double2 u0 = ld(Uc, J), um = ld(Uc, J - 1), v = ld(U2, J);
      const double den = 2. * dy[J];
      return make_double2(((dyf[J] * u0.x + dyf[J - 1] * um.x) / den) * v.x,
                          ((dyf[J] * u0.y + dyf[J - 1] * um.y) / den) * v.y);
    } else {  // d/dy (U2 U2)
      double2 vp = ld(U2, J + 1), v0 = ld(U2, J), vm = ld(U2, J - 1);
      double ax = v0.x + vp.x, bx = v0.x + vm.x, ay = v0.y + vp.y, by = v0.y + vm.y;
      return make_double2((0.25 * (ax * ax) - 0.25 * (bx * bx)) / dy[J], (0.25 * (ay * ay) - 0.25 * (by * by)) / dy[J]);
    }
  } else {
    const double *TH = a.uphys + (long long)c * a.fs;
    if (p == 0) {
      double2 t = ld(TH, J), u = ld(U1, J);
      return make_double2(t.x * u.x, t.y * u.y);
    } else if (p == 1) {
      double2 t = ld(TH, J), u = ld(U3, J);
      return make_double2(t.x * u.x, t.y * u.y);
    } else {
      double2 tp = ld(TH, J + 1), t0 = ld(TH, J), tm = ld(TH, J - 1);
      double2 vp = ld(U2, J + 1), v0 = ld(U2, J);
      const double den = 2.0 * dyf[J];
      return make_double2((tp.x * vp.x + t0.x * vp.x - t0.x * v0.x - tm.x * v0.x) / den,
                          (tp.y * vp.y + t0.y * vp.y - t0.y * v0.y - tm.y * v0.y) / den);
    }
  }
}

// Explicit Fourier terms of CFc before the nonlinear products
// (channel.f:175-243 for velocity, 274-290 for scalars).
__device__ __forceinline__ double2 cf_init(const StepArgs &a, int c, int J, int kx, int kzp, int m) {
  const Geo &g = a.g;
  const long long o = (long long)J * g.pf + m;
  const double k2 = g.kx2[kx] + g.kz2[kzp];
  if (c < 3) {
    const double f = a.nu * kpow(k2, a.beta);
    double2 u = a.cu[c][o];
    double2 cf = make_double2(-(f * u.x), -(f * u.y));
    if (c == 0) {
      double2 w = a.cu[2][o];
      cf = make_double2(cf.x + a.i_ro * w.x, cf.y + a.i_ro * w.y);
    } else if (c == 2) {
      double2 w = a.cu[0][o];
      cf = make_double2(cf.x - a.i_ro * w.x, cf.y - a.i_ro * w.y);
    }
    for (int n = 0; n < a.n_th; ++n) {
      const double2 *CTH = a.cth + n * a.cfs;
      if (c == 1) {
        if (kx == 0 && kzp == 0) continue;  // plane mean skipped (channel.f:217-226)
        double2 t0 = CTH[o], tm = CTH[o - g.pf];
        const double s = a.ri[n] * a.grav_y;
        const double den = 2.0 * g.dy[J];
        cf = make_double2(cf.x + (s * (t0.x * g.dyf[J - 1] + tm.x * g.dyf[J])) / den,
                          cf.y + (s * (t0.y * g.dyf[J - 1] + tm.y * g.dyf[J])) / den);
      } else {
        const double s = a.ri[n] * ((c == 0) ? a.grav_x : a.grav_z);
        double2 t = CTH[o];
        cf = make_double2(cf.x + s * t.x, cf.y + s * t.y);
      }
    }
    return cf;
  } else {
    const int n = c - 3;
    const double f = (a.nu / a.pr[n]) * kpow(k2, a.beta);
    double2 t = a.cth[n * a.cfs + o];
    double2 cf = make_double2(-(f * t.x), -(f * t.y));
    double2 u1 = a.cu[0][o], u3 = a.cu[2][o];
    cf = make_double2(cf.x - u1.x * a.drhodx[n] - u3.x * a.drhodz[n], cf.y - u1.y * a.drhodx[n] - u3.y * a.drhodz[n]);
    return cf;
  }
}

// Fused explicit pass for plane J (1..NY) and component c (U1,U2,U3,TH...).
__global__ void __launch_bounds__(512) k_nonlinear(StepArgs a) {
  extern __shared__ double2 A[];
  const Geo &g = a.g;
  const int J = blockIdx.x + 1, c = blockIdx.y;
  int jlo = 2, jhi = g.ny;
  if (c >= 3) {
    jlo = a.jstart_th[c - 3];
    jhi = a.jend_th[c - 3];
  }
  if (J < jlo || J > jhi) return;
  const int rk = a.rk;
  const double hb = a.clk->h_bar[rk];
  const double T3 = a.clk->zeta_bar[rk] * hb;
  const double T4 = hb;
  const double T5 = a.clk->beta_bar[rk] * hb;
  const double time = a.clk->time;
  const int nthr = blockDim.x;

  double2 acc[kMaxQ];
#pragma unroll
  for (int q = 0; q < kMaxQ; ++q) {
    const int m = threadIdx.x + q * nthr;
    if (m < g.pf) {
      const int kzp = m / g.nkx1, kx = m - kzp * g.nkx1;
      acc[q] = cf_init(a, c, J, kx, kzp, m);
    }
  }
  // three products, forward transformed and accumulated in statement order
  const int npair = g.nz * g.lw;
  for (int p = 0; p < 3; ++p) {
    for (int e = threadIdx.x; e < npair; e += nthr) {
      const int z = e / g.lw, xp = e - z * g.lw;
      A[e] = product_pair(a, c, p, J, (long long)z * g.nx + 2 * xp);
    }
    __syncthreads();
    plane_rows_fwd(A, g);
    plane_cols_fwd(A, g);
    // derivative factor: c==0 takes (ikz, ikx, 1); all others (ikx, ikz, 1)
#pragma unroll
    for (int q = 0; q < kMaxQ; ++q) {
      const int m = threadIdx.x + q * nthr;
      if (m < g.pf) {
        const int kzp = m / g.nkx1, kx = m - kzp * g.nkx1;
        double2 s = A[fwd_slot(g, kx, kzp)];
        s = make_double2(s.x / g.nz, s.y / g.nz);
        if (p == 2) {
          acc[q] = csub(acc[q], s);
        } else {
          const bool use_kz = (c == 0) ? (p == 0) : (p == 1);
          const double k = use_kz ? g.kz[kzp] : g.kx[kx];
          acc[q] = csub(acc[q], cmul_ik(k, s));
        }
      }
    }
    __syncthreads();
  }
  // RK combination (channel.f:88-167, 581-595; scalars 254-270, 723-729),
  // new CF stored, CR scattered into LDS for the inverse transform.
  for (int e = threadIdx.x; e < g.nz * g.nkx1; e += nthr) {
    const int zr = e / g.nkx1, kx = e - zr * g.nkx1;
    if (zr > g.nkz && zr < g.nz - g.nkz) A[zr * g.lw + kx] = make_double2(0.0, 0.0);
  }
  double2 *CF = (c < 3) ? a.cf[c] : a.cfth + (c - 3) * a.cfs;
  const long long base = (long long)J * g.pf;
#pragma unroll
  for (int q = 0; q < kMaxQ; ++q) {
    const int m = threadIdx.x + q * nthr;
    if (m < g.pf) {
      const int kzp = m / g.nkx1, kx = m - kzp * g.nkx1;
      const long long o = base + m;
      double2 cr = (c < 3) ? a.cu[c][o] : a.cth[(c - 3) * a.cfs + o];
      if (rk > 0) {
        double2 old = CF[o];
        cr = make_double2(cr.x + T3 * old.x, cr.y + T3 * old.y);
      }
      if (c == 0 || c == 2) {
        const double k = (c == 0) ? g.kx[kx] : g.kz[kzp];
        double2 ikp = cmul_ik(k, a.cp[o]);
        cr = make_double2(cr.x - T4 * ikp.x, cr.y - T4 * ikp.y);
        if (c == 0 && kx == 0 && kzp == 0) {
          if (a.f_type == 1)
            cr.x = cr.x - T4 * a.px0;
          else if (a.f_type == 2)
            cr.x = cr.x - T4 * (a.px0 + a.amp_omega0 * cos(a.omega0 * time));
        }
      } else if (c == 1) {
        double2 p0 = a.cp[o], pm = a.cp[o - g.pf];
        double2 cs = make_double2((p0.x - pm.x) / g.dy[J], (p0.y - pm.y) / g.dy[J]);
        cr = make_double2(cr.x - T4 * cs.x, cr.y - T4 * cs.y);
      }
      const double2 cfn = acc[q];
      CF[o] = cfn;
      cr = make_double2(cr.x + T5 * cfn.x, cr.y + T5 * cfn.y);
      A[kz_row(g, kzp) * g.lw + kx] = cr;
    }
  }
  __syncthreads();
  plane_cols_inv(A, g);
  plane_rows_inv(A, g);
  double2 *dst = reinterpret_cast<double2 *>(a.rphys + (long long)c * a.fs + (long long)J * g.pp);
  for (int e = threadIdx.x; e < npair; e += nthr) {
    const int z = e / g.lw, m = e - z * g.lw;
    dst[e] = A[inv_slot(g, z, m)];
  }
}

// ---------------------------------------------------------------------------
// Real Crank-Nicolson solves (channel.f:606-626, 739-747 explicit part;
// 771-1029 implicit part with APPLY_BC_* rows and THOMAS_*_REAL_MPI).
// ---------------------------------------------------------------------------
struct SolveArgs {
  Geo g;
  Clock *clk;
  int rk;
  double *uphys;
  const double *rphys;
  long long fs;
  RealFactors fac;
  int nfields;
  int wlo[kMaxFields], whi[kMaxFields];
  double nu, nu_v_scale, pr[DH_MAX_TH];
  int courant;  // 1: accumulate the CFL minimum of U1..U3
  double cfl, dx, dz;
};

__global__ void __launch_bounds__(256) k_solve(SolveArgs a) {
  const Geo &g = a.g;
  const int f = blockIdx.y;
  const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long pp = g.pp;
  const int NY = g.ny;
  const int rk = a.rk;
  const double hb = a.clk->h_bar[rk];
  const double T1 = a.nu * hb / 2.0;
  const double coef = (f < 3) ? T1 * a.nu_v_scale : (T1 / a.pr[f - 3]) * a.nu_v_scale;
  const int nrow = NY + 2;
  const double *Ap = a.fac.ap + ((long long)rk * a.nfields + f) * nrow;
  const double *Bp = a.fac.bp + ((long long)rk * a.nfields + f) * nrow;
  const double *Cc = a.fac.cc + ((long long)rk * a.nfields + f) * nrow;
  const int *isbc = a.fac.isbc + (long long)f * nrow;
  const double *vc = a.fac.vconst + (long long)f * nrow;
  const double *dy = g.dy, *dyf = g.dyf;
  const int wlo = a.wlo[f];
  double dtmin = __longlong_as_double(0x7FF0000000000000ll);
  if (col < pp) {
    double *U = a.uphys + (long long)f * a.fs + col;
    const double *R = a.rphys + (long long)f * a.fs + col;
    double um = 0.0, u0 = U[0], up;
    double gprev = 0.0;
    for (int J = 0; J <= NY + 1; ++J) {
      up = (J <= NY) ? U[(long long)(J + 1) * pp] : 0.0;
      double rhs;
      if (isbc[J]) {
        rhs = vc[J];
      } else {
        double ex;
        if (f == 1)
          ex = ((up - u0) / dyf[J] - (u0 - um) / dyf[J - 1]) / dy[J];
        else
          ex = ((up - u0) / dy[J + 1] - (u0 - um) / dy[J]) / dyf[J];
        rhs = R[(long long)J * pp] + coef * ex;
      }
      const double gj = (J == 0) ? rhs : rhs + Ap[J] * gprev;
      if (J >= wlo) U[(long long)J * pp] = gj;
      gprev = gj;
      um = u0;
      u0 = up;
    }
    double x = gprev / Bp[NY + 1];
    U[(long long)(NY + 1) * pp] = x;
    const bool cour = a.courant && f < 3;
    const double dxf = (f == 0) ? a.dx : a.dz;
    for (int J = NY; J >= wlo; --J) {
      const double gj = U[(long long)J * pp];
      const double b = Bp[J];
      x = (b != 0.0) ? (gj - Cc[J] * x) / b : 0.0;
      U[(long long)J * pp] = x;
      if (cour && J >= 1) {
        const double d = (f == 1) ? dy[J] : dxf;
        const double t = a.cfl * d / fabs(x);
        dtmin = fmin(dtmin, t);
      }
    }
  }
  if (a.courant && f < 3) {
    __shared__ double red[256];
    red[threadIdx.x] = dtmin;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      const double v = red[0];
      if (v >= 0.0) atomicMin(&a.clk->cfl_min_bits, (unsigned long long)__double_as_longlong(v));
    }
  }
}

// ---------------------------------------------------------------------------
// Factor tables of the real CN matrices (matrix build of channel.f:771-1005,
// BC rows of channel.f:1660-2282, forward elimination of mpi.f:630-703).
// ---------------------------------------------------------------------------
struct FactorIn {
  int ny, nfields;
  double nu, nu_v_scale, pr[DH_MAX_TH];
  int jstart, jend, jstart_th[DH_MAX_TH], jend_th[DH_MAX_TH];
  int bcmin[kMaxFields], bcmax[kMaxFields];
  double c1min[kMaxFields], c1max[kMaxFields];
  const double *dy, *dyf;
};

// One field, one RK substep.  L,D,Uu,V are caller scratch of NY+2 rows.
__host__ __device__ inline void build_real_system(const FactorIn &in, int f, double temp1, double *L, double *D,
                                                  double *Uu, double *V, int *isbc) {
  const int NY = in.ny;
  const double *DY = in.dy, *DYF = in.dyf;
  for (int j = 0; j <= NY + 1; ++j) {
    L[j] = 0.;
    D[j] = 1.;
    Uu[j] = 0.;
    V[j] = 0.;
    isbc[j] = 1;
  }
  if (f == 1) {  // U2, channel.f:856-862
    for (int J = 2; J <= NY; ++J) {
      L[J] = -temp1 * in.nu_v_scale / (DYF[J - 1] * DY[J]);
      D[J] = 1. + temp1 * in.nu_v_scale / (DYF[J] * DY[J]) + temp1 * in.nu_v_scale / (DYF[J - 1] * DY[J]);
      Uu[J] = -temp1 * in.nu_v_scale / (DYF[J] * DY[J]);
      isbc[J] = 0;
    }
    // APPLY_BC_2_LOWER / _UPPER
    if (in.bcmin[f] == 0) {
      for (int J = 1; J <= 2; ++J) {
        L[J] = 0.; D[J] = 1.; Uu[J] = 0.; V[J] = in.c1min[f]; isbc[J] = 1;
      }
    } else if (in.bcmin[f] == 1) {
      D[1] = -1.; Uu[1] = 1.; L[1] = 0.; V[1] = DYF[1] * in.c1min[f]; isbc[1] = 1;
    }
    L[0] = 0.; D[0] = 1.; Uu[0] = 0.; V[0] = 0.; isbc[0] = 1;
    if (in.bcmax[f] == 0) {
      L[NY + 1] = 0.; D[NY + 1] = 1.; Uu[NY + 1] = 0.; V[NY + 1] = in.c1max[f]; isbc[NY + 1] = 1;
      L[NY] = 0.; D[NY] = 1.; Uu[NY] = 0.; V[NY] = in.c1max[f]; isbc[NY] = 1;
    } else if (in.bcmax[f] == 1) {
      L[NY + 1] = -1.; D[NY + 1] = 1.; Uu[NY + 1] = 0.; V[NY + 1] = DYF[NY] * in.c1max[f]; isbc[NY + 1] = 1;
    }
    return;
  }
  int js, je;
  double c;
  if (f < 3) {  // U1, U3: channel.f:914-920, 978-984
    js = in.jstart;
    je = in.jend;
    for (int J = js; J <= je; ++J) {
      L[J] = -temp1 * in.nu_v_scale / (DY[J] * DYF[J]);
      D[J] = 1. - temp1 * in.nu_v_scale * (-1. / (DY[J + 1] * DYF[J]) - 1. / (DY[J] * DYF[J]));
      Uu[J] = -temp1 * in.nu_v_scale / (DY[J + 1] * DYF[J]);
      isbc[J] = 0;
    }
  } else {  // TH: channel.f:783-791
    const int n = f - 3;
    js = in.jstart_th[n];
    je = in.jend_th[n];
    c = temp1 / in.pr[n] * in.nu_v_scale;
    for (int J = js; J <= je; ++J) {
      L[J] = -c / (DY[J] * DYF[J]);
      D[J] = 1. + c / (DY[J + 1] * DYF[J]) + c / (DY[J] * DYF[J]);
      Uu[J] = -c / (DY[J + 1] * DYF[J]);
      isbc[J] = 0;
    }
  }
  // APPLY_BC_{1,3,TH}_LOWER / _UPPER
  if (in.bcmin[f] == 0) {
    L[0] = 0.; D[0] = 1.; Uu[0] = 0.; V[0] = 0.; isbc[0] = 1;
    L[1] = 0.; D[1] = 1.; Uu[1] = 0.; V[1] = in.c1min[f]; isbc[1] = 1;
  } else {
    L[0] = 0.; D[0] = 1.; Uu[0] = 0.; V[0] = 0.; isbc[0] = 1;
    L[1] = 0.; D[1] = -1.; Uu[1] = 1.; V[1] = DY[2] * in.c1min[f]; isbc[1] = 1;
  }
  if (in.bcmax[f] == 0) {
    L[NY + 1] = 0.; D[NY + 1] = 1.; Uu[NY + 1] = 0.; V[NY + 1] = 0.; isbc[NY + 1] = 1;
    L[NY] = 0.; D[NY] = 1.; Uu[NY] = 0.; V[NY] = in.c1max[f]; isbc[NY] = 1;
  } else {
    L[NY] = -1.; D[NY] = 1.; Uu[NY] = 0.; V[NY] = DY[NY] * in.c1max[f]; isbc[NY] = 1;
    L[NY + 1] = 0.; D[NY + 1] = 1.; Uu[NY + 1] = 0.; V[NY + 1] = 0.; isbc[NY + 1] = 1;
  }
}

// Forward elimination of the matrix only (mpi.f:668-702 on one rank):
// A(J) <- -A(J)/B(J-1), B(J) <- B(J) + A(J) C(J-1), J = 1..NY+1.
__host__ __device__ inline void factor_real(int NY, double *L, double *D, const double *Uu) {
  for (int J = 1; J <= NY + 1; ++J) {
    L[J] = -L[J] / D[J - 1];
    D[J] = D[J] + L[J] * Uu[J - 1];
  }
}

// COURANT (courant.f:45-62) after the CFL minimum of the solve kernel, then
// the factor tables of all three substeps for the new DELTA_T: one thread
// per (rk, field) chain.  V/isbc rows of the builder go to LDS scratch.
__global__ void k_courant_fin(Clock *clk, double dt_cap, int variable_dt, int update_dt, FactorIn in,
                              RealFactors fac) {
  extern __shared__ double fsm[];
  __shared__ int changed;
  if (threadIdx.x == 0) {
    changed = 0;
    Clock &c = *clk;
    // TEMP4 of this substep was fixed before COURANT (channel.f:80); the
    // pressure update after REM_DIV (channel.f:1059-1065) still divides by it
    c.h_cp = c.h_bar[2];
    if (variable_dt && (c.time_step % update_dt) == 0) {
      double dt = dt_cap;
      const double m = __longlong_as_double((long long)c.cfl_min_bits);
      if (m < dt) dt = m;
      double DT;
      if (dt <= 0)
        DT = 0.0001;
      else if (dt >= 999.)
        DT = 999.0;
      else
        DT = dt;
      c.dt = DT;
      c.h_bar[0] = DT * (8.0 / 15.0);
      c.h_bar[1] = DT * (2.0 / 15.0);
      c.h_bar[2] = DT * (5.0 / 15.0);
      changed = 1;
    }
    c.cfl_min_bits = 0x7FF0000000000000ull;
  }
  __syncthreads();
  if (!changed) return;
  const int nrow = in.ny + 2;
  const int t = threadIdx.x;
  if (t < 3 * in.nfields) {
    const int rk = t / in.nfields, f = t - rk * in.nfields;
    const double temp1 = in.nu * clk->h_bar[rk] / 2.0;
    double *L = fac.ap + ((long long)rk * in.nfields + f) * nrow;
    double *D = fac.bp + ((long long)rk * in.nfields + f) * nrow;
    double *Uu = fac.cc + ((long long)rk * in.nfields + f) * nrow;
    double *Vs = fsm + (long long)t * nrow;
    int *Is = reinterpret_cast<int *>(fsm + (long long)3 * in.nfields * nrow) + (long long)t * nrow;
    build_real_system(in, f, temp1, L, D, Uu, Vs, Is);
    factor_real(in.ny, L, D, Uu);
  }
}

void build_real_factors_host(const Ctx &c, double dt, std::vector<double> &ap, std::vector<double> &bp,
                             std::vector<double> &cc, std::vector<int> &isbc, std::vector<double> &vconst);

// ---------------------------------------------------------------------------
// REM_DIV_CHAN (channel.f:1103-1214) + CP update (channel.f:1059-1065).
// One thread per (kx, kz) mode; the pivots and modified right-hand sides of
// the forward sweep are parked in [J][mode] scratch for the backward sweep.
// ---------------------------------------------------------------------------
struct RemDivArgs {
  Geo g;
  Clock *clk;
  int rk;       // <0: no CP update
  int cp_old;   // 1: COURANT ran in this substep, divide by the saved H_BAR
  double2 *cu[3];
  double2 *cp;
  double2 *phi;
  double *sbp;
  double2 *sg;
  const double *rl, *ru;  // 1/(DY(J)DYF(J)), 1/(DY(J+1)DYF(J))
};

__global__ void __launch_bounds__(64) k_rem_div(RemDivArgs a) {
  const Geo &g = a.g;
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= g.pf) return;
  const int kzp = m / g.nkx1, kx = m - kzp * g.nkx1;
  const int NY = g.ny;
  const long long pf = g.pf;
  const double kxv = g.kx[kx], kzv = g.kz[kzp], kx2 = g.kx2[kx], kz2 = g.kz2[kzp];
  const bool mean = (kx == 0 && kzp == 0);
  double2 *CU1 = a.cu[0] + m, *CU2 = a.cu[1] + m, *CU3 = a.cu[2] + m;
  // row 1 (APPLY_BC_REM_DIV_MPI, mpi.f:1244-1262)
  double bprev = 1.0;
  double uprev = mean ? 0.0 : -1.0;
  double2 gprev = make_double2(0.0, 0.0);
  double2 cu2_j = CU2[pf * 2];
  for (int J = 2; J <= NY; ++J) {
    double L, D, Uu;
    double2 V;
    if (J < NY) {
      L = a.rl[J];
      Uu = a.ru[J];
      D = -kx2 - kz2 - Uu - L;
      const double2 c1 = CU1[pf * J], c3 = CU3[pf * J];
      const double2 c2p = CU2[pf * (J + 1)];
      const double2 t1 = cmul_ik(kxv, c1), t3 = cmul_ik(kzv, c3);
      const double d = g.dyf[J];
      V = make_double2(t1.x + (c2p.x - cu2_j.x) / d + t3.x, t1.y + (c2p.y - cu2_j.y) / d + t3.y);
      cu2_j = c2p;
    } else {  // row NY (mpi.f:1265-1272)
      L = 1.;
      D = -1.;
      Uu = 0.;
      V = make_double2(0.0, 0.0);
    }
    const double A = (bprev != 0.0) ? -L / bprev : 0.0;
    const double B = D + A * uprev;
    const double2 G = make_double2(V.x + A * gprev.x, V.y + A * gprev.y);
    a.sbp[pf * J + m] = B;
    a.sg[pf * J + m] = G;
    bprev = B;
    uprev = Uu;
    gprev = G;
  }
  // backward sweep with projection (channel.f:1196-1211) and CP update
  const double hb = (a.rk < 0) ? 1.0 : (a.cp_old ? a.clk->h_cp : a.clk->h_bar[a.rk]);
  double2 phi_next = make_double2(0.0, 0.0);
  for (int J = NY; J >= 1; --J) {
    double2 phi;
    if (J == NY) {
      phi = (bprev != 0.0) ? make_double2(gprev.x / bprev, gprev.y / bprev) : make_double2(0.0, 0.0);
    } else {
      double B, Uu;
      double2 G;
      if (J >= 2) {
        B = a.sbp[pf * J + m];
        G = a.sg[pf * J + m];
        Uu = a.ru[J];
      } else {
        B = 1.0;
        G = make_double2(0.0, 0.0);
        Uu = mean ? 0.0 : -1.0;
      }
      phi = (B != 0.0) ? make_double2((G.x - Uu * phi_next.x) / B, (G.y - Uu * phi_next.y) / B)
                       : make_double2(0.0, 0.0);
    }
    a.phi[pf * J + m] = phi;
    // CU1, CU3 at J
    {
      double2 t = cmul_ik(kxv, phi);
      double2 u = CU1[pf * J];
      CU1[pf * J] = make_double2(u.x - t.x, u.y - t.y);
      t = cmul_ik(kzv, phi);
      u = CU3[pf * J];
      CU3[pf * J] = make_double2(u.x - t.x, u.y - t.y);
    }
    // CU2 at J+1 (needs phi(J+1), phi(J)), J+1 in 2..NY
    if (J + 1 <= NY) {
      const double d = g.dy[J + 1];
      double2 u = CU2[pf * (J + 1)];
      CU2[pf * (J + 1)] = make_double2(u.x - (phi_next.x - phi.x) / d, u.y - (phi_next.y - phi.y) / d);
    }
    if (a.rk >= 0 && J >= 2) {
      double2 p = a.cp[pf * J + m];
      a.cp[pf * J + m] = make_double2(p.x + phi.x / hb, p.y + phi.y / hb);
    }
    phi_next = phi;
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static StepArgs make_step_args(Ctx &c, int rk) {
  StepArgs a{};
  a.g = c.g;
  a.clk = c.clock;
  a.rk = rk;
  for (int i = 0; i < 3; ++i) {
    a.cu[i] = c.cu[i];
    a.cf[i] = c.cf[i];
  }
  a.cp = c.cp;
  a.cth = c.cth;
  a.cfth = c.cfth;
  a.uphys = c.uphys;
  a.rphys = c.rphys;
  a.fs = (long long)(c.g.ny + 2) * c.g.pp;
  a.cfs = (long long)(c.g.ny + 2) * c.g.pf;
  const dh_params &p = c.p;
  a.nu = p.nu;
  a.beta = p.beta;
  a.i_ro = p.i_ro;
  a.grav_x = p.grav_x;
  a.grav_y = p.grav_y;
  a.grav_z = p.grav_z;
  a.px0 = p.px0;
  a.omega0 = p.omega0;
  a.amp_omega0 = p.amp_omega0;
  a.f_type = p.f_type;
  a.n_th = p.n_th;
  for (int n = 0; n < DH_MAX_TH; ++n) {
    a.ri[n] = p.ri[n];
    a.pr[n] = p.pr[n];
    a.drhodx[n] = p.drhodx[n];
    a.drhodz[n] = p.drhodz[n];
    a.jstart_th[n] = c.jstart_th[n];
    a.jend_th[n] = c.jend_th[n];
  }
  return a;
}

static FactorIn make_factor_in(const Ctx &c, const double *dy, const double *dyf) {
  FactorIn in{};
  in.ny = c.g.ny;
  in.nfields = c.nfields;
  in.nu = c.p.nu;
  in.nu_v_scale = c.p.nu_v_scale;
  in.jstart = c.jstart;
  in.jend = c.jend;
  for (int n = 0; n < DH_MAX_TH; ++n) {
    in.pr[n] = c.p.pr[n];
    in.jstart_th[n] = c.jstart_th[n];
    in.jend_th[n] = c.jend_th[n];
  }
  const dh_params &p = c.p;
  in.bcmin[0] = p.u_bc_ymin; in.bcmax[0] = p.u_bc_ymax; in.c1min[0] = p.u_bc_ymin_c1; in.c1max[0] = p.u_bc_ymax_c1;
  in.bcmin[1] = p.v_bc_ymin; in.bcmax[1] = p.v_bc_ymax; in.c1min[1] = p.v_bc_ymin_c1; in.c1max[1] = p.v_bc_ymax_c1;
  in.bcmin[2] = p.w_bc_ymin; in.bcmax[2] = p.w_bc_ymax; in.c1min[2] = p.w_bc_ymin_c1; in.c1max[2] = p.w_bc_ymax_c1;
  for (int n = 0; n < p.n_th; ++n) {
    in.bcmin[3 + n] = p.th_bc_ymin[n];
    in.bcmax[3 + n] = p.th_bc_ymax[n];
    in.c1min[3 + n] = p.th_bc_ymin_c1[n];
    in.c1max[3 + n] = p.th_bc_ymax_c1[n];
  }
  in.dy = dy;
  in.dyf = dyf;
  return in;
}

void build_real_factors_host(const Ctx &c, double dt, std::vector<double> &ap, std::vector<double> &bp,
                             std::vector<double> &cc, std::vector<int> &isbc, std::vector<double> &vconst) {
  const int nrow = c.g.ny + 2, nf = c.nfields;
  FactorIn in = make_factor_in(c, c.dy.data(), c.dyf.data());
  ap.assign((size_t)3 * nf * nrow, 0.0);
  bp.assign((size_t)3 * nf * nrow, 0.0);
  cc.assign((size_t)3 * nf * nrow, 0.0);
  isbc.assign((size_t)nf * nrow, 1);
  vconst.assign((size_t)nf * nrow, 0.0);
  const double hbar[3] = {dt * (8.0 / 15.0), dt * (2.0 / 15.0), dt * (5.0 / 15.0)};
  std::vector<double> V(nrow);
  std::vector<int> I(nrow);
  for (int rk = 0; rk < 3; ++rk)
    for (int f = 0; f < nf; ++f) {
      const double temp1 = c.p.nu * hbar[rk] / 2.0;
      double *L = &ap[((size_t)rk * nf + f) * nrow];
      double *D = &bp[((size_t)rk * nf + f) * nrow];
      double *U = &cc[((size_t)rk * nf + f) * nrow];
      build_real_system(in, f, temp1, L, D, U, V.data(), I.data());
      factor_real(c.g.ny, L, D, U);
      if (rk == 0)
        for (int j = 0; j < nrow; ++j) {
          isbc[(size_t)f * nrow + j] = I[j];
          vconst[(size_t)f * nrow + j] = V[j];
        }
    }
}

hipError_t launch_state_to_phys(Ctx &c);
hipError_t launch_state_to_four(Ctx &c);

static hipError_t launch_nonlinear(Ctx &c, int rk) {
  StepArgs a = make_step_args(c, rk);
  hipLaunchKernelGGL(k_nonlinear, dim3(c.g.ny, c.nfields), dim3(c.plane_threads), c.lds_bytes, c.stream, a);
  return hipGetLastError();
}

static hipError_t launch_solve(Ctx &c, int rk, int courant) {
  SolveArgs a{};
  a.g = c.g;
  a.clk = c.clock;
  a.rk = rk;
  a.uphys = c.uphys;
  a.rphys = c.rphys;
  a.fs = (long long)(c.g.ny + 2) * c.g.pp;
  a.fac = c.fac;
  a.nfields = c.nfields;
  for (int f = 0; f < c.nfields; ++f) {
    a.wlo[f] = c.wlo[f];
    a.whi[f] = c.whi[f];
  }
  a.nu = c.p.nu;
  a.nu_v_scale = c.p.nu_v_scale;
  for (int n = 0; n < DH_MAX_TH; ++n) a.pr[n] = c.p.pr[n];
  a.courant = courant;
  a.cfl = c.p.cfl;
  a.dx = c.p.lx / c.g.nx;
  a.dz = c.p.lz / c.g.nz;
  const int nb = (int)((c.g.pp + 255) / 256);
  hipLaunchKernelGGL(k_solve, dim3(nb, c.nfields), dim3(256), 0, c.stream, a);
  return hipGetLastError();
}

hipError_t launch_rem_div(Ctx &c, int rk) {
  RemDivArgs a{};
  a.g = c.g;
  a.clk = c.clock;
  a.rk = rk;
  a.cp_old = (c.p.variable_dt && rk == 2) ? 1 : 0;
  for (int i = 0; i < 3; ++i) a.cu[i] = c.cu[i];
  a.cp = c.cp;
  a.phi = c.phi;
  a.sbp = c.rd_bp;
  a.sg = c.rd_g;
  a.rl = c.rd_l;
  a.ru = c.rd_u;
  const int nb = (c.g.pf + 63) / 64;
  hipLaunchKernelGGL(k_rem_div, dim3(nb), dim3(64), 0, c.stream, a);
  return hipGetLastError();
}

static hipError_t launch_courant_fin(Ctx &c) {
  FactorIn in = make_factor_in(c, c.d_dy, c.d_dyf);
  const int nrow = c.g.ny + 2;
  const size_t shm = (size_t)3 * c.nfields * nrow * (sizeof(double) + sizeof(int));
  hipLaunchKernelGGL(k_courant_fin, dim3(1), dim3(64), shm, c.stream, c.clock, c.dt_cap, c.p.variable_dt,
                     c.p.update_dt, in, c.fac);
  return hipGetLastError();
}

// Kernels of one substep in launch order (channel.f:386-1065).
const char *const kPhaseNames[kNumPhases] = {"to_physical", "nonlinear", "cn_solve", "courant", "to_fourier",
                                             "rem_div"};

hipError_t launch_phase(Ctx &c, int rk, int phase) {
  const int courant = (c.p.variable_dt && rk == 2) ? 1 : 0;
  switch (phase) {
    case 0: return launch_state_to_phys(c);
    case 1: return launch_nonlinear(c, rk);
    case 2: return launch_solve(c, rk, courant);
    case 3: return courant ? launch_courant_fin(c) : hipSuccess;
    case 4: return launch_state_to_four(c);
    case 5: return launch_rem_div(c, rk);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_substep(Ctx &c, int rk) {
  for (int ph = 0; ph < kNumPhases; ++ph) {
    hipError_t e = launch_phase(c, rk, ph);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

__global__ void k_step_begin(Clock *clk) {
  if (threadIdx.x == 0) clk->time_step += 1;
}
__global__ void k_step_end(Clock *clk) {
  if (threadIdx.x == 0) clk->time = clk->time + clk->dt;
}

hipError_t launch_step_begin(Ctx &c) {
  hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(64), 0, c.stream, c.clock);
  return hipGetLastError();
}
hipError_t launch_end_step(Ctx &c) {
  hipLaunchKernelGGL(k_step_end, dim3(1), dim3(64), 0, c.stream, c.clock);
  return hipGetLastError();
}

hipError_t prepare_step_kernels(Ctx &c) {
  return hipFuncSetAttribute((const void *)k_nonlinear, hipFuncAttributeMaxDynamicSharedMemorySize, c.lds_bytes);
}

}  // namespace dh
