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

constexpr int kMaxQ = 4;  // Fourier modes owned per thread in k_nonlinear (pf <= kMaxQ*threads)

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
__global__ void __launch_bounds__(1024) k_nonlinear(StepArgs a) {
  extern __shared__ double2 A[];
  const Geo &g = a.g;
  const Tables T = load_tables(A, g);
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
      const int kzp = (int)fdiv((unsigned)m, g.f_nkx1), kx = m - kzp * g.nkx1;
      acc[q] = cf_init(a, c, J, kx, kzp, m);
    }
  }
  // three products, forward transformed and accumulated in statement order
  const int npair = g.nz * g.lw;
  for (int p = 0; p < 3; ++p) {
    constexpr int PB = 2;  // products per thread per batch: loads issued before LDS stores
    for (int base = threadIdx.x; base < npair; base += PB * nthr) {
      double2 v[PB];
      int sl[PB];
#pragma unroll
      for (int b = 0; b < PB; ++b) {
        const int e = base + b * nthr;
        sl[b] = -1;
        if (e < npair) {
          const int z = (int)fdiv((unsigned)e, g.f_lw), xp = e - z * g.lw;
          sl[b] = phys_slot(T, g, z, xp);
          v[b] = product_pair(a, c, p, J, (long long)z * g.nx + 2 * xp);
        }
      }
#pragma unroll
      for (int b = 0; b < PB; ++b)
        if (sl[b] >= 0) A[sl[b]] = v[b];
    }
    __syncthreads();
    plane_rows_fwd(A, g, T);
    plane_cols_fwd(A, g, T);
    // derivative factor: c==0 takes (ikz, ikx, 1); all others (ikx, ikz, 1)
#pragma unroll
    for (int q = 0; q < kMaxQ; ++q) {
      const int m = threadIdx.x + q * nthr;
      if (m < g.pf) {
        const int kzp = (int)fdiv((unsigned)m, g.f_nkx1), kx = m - kzp * g.nkx1;
        double2 s = A[fwd_slot(g, T, kx, kzp)];
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
    const int zr = (int)fdiv((unsigned)e, g.f_nkx1), kx = e - zr * g.nkx1;
    if (zr > g.nkz && zr < g.nz - g.nkz) A[zr * g.lw + kx] = make_double2(0.0, 0.0);
  }
  double2 *CF = (c < 3) ? a.cf[c] : a.cfth + (c - 3) * a.cfs;
  const long long base = (long long)J * g.pf;
#pragma unroll
  for (int q = 0; q < kMaxQ; ++q) {
    const int m = threadIdx.x + q * nthr;
    if (m < g.pf) {
      const int kzp = (int)fdiv((unsigned)m, g.f_nkx1), kx = m - kzp * g.nkx1;
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
  plane_cols_inv(A, g, T);
  plane_rows_inv<8>(A, g, T);
  double2 *dst = reinterpret_cast<double2 *>(a.rphys + (long long)c * a.fs + (long long)J * g.pp);
  for (int e = threadIdx.x; e < npair; e += nthr) {
    const int z = (int)fdiv((unsigned)e, g.f_lw), m = e - z * g.lw;
    dst[e] = A[inv_slot(g, T, z, m)];
  }
}
// ---------------------------------------------------------------------------
// Real Crank-Nicolson systems (matrix build of channel.f:771-1005, BC rows
// of APPLY_BC_{1,2,3,TH}_{LOWER,UPPER} channel.f:1660-2282).  With LES off
// the matrix of each field depends only on j, so all (x,z) columns share it.
// ---------------------------------------------------------------------------
struct FactorIn {
  int ny, nfields;
  double nu, nu_v_scale, pr[DH_MAX_TH];
  int jstart, jend, jstart_th[DH_MAX_TH], jend_th[DH_MAX_TH];
  int bcmin[kMaxFields], bcmax[kMaxFields];
  double c1min[kMaxFields], c1max[kMaxFields];
  const double *dy, *dyf;
};

// Row J (0..NY+1) of the system of field f: lower L, diagonal D, upper U,
// constant right-hand side V of a boundary row (isbc=1), exactly as the
// reference fills MATL/MATD/MATU/VEC: initialise (0,1,0,0), interior rows,
// then the BC rows overwrite.
__host__ __device__ inline void real_row(const FactorIn &in, int f, double temp1, int J, double &L, double &D,
                                         double &U, double &V, int &isbc) {
  const int NY = in.ny;
  const double *DY = in.dy, *DYF = in.dyf;
  L = 0.;
  D = 1.;
  U = 0.;
  V = 0.;
  isbc = 1;
  if (f == 1) {  // U2, channel.f:856-862; APPLY_BC_2_LOWER/UPPER
    if (J >= 2 && J <= NY) {
      L = -temp1 * in.nu_v_scale / (DYF[J - 1] * DY[J]);
      D = 1. + temp1 * in.nu_v_scale / (DYF[J] * DY[J]) + temp1 * in.nu_v_scale / (DYF[J - 1] * DY[J]);
      U = -temp1 * in.nu_v_scale / (DYF[J] * DY[J]);
      isbc = 0;
    }
    if (in.bcmin[f] == 0 && (J == 1 || J == 2)) {
      L = 0.; D = 1.; U = 0.; V = in.c1min[f]; isbc = 1;
    } else if (in.bcmin[f] == 1 && J == 1) {
      D = -1.; U = 1.; L = 0.; V = DYF[1] * in.c1min[f]; isbc = 1;
    }
    if (J == 0) {
      L = 0.; D = 1.; U = 0.; V = 0.; isbc = 1;
    }
    if (in.bcmax[f] == 0 && (J == NY || J == NY + 1)) {
      L = 0.; D = 1.; U = 0.; V = in.c1max[f]; isbc = 1;
    } else if (in.bcmax[f] == 1 && J == NY + 1) {
      L = -1.; D = 1.; U = 0.; V = DYF[NY] * in.c1max[f]; isbc = 1;
    }
    return;
  }
  if (f < 3) {  // U1, U3: channel.f:914-920, 978-984
    if (J >= in.jstart && J <= in.jend) {
      L = -temp1 * in.nu_v_scale / (DY[J] * DYF[J]);
      D = 1. - temp1 * in.nu_v_scale * (-1. / (DY[J + 1] * DYF[J]) - 1. / (DY[J] * DYF[J]));
      U = -temp1 * in.nu_v_scale / (DY[J + 1] * DYF[J]);
      isbc = 0;
    }
  } else {  // TH: channel.f:783-791
    const int n = f - 3;
    if (J >= in.jstart_th[n] && J <= in.jend_th[n]) {
      const double c = temp1 / in.pr[n] * in.nu_v_scale;
      L = -c / (DY[J] * DYF[J]);
      D = 1. + c / (DY[J + 1] * DYF[J]) + c / (DY[J] * DYF[J]);
      U = -c / (DY[J + 1] * DYF[J]);
      isbc = 0;
    }
  }
  // APPLY_BC_{1,3,TH}_LOWER / _UPPER
  if (J == 0) {
    L = 0.; D = 1.; U = 0.; V = 0.; isbc = 1;
  } else if (J == 1) {
    if (in.bcmin[f] == 0) {
      L = 0.; D = 1.; U = 0.; V = in.c1min[f];
    } else {
      L = 0.; D = -1.; U = 1.; V = DY[2] * in.c1min[f];
    }
    isbc = 1;
  }
  if (J == NY + 1) {
    L = 0.; D = 1.; U = 0.; V = 0.; isbc = 1;
  } else if (J == NY) {
    if (in.bcmax[f] == 0) {
      L = 0.; D = 1.; U = 0.; V = in.c1max[f];
    } else {
      L = -1.; D = 1.; U = 0.; V = DY[NY] * in.c1max[f];
    }
    isbc = 1;
  }
}

// Parallel cyclic reduction tables of the shared matrix of (rk, field):
// step s (distance 2^s) eliminates x_{i-2^s}, x_{i+2^s} from row i with the
// multipliers k1 = a_i/b_{i-d}, k2 = c_i/b_{i+d}; after all steps x_i =
// d_i / b_i.  Rows outside 0..NY+1 are the identity.  One block per (rk,
// field); dynamic LDS 3*(NY+2) doubles.  The constant right-hand sides of
// the boundary rows (dt independent) are written by the rk=0 blocks.
__global__ void __launch_bounds__(256) k_pcr_tables(const Clock *clk, FactorIn in, RealFactors fac) {
  extern __shared__ double sm[];
  const int nrow = in.ny + 2;
  double *a = sm, *b = sm + nrow, *c = sm + 2 * nrow;
  const int rk = blockIdx.x / in.nfields, f = blockIdx.x - rk * in.nfields;
  const double temp1 = in.nu * clk->h_bar[rk] / 2.0;
  for (int i = threadIdx.x; i < nrow; i += blockDim.x) {
    double L, D, U, V;
    int bc;
    real_row(in, f, temp1, i, L, D, U, V, bc);
    a[i] = L;
    b[i] = D;
    c[i] = U;
    if (rk == 0) {
      fac.isbc[(long long)f * nrow + i] = bc;
      fac.vconst[(long long)f * nrow + i] = V;
    }
  }
  __syncthreads();
  const long long tb = ((long long)rk * in.nfields + f) * (long long)kMaxPcr * nrow;
  int s = 0;
  for (int d = 1; d < nrow; d *= 2, ++s) {
    double na[8], nb[8], nc[8];
    int q = 0;
    for (int i = threadIdx.x; i < nrow; i += blockDim.x, ++q) {
      const bool lo = i - d >= 0, hi = i + d < nrow;
      const double k1 = lo ? a[i] / b[i - d] : 0.0;
      const double k2 = hi ? c[i] / b[i + d] : 0.0;
      na[q] = lo ? -a[i - d] * k1 : 0.0;
      nc[q] = hi ? -c[i + d] * k2 : 0.0;
      nb[q] = b[i] - (lo ? c[i - d] * k1 : 0.0) - (hi ? a[i + d] * k2 : 0.0);
      fac.k1[tb + (long long)s * nrow + i] = k1;
      fac.k2[tb + (long long)s * nrow + i] = k2;
    }
    __syncthreads();
    q = 0;
    for (int i = threadIdx.x; i < nrow; i += blockDim.x, ++q) {
      a[i] = na[q];
      b[i] = nb[q];
      c[i] = nc[q];
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < nrow; i += blockDim.x)
    fac.bfin[((long long)rk * in.nfields + f) * nrow + i] = b[i];
}

// ---------------------------------------------------------------------------
// Real Crank-Nicolson solves of every (x,z) column (channel.f:606-626 and
// 739-747 explicit half; 771-1029 implicit half), solved by applying the
// shared PCR tables to the right-hand sides held in LDS; fused CFL minimum
// of COURANT (courant.f:34-43).  Block: kSolveCols consecutive x columns of
// one field; LDS [row][col].
// ---------------------------------------------------------------------------
constexpr int kSolveCols = 32;
constexpr int kSolveThreads = 1024;
constexpr int kSolveGroups = kSolveThreads / kSolveCols;  // row groups
constexpr int kSolveRows = 9;  // contiguous rows per group: NY+2 <= 9 * 32

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

__global__ void __launch_bounds__(kSolveThreads) k_solve(SolveArgs a) {
  extern __shared__ double sd[];
  __shared__ double red[kSolveThreads / 64];
  const Geo &g = a.g;
  const int f = blockIdx.y;
  const long long pp = g.pp;
  const int NY = g.ny, nrow = NY + 2;
  const int rk = a.rk;
  const int nst = a.fac.nsteps;
  const double hb = a.clk->h_bar[rk];
  const double T1 = a.nu * hb / 2.0;
  const double coef = (f < 3) ? T1 * a.nu_v_scale : (T1 / a.pr[f - 3]) * a.nu_v_scale;
  // LDS: right-hand sides [row][col], then this (rk, field)'s tables
  double *D = sd;
  double *K1 = D + nrow * kSolveCols, *K2 = K1 + nst * nrow;
  double *BF = K2 + nst * nrow, *VC = BF + nrow, *DYs = VC + nrow, *DYFs = DYs + nrow;
  int *IB = reinterpret_cast<int *>(DYFs + nrow);
  const long long tb = ((long long)rk * a.nfields + f) * (long long)kMaxPcr * nrow;
  for (int i = threadIdx.x; i < nst * nrow; i += blockDim.x) {
    K1[i] = a.fac.k1[tb + i];
    K2[i] = a.fac.k2[tb + i];
  }
  for (int i = threadIdx.x; i < nrow; i += blockDim.x) {
    BF[i] = a.fac.bfin[((long long)rk * a.nfields + f) * nrow + i];
    VC[i] = a.fac.vconst[(long long)f * nrow + i];
    IB[i] = a.fac.isbc[(long long)f * nrow + i];
    DYs[i] = g.dy[i];
    DYFs[i] = g.dyf[i];
  }
  const int lane = threadIdx.x % kSolveCols, rg = threadIdx.x / kSolveCols;
  const int rows = (nrow + kSolveGroups - 1) / kSolveGroups;
  const int i0 = rg * rows;
  const long long col = (long long)blockIdx.x * kSolveCols + lane;
  const bool live = col < pp;
  double *U = a.uphys + (long long)f * a.fs + col;
  const double *R = a.rphys + (long long)f * a.fs + col;
  // U_old rows i0-1 .. i0+rows and R rows i0 .. i0+rows-1, all loads in flight
  double u[kSolveRows + 2], r[kSolveRows];
#pragma unroll
  for (int q = 0; q < kSolveRows + 2; ++q) {
    const int i = i0 - 1 + q;
    u[q] = (live && q <= rows + 1 && i >= 0 && i < nrow) ? U[(long long)i * pp] : 0.0;
  }
#pragma unroll
  for (int q = 0; q < kSolveRows; ++q) {
    const int i = i0 + q;
    r[q] = (live && q < rows && i < nrow) ? R[(long long)i * pp] : 0.0;
  }
  __syncthreads();
  // right-hand sides: boundary constants or R + explicit CN term of U_old
#pragma unroll
  for (int q = 0; q < kSolveRows; ++q) {
    const int i = i0 + q;
    if (q < rows && i < nrow) {
      double d = 0.0;
      if (IB[i]) {
        d = VC[i];
      } else {
        const double um = u[q], u0 = u[q + 1], up = u[q + 2];
        double ex;
        if (f == 1)
          ex = ((up - u0) / DYFs[i] - (u0 - um) / DYFs[i - 1]) / DYs[i];
        else
          ex = ((up - u0) / DYs[i + 1] - (u0 - um) / DYs[i]) / DYFs[i];
        d = r[q] + coef * ex;
      }
      D[i * kSolveCols + lane] = d;
    }
  }
  __syncthreads();
  for (int s = 0, dd = 1; s < nst; ++s, dd *= 2) {
    const double *k1 = K1 + s * nrow, *k2 = K2 + s * nrow;
    double nd[kSolveRows];
#pragma unroll
    for (int q = 0; q < kSolveRows; ++q) {
      const int i = i0 + q;
      if (q < rows && i < nrow) {
        double v = D[i * kSolveCols + lane];
        if (i - dd >= 0) v -= k1[i] * D[(i - dd) * kSolveCols + lane];
        if (i + dd < nrow) v -= k2[i] * D[(i + dd) * kSolveCols + lane];
        nd[q] = v;
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kSolveRows; ++q) {
      const int i = i0 + q;
      if (q < rows && i < nrow) D[i * kSolveCols + lane] = nd[q];
    }
    __syncthreads();
  }
  const bool cour = a.courant && f < 3;
  const double dxf = (f == 0) ? a.dx : a.dz;
  double dtmin = __longlong_as_double(0x7FF0000000000000ll);
  const int wlo = a.wlo[f], whi = a.whi[f];
#pragma unroll
  for (int q = 0; q < kSolveRows; ++q) {
    const int i = i0 + q;
    if (!live || q >= rows || i >= nrow || i < wlo || i > whi) continue;
    const double x = D[i * kSolveCols + lane] / BF[i];
    U[(long long)i * pp] = x;
    if (cour && i >= 1 && i <= NY) {
      const double dl = (f == 1) ? DYs[i] : dxf;
      dtmin = fmin(dtmin, a.cfl * dl / fabs(x));
    }
  }
  if (cour) {
    // wave minimum by shuffles, then one LDS slot per wave
    for (int o = 32; o > 0; o >>= 1) dtmin = fmin(dtmin, __shfl_down(dtmin, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = dtmin;
    __syncthreads();
    if (threadIdx.x == 0) {
      double v = red[0];
      for (int w = 1; w < (int)(blockDim.x >> 6); ++w) v = fmin(v, red[w]);
      if (v >= 0.0) atomicMin(&a.clk->cfl_min_bits, (unsigned long long)__double_as_longlong(v));
    }
  }
}

// COURANT (courant.f:45-62) from the CFL minimum of the solve kernel; the
// PCR tables for the new DELTA_T are rebuilt right after by k_pcr_tables.
__global__ void k_courant_dt(Clock *clk, double dt_cap, int variable_dt, int update_dt) {
  if (threadIdx.x != 0) return;
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
  }
  c.cfl_min_bits = 0x7FF0000000000000ull;
}

// ---------------------------------------------------------------------------
// REM_DIV_CHAN (channel.f:1103-1214 with APPLY_BC_REM_DIV_MPI mpi.f:1239-1278)
// and the CP update (channel.f:1059-1065).  Each block holds the Poisson
// systems of kRdModes consecutive (kx,kz) modes, rows 1..NY, in LDS and
// solves them by parallel cyclic reduction; then projects CU and updates CP.
// ---------------------------------------------------------------------------
constexpr int kRdModes = 8;  // rows per thread <= 9: NY <= 9 * 256 / kRdModes

struct RemDivArgs {
  Geo g;
  Clock *clk;
  int rk;       // <0: no CP update
  int cp_old;   // 1: COURANT ran in this substep, divide by the saved H_BAR
  double2 *cu[3];
  double2 *cp;
  double2 *phi;
  const double *rl, *ru;  // 1/(DY(J)DYF(J)), 1/(DY(J+1)DYF(J))
};

__global__ void __launch_bounds__(256) k_rem_div(RemDivArgs a) {
  extern __shared__ double sr[];
  const Geo &g = a.g;
  const int NY = g.ny;
  const int n = NY;  // rows 1..NY stored at 0..NY-1
  double *A = sr, *B = sr + n * kRdModes, *C = sr + 2 * n * kRdModes;
  double2 *D = reinterpret_cast<double2 *>(sr + 3 * n * kRdModes);
  const int lane = threadIdx.x % kRdModes, rg = threadIdx.x / kRdModes;
  const int ngrp = blockDim.x / kRdModes;
  const int m = blockIdx.x * kRdModes + lane;
  const bool live = m < g.pf;
  const long long pf = g.pf;
  int kx = 0, kzp = 0;
  if (live) {
    kzp = (int)fdiv((unsigned)m, g.f_nkx1);
    kx = m - kzp * g.nkx1;
  }
  const double kxv = g.kx[kx], kzv = g.kz[kzp], kx2 = g.kx2[kx], kz2 = g.kz2[kzp];
  const bool mean = (kx == 0 && kzp == 0);
  // system rows J = 1..NY (channel.f:1126-1143, BC rows mpi.f:1244-1272)
  for (int r = rg; r < n; r += ngrp) {
    const int J = r + 1;
    double L, Dg, Uu;
    double2 V = make_double2(0.0, 0.0);
    if (J == 1) {
      L = 0.; Dg = 1.; Uu = mean ? 0. : -1.;
    } else if (J == NY) {
      L = 1.; Dg = -1.; Uu = 0.;
    } else {
      L = a.rl[J];
      Uu = a.ru[J];
      Dg = -kx2 - kz2 - Uu - L;
      if (live) {
        const double2 c1 = a.cu[0][pf * J + m], c3 = a.cu[2][pf * J + m];
        const double2 c2 = a.cu[1][pf * J + m], c2p = a.cu[1][pf * (J + 1) + m];
        const double2 t1 = cmul_ik(kxv, c1), t3 = cmul_ik(kzv, c3);
        const double d = g.dyf[J];
        V = make_double2(t1.x + (c2p.x - c2.x) / d + t3.x, t1.y + (c2p.y - c2.y) / d + t3.y);
      }
    }
    A[r * kRdModes + lane] = L;
    B[r * kRdModes + lane] = Dg;
    C[r * kRdModes + lane] = Uu;
    D[r * kRdModes + lane] = V;
  }
  __syncthreads();
  for (int dd = 1; dd < n; dd *= 2) {
    double na[9], nb[9], nc[9];
    double2 nd[9];
    int q = 0;
    for (int r = rg; r < n; r += ngrp, ++q) {
      const int o = r * kRdModes + lane;
      const bool lo = r - dd >= 0, hi = r + dd < n;
      const int ol = (r - dd) * kRdModes + lane, oh = (r + dd) * kRdModes + lane;
      const double k1 = lo ? A[o] / B[ol] : 0.0;
      const double k2 = hi ? C[o] / B[oh] : 0.0;
      na[q] = lo ? -A[ol] * k1 : 0.0;
      nc[q] = hi ? -C[oh] * k2 : 0.0;
      nb[q] = B[o] - (lo ? C[ol] * k1 : 0.0) - (hi ? A[oh] * k2 : 0.0);
      double2 v = D[o];
      if (lo) v = make_double2(v.x - k1 * D[ol].x, v.y - k1 * D[ol].y);
      if (hi) v = make_double2(v.x - k2 * D[oh].x, v.y - k2 * D[oh].y);
      nd[q] = v;
    }
    __syncthreads();
    q = 0;
    for (int r = rg; r < n; r += ngrp, ++q) {
      const int o = r * kRdModes + lane;
      A[o] = na[q];
      B[o] = nb[q];
      C[o] = nc[q];
      D[o] = nd[q];
    }
    __syncthreads();
  }
  // phi = d / b (held in D), then the projection and pressure update
  for (int r = rg; r < n; r += ngrp) {
    const int o = r * kRdModes + lane;
    const double b = B[o];
    const double2 v = D[o];
    D[o] = (b != 0.0) ? make_double2(v.x / b, v.y / b) : make_double2(0.0, 0.0);
  }
  __syncthreads();
  if (!live) return;
  const double hb = (a.rk < 0) ? 1.0 : (a.cp_old ? a.clk->h_cp : a.clk->h_bar[a.rk]);
  for (int r = rg; r < n; r += ngrp) {
    const int J = r + 1;
    const double2 phi = D[r * kRdModes + lane];
    a.phi[pf * J + m] = phi;
    double2 t = cmul_ik(kxv, phi), u = a.cu[0][pf * J + m];
    a.cu[0][pf * J + m] = make_double2(u.x - t.x, u.y - t.y);
    t = cmul_ik(kzv, phi);
    u = a.cu[2][pf * J + m];
    a.cu[2][pf * J + m] = make_double2(u.x - t.x, u.y - t.y);
    if (J >= 2) {
      const double2 pm = D[(r - 1) * kRdModes + lane];
      const double d = g.dy[J];
      u = a.cu[1][pf * J + m];
      a.cu[1][pf * J + m] = make_double2(u.x - (phi.x - pm.x) / d, u.y - (phi.y - pm.y) / d);
      if (a.rk >= 0) {
        double2 p = a.cp[pf * J + m];
        a.cp[pf * J + m] = make_double2(p.x + phi.x / hb, p.y + phi.y / hb);
      }
    }
  }
}

// REM_DIV_CHAN with the reference's Thomas sweeps (THOMAS_FORWARD/BACKWARD_
// COMPLEX_MPI, mpi.f:708-865) in their operation order, one thread per
// (kx,kz) mode: all threads stage the right-hand sides of kRdtModes modes
// into LDS with coalesced loads, one thread per mode runs the sequential
// sweeps out of LDS, and all threads apply the projection and the CP update.
constexpr int kRdtModes = 32;

__global__ void __launch_bounds__(256) k_rem_div_thomas(RemDivArgs a) {
  extern __shared__ double st[];
  const Geo &g = a.g;
  const int NY = g.ny;
  double2 *Gs = reinterpret_cast<double2 *>(st);                  // [J][mode], J = 0..NY
  double *Bs = st + 2 * (NY + 1) * kRdtModes;                       // [J][mode]
  double *rl = Bs + (NY + 1) * kRdtModes, *ru = rl + (NY + 1);       // [J]
  const int lane = threadIdx.x % kRdtModes, rg = threadIdx.x / kRdtModes;
  const int ngrp = blockDim.x / kRdtModes;
  const int m = blockIdx.x * kRdtModes + lane;
  const bool live = m < g.pf;
  const long long pf = g.pf;
  int kx = 0, kzp = 0;
  if (live) {
    kzp = (int)fdiv((unsigned)m, g.f_nkx1);
    kx = m - kzp * g.nkx1;
  }
  const double kxv = g.kx[kx], kzv = g.kz[kzp], kx2 = g.kx2[kx], kz2 = g.kz2[kzp];
  const bool mean = (kx == 0 && kzp == 0);
  for (int j = threadIdx.x; j <= NY; j += blockDim.x) {
    rl[j] = a.rl[j];
    ru[j] = a.ru[j];
  }
  // right-hand sides of rows 2..NY-1 (channel.f:1137-1143); rows 1, NY are BC
  // rows.  Batches of 4 rows per thread keep 16 loads in flight.
  for (int J0 = 2 + rg; J0 < NY; J0 += 4 * ngrp) {
    double2 c1[4], c2[4], c2p[4], c3[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int J = J0 + b * ngrp;
      if (live && J < NY) {
        c1[b] = a.cu[0][pf * J + m];
        c3[b] = a.cu[2][pf * J + m];
        c2[b] = a.cu[1][pf * J + m];
        c2p[b] = a.cu[1][pf * (J + 1) + m];
      }
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int J = J0 + b * ngrp;
      if (J < NY) {
        double2 V = make_double2(0.0, 0.0);
        if (live) {
          const double2 t1 = cmul_ik(kxv, c1[b]), t3 = cmul_ik(kzv, c3[b]);
          const double d = g.dyf[J];
          V = make_double2(t1.x + (c2p[b].x - c2[b].x) / d + t3.x, t1.y + (c2p[b].y - c2[b].y) / d + t3.y);
        }
        Gs[J * kRdtModes + lane] = V;
      }
    }
  }
  __syncthreads();
  if (rg == 0 && live) {
    // forward, J = 2..NY (mpi.f:736-747); row 1: D=1, U=0 (mean) or -1, G=0
    double bprev = 1.0, uprev = mean ? 0.0 : -1.0;
    double2 gprev = make_double2(0.0, 0.0);
    for (int J = 2; J <= NY; ++J) {
      double L, D, U;
      double2 V;
      if (J < NY) {
        L = rl[J];
        U = ru[J];
        D = -kx2 - kz2 - U - L;
        V = Gs[J * kRdtModes + lane];
      } else {
        L = 1.;
        D = -1.;
        U = 0.;
        V = make_double2(0.0, 0.0);
      }
      const double A = (bprev != 0.0) ? -L / bprev : 0.0;
      const double B = D + A * uprev;
      const double2 G = make_double2(V.x + A * gprev.x, V.y + A * gprev.y);
      Bs[J * kRdtModes + lane] = B;
      Gs[J * kRdtModes + lane] = G;
      bprev = B;
      uprev = U;
      gprev = G;
    }
    // backward, J = NY..1 (mpi.f:838-856)
    double2 x = (bprev != 0.0) ? make_double2(gprev.x / bprev, gprev.y / bprev) : make_double2(0.0, 0.0);
    Gs[NY * kRdtModes + lane] = x;
    for (int J = NY - 1; J >= 1; --J) {
      double B, U;
      double2 G;
      if (J >= 2) {
        B = Bs[J * kRdtModes + lane];
        G = Gs[J * kRdtModes + lane];
        U = ru[J];
      } else {
        B = 1.0;
        G = make_double2(0.0, 0.0);
        U = mean ? 0.0 : -1.0;
      }
      x = (B != 0.0) ? make_double2((G.x - U * x.x) / B, (G.y - U * x.y) / B) : make_double2(0.0, 0.0);
      Gs[J * kRdtModes + lane] = x;
    }
  }
  __syncthreads();
  if (!live) return;
  const double hb = (a.rk < 0) ? 1.0 : (a.cp_old ? a.clk->h_cp : a.clk->h_bar[a.rk]);
  for (int J0 = 1 + rg; J0 <= NY; J0 += 4 * ngrp) {
    double2 u1[4], u2[4], u3[4], pc[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int J = J0 + b * ngrp;
      if (J <= NY) {
        u1[b] = a.cu[0][pf * J + m];
        u3[b] = a.cu[2][pf * J + m];
        u2[b] = a.cu[1][pf * J + m];
        if (a.rk >= 0) pc[b] = a.cp[pf * J + m];
      }
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int J = J0 + b * ngrp;
      if (J > NY) continue;
      const double2 phi = Gs[J * kRdtModes + lane];
      a.phi[pf * J + m] = phi;
      double2 t = cmul_ik(kxv, phi);
      a.cu[0][pf * J + m] = make_double2(u1[b].x - t.x, u1[b].y - t.y);
      t = cmul_ik(kzv, phi);
      a.cu[2][pf * J + m] = make_double2(u3[b].x - t.x, u3[b].y - t.y);
      if (J >= 2) {
        const double2 pm = Gs[(J - 1) * kRdtModes + lane];
        const double d = g.dy[J];
        a.cu[1][pf * J + m] = make_double2(u2[b].x - (phi.x - pm.x) / d, u2[b].y - (phi.y - pm.y) / d);
        if (a.rk >= 0) a.cp[pf * J + m] = make_double2(pc[b].x + phi.x / hb, pc[b].y + phi.y / hb);
      }
    }
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

static FactorIn make_factor_in(const Ctx &c) {
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
  in.dy = c.d_dy;
  in.dyf = c.d_dyf;
  return in;
}

hipError_t launch_pcr_tables(Ctx &c) {
  FactorIn in = make_factor_in(c);
  const size_t shm = (size_t)3 * (c.g.ny + 2) * sizeof(double);
  hipLaunchKernelGGL(k_pcr_tables, dim3(3 * c.nfields), dim3(256), shm, c.stream, c.clock, in, c.fac);
  return hipGetLastError();
}

static hipError_t launch_nonlinear(Ctx &c, int rk) {
  StepArgs a = make_step_args(c, rk);
  hipLaunchKernelGGL(k_nonlinear, dim3(c.g.ny, c.nfields), dim3(c.plane_threads), c.lds_bytes, c.stream, a);
  return hipGetLastError();
}

static size_t solve_lds(const Ctx &c) {
  const size_t nrow = c.g.ny + 2;
  return (nrow * kSolveCols + 2 * (size_t)c.fac.nsteps * nrow + 4 * nrow) * sizeof(double) + nrow * sizeof(int);
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
  const int nb = (int)((c.g.pp + kSolveCols - 1) / kSolveCols);
  hipLaunchKernelGGL(k_solve, dim3(nb, c.nfields), dim3(kSolveThreads), solve_lds(c), c.stream, a);
  return hipGetLastError();
}

static size_t rem_div_thomas_lds(const Ctx &c) {
  return (size_t)(c.g.ny + 1) * kRdtModes * 3 * sizeof(double) + 2 * (size_t)(c.g.ny + 1) * sizeof(double);
}

hipError_t launch_rem_div(Ctx &c, int rk) {
  RemDivArgs a{};
  a.g = c.g;
  a.clk = c.clock;
  a.rk = rk;
  a.cp_old = (rk == 2 && c.p.variable_dt) ? 1 : 0;
  for (int i = 0; i < 3; ++i) a.cu[i] = c.cu[i];
  a.cp = c.cp;
  a.phi = c.phi;
  a.rl = c.rd_l;
  a.ru = c.rd_u;
  if (c.rd_pcr) {
    const int nb = (c.g.pf + kRdModes - 1) / kRdModes;
    const size_t shm = (size_t)c.g.ny * kRdModes * 5 * sizeof(double);
    hipLaunchKernelGGL(k_rem_div, dim3(nb), dim3(256), shm, c.stream, a);
  } else {
    const int nb = (c.g.pf + kRdtModes - 1) / kRdtModes;
    hipLaunchKernelGGL(k_rem_div_thomas, dim3(nb), dim3(256), rem_div_thomas_lds(c), c.stream, a);
  }
  return hipGetLastError();
}

static hipError_t launch_courant(Ctx &c) {
  hipLaunchKernelGGL(k_courant_dt, dim3(1), dim3(64), 0, c.stream, c.clock, c.dt_cap, c.p.variable_dt,
                     c.p.update_dt);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_pcr_tables(c);
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
    case 3: return courant ? launch_courant(c) : hipSuccess;
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
  hipError_t e =
      hipFuncSetAttribute((const void *)k_nonlinear, hipFuncAttributeMaxDynamicSharedMemorySize, c.lds_bytes);
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute((const void *)k_solve, hipFuncAttributeMaxDynamicSharedMemorySize, (int)solve_lds(c));
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute((const void *)k_rem_div_thomas, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)rem_div_thomas_lds(c));
}

}  // namespace dh
