// Kernels of one RK3/Crank-Nicolson substep (RK_CHAN_1, channel.f:62-1073)
// other than the plain plane transforms:
//   k_nonlinear   explicit Fourier right-hand side, nonlinear products and
//                 their forward transforms, RK combination, explicit half of
//                 the y-viscous term and the inverse transform of the
//                 right-hand side (channel.f:77-626);
//   k_solve_seg   real tridiagonal Crank-Nicolson solve of every (x,z)
//                 column, and the CFL minimum (channel.f:771-1029,
//                 courant.f:34-43);
//   k_rem_div_thomas
//                 divergence, pressure-Poisson solve per (kx,kz), projection
//                 and pressure update (channel.f:1103-1214, 1059-1065);
//   k_courant_dt  the rest of COURANT (courant.f:45-62) on the device.
#include <type_traits>

#include "dh_fft_dev.h"

namespace dh {

constexpr int kMaxQ = 4;  // Fourier modes owned per thread in the 1024-thread k_nonlinear
constexpr int kNlRegThreads = 512, kNlRegQ = 8;  // the register-row variant

struct StepArgs {
  Geo g;
  Clock *clk;
  int rk;  // 0..2
  const double2 *cu[3];
  const double2 *cp;
  double2 *cf[3];
  const double2 *cth;
  double2 *cfth;
  const void *uphys;  // physical U and R: double, or float in the mixed form
  void *rphys;
  long long fs;   // physical field stride (NY+2)*NX*NZ
  long long cfs;  // Fourier field stride (NY+2)*pf
  int c0, j0;     // component of blockIdx.y = 0 and plane of the first block row
  int rows32;     // FP64 LDS variant of <= 512 threads: x rows of NX = 64 in registers (NZ % 16 == 0)
  double nu, nu_v_scale, beta, i_ro, grav_x, grav_y, grav_z, px0, omega0, amp_omega0;
  int f_type, n_th;
  double ri[DH_MAX_TH], pr[DH_MAX_TH], drhodx[DH_MAX_TH], drhodz[DH_MAX_TH];
  int jstart_th[DH_MAX_TH], jend_th[DH_MAX_TH];
  const double2 *sgs;  // LES: subgrid increments of CF1..3, CFTH (null without LES)
};

__device__ __forceinline__ double kpow(double x, double beta) { return (beta == 1.0) ? x : pow(x, beta); }

// Physical product p (0..2) of component c at plane J (channel.f:410-571
// for c<3, 669-720 for scalars), resolved once per block into one of five
// forms over up to five planes, each evaluated for the two samples stored in
// one complex slot (x = 2m, 2m+1):
//   0: s0 s1                                  U3U1, TH U1, TH U3
//   1: s0 s0                                  U1U1, U3U3
//   2: (s0 s3 + s1 s3 - s1 s4 - s2 s4)/(2 DYF(J))
//      with s0..s2 = C(J+1), C(J), C(J-1), s3, s4 = U2(J+1), U2(J):
//      d/dy(C U2) on the staggered grid, C = U1, U3 or TH
//   3: ((DYF(J) s0 + DYF(J-1) s1) / (2 DY(J))) s2   interp_y(U1 or U3) U2
//   4: (0.25 (s1+s0)^2 - 0.25 (s1+s2)^2) / DY(J)    d/dy(U2 U2)
// Planes hold complex pairs C (double2, or float2 in the mixed form) whose
// real type S = Sc<C> is also the arithmetic type of the forms.
template <class C>
struct Prod {
  const Sc<C> *s[5];
  int kind;
  Sc<C> w0, w1, w2;
  Sc<C> d, r;  // divisor of forms 2-4 and its reciprocal (correctly rounded in FP64)
};

template <class C>
__device__ __forceinline__ Prod<C> make_prod(const StepArgs &a, int c, int p, int J) {
  using S = Sc<C>;
  const long long pp = a.g.pp;
  const S *up = static_cast<const S *>(a.uphys);
  const S *U1 = up, *U2 = up + a.fs, *U3 = up + 2 * a.fs;
  auto pl = [&](const S *F, int j) { return F + (long long)j * pp; };
  Prod<C> q{};
  if (c == 0 || c == 2) {
    const S *Uc = (c == 0) ? U1 : U3;
    if (p == 0) {
      q.kind = 0; q.s[0] = pl(U3, J); q.s[1] = pl(U1, J);
    } else if (p == 1) {
      q.kind = 1; q.s[0] = pl(Uc, J);
    } else {
      q.kind = 2; q.s[0] = pl(Uc, J + 1); q.s[1] = pl(Uc, J); q.s[2] = pl(Uc, J - 1);
      q.s[3] = pl(U2, J + 1); q.s[4] = pl(U2, J); q.w0 = 2.0 * a.g.dyf[J];
    }
  } else if (c == 1) {
    if (p <= 1) {
      const S *Uc = (p == 0) ? U1 : U3;
      q.kind = 3; q.s[0] = pl(Uc, J); q.s[1] = pl(Uc, J - 1); q.s[2] = pl(U2, J);
      q.w0 = a.g.dyf[J]; q.w1 = a.g.dyf[J - 1]; q.w2 = 2. * a.g.dy[J];
    } else {
      q.kind = 4; q.s[0] = pl(U2, J + 1); q.s[1] = pl(U2, J); q.s[2] = pl(U2, J - 1); q.w0 = a.g.dy[J];
    }
  } else {
    const S *TH = up + (long long)c * a.fs;
    if (p <= 1) {
      q.kind = 0; q.s[0] = pl(TH, J); q.s[1] = pl(p == 0 ? U1 : U3, J);
    } else {
      q.kind = 2; q.s[0] = pl(TH, J + 1); q.s[1] = pl(TH, J); q.s[2] = pl(TH, J - 1);
      q.s[3] = pl(U2, J + 1); q.s[4] = pl(U2, J); q.w0 = 2.0 * a.g.dyf[J];
    }
  }
  // reciprocals from the tables (a power-of-two factor scales RN(1/x) exactly)
  q.d = (q.kind == 3) ? q.w2 : q.w0;
  q.r = (S)((q.kind == 2) ? 0.5 * a.g.rdyf[J] : (q.kind == 3) ? 0.5 * a.g.rdy[J] : (q.kind == 4) ? a.g.rdy[J] : 0.0);
  return q;
}

// The samples a form reads (planes s[0..n-1] at one offset), loaded apart
// from the arithmetic so that a caller can issue many loads ahead of it.
template <int KIND, class C>
struct ProdRaw {
  static constexpr int n = (KIND == 1) ? 1 : (KIND == 0) ? 2 : (KIND == 2) ? 5 : 3;
  C s[n];
};

template <int KIND, class C>
__device__ __forceinline__ ProdRaw<KIND, C> fetch_prod(const Prod<C> &q, long long off) {
  ProdRaw<KIND, C> r;
#pragma unroll
  for (int i = 0; i < ProdRaw<KIND, C>::n; ++i) r.s[i] = *reinterpret_cast<const C *>(q.s[i] + off);
  return r;
}

template <int KIND, class C>
__device__ __forceinline__ C eval_raw(const Prod<C> &q, const ProdRaw<KIND, C> &r) {
  using S = Sc<C>;
  if constexpr (KIND == 0) {
    const C u = r.s[0], v = r.s[1];
    return mkc<C>(u.x * v.x, u.y * v.y);
  } else if constexpr (KIND == 1) {
    const C u = r.s[0];
    return mkc<C>(u.x * u.x, u.y * u.y);
  } else if constexpr (KIND == 2) {
    const C up = r.s[0], u0 = r.s[1], um = r.s[2], vp = r.s[3], v0 = r.s[4];
    return mkc<C>(div_by_recip(up.x * vp.x + u0.x * vp.x - u0.x * v0.x - um.x * v0.x, q.d, q.r),
                  div_by_recip(up.y * vp.y + u0.y * vp.y - u0.y * v0.y - um.y * v0.y, q.d, q.r));
  } else if constexpr (KIND == 3) {
    const C u0 = r.s[0], um = r.s[1], v = r.s[2];
    return mkc<C>(div_by_recip(q.w0 * u0.x + q.w1 * um.x, q.d, q.r) * v.x,
                  div_by_recip(q.w0 * u0.y + q.w1 * um.y, q.d, q.r) * v.y);
  } else {
    const C vp = r.s[0], v0 = r.s[1], vm = r.s[2];
    const S ax = v0.x + vp.x, bx = v0.x + vm.x, ay = v0.y + vp.y, by = v0.y + vm.y;
    const S q4 = S(0.25);
    return mkc<C>(div_by_recip(q4 * (ax * ax) - q4 * (bx * bx), q.d, q.r),
                  div_by_recip(q4 * (ay * ay) - q4 * (by * by), q.d, q.r));
  }
}

template <int KIND, class C>
__device__ __forceinline__ C eval_prod(const Prod<C> &q, long long off) {
  return eval_raw<KIND, C>(q, fetch_prod<KIND, C>(q, off));
}

template <class C>
__device__ __forceinline__ C eval_prod(const Prod<C> &q, long long off) {
  switch (q.kind) {
    case 0: return eval_prod<0, C>(q, off);
    case 1: return eval_prod<1, C>(q, off);
    case 2: return eval_prod<2, C>(q, off);
    case 3: return eval_prod<3, C>(q, off);
    default: return eval_prod<4, C>(q, off);
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

// cf_init for the NQ modes m = threadIdx.x + q nthr of one thread, into
// accumulators of type C (computed in FP64, rounded to FP32 in the mixed
// form).  For the velocity components without scalars and with BETA = 1
// (the channel configuration) the loads of a batch of QB modes are issued
// before their arithmetic, so the modes' memory latencies overlap; other
// configurations take cf_init mode by mode.  Indices beyond the plane are
// clamped for the loads and their results never stored.
// SGS: LES, whose subgrid increments (les.f:280-336; channel.f:346-372
// for the scalars) join the explicit terms before the products, as
// les_chan adds them to CF ahead of the nonlinear terms.
template <int NQ, bool SGS, class C>
__device__ __forceinline__ void cf_init_all(const StepArgs &a, int c, int J, int nthr, C (&acc)[NQ]) {
  const Geo &g = a.g;
  if (SGS || c >= 3 || a.n_th > 0 || a.beta != 1.0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int m = threadIdx.x + q * nthr;
      if (m < g.pf) {
        const int kzp = (int)fdiv((unsigned)m, g.f_nkx1), kx = m - kzp * g.nkx1;
        double2 cf = cf_init(a, c, J, kx, kzp, m);
        if constexpr (SGS) cf = cadd(cf, a.sgs[c * a.cfs + (long long)J * g.pf + m]);
        acc[q] = cvt<C>(cf);
      }
    }
    return;
  }
  const double2 *U = a.cu[c], *W = a.cu[c == 0 ? 2 : 0];
  const long long base = (long long)J * g.pf;
  // all modes at once in FP64; half of them at a time in the mixed form,
  // whose kernel has half the registers
  constexpr int QB = (sizeof(C) == sizeof(double2) || NQ < 2) ? NQ : NQ / 2;
#pragma unroll
  for (int q0 = 0; q0 < NQ; q0 += QB) {
    double2 u[QB], w[QB];
    double k2[QB];
#pragma unroll
    for (int b = 0; b < QB; ++b) {
      const int m = min((int)threadIdx.x + (q0 + b) * nthr, g.pf - 1);
      const int kzp = (int)fdiv((unsigned)m, g.f_nkx1), kx = m - kzp * g.nkx1;
      u[b] = U[base + m];
      k2[b] = g.kx2[kx] + g.kz2[kzp];
    }
    if (c != 1) {
#pragma unroll
      for (int b = 0; b < QB; ++b) w[b] = W[base + min((int)threadIdx.x + (q0 + b) * nthr, g.pf - 1)];
    }
#pragma unroll
    for (int b = 0; b < QB; ++b) {
      const double f = a.nu * k2[b];  // kpow(k2, 1) = k2
      double2 cf = make_double2(-(f * u[b].x), -(f * u[b].y));
      if (c == 0) {
        cf = make_double2(cf.x + a.i_ro * w[b].x, cf.y + a.i_ro * w[b].y);
      } else if (c == 2) {
        cf = make_double2(cf.x - a.i_ro * w[b].x, cf.y - a.i_ro * w[b].y);
      }
      acc[q0 + b] = cvt<C>(cf);
    }
  }
}

// Fused explicit pass for plane J (1..NY) and component c (U1,U2,U3,TH...).
// REG: row transforms in registers, 512 threads (twice the register budget
// of the 1024-thread LDS variant), NQ Fourier modes owned per thread.
// The mixed form transforms only the fluctuations: every field without its
// (kx,kz) = (0,0) coefficient, the plane mean, which stays FP64.  The three
// products of a component are bilinear, so their terms with a plane-mean
// factor are mode-wise products of FP64 Fourier coefficients and need no
// transform.  mean_terms returns those terms of f0 P0 + f1 P1 + P2 for one
// mode (f = ikx or ikz as the pass applies them, P in the forms of
// make_prod): xm, x0, xp are the component's coefficients at J-1, J, J+1,
// mu[f][j] the means of U1, U2, U3 and the component's scalar at J-1+j.  At
// the (0,0) mode the fluctuations vanish and only the mean x mean part of
// the y flux P2 remains.
__device__ __forceinline__ double2 cscale(double s, double2 z) { return make_double2(s * z.x, s * z.y); }

__device__ double2 mean_terms(const StepArgs &a, int c, int J, long long o, bool m00, double kxv, double kzv,
                              double2 xm, double2 x0, double2 xp, const double (&mu)[4][3]) {
  const Geo &g = a.g;
  const long long pf = g.pf;
  const double2 zero = make_double2(0.0, 0.0);
  if (m00) xm = x0 = xp = zero;
  auto ld = [&](int f, int dj) { return m00 ? zero : a.cu[f][o + dj * pf]; };
  double2 L0, L1, L2;
  double mm;  // mean x mean part of P2
  if (c == 1) {  // interp(U1) U2, interp(U3) U2, d/dy(U2 U2)
    const double2 A0 = ld(0, 0), Am = ld(0, -1), B0 = ld(2, 0), Bm = ld(2, -1);
    const double w0 = g.dyf[J], w1 = g.dyf[J - 1], r = 0.5 * g.rdy[J];
    const double2 I1 = cscale(r, cadd(cscale(w0, A0), cscale(w1, Am)));
    const double2 I3 = cscale(r, cadd(cscale(w0, B0), cscale(w1, Bm)));
    const double i1 = (w0 * mu[0][1] + w1 * mu[0][0]) * r, i3 = (w0 * mu[2][1] + w1 * mu[2][0]) * r;
    L0 = cadd(cscale(i1, x0), cscale(mu[1][1], I1));
    L1 = cadd(cscale(i3, x0), cscale(mu[1][1], I3));
    const double am = mu[1][1] + mu[1][2], bm = mu[1][1] + mu[1][0];
    L2 = cscale(g.rdy[J], csub(cscale(0.5 * am, cadd(x0, xp)), cscale(0.5 * bm, cadd(x0, xm))));
    mm = (0.25 * am * am - 0.25 * bm * bm) * g.rdy[J];
  } else {
    // own field C (U1, U3 or TH) with the y flux d/dy(C U2); P0 and P1 are
    // U3 U1 and U_c U_c for velocity, TH U1 and TH U3 for a scalar
    const double *uo = (c < 3) ? mu[c] : mu[3];
    const double2 V0 = ld(1, 0), Vp = ld(1, 1);
    if (c < 3) {
      const int other = 2 - c;
      const double2 Y0 = ld(other, 0);
      L0 = cadd(cscale(mu[other][1], x0), cscale(uo[1], Y0));
      L1 = cscale(2.0 * uo[1], x0);
    } else {
      const double2 A0 = ld(0, 0), B0 = ld(2, 0);
      L0 = cadd(cscale(uo[1], A0), cscale(mu[0][1], x0));
      L1 = cadd(cscale(uo[1], B0), cscale(mu[2][1], x0));
    }
    const double r2 = 0.5 * g.rdyf[J];
    double2 s = cadd(cscale(uo[2], Vp), cscale(mu[1][2], xp));
    s = cadd(s, cadd(cscale(uo[1], Vp), cscale(mu[1][2], x0)));
    s = csub(s, cadd(cscale(uo[1], V0), cscale(mu[1][1], x0)));
    s = csub(s, cadd(cscale(uo[0], V0), cscale(mu[1][1], xm)));
    L2 = cscale(r2, s);
    mm = (uo[2] * mu[1][2] + uo[1] * mu[1][2] - uo[1] * mu[1][1] - uo[0] * mu[1][1]) * r2;
  }
  if (m00) return make_double2(mm, 0.0);
  const double f0 = (c == 0) ? kzv : kxv, f1 = (c == 0) ? kxv : kzv;
  return cadd(cadd(cmul_ik(f0, L0), cmul_ik(f1, L1)), L2);
}

// C: complex type of the plane and the physical arrays (double2, or float2
// in the mixed form, whose products and transforms run in FP32 while the
// Fourier state, the accumulation of CF and the RK combination stay FP64).
// R32: the LDS variant may take the x rows of NX = 64 in registers (when
// a.rows32 is set); compiled only where the register budget holds them.
template <bool REG, int NQ, class C, bool SGS = false, bool R32 = false>
__device__ __forceinline__ void nonlinear_pass(const StepArgs &a, C *A) {
  const Geo &g = a.g;
  const auto T = load_tables(A, g);
  // Work item of this block: workgroups are dealt to the 8 XCDs round-robin,
  // so give XCD x a contiguous range of (plane, component) items, planes
  // outer: the blocks sharing an XCD's L2 then read the same U planes.
  int J, c;
  {
    const int nblk = gridDim.x * gridDim.y, lin = blockIdx.x + blockIdx.y * gridDim.x;
    const int x = lin & 7, q = nblk >> 3, r = nblk & 7;
    const int w = x * q + min(x, r) + (lin >> 3);
    const int jr = w / (int)gridDim.y;
    J = __builtin_amdgcn_readfirstlane(a.j0 + jr);
    c = __builtin_amdgcn_readfirstlane(a.c0 + w - jr * (int)gridDim.y);
  }
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

  C acc[NQ];  // CF of this thread's modes (FP64, or FP32 in the mixed form)
  cf_init_all<NQ, SGS>(a, c, J, nthr, acc);
  constexpr bool DELTA = sizeof(C) != sizeof(double2);  // the mixed form transforms increments
  if constexpr (DELTA) {
    // the product terms with a plane-mean factor, in FP64, before the
    // transformed fluctuation products (few registers are live here)
    const double2 *CCm = (c < 3) ? a.cu[c] : a.cth + (c - 3) * a.cfs;
    double mu[4][3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const long long pl = (long long)(J - 1 + j) * g.pf;
      mu[0][j] = a.cu[0][pl].x;
      mu[1][j] = a.cu[1][pl].x;
      mu[2][j] = a.cu[2][pl].x;
      mu[3][j] = (c >= 3) ? CCm[pl].x : 0.0;
    }
    const long long base0 = (long long)J * g.pf;
#pragma unroll 2
    for (int q = 0; q < NQ; ++q) {
      const int m = threadIdx.x + q * nthr;
      if (m < g.pf) {
        const int kzp = (int)fdiv((unsigned)m, g.f_nkx1), kx = m - kzp * g.nkx1;
        const long long o = base0 + m;
        const double2 L = mean_terms(a, c, J, o, kx == 0 && kzp == 0, g.kx[kx], g.kz[kzp], CCm[o - g.pf], CCm[o],
                                     CCm[o + g.pf], mu);
        acc[q] = mkc<C>((Sc<C>)(acc[q].x - L.x), (Sc<C>)(acc[q].y - L.y));
      }
    }
  }
  // the three products, forward transformed and subtracted with their
  // derivative factors: c==0 takes (ikz, ikx, 1), all others (ikx, ikz, 1)
  // (statement order of channel.f:410-571, 669-720)
  const LoadTile lt(g);
  const int nw = nthr >> 6;
  // quotients by per-block constants take their correctly rounded
  // reciprocals (div_by_recip): the same doubles as the divisions
  using S = Sc<C>;
  const S nzd = S(g.dnz), rnz = S(g.rnz);
  // acc -= f (plane / NZ) after the z transform, f = ikz (mode 0), ikx (1)
  // or 1 (2); the wavenumbers are loaded ahead of the LDS reads
  auto accumulate = [&](int mode) {
    S kq[NQ];
    if (mode != 2) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const int m = min((int)threadIdx.x + q * nthr, g.pf - 1);
        const int kzp = (int)fdiv((unsigned)m, g.f_nkx1), kx = m - kzp * g.nkx1;
        kq[q] = S((mode == 0) ? g.kz[kzp] : g.kx[kx]);
      }
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int m = threadIdx.x + q * nthr;
      if (m < g.pf) {
        const int kzp = (int)fdiv((unsigned)m, g.f_nkx1), kx = m - kzp * g.nkx1;
        const C s0 = A[fwd_slot(g, T, kx, kzp)];
        const C s = mkc<C>(div_by_recip(s0.x, nzd, rnz), div_by_recip(s0.y, nzd, rnz));
        acc[q] = csub(acc[q], (mode == 2) ? s : cmul_ik(kq[q], s));
      }
    }
    lds_barrier();
  };
  const int pz = (c == 0) ? 0 : 1;  // the ikz product; the other of 0, 1 takes ikx
  if constexpr (REG) {
    // x transforms in registers.  The ikz product has its own z transform;
    // the ikx product and the plain y-flux product (p = 2) share one: the
    // x-transformed ikx product is scaled by i KX and added to the y flux's
    // (both transforms are linear and KX is constant along z).
    auto rows = [&](int p, auto merge) {
      constexpr bool MERGE = decltype(merge)::value;
      const Prod<C> pr = make_prod<C>(a, c, p, J);
      // the form is resolved once so that each slab's loads are
      // straight-line code ahead of the arithmetic
      auto run = [&](auto K) {
        constexpr int KIND = decltype(K)::value;
        rows_fwd_reg64<MERGE>(
            A, g, T, [&](int z, int xp) { return fetch_prod<KIND, C>(pr, (long long)z * g.nx + 2 * xp); },
            [&](const ProdRaw<KIND, C> &r) { return eval_raw<KIND, C>(pr, r); }, g.kx);
      };
      // forms: the y fluxes (p = 2) are 2 or 4, the others 0, 1 or 3
      if constexpr (MERGE) {
        if (pr.kind == 2) run(std::integral_constant<int, 2>{});
        else run(std::integral_constant<int, 4>{});
      } else {
        if (pr.kind == 0) run(std::integral_constant<int, 0>{});
        else if (pr.kind == 1) run(std::integral_constant<int, 1>{});
        else run(std::integral_constant<int, 3>{});
      }
    };
    rows(pz, std::false_type{});
    plane_cols_fwd(A, g, T);
    accumulate(0);
    rows(1 - pz, std::false_type{});
    rows(2, std::true_type{});
    plane_cols_fwd(A, g, T);
    accumulate(2);
  } else {
    for (int p = 0; p < 3; ++p) {
      const Prod<C> pr = make_prod<C>(a, c, p, J);
      if (R32 && a.rows32) {
        rows_fwd_reg32(A, g, T, [&](int z, int xp) { return eval_prod(pr, (long long)z * g.nx + 2 * xp); });
      } else {
        for (int t0 = threadIdx.x >> 6; t0 < lt.ntiles; t0 += nw) {
          int z, xp;
          if (lt.at(g, t0, z, xp)) A[phys_slot(T, g, z, xp)] = eval_prod(pr, (long long)z * g.nx + 2 * xp);
        }
        lds_barrier();
        plane_rows_fwd(A, g, T);
      }
      plane_cols_fwd(A, g, T);
      accumulate(p == 2 ? 2 : (p == pz ? 0 : 1));
    }
  }
  // RK combination (channel.f:88-167, 581-595; scalars 254-270, 723-729),
  // new CF stored, CR scattered into LDS for the inverse transform.
  for (int e = threadIdx.x; e < g.nz * g.nkx1; e += nthr) {
    const int zr = (int)fdiv((unsigned)e, g.f_nkx1), kx = e - zr * g.nkx1;
    if (zr > g.nkz && zr < g.nz - g.nkz) A[inv_row(T, zr) * g.ls + kx] = mkc<C>(0, 0);
  }
  double2 *CF = (c < 3) ? a.cf[c] : a.cfth + (c - 3) * a.cfs;
  const double2 *CC = (c < 3) ? a.cu[c] : a.cth + (c - 3) * a.cfs;
  const long long base = (long long)J * g.pf;
  // explicit half of the Crank-Nicolson y-viscous term (channel.f:606-626,
  // 739-747), applied to the Fourier coefficients: the y-difference operator
  // commutes with the x-z transform, so adding it to CR before the inverse
  // transform gives R + T1 NU_V_SCALE d/dy(d/dy U) of the reference
  const double coefv = (c < 3) ? (a.nu * hb / 2.0) * a.nu_v_scale : ((a.nu * hb / 2.0) / a.pr[c - 3]) * a.nu_v_scale;
  double ey0, ey1, ey2;  // weights of C(J+1)-C(J), C(J)-C(J-1) and the outer divisor
  double rey0, rey1, rey2;
  if (c == 1) {
    ey0 = g.dyf[J];
    ey1 = g.dyf[J - 1];
    ey2 = g.dy[J];
    rey0 = g.rdyf[J];
    rey1 = g.rdyf[J - 1];
    rey2 = g.rdy[J];
  } else {
    ey0 = g.dy[J + 1];
    ey1 = g.dy[J];
    ey2 = g.dyf[J];
    rey0 = g.rdy[J + 1];
    rey1 = g.rdy[J];
    rey2 = g.rdyf[J];
  }
  // modes in batches of NB: the batch's loads (indices clamped to the plane)
  // are all issued before its arithmetic (smaller batches in the
  // 1024-thread variant, whose register budget is half as large)
  constexpr int NB = (REG && sizeof(C) == sizeof(double2)) ? 4 : (sizeof(C) == sizeof(double2) ? 2 : 1);
  const double dyJ = g.dy[J], rdyJ = g.rdy[J];
  static_assert(NQ % NB == 0, "mode batches must tile the modes of a thread");
#pragma unroll
  for (int q0 = 0; q0 < NQ; q0 += NB) {
    double2 cm1[NB], cp1[NB], cc0[NB], old[NB], p0[NB], pm[NB];
    double kk[NB];  // KX (U1) or KZ (U3) of the pressure gradient
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int mc = min((int)threadIdx.x + (q0 + b) * nthr, g.pf - 1);
      const long long o = base + mc;
      cm1[b] = CC[o - g.pf];
      cp1[b] = CC[o + g.pf];
      cc0[b] = CC[o];
      if (c == 0 || c == 2) {
        const int kzp = (int)fdiv((unsigned)mc, g.f_nkx1), kx = mc - kzp * g.nkx1;
        kk[b] = (c == 0) ? g.kx[kx] : g.kz[kzp];
      }
    }
    if (rk > 0) {
#pragma unroll
      for (int b = 0; b < NB; ++b) old[b] = CF[base + min((int)threadIdx.x + (q0 + b) * nthr, g.pf - 1)];
    }
    if (c < 3) {
#pragma unroll
      for (int b = 0; b < NB; ++b) p0[b] = a.cp[base + min((int)threadIdx.x + (q0 + b) * nthr, g.pf - 1)];
    }
    if (c == 1) {
#pragma unroll
      for (int b = 0; b < NB; ++b) pm[b] = a.cp[base - g.pf + min((int)threadIdx.x + (q0 + b) * nthr, g.pf - 1)];
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int m = threadIdx.x + (q0 + b) * nthr;
      if (m < g.pf) {
        const int kzp = (int)fdiv((unsigned)m, g.f_nkx1), kx = m - kzp * g.nkx1;
        const long long o = base + m;
        // the mixed form leaves out CU and takes the explicit y-viscous half
        // twice: CR - CU + T1 NU_V_SCALE d/dy(d/dy CU) is the right-hand side of
        // the increment U_new - U_old under the same implicit operator
        const double2 c0 = cc0[b];
        double2 cr = DELTA ? make_double2(0.0, 0.0) : c0;
        const double2 ev = make_double2(
            coefv * div_by_recip(div_by_recip(cp1[b].x - c0.x, ey0, rey0) - div_by_recip(c0.x - cm1[b].x, ey1, rey1),
                                 ey2, rey2),
            coefv * div_by_recip(div_by_recip(cp1[b].y - c0.y, ey0, rey0) - div_by_recip(c0.y - cm1[b].y, ey1, rey1),
                                 ey2, rey2));
        if (rk > 0) cr = make_double2(cr.x + T3 * old[b].x, cr.y + T3 * old[b].y);
        if (c == 0 || c == 2) {
          double2 ikp = cmul_ik(kk[b], p0[b]);
          cr = make_double2(cr.x - T4 * ikp.x, cr.y - T4 * ikp.y);
          if (c == 0 && kx == 0 && kzp == 0) {
            if (a.f_type == 1)
              cr.x = cr.x - T4 * a.px0;
            else if (a.f_type == 2)
              cr.x = cr.x - T4 * (a.px0 + a.amp_omega0 * cos(a.omega0 * time));
          }
        } else if (c == 1) {
          double2 cs = make_double2(div_by_recip(p0[b].x - pm[b].x, dyJ, rdyJ), div_by_recip(p0[b].y - pm[b].y, dyJ, rdyJ));
          cr = make_double2(cr.x - T4 * cs.x, cr.y - T4 * cs.y);
        }
        const double2 cfn = make_double2(acc[q0 + b].x, acc[q0 + b].y);
        CF[o] = cfn;
        cr = make_double2(cr.x + T5 * cfn.x, cr.y + T5 * cfn.y);
        // LES takes the explicit y-viscous half in physical space, from U with
        // les_chan's BCs (channel.f:606-626), in its column solve
        if constexpr (!SGS) cr = make_double2(cr.x + ev.x, cr.y + ev.y);
        if (DELTA) cr = make_double2(cr.x + ev.x, cr.y + ev.y);
        A[inv_row(T, kz_row(g, kzp)) * g.ls + kx] = cvt<C>(cr);
      }
    }
  }
  lds_barrier();
  plane_cols_inv(A, g, T);
  C *dst = reinterpret_cast<C *>(static_cast<Sc<C> *>(a.rphys) + (long long)c * a.fs + (long long)J * g.pp);
  if constexpr (REG) {
    rows_inv_reg64(A, g, T, [&](int z, int m, C v) { dst[z * 64 + m] = v; });
  } else if (R32 && a.rows32) {
    rows_inv_reg32(A, g, T, [&](int z, int m, C v) { dst[z * 32 + m] = v; });
  } else {
    plane_rows_inv(A, g, T);
    plane_store_phys(A, g, T, dst);
  }
}
// The nonlinear pass for planes too large for LDS (FP64, no LES): the same
// terms in the same order as nonlinear_pass's LDS variant, with the
// transforms staged through the LDS buffer `buf` (gp_rows_fwd and its
// kin), the plane between their passes in Q and the CF accumulators of the
// PF modes in ACC (device memory), modes taken m = threadIdx.x,
// + blockDim.x, ... in every loop.
__device__ void nonlinear_pass_gp(const StepArgs &a, int J, int c, double2 *Q, double2 *ACC, const Tables &T,
                                  double2 *buf, int stage) {
  const Geo &g = a.g;
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
  for (int m = threadIdx.x; m < g.pf; m += nthr) {
    const int kzp = (int)fdiv((unsigned)m, g.f_nkx1), kx = m - kzp * g.nkx1;
    ACC[m] = cf_init(a, c, J, kx, kzp, m);
  }
  // the three products, forward transformed and subtracted with their
  // derivative factors: c==0 takes (ikz, ikx, 1), all others (ikx, ikz, 1)
  const int pz = (c == 0) ? 0 : 1;
  const double nzd = g.dnz, rnz = g.rnz;
  __syncthreads();  // ACC initialised before the first product's modes are taken from it
#pragma unroll 1
  for (int p = 0; p < 3; ++p) {
    const Prod<double2> pr = make_prod<double2>(a, c, p, J);
    gp_rows_fwd<1>(Q, g, T, buf, stage, [&](int z, int xp) { return eval_prod(pr, (long long)z * g.nx + 2 * xp); });
    __syncthreads();  // Q complete before the column pass
    const int mode = (p == 2) ? 2 : (p == pz ? 0 : 1);
    gp_cols_fwd(Q, g, T, buf, stage, [&](int kx, int kzp, double2 s0) {
      const int m = kzp * g.nkx1 + kx;
      const double2 sv = make_double2(div_by_recip(s0.x, nzd, rnz), div_by_recip(s0.y, nzd, rnz));
      ACC[m] = csub(ACC[m], (mode == 2) ? sv : cmul_ik((mode == 0) ? g.kz[kzp] : g.kx[kx], sv));
    });
    __syncthreads();  // the next product is written over Q; ACC's modes change threads
  }
  // RK combination (channel.f:88-167, 581-595; scalars 254-270, 723-729):
  // new CF stored, CR written over ACC for the inverse transform
  double2 *CF = (c < 3) ? a.cf[c] : a.cfth + (c - 3) * a.cfs;
  const double2 *CC = (c < 3) ? a.cu[c] : a.cth + (c - 3) * a.cfs;
  const long long base = (long long)J * g.pf;
  // explicit half of the Crank-Nicolson y-viscous term (channel.f:606-626,
  // 739-747) in Fourier space, as in nonlinear_pass
  const double coefv = (c < 3) ? (a.nu * hb / 2.0) * a.nu_v_scale : ((a.nu * hb / 2.0) / a.pr[c - 3]) * a.nu_v_scale;
  double ey0, ey1, ey2, rey0, rey1, rey2;
  if (c == 1) {
    ey0 = g.dyf[J];
    ey1 = g.dyf[J - 1];
    ey2 = g.dy[J];
    rey0 = g.rdyf[J];
    rey1 = g.rdyf[J - 1];
    rey2 = g.rdy[J];
  } else {
    ey0 = g.dy[J + 1];
    ey1 = g.dy[J];
    ey2 = g.dyf[J];
    rey0 = g.rdy[J + 1];
    rey1 = g.rdy[J];
    rey2 = g.rdyf[J];
  }
  const double dyJ = g.dy[J], rdyJ = g.rdy[J];
  for (int m = threadIdx.x; m < g.pf; m += nthr) {
    const int kzp = (int)fdiv((unsigned)m, g.f_nkx1), kx = m - kzp * g.nkx1;
    const long long o = base + m;
    const double2 c0 = CC[o], cm1 = CC[o - g.pf], cp1 = CC[o + g.pf];
    double2 cr = c0;
    const double2 ev = make_double2(
        coefv * div_by_recip(div_by_recip(cp1.x - c0.x, ey0, rey0) - div_by_recip(c0.x - cm1.x, ey1, rey1), ey2, rey2),
        coefv * div_by_recip(div_by_recip(cp1.y - c0.y, ey0, rey0) - div_by_recip(c0.y - cm1.y, ey1, rey1), ey2, rey2));
    if (rk > 0) {
      const double2 old = CF[o];
      cr = make_double2(cr.x + T3 * old.x, cr.y + T3 * old.y);
    }
    if (c == 0 || c == 2) {
      const double2 ikp = cmul_ik((c == 0) ? g.kx[kx] : g.kz[kzp], a.cp[o]);
      cr = make_double2(cr.x - T4 * ikp.x, cr.y - T4 * ikp.y);
      if (c == 0 && kx == 0 && kzp == 0) {
        if (a.f_type == 1)
          cr.x = cr.x - T4 * a.px0;
        else if (a.f_type == 2)
          cr.x = cr.x - T4 * (a.px0 + a.amp_omega0 * cos(a.omega0 * time));
      }
    } else if (c == 1) {
      const double2 p0 = a.cp[o], pm = a.cp[o - g.pf];
      const double2 cs = make_double2(div_by_recip(p0.x - pm.x, dyJ, rdyJ), div_by_recip(p0.y - pm.y, dyJ, rdyJ));
      cr = make_double2(cr.x - T4 * cs.x, cr.y - T4 * cs.y);
    }
    const double2 cfn = ACC[m];
    CF[o] = cfn;
    cr = make_double2(cr.x + T5 * cfn.x, cr.y + T5 * cfn.y);
    cr = make_double2(cr.x + ev.x, cr.y + ev.y);
    ACC[m] = cr;
  }
  __syncthreads();
  gp_cols_inv(Q, g, T, buf, stage, [&](int kx, int zr) {
    int kzp = -1;
    if (zr <= g.nkz)
      kzp = zr;
    else if (zr >= g.nz - g.nkz)
      kzp = zr - (g.nz - 1 - g.nkz) + g.nkz;
    return kzp >= 0 ? ACC[kzp * g.nkx1 + kx] : make_double2(0.0, 0.0);
  });
  __syncthreads();
  double2 *dst = reinterpret_cast<double2 *>(static_cast<double *>(a.rphys) + (long long)c * a.fs + (long long)J * g.pp);
  gp_rows_inv(Q, g, T, buf, stage, [&](int z, int m, double2 v) { dst[z * g.lw + m] = v; });
}

// Persistent over the (plane, component) items of a launch: block b on
// its own scratch slot (Q, then the accumulators) at gp + b gp_slot.
__global__ void __launch_bounds__(kGpThreads, 4) k_nonlinear_gp(StepArgs a, double2 *gp, long long gp_slot, int ncomp,
                                                                int nitems) {
  extern __shared__ double2 smem[];
  const Tables T = load_tables_at(smem, a.g);
  double2 *buf = smem + gp_table_elems(a.g);
  const int stage = gp_stage_elems(a.g);
  double2 *Q = gp + (long long)blockIdx.x * gp_slot;
  double2 *ACC = Q + (long long)a.g.nz * a.g.nkx1;
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const int jr = it / ncomp;
    nonlinear_pass_gp(a, a.j0 + jr, a.c0 + (it - jr * ncomp), Q, ACC, T, buf, stage);
    __syncthreads();  // the next item writes over Q and ACC
  }
}

// LB: the block-size bound.  The LDS variant takes any block size up to
// 1024; its instantiation for blocks of at most 512 threads has twice the
// register budget (256 VGPRs) and no spills, and takes the NX = 64 rows in
// registers.
template <bool REG, int NQ, bool SGS = false, int LB = (REG ? 512 : 1024)>
__global__ void __launch_bounds__(LB) k_nonlinear(StepArgs a) {
  extern __shared__ double2 smem[];
  nonlinear_pass<REG, NQ, double2, SGS, !REG && LB <= 512>(a, smem);
}

// The mixed form: the FP32 plane takes half the LDS and half the registers
// of a complex value, so two 512-thread blocks share a CU (4 waves per SIMD).
template <bool REG, int NQ>
__global__ void __launch_bounds__(REG ? 512 : 1024) __attribute__((amdgpu_waves_per_eu(4)))
k_nonlinear32(StepArgs a) {
  extern __shared__ double2 smem[];
  nonlinear_pass<REG, NQ, float2>(a, reinterpret_cast<float2 *>(smem));
}
// batches of contexts of one shape: argument structs in device memory,
// member blockIdx.z
template <bool REG, int NQ, int LB = (REG ? 512 : 1024)>
__global__ void __launch_bounds__(LB) k_nonlinear_b(const StepArgs *__restrict__ av) {
  extern __shared__ double2 smem[];
  nonlinear_pass<REG, NQ, double2, false, !REG && LB <= 512>(av[blockIdx.z], smem);
}
template <bool REG, int NQ>
__global__ void __launch_bounds__(REG ? 512 : 1024) __attribute__((amdgpu_waves_per_eu(4)))
k_nonlinear32_b(const StepArgs *__restrict__ av) {
  extern __shared__ double2 smem[];
  nonlinear_pass<REG, NQ, float2>(av[blockIdx.z], reinterpret_cast<float2 *>(smem));
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
  int seg_rows;  // rows per register segment of the solve
};

// Row J (0..NY+1) of the system of field f: lower L, diagonal D, upper U,
// constant right-hand side V of a boundary row (isbc=1), exactly as the
// reference fills MATL/MATD/MATU/VEC: initialise (0,1,0,0), interior rows,
// then the BC rows overwrite.
__host__ __device__ inline void real_row(const FactorIn &in, int f, double temp1, int J, double &L, double &D,
                                         double &U, double &V, int &isbc) {
#pragma clang fp contract(off)
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

// Tables of the shared matrix of (rk, field), one block each: rows from
// real_row, the forward elimination of THOMAS_REAL_MPI (mpi.f:630-703) in
// its operation order (multipliers TL, pivots TB, upper diagonal TC, 1/TB),
// and the carry factors of the segmented sweeps of k_solve_seg:
//   PF_J = prod_{i=a..J} TL_i          (a: first row of J's segment)
//   PB_J = prod_{i=J..b-1} -TC_i/TB_i  (b-1: last row of J's segment)
// The constant right-hand sides of the boundary rows and the rows'
// coefficients (dt independent) are written by the rk=0 blocks.
__device__ __forceinline__ void cn_tables_body(const Clock *clk, const FactorIn &in, const RealFactors &fac, int blk) {
#pragma clang fp contract(off)
  extern __shared__ double sm[];
  const int nrow = in.ny + 2;
  double *a = sm, *b = sm + nrow, *c = sm + 2 * nrow;
  const int rk = blk / in.nfields, f = blk - rk * in.nfields;
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
      fac.bcl[(long long)f * nrow + i] = bc ? L : 0.0;
      fac.bcd[(long long)f * nrow + i] = bc ? D : 0.0;
      fac.bcu[(long long)f * nrow + i] = bc ? U : 0.0;
    }
  }
  __syncthreads();
  // the forward elimination is a recurrence: one thread, operands in LDS.
  // Each row's operands are read one row ahead, so that their LDS latency
  // overlaps the previous row's division instead of adding to it.
  double *tl = sm + 3 * nrow, *tb = sm + 4 * nrow, *tr = sm + 5 * nrow, *pf = sm + 6 * nrow, *pb = sm + 7 * nrow;
  const int sr = in.seg_rows;
  if (threadIdx.x == 0) {
    double bp = b[0], pfv = 0.0;
    tl[0] = 0.0;
    tb[0] = bp;
    pf[0] = 0.0;
    double an = a[1], bn = b[1], cn = c[0];
    int js = (sr == 1) ? 0 : 1;  // J % sr
    for (int J = 1; J < nrow; ++J) {
      const double aj = an, bj = bn, cj = cn;
      if (J + 1 < nrow) {
        an = a[J + 1];
        bn = b[J + 1];
        cn = c[J];
      }
      const double m = -aj / bp;
      bp = bj + m * cj;
      tl[J] = m;
      tb[J] = bp;
      pfv = (js == 0) ? m : pfv * m;
      pf[J] = pfv;
      js = (js + 1 == sr) ? 0 : js + 1;
    }
  }
  __syncthreads();
  // the backward factors' quotients do not depend on one another: every
  // thread takes some, and one thread chains the products
  for (int J = threadIdx.x; J < nrow; J += blockDim.x) tr[J] = (tb[J] != 0.0) ? -c[J] / tb[J] : 0.0;
  __syncthreads();
  if (threadIdx.x == 0) {
    double pbv = 0.0;
    int js = (nrow - 1) % sr;  // J % sr, counted down with J
    for (int J = nrow - 1; J >= 0; --J) {
      const double fj = tr[J];
      pbv = (js == sr - 1 || J == nrow - 1) ? fj : fj * pbv;
      pb[J] = pbv;
      js = (js == 0) ? sr - 1 : js - 1;
    }
  }
  __syncthreads();
  const long long o = ((long long)rk * in.nfields + f) * nrow;
  for (int i = threadIdx.x; i < nrow; i += blockDim.x) {
    fac.tl[o + i] = tl[i];
    fac.tb[o + i] = tb[i];
    fac.tc[o + i] = c[i];
    fac.tr[o + i] = 1.0 / tb[i];
    fac.pf[o + i] = pf[i];
    fac.pb[o + i] = pb[i];
  }
}
struct CnArgs {
  const Clock *clk;
  FactorIn in;
  RealFactors fac;
};
__global__ void __launch_bounds__(256) k_cn_tables(CnArgs a) { cn_tables_body(a.clk, a.in, a.fac, blockIdx.x); }
__global__ void __launch_bounds__(256) k_cn_tables_b(const CnArgs *__restrict__ av) {
  const CnArgs &a = av[blockIdx.z];
  cn_tables_body(a.clk, a.in, a.fac, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Real Crank-Nicolson solves (channel.f:771-1029; the explicit half of the
// y-viscous term is already in R, added in Fourier space by k_nonlinear) and
// the CFL minimum of COURANT (courant.f:34-43).  THOMAS_REAL_MPI's sweeps
// (mpi.f:630-703, 767-809) run on columns split into NS segments of
// seg_rows rows, each held in the registers of one lane (lane = segment *
// 64/NS + column):
//   forward:  each segment sweeps from a zero carry, h_J = d_J + TL_J h_{J-1};
//             the carries g_{a-1} pass up the segments by shuffles, and
//             g_J = h_J + PF_J g_{a-1};
//   backward: each segment sweeps from a zero carry, y_J = (g_J - TC_J
//             y_{J+1}) / TB_J (quotient from 1/TB_J, the same double as the
//             division); the carries x_b pass down, and x_J = y_J + PB_J x_b.
// The first segment's forward values and the last segment's backward values
// are the reference's own; the others differ from a single sweep by the
// rounding of the carry terms.
// ---------------------------------------------------------------------------
struct SolveArgs {
  Geo g;
  Clock *clk;
  int rk;
  void *dst[kMaxFields];  // solution of field f (may be the field's own R: each lane reads its rows first)
  const void *rphys;       // physical R and U: double, or float in the mixed form (FP64 arithmetic either way)
  void *uold;              // mixed form: the fluctuations at the start of the substep (float)
  const double2 *four[kMaxFields];  // mixed form: Fourier fields, whose (0,0) coefficients are the plane means
  int f0;                   // field of blockIdx.y = 0
  long long fs;
  RealFactors fac;
  int nfields;
  int wlo[kMaxFields], whi[kMaxFields];
  int courant;  // 1: accumulate the CFL minimum of U1..U3
  double cfl, dx, dz;
};

constexpr int kSegThreads = 256;

template <int NS, class P>
__device__ __forceinline__ void solve_seg_body(const SolveArgs &a) {
#pragma clang fp contract(off)
  constexpr int CW = 64 / NS;  // columns per wave
  constexpr int SR = kSegRows;
  // tables in LDS, padded to NS * seg_rows rows
  __shared__ double TL[NS * SR], TB[NS * SR], TC[NS * SR], TR[NS * SR], PF[NS * SR], PB[NS * SR], VC[NS * SR];
  __shared__ int IB[NS * SR];
  __shared__ double red[kSegThreads / 64];
  const Geo &g = a.g;
  const int f = a.f0 + blockIdx.y;
  const long long pp = g.pp;
  const int NY = g.ny, nrow = NY + 2;
  const int sr = a.fac.seg_rows;
  const long long to = ((long long)a.rk * a.nfields + f) * nrow;
  for (int i = threadIdx.x; i < NS * sr; i += blockDim.x) {
    const bool in = i < nrow;
    TL[i] = in ? a.fac.tl[to + i] : 0.0;
    TB[i] = in ? a.fac.tb[to + i] : 1.0;
    TC[i] = in ? a.fac.tc[to + i] : 0.0;
    TR[i] = in ? a.fac.tr[to + i] : 1.0;
    PF[i] = in ? a.fac.pf[to + i] : 0.0;
    PB[i] = in ? a.fac.pb[to + i] : 0.0;
    VC[i] = in ? a.fac.vconst[(long long)f * nrow + i] : 0.0;
    IB[i] = in ? a.fac.isbc[(long long)f * nrow + i] : 1;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int seg = lane / CW, cl = lane - seg * CW;
  const long long col = ((long long)blockIdx.x * (blockDim.x >> 6) + w) * CW + cl;
  const bool live = col < pp;
  const int r0 = seg * sr;
  const int nr = min(sr, max(nrow - r0, 0));
  // rows are addressed from the field's (uniform) base with 32-bit element
  // offsets i pp + col (a field is below 2^31 elements)
  P *const Ub = static_cast<P *>(a.dst[f]);
  const P *const Rb = static_cast<const P *>(a.rphys) + (long long)f * a.fs;
  const unsigned ppu = (unsigned)pp, colu = (unsigned)col;
  auto ro_ = [&](int i) { return (unsigned)i * ppu + colu; };
  // the mixed form solves for the increment d = U_new - U_old: every row's
  // right-hand side is in R (interior rows from the nonlinear pass,
  // boundary rows from k_bc_increment)
  constexpr bool DELTA = sizeof(P) == sizeof(float);
  P *const Uob = DELTA ? static_cast<P *>(a.uold) + (long long)f * a.fs : nullptr;
  // right-hand sides: the boundary constant or R
  double v[SR];
#pragma unroll
  for (int q = 0; q < SR; ++q) {
    const int i = r0 + q;
    v[q] = (live && q < nr && (DELTA || !IB[i])) ? Rb[ro_(i)] : 0.0;
    // FP32 rows are converted in chunks, so that the loads in flight and
    // their FP64 values do not both hold all rows' registers
    if (DELTA && q % 12 == 11) __builtin_amdgcn_sched_barrier(0);
  }
  if constexpr (!DELTA) {
#pragma unroll
    for (int q = 0; q < SR; ++q) {
      const int i = r0 + q;
      if (q < nr && IB[i]) v[q] = VC[i];
    }
  }
  // forward sweep within the segment from a zero carry
  double hp = 0.0, hend = 0.0;
#pragma unroll
  for (int q = 0; q < SR; ++q) {
    const int i = r0 + q;
    const double h = v[q] + TL[i] * hp;
    v[q] = h;
    hp = h;
    hend = (q == nr - 1) ? h : hend;
  }
  // carries up the segments: g_in(s) = g_end(s-1), g_end = h_end + PF_end g_in
  const double pfe = (nr > 0) ? PF[r0 + nr - 1] : 0.0;
  double gin = 0.0, gend = hend;
#pragma unroll
  for (int t = 1; t < NS; ++t) {
    const double prev = __shfl_up(gend, CW);
    if (seg == t) {
      gin = prev;
      gend = hend + pfe * gin;
    }
  }
#pragma unroll
  for (int q = 0; q < SR; ++q) v[q] = (seg == 0) ? v[q] : v[q] + PF[r0 + q] * gin;
  // backward sweep within the segment from a zero carry
  double yn = 0.0, ystart = 0.0;
#pragma unroll
  for (int q = SR - 1; q >= 0; --q) {
    const int i = r0 + q;
    const double b = TB[i];
    const double y = (b != 0.0) ? div_by_recip(v[q] - TC[i] * yn, b, TR[i]) : 0.0;
    const bool in = q < nr;
    v[q] = y;
    yn = in ? y : 0.0;
  }
  ystart = v[0];
  // carries down the segments: x_in(s) = x_start(s+1), x_start = y_start + PB_start x_in
  const double pbs = PB[r0];
  double xin = 0.0, xstart = ystart;
#pragma unroll
  for (int t = NS - 2; t >= 0; --t) {
    const double nxt = __shfl_down(xstart, CW);
    if (seg == t) {
      xin = nxt;
      xstart = ystart + pbs * xin;
    }
  }
  // write-back rows and the CFL minimum
  const bool cour = a.courant && f < 3;
  const double dxf = (f == 0) ? a.dx : a.dz;
  double dtmin = __longlong_as_double(0x7FF0000000000000ll);
  const int wlo = a.wlo[f], whi = a.whi[f];
#pragma unroll
  for (int q = 0; q < SR; ++q) {
    const int i = r0 + q;
    const double x = (seg == NS - 1) ? v[q] : v[q] + PB[i] * xin;
    if (live && q < nr && i >= wlo && i <= whi) {
      Ub[ro_(i)] = x;
      if (!DELTA && cour && i >= 1 && i <= NY) {
        const double dl = (f == 1) ? g.dy[i] : dxf;
        dtmin = fmin(dtmin, a.cfl * dl / fabs(x));
      }
    }
  }
  // the mixed form: the new values U_old + d, for COURANT and for the
  // physical field of a scalar (which persists); in a loop of its own that
  // reads d back from U (this lane's own stores), so that the rows' values
  // need not stay in registers
  if constexpr (DELTA) {
    if (cour || f >= 3) {
#pragma unroll 6
      for (int q = 0; q < SR; ++q) {
        const int i = r0 + q;
        if (live && q < nr && i >= wlo && i <= whi) {
          const double x = (double)Ub[ro_(i)];
          const double un = (a.four[f][(long long)i * g.pf].x + (double)Uob[ro_(i)]) + x;
          if (f >= 3) Uob[ro_(i)] = (P)un;
          if (cour && i >= 1 && i <= NY) {
            const double dl = (f == 1) ? g.dy[i] : dxf;
            dtmin = fmin(dtmin, a.cfl * dl / fabs(un));
          }
        }
      }
    }
  }
  if (cour) {
    for (int o = 32; o > 0; o >>= 1) dtmin = fmin(dtmin, __shfl_down(dtmin, o));
    if (lane == 0) red[w] = dtmin;
    __syncthreads();
    if (threadIdx.x == 0) {
      double m = red[0];
      for (int q = 1; q < (int)(blockDim.x >> 6); ++q) m = fmin(m, red[q]);
      if (m >= 0.0) atomicMin(&a.clk->cfl_min_bits, (unsigned long long)__double_as_longlong(m));
    }
  }
}
template <int NS, class P>
__global__ void __launch_bounds__(kSegThreads) k_solve_seg(SolveArgs a) {
  solve_seg_body<NS, P>(a);
}
template <int NS, class P>
__global__ void __launch_bounds__(kSegThreads) k_solve_seg_b(const SolveArgs *__restrict__ av) {
  solve_seg_body<NS, P>(av[blockIdx.z]);
}

// COURANT (courant.f:45-62) from the CFL minimum of the solve kernel; the
// solve tables for the new DELTA_T are rebuilt right after by k_cn_tables.
__device__ __forceinline__ void courant_dt_body(Clock *clk, double dt_cap, int variable_dt, int update_dt) {
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
struct CourantArgs {
  Clock *clk;
  double dt_cap;
  int variable_dt, update_dt;
};
__global__ void k_courant_dt(CourantArgs a) { courant_dt_body(a.clk, a.dt_cap, a.variable_dt, a.update_dt); }
__global__ void k_courant_dt_b(const CourantArgs *__restrict__ av) {
  const CourantArgs &a = av[blockIdx.z];
  courant_dt_body(a.clk, a.dt_cap, a.variable_dt, a.update_dt);
}

// ---------------------------------------------------------------------------
// REM_DIV_CHAN (channel.f:1103-1214): divergence, pressure-Poisson solve per
// (kx,kz) mode, projection; plus the CP update of channel.f:1059-1065.
// ---------------------------------------------------------------------------
struct RemDivArgs {
  Geo g;
  Clock *clk;
  int rk;       // <0: no CP update
  int cp_old;   // 1: COURANT ran in this substep, divide by the saved H_BAR
  const double2 *cin[3];  // velocity before the projection (may equal cu)
  double2 *cu[3];         // projected velocity
  double2 *cp;
  double2 *phi;
  const double *rl, *ru;  // 1/(DY(J)DYF(J)), 1/(DY(J+1)DYF(J))
  double *ta, *tb, *tr;   // sweep multipliers, pivots, 1/pivots [J][mode]
  int seg;                // rows per sweep segment; 0: one sequential sweep
  // the step boundary after a step's last substep (diablo.f:66-110): TIME
  // += DELTA_T (bnd_end), TIME_STEP += 1 (bnd_begin), by one thread; no
  // thread of the launch reads TIME or TIME_STEP
  int bnd_end, bnd_begin;
};

// Pivots of the pressure-Poisson systems (channel.f:1112-1165 with the BC
// rows of APPLY_BC_REM_DIV_MPI, mpi.f:1239-1278): the matrix of mode (kx,kz)
// depends only on the grid and k^2, so the multipliers and pivots of the
// forward sweep of THOMAS_FORWARD_COMPLEX_MPI (mpi.f:708-764) are computed
// once per grid, in the same operation order.  ta/tb: [J][mode], J = 2..NY.
__global__ void __launch_bounds__(256) k_rem_div_factors(RemDivArgs a) {
#pragma clang fp contract(off)
  const Geo &g = a.g;
  const int NY = g.ny;
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= g.pf) return;
  const long long pf = g.pf;
  const int kzp = (int)fdiv((unsigned)m, g.f_nkx1), kx = m - kzp * g.nkx1;
  const double kx2 = g.kx2[kx], kz2 = g.kz2[kzp];
  const bool mean = (kx == 0 && kzp == 0);
  // row 1: D=1, U=0 (mean mode) or D=1, U=-1
  double bprev = 1.0, uprev = mean ? 0.0 : -1.0;
  for (int J = 2; J <= NY; ++J) {
    double L, D, U;
    if (J < NY) {
      L = a.rl[J];
      U = a.ru[J];
      D = -kx2 - kz2 - U - L;
    } else {  // row NY: L=1, D=-1, U=0
      L = 1.;
      D = -1.;
      U = 0.;
    }
    const double A = (bprev != 0.0) ? -L / bprev : 0.0;
    const double B = D + A * uprev;
    a.ta[pf * J + m] = A;
    a.tb[pf * J + m] = B;
    a.tr[pf * J + m] = 1.0 / B;
    bprev = B;
    uprev = U;
  }
}

// One pressure-Poisson solve per (kx,kz) mode.  All threads stage the
// divergence (channel.f:1137-1143) and the mode's multipliers and pivots of
// kRdtModes modes into LDS; one lane per mode runs the forward and backward
// sweeps (mpi.f:736-747, 838-856); then all threads project U and update CP
// (channel.f:1185-1210, 1059-1065).  The sweeps read their operands
// kRdtAhead rows ahead without bounds tests (kRdtPad rows of padding below
// and above rows 0..NY), so nothing in the recurrence waits on LDS.
constexpr int kRdtModes = 8;

// The sweeps of one mode split into kRdsSeg segments of `seg` rows
// [2 + s seg, 1 + (s+1) seg] (capped at NY), segment s on lane s*kRdtModes
// + mode of the first wave.  Each segment runs the recurrence from a zero
// carry in registers; the carries G_{r0-1} (up) and X_{r1+1} (down) then pass
// between the segments by shuffles and are added to every row:
//   G_J = h_J + PF_J G_{r0-1},  PF_J = A_{r0} ... A_J;
//   X_J = y_J + PB_J X_{r1+1},  PB_J = (-U_J/B_J) ... (-U_{r1}/B_{r1}),
// PB = 0 at and below a zero pivot (the sweep sets X = 0 there), U_NY = 0.
// The factors are products of the tabulated multipliers and of quotients
// formed from the tabulated reciprocal pivots.  The top segment takes no
// carry down, so its rows equal the sequential sweep's; the others differ
// from it by rounding.  Row 1 is rewritten last.
__device__ __forceinline__ void rem_div_sweep_seg(int seg, int s, int lane, int NY, bool mean, double2 *Gs,
                                                  const double *As, const double *Bs, const double *Rs,
                                                  const double *ru) {
#pragma clang fp contract(off)
  const int r0 = 2 + s * seg, r1 = min(NY, r0 + seg - 1);
  const bool any = r0 <= r1;
  double2 h[kRdsRows];
  double2 hp = make_double2(0.0, 0.0);
  double pfe = 1.0;
#pragma unroll
  for (int q = 0; q < kRdsRows; ++q) {
    const int J = r0 + q;
    if (J <= r1) {
      const double2 v = Gs[J * kRdtModes + lane];
      const double A = As[J * kRdtModes + lane];
      hp = make_double2(v.x + A * hp.x, v.y + A * hp.y);
      pfe = (q == 0) ? A : pfe * A;
    }
    h[q] = hp;
  }
  const double2 hend = any ? hp : make_double2(0.0, 0.0);
  double2 gin = make_double2(0.0, 0.0), gend = hend;  // G_1 = 0 below segment 0
#pragma unroll
  for (int t = 1; t < kRdsSeg; ++t) {
    const double px = __shfl_up(gend.x, kRdtModes), py = __shfl_up(gend.y, kRdtModes);
    if (s == t) {
      gin = make_double2(px, py);
      gend = make_double2(hend.x + pfe * gin.x, hend.y + pfe * gin.y);
    }
  }
  double pf = 1.0;
#pragma unroll
  for (int q = 0; q < kRdsRows; ++q) {
    const int J = r0 + q;
    if (J <= r1) {
      const double A = As[J * kRdtModes + lane];
      pf = (q == 0) ? A : pf * A;
      h[q] = make_double2(h[q].x + pf * gin.x, h[q].y + pf * gin.y);
    }
  }
  // backward from a zero carry: X_J = (G_J - U_J X_{J+1}) / B_J, X = 0 where
  // B = 0; at J = NY the carry term U X_{NY+1} is 0 and X_NY = G_NY / B_NY
  auto pb_step = [&](int J, double B, double R, double U, double pb) {
    const double qv = (B != 0.0) ? div_by_recip(-(J < NY ? U : 0.0), B, R) : 0.0;
    return (J == r1) ? qv : ((B != 0.0) ? qv * pb : 0.0);
  };
  double2 yn = make_double2(0.0, 0.0);
  double pbs = 1.0;
#pragma unroll
  for (int q = kRdsRows - 1; q >= 0; --q) {
    const int J = r0 + q;
    if (J <= r1) {
      const double B = Bs[J * kRdtModes + lane], R = Rs[J * kRdtModes + lane], U = ru[J];
      const double xr = div_by_recip(h[q].x - U * yn.x, B, R);
      const double xi = div_by_recip(h[q].y - U * yn.y, B, R);
      yn.x = (B != 0.0) ? xr : 0.0;
      yn.y = (B != 0.0) ? xi : 0.0;
      h[q] = yn;
      pbs = pb_step(J, B, R, U, pbs);
    }
  }
  const double2 ystart = any ? h[0] : make_double2(0.0, 0.0);
  double2 xin = make_double2(0.0, 0.0), xstart = ystart;  // nothing above row NY
#pragma unroll
  for (int t = kRdsSeg - 2; t >= 0; --t) {
    const double nx = __shfl_down(xstart.x, kRdtModes), ny = __shfl_down(xstart.y, kRdtModes);
    if (s == t) {
      xin = make_double2(nx, ny);
      xstart = make_double2(ystart.x + pbs * xin.x, ystart.y + pbs * xin.y);
    }
  }
  const double U1 = mean ? 0.0 : -1.0;
  double pb = 1.0;
#pragma unroll
  for (int q = kRdsRows - 1; q >= 0; --q) {
    const int J = r0 + q;
    if (J <= r1) {
      pb = pb_step(J, Bs[J * kRdtModes + lane], Rs[J * kRdtModes + lane], ru[J], pb);
      const double2 x = make_double2(h[q].x + pb * xin.x, h[q].y + pb * xin.y);
      Gs[J * kRdtModes + lane] = x;
      // row 1: B=1, G=0, U = 0 (mean mode) or -1
      if (J == 2) Gs[1 * kRdtModes + lane] = make_double2((0.0 - U1 * x.x) / 1.0, (0.0 - U1 * x.y) / 1.0);
    }
  }
}

constexpr int kRdtAhead = 4;
constexpr int kRdtPad = 16;  // >= 4 * kRdtAhead - 1: the sweeps read up to 4 kRdtAhead - 1 rows past their range

// Rows of the velocity kept in registers between the divergence and the
// projection: J = 1 + rg + b ngrp for b < kRdtRegRows (all rows while NY <=
// kRdtRegRows ngrp); rows beyond are read again for the projection.
constexpr int kRdtRegRows = 5;

// nb blocks of modes, this one lin
__device__ __forceinline__ void rem_div_thomas_body(const RemDivArgs &a, int nb, int lin) {
#pragma clang fp contract(off)
  extern __shared__ double st[];
  if (lin == 0 && threadIdx.x == 0) {
    if (a.bnd_end) a.clk->time = a.clk->time + a.clk->dt;
    if (a.bnd_begin) a.clk->time_step += 1;
  }
  const Geo &g = a.g;
  const int NY = g.ny, ntot = NY + 1 + 2 * kRdtPad;
  // LDS, row J at index J + kRdtPad: G [row][mode] complex, A, B and 1/B
  // [row][mode], U [row]
  double2 *Gs = reinterpret_cast<double2 *>(st) + kRdtPad * kRdtModes;
  double *As = st + 2 * ntot * kRdtModes + kRdtPad * kRdtModes;
  double *Bs = As + ntot * kRdtModes;
  double *Rs = Bs + ntot * kRdtModes;
  double *ru = st + 5 * ntot * kRdtModes + kRdtPad;
  const int lane = threadIdx.x % kRdtModes, rg = threadIdx.x / kRdtModes;
  const int ngrp = blockDim.x / kRdtModes;
  // mode groups dealt so that XCD x (blocks x, x+8, ...) takes a contiguous
  // range: neighbouring groups share the cache lines their rows straddle
  int grp;
  {
    const int x = lin & 7, q = nb >> 3, r = nb & 7;
    grp = x * q + min(x, r) + (lin >> 3);
  }
  const int m = grp * kRdtModes + lane;
  const bool live = m < g.pf;
  const long long pf = g.pf;
  int kx = 0, kzp = 0;
  if (live) {
    kzp = (int)fdiv((unsigned)m, g.f_nkx1);
    kx = m - kzp * g.nkx1;
  }
  const double kxv = g.kx[kx], kzv = g.kz[kzp];
  const bool mean = (kx == 0 && kzp == 0);
  for (int j = (int)threadIdx.x - kRdtPad; j <= NY + kRdtPad; j += blockDim.x)
    ru[j] = (j >= 0 && j <= NY) ? a.ru[j] : 0.0;
  // Each thread's rows J = 1 + rg + b ngrp: the velocity (kept for the
  // projection), the pressure, CU2 at J+1 and the sweep tables, all loads
  // issued before any arithmetic.  Right-hand sides of rows 2..NY-1 (rows 1,
  // NY are BC rows with zero right-hand side) and the tables of rows 2..NY
  // go to LDS.
  const bool upd = a.rk >= 0;
  double2 u1[kRdtRegRows], u2[kRdtRegRows], u3[kRdtRegRows], pc[kRdtRegRows];
  auto stage_row = [&](int J, double2 c1, double2 c2, double2 c3, double2 c2p, double ta, double tb, double tr) {
    double2 V = make_double2(0.0, 0.0);
    if (live && J < NY) {
      const double2 t1 = cmul_ik(kxv, c1), t3 = cmul_ik(kzv, c3);
      const double d = g.dyf[J], rd = g.rdyf[J];
      V = make_double2(t1.x + div_by_recip(c2p.x - c2.x, d, rd) + t3.x, t1.y + div_by_recip(c2p.y - c2.y, d, rd) + t3.y);
    }
    Gs[J * kRdtModes + lane] = V;
    As[J * kRdtModes + lane] = live ? ta : 0.0;
    Bs[J * kRdtModes + lane] = live ? tb : 1.0;
    Rs[J * kRdtModes + lane] = live ? tr : 1.0;
  };
  {
    double2 c2p[kRdtRegRows];
    double ta[kRdtRegRows], tb[kRdtRegRows], tr[kRdtRegRows];
#pragma unroll
    for (int b = 0; b < kRdtRegRows; ++b) {
      const int J = 1 + rg + b * ngrp;
      if (live && J <= NY) {
        u1[b] = a.cin[0][pf * J + m];
        u3[b] = a.cin[2][pf * J + m];
        u2[b] = a.cin[1][pf * J + m];
        if (J >= 2) {
          ta[b] = a.ta[pf * J + m];
          tb[b] = a.tb[pf * J + m];
          tr[b] = a.tr[pf * J + m];
          if (upd) pc[b] = a.cp[pf * J + m];
          if (J < NY) c2p[b] = a.cin[1][pf * (J + 1) + m];
        }
      }
    }
#pragma unroll
    for (int b = 0; b < kRdtRegRows; ++b) {
      const int J = 1 + rg + b * ngrp;
      if (J >= 2 && J <= NY) stage_row(J, u1[b], u2[b], u3[b], c2p[b], ta[b], tb[b], tr[b]);
    }
  }
  for (int J = 1 + rg + kRdtRegRows * ngrp; J <= NY; J += ngrp) {  // rows beyond the registers
    double2 c1 = make_double2(0.0, 0.0), c2 = c1, c3 = c1, c2p = c1;
    double ta = 0.0, tb = 1.0, tr = 1.0;
    if (live) {
      c1 = a.cin[0][pf * J + m];
      c3 = a.cin[2][pf * J + m];
      c2 = a.cin[1][pf * J + m];
      ta = a.ta[pf * J + m];
      tb = a.tb[pf * J + m];
      tr = a.tr[pf * J + m];
      if (J < NY) c2p = a.cin[1][pf * (J + 1) + m];
    }
    stage_row(J, c1, c2, c3, c2p, ta, tb, tr);
  }
  lds_barrier();
  if (a.seg > 0) {
    if (rg < kRdsSeg) rem_div_sweep_seg(a.seg, rg, lane, NY, mean, Gs, As, Bs, Rs, ru);
  } else if (rg == 0) {
    // forward: G_J = V_J + A_J G_{J-1}, J = 2..NY, G_1 = 0.  Two register
    // sets of kRdtAhead rows alternate: one set's rows are computed while
    // the other's LDS reads are in flight (a rotation of one set would make
    // every read wait before the first row that follows it).
    double2 gp = make_double2(0.0, 0.0), gny = gp;
    double2 va[kRdtAhead], vb[kRdtAhead];
    double aa[kRdtAhead], ab[kRdtAhead];
    // the empty asm after each read group keeps the compiler from sinking the
    // reads to their first use in the next group
    auto fwd_load = [&](double2(&v)[kRdtAhead], double(&ac)[kRdtAhead], int J0) {
#pragma unroll
      for (int q = 0; q < kRdtAhead; ++q) {
        v[q] = Gs[(J0 + q) * kRdtModes + lane];
        ac[q] = As[(J0 + q) * kRdtModes + lane];
      }
      asm volatile("" ::: "memory");
    };
    auto fwd_rows = [&](const double2(&v)[kRdtAhead], const double(&ac)[kRdtAhead], int J0) {
#pragma unroll
      for (int q = 0; q < kRdtAhead; ++q) {
        const int J = J0 + q;
        gp = make_double2(v[q].x + ac[q] * gp.x, v[q].y + ac[q] * gp.y);
        Gs[J * kRdtModes + lane] = gp;  // rows above NY land in the padding
        // component-wise selects: a select of double2 values goes through scratch
        gny.x = (J == NY) ? gp.x : gny.x;
        gny.y = (J == NY) ? gp.y : gny.y;
      }
    };
    fwd_load(va, aa, 2);
    fwd_load(vb, ab, 2 + kRdtAhead);
    for (int J0 = 2; J0 <= NY; J0 += 2 * kRdtAhead) {
      fwd_rows(va, aa, J0);
      fwd_load(va, aa, J0 + 2 * kRdtAhead);
      fwd_rows(vb, ab, J0 + kRdtAhead);
      fwd_load(vb, ab, J0 + 3 * kRdtAhead);
    }
    // backward: X_NY = G_NY / B_NY; X_J = (G_J - U_J X_{J+1}) / B_J,
    // J = NY-1..2; row 1: B=1, G=0, U = 0 (mean mode) or -1.  The rows the
    // forward sweep wrote are complete before these reads (in-order LDS of
    // one lane).
    const double bny = Bs[NY * kRdtModes + lane];
    double2 x;
    x.x = (bny != 0.0) ? gny.x / bny : 0.0;
    x.y = (bny != 0.0) ? gny.y / bny : 0.0;
    Gs[NY * kRdtModes + lane] = x;
    double2 x2 = x;
    double2 wa[kRdtAhead], wb[kRdtAhead];
    double ba[kRdtAhead], bb[kRdtAhead], ra[kRdtAhead], rb[kRdtAhead], ua[kRdtAhead], ub[kRdtAhead];
    auto bwd_load = [&](double2(&w)[kRdtAhead], double(&b)[kRdtAhead], double(&r)[kRdtAhead],
                        double(&u)[kRdtAhead], int J0) {
#pragma unroll
      for (int q = 0; q < kRdtAhead; ++q) {
        const int J = J0 - q;
        w[q] = Gs[J * kRdtModes + lane];
        b[q] = Bs[J * kRdtModes + lane];
        r[q] = Rs[J * kRdtModes + lane];
        u[q] = ru[J];
      }
      asm volatile("" ::: "memory");
    };
    auto bwd_rows = [&](const double2(&w)[kRdtAhead], const double(&b)[kRdtAhead], const double(&r)[kRdtAhead],
                        const double(&u)[kRdtAhead], int J0) {
#pragma unroll
      for (int q = 0; q < kRdtAhead; ++q) {
        const int J = J0 - q;
        const double B = b[q], U = u[q];
        // (G_J - U_J X_{J+1}) / B_J, quotients formed from the tabulated 1/B_J
        const double xr = div_by_recip(w[q].x - U * x.x, B, r[q]);
        const double xi = div_by_recip(w[q].y - U * x.y, B, r[q]);
        x.x = (B != 0.0) ? xr : 0.0;
        x.y = (B != 0.0) ? xi : 0.0;
        Gs[max(J, 1) * kRdtModes + lane] = x;  // row 1 is rewritten below
        x2.x = (J == 2) ? x.x : x2.x;
        x2.y = (J == 2) ? x.y : x2.y;
      }
    };
    bwd_load(wa, ba, ra, ua, NY - 1);
    bwd_load(wb, bb, rb, ub, NY - 1 - kRdtAhead);
    for (int J0 = NY - 1; J0 >= 2; J0 -= 2 * kRdtAhead) {
      bwd_rows(wa, ba, ra, ua, J0);
      bwd_load(wa, ba, ra, ua, J0 - 2 * kRdtAhead);
      bwd_rows(wb, bb, rb, ub, J0 - kRdtAhead);
      bwd_load(wb, bb, rb, ub, J0 - 3 * kRdtAhead);
    }
    const double U1 = mean ? 0.0 : -1.0;
    Gs[1 * kRdtModes + lane] = make_double2((0.0 - U1 * x2.x) / 1.0, (0.0 - U1 * x2.y) / 1.0);
  }
  lds_barrier();
  if (!live) return;
  const double hb = upd ? (a.cp_old ? a.clk->h_cp : a.clk->h_bar[a.rk]) : 1.0;
  const double rhb = 1.0 / hb;
  auto project = [&](int J, double2 c1, double2 c2, double2 c3, double2 p) {
    const double2 phi = Gs[J * kRdtModes + lane];
    if (!upd) a.phi[pf * J + m] = phi;  // REM_DIV alone leaves phi (CR1) behind
    double2 t = cmul_ik(kxv, phi);
    a.cu[0][pf * J + m] = make_double2(c1.x - t.x, c1.y - t.y);
    t = cmul_ik(kzv, phi);
    a.cu[2][pf * J + m] = make_double2(c3.x - t.x, c3.y - t.y);
    if (J >= 2) {
      const double2 pm = Gs[(J - 1) * kRdtModes + lane];
      const double d = g.dy[J], rd = g.rdy[J];
      a.cu[1][pf * J + m] =
          make_double2(c2.x - div_by_recip(phi.x - pm.x, d, rd), c2.y - div_by_recip(phi.y - pm.y, d, rd));
      if (upd) a.cp[pf * J + m] = make_double2(p.x + div_by_recip(phi.x, hb, rhb), p.y + div_by_recip(phi.y, hb, rhb));
    }
  };
#pragma unroll
  for (int b = 0; b < kRdtRegRows; ++b) {
    const int J = 1 + rg + b * ngrp;
    if (J <= NY) project(J, u1[b], u2[b], u3[b], pc[b]);
  }
  for (int J = 1 + rg + kRdtRegRows * ngrp; J <= NY; J += ngrp) {
    const double2 c1 = a.cin[0][pf * J + m], c2 = a.cin[1][pf * J + m], c3 = a.cin[2][pf * J + m];
    const double2 p = upd ? a.cp[pf * J + m] : make_double2(0.0, 0.0);
    project(J, c1, c2, c3, p);
  }
  // planes the projection leaves alone (the ghost planes 0 and NY+1, and
  // plane 1 of CU2) pass from the transformed velocity unchanged
  if (a.cin[0] != a.cu[0] && rg == 0) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const long long o = (long long)pf * (i ? NY + 1 : 0) + m;
      a.cu[0][o] = a.cin[0][o];
      a.cu[1][o] = a.cin[1][o];
      a.cu[2][o] = a.cin[2][o];
    }
    a.cu[1][pf + m] = a.cin[1][pf + m];
  }
}
// Two blocks per CU (the velocity rows held across the sweep take the
// registers of a third): 512 resident blocks cover the 457 of a 128 x 128
// plane's modes at once.
__global__ void __launch_bounds__(256) k_rem_div_thomas(RemDivArgs a) { rem_div_thomas_body(a, gridDim.x, blockIdx.x); }
__global__ void __launch_bounds__(256) k_rem_div_thomas_b(const RemDivArgs *__restrict__ av) {
  rem_div_thomas_body(av[blockIdx.z], gridDim.x, blockIdx.x);
}
// REM_DIV with the CN tables of a new DELTA_T in the grid's last ncn blocks
// (the substep after COURANT: the tables serve the next substep's solve,
// and REM_DIV reads none of what they write).
__global__ void __launch_bounds__(256) k_rem_div_thomas_cn(RemDivArgs a, CnArgs cn, int ncn) {
  const int nrd = (int)gridDim.x - ncn;
  if ((int)blockIdx.x >= nrd) {
    cn_tables_body(cn.clk, cn.in, cn.fac, (int)blockIdx.x - nrd);
    return;
  }
  rem_div_thomas_body(a, nrd, blockIdx.x);
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
  a.uphys = c.mixed ? (const void *)c.uphys32 : (const void *)c.uphys;
  a.rphys = c.mixed ? (void *)c.rphys32 : (void *)c.rphys;
  a.fs = (long long)(c.g.ny + 2) * c.g.pp;
  a.cfs = (long long)(c.g.ny + 2) * c.g.pf;
  a.c0 = 0;
  a.j0 = 1;
  a.sgs = c.les ? c.sgs : nullptr;
  a.rows32 = (c.reg_rows && c.g.lw == 32 && !c.mixed && c.plane_threads <= 512) ? 1 : 0;
  const dh_params &p = c.p;
  a.nu = p.nu;
  a.nu_v_scale = p.nu_v_scale;
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
  in.seg_rows = c.fac.seg_rows;
  return in;
}

hipError_t launch_cn_tables(Ctx &c, hipStream_t st) {
  FactorIn in = make_factor_in(c);
  const size_t shm = (size_t)8 * (c.g.ny + 2) * sizeof(double);
  const CnArgs ca{c.clock, in, c.fac};
  return klaunch(c, (const void *)k_cn_tables, (const void *)k_cn_tables_b, dim3(3 * c.nfields), dim3(256),
                 (unsigned)shm, st, ca);
}

// Nonlinear pass of components c0..c1-1 on stream st: one block per plane
// J of the components' ranges (JSTART..JEND, channel.f:77-626) and component.
// The nonlinear pass's kernel for the context's precision, model, row
// variant and block size (planes in LDS), with its block and LDS sizes.
struct NlKernel {
  const void *fn, *fnb;  // the plain and the batched form (nullptr: none)
  int threads;
  unsigned lds;
};
static NlKernel nonlinear_kernel(const Ctx &c) {
  if (c.mixed) {
    if (c.reg_rows_nl)
      return {(const void *)k_nonlinear32<true, kNlRegQ>, (const void *)k_nonlinear32_b<true, kNlRegQ>, kNlRegThreads,
              (unsigned)c.lds_bytes32};
    return {(const void *)k_nonlinear32<false, kMaxQ>, (const void *)k_nonlinear32_b<false, kMaxQ>, c.plane_threads,
            (unsigned)c.lds_bytes32};
  }
  const bool half = c.plane_threads <= 512;  // the LDS variant's 512-thread instantiation
  const unsigned lds = (unsigned)c.lds_bytes;
  if (c.les) {  // FP64 only; the LES kernels have no batched forms
    if (c.reg_rows_nl) return {(const void *)k_nonlinear<true, kNlRegQ, true>, nullptr, kNlRegThreads, lds};
    return {half ? (const void *)k_nonlinear<false, kMaxQ, true, 512> : (const void *)k_nonlinear<false, kMaxQ, true, 1024>,
            nullptr, c.plane_threads, lds};
  }
  if (c.reg_rows_nl)
    return {(const void *)k_nonlinear<true, kNlRegQ>, (const void *)k_nonlinear_b<true, kNlRegQ>, kNlRegThreads, lds};
  if (half)
    return {(const void *)k_nonlinear<false, kMaxQ, false, 512>, (const void *)k_nonlinear_b<false, kMaxQ, 512>,
            c.plane_threads, lds};
  return {(const void *)k_nonlinear<false, kMaxQ, false, 1024>, (const void *)k_nonlinear_b<false, kMaxQ, 1024>,
          c.plane_threads, lds};
}

// Planes of the nonlinear pass over fields c0..c1-1: rows jlo..jhi (empty
// when jhi < jlo).
static void nonlinear_rows(const Ctx &c, int c0, int c1, int &jlo, int &jhi) {
  jlo = c.g.ny;
  jhi = 1;
  for (int k = c0; k < c1; ++k) {
    jlo = min(jlo, k < 3 ? 2 : c.jstart_th[k - 3]);
    jhi = max(jhi, k < 3 ? c.g.ny : c.jend_th[k - 3]);
  }
}

static hipError_t launch_nonlinear_range(Ctx &c, int rk, int c0, int c1, hipStream_t st) {
  if (c1 <= c0) return hipSuccess;
  StepArgs a = make_step_args(c, rk);
  int jlo, jhi;
  nonlinear_rows(c, c0, c1, jlo, jhi);
  if (jhi < jlo) return hipSuccess;
  a.c0 = c0;
  a.j0 = jlo;
  const dim3 grid(jhi - jlo + 1, c1 - c0);
  if (c.gp) {  // planes beyond LDS (FP64, no LES)
    if (c.recording) return hipErrorNotSupported;
    double2 *gp = gp_base(c, st);
    long long slot = c.gp_slot;
    int ncomp = c1 - c0, nitems = (int)(grid.x * grid.y);
    void *args[] = {&a, &gp, &slot, &ncomp, &nitems};
    return hipLaunchKernel((const void *)k_nonlinear_gp, dim3(nitems < kGpSlots ? nitems : kGpSlots),
                           dim3(kGpThreads), args, (unsigned)kGpLds, st);
  }
  const NlKernel k = nonlinear_kernel(c);
  return klaunch(c, k.fn, k.fnb, grid, dim3(k.threads), k.lds, st, a);
}

// Whether the substep runs flat: each phase one launch over all fields on
// the context stream.  Planes in LDS do: with U3's nonlinear pass, solve
// and transform on a branch beside U1's and U2's, the branch's solve and
// transform wait for its pass, and the fork and join cost 6-10 us each
// (config 1: 0.638 ms per step flat against 0.657 with the branch).  The
// large-plane kernels keep the branch (256 x 193 x 256: 7.2 ms per step
// against 7.55 flat).
static bool flat_substep(const Ctx &c) { return !c.gp; }


static hipError_t launch_nonlinear(Ctx &c, int rk) { return launch_nonlinear_range(c, rk, 0, c.nfields, c.stream); }

// The mixed form's boundary rows: the right-hand side of the increment
// d = U_new - U_old on a boundary row is V - (the row applied to U_old), with
// U_old = plane mean (FP64) + fluctuation (FP32).  Boundary rows are among
// 0, 1, 2, NY, NY+1 (real_row); one thread per (x,z) column of field
// f0 + blockIdx.y writes them into the field's R array, after the
// nonlinear pass (which writes R on its own planes, some of them boundary
// rows) and before the solve.
__device__ __forceinline__ void bc_increment_body(const SolveArgs &a) {
  const Geo &g = a.g;
  const int f = a.f0 + blockIdx.y;
  const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= g.pp) return;
  const int NY = g.ny, nrow = NY + 2;
  const long long pp = g.pp, ro = (long long)f * nrow;
  const float *Uo = static_cast<const float *>(a.uold) + (long long)f * a.fs + col;
  float *R = static_cast<float *>(const_cast<void *>(a.rphys)) + (long long)f * a.fs + col;
  auto uold = [&](int i) { return a.four[f][(long long)i * g.pf].x + (double)Uo[(long long)i * pp]; };
  for (int k = 0; k < 5; ++k) {
    const int i = (k < 3) ? k : NY - 3 + k;
    if (!a.fac.isbc[ro + i]) continue;
    const double cl = a.fac.bcl[ro + i], cd = a.fac.bcd[ro + i], cu = a.fac.bcu[ro + i];
    double s = 0.0;
    if (cl != 0.0 && i >= 1) s += cl * uold(i - 1);
    if (cd != 0.0) s += cd * uold(i);
    if (cu != 0.0 && i + 1 < nrow) s += cu * uold(i + 1);
    R[(long long)i * pp] = (float)(a.fac.vconst[ro + i] - s);
  }
}
__global__ void __launch_bounds__(256) k_bc_increment(SolveArgs a) { bc_increment_body(a); }
__global__ void __launch_bounds__(256) k_bc_increment_b(const SolveArgs *__restrict__ av) { bc_increment_body(av[blockIdx.z]); }

// Crank-Nicolson solve of fields f0..f1-1 on stream st.  A velocity
// field's solution overwrites its own R (other components' nonlinear
// passes may still read the physical velocity); a scalar's goes to its
// physical field, which only its own nonlinear pass reads.
static hipError_t launch_solve_range(Ctx &c, int rk, int courant, int f0, int f1, hipStream_t st) {
  if (f1 <= f0) return hipSuccess;
  SolveArgs a{};
  a.g = c.g;
  a.clk = c.clock;
  a.rk = rk;
  a.fs = (long long)(c.g.ny + 2) * c.g.pp;
  if (c.mixed) {
    // increments of every field over their R arrays; scalars' new values
    // also into their physical fields
    a.rphys = c.rphys32;
    a.uold = c.uphys32;
    const long long cfs = (long long)(c.g.ny + 2) * c.g.pf;
    for (int f = 0; f < c.nfields; ++f) {
      a.dst[f] = c.rphys32 + f * a.fs;
      a.four[f] = (f < 3) ? c.cu[f] : c.cth + (f - 3) * cfs;
    }
  } else {
    a.rphys = c.rphys;
    for (int f = 0; f < c.nfields; ++f) a.dst[f] = (f < 3 ? c.rphys : c.uphys) + f * a.fs;
  }
  a.f0 = f0;
  a.fac = c.fac;
  a.nfields = c.nfields;
  for (int f = 0; f < c.nfields; ++f) {
    a.wlo[f] = c.wlo[f];
    a.whi[f] = c.whi[f];
  }
  a.courant = courant;
  a.cfl = c.p.cfl;
  a.dx = c.p.lx / c.g.nx;
  a.dz = c.p.lz / c.g.nz;
  const int cols = (kSegThreads / 64) * (64 / c.fac.nseg);  // columns per block
  const int nb = (int)((c.g.pp + cols - 1) / cols);
  const dim3 grid(nb, f1 - f0);
  if (c.mixed) {
    hipError_t e = klaunch(c, (const void *)k_bc_increment, (const void *)k_bc_increment_b,
                           dim3((unsigned)((c.g.pp + 255) / 256), f1 - f0), dim3(256), 0, st, a);
    if (e != hipSuccess) return e;
    if (c.fac.nseg == 4)
      return klaunch(c, (const void *)k_solve_seg<4, float>, (const void *)k_solve_seg_b<4, float>, grid,
                     dim3(kSegThreads), 0, st, a);
    return klaunch(c, (const void *)k_solve_seg<8, float>, (const void *)k_solve_seg_b<8, float>, grid,
                   dim3(kSegThreads), 0, st, a);
  }
  if (c.fac.nseg == 4)
    return klaunch(c, (const void *)k_solve_seg<4, double>, (const void *)k_solve_seg_b<4, double>, grid,
                   dim3(kSegThreads), 0, st, a);
  return klaunch(c, (const void *)k_solve_seg<8, double>, (const void *)k_solve_seg_b<8, double>, grid,
                 dim3(kSegThreads), 0, st, a);
}

static hipError_t launch_solve(Ctx &c, int rk, int courant) {
  return launch_solve_range(c, rk, courant, 0, c.nfields, c.stream);
}

// LES: the implicit solves with the eddy viscosity (channel.f:631-659 and
// 748-760 explicit, 793-807, 862-875, 927-941, 991-1005 implicit): with
// NU_T varying over (x, z) every column has its own matrix, so one thread
// builds and sweeps one column (THOMAS_FORWARD/BACKWARD_REAL_MPI on one
// rank, mpi.f:630-703, 767-809, with the B = 0 guard), keeping the pivots,
// right-hand sides and upper diagonal in work arrays between the sweeps.
// Rows are those of real_row; interior rows (the ones no BC overwrites) add
// the NU_T parts of the reference's MATL/MATD/MATU, NU_T interpolated to GYF
// for U2, KAPPA_T = NU_T for the scalars.  Their right-hand sides take the
// whole explicit half here, in physical space as the reference forms it
// (the nonlinear pass leaves it out with LES): the molecular part from U
// after les_chan's physical BCs, then the eddy-viscosity part.
struct LesSolveArgs {
  Geo g;
  Clock *clk;
  int rk, f0;
  FactorIn in;
  const double *rphys;  // R of every field (FP64)
  const double *uphys;  // U and TH at the start of the substep
  double *dst[kMaxFields];
  const double *nut;
  double *wk;           // [3][nfields][NY+2][NZ][NX]
  long long fs;
  int wlo[kMaxFields], whi[kMaxFields];
  int courant;
  double cfl, dx, dz;
};

// One wave per block: the columns of a field spread over as many CUs as
// there are waves (4096 columns: 64 CUs per field).
constexpr int kLesSolveThreads = 64;
constexpr int kLesRowMax = 8 * kSegRows;  // rows per column (NY+2, as the segmented solve)
__global__ void __launch_bounds__(kLesSolveThreads) k_solve_les(LesSolveArgs a) {
#pragma clang fp contract(off)
  __shared__ double red[kLesSolveThreads / 64];
  // per-row values shared by every column of the field: real_row's entries,
  // the metric reciprocals of the explicit half, and the eddy-viscosity
  // coefficients T2/(DY DYF) (halved for U2), so that a column's rows take
  // multiplications where the reference divides (the same values to
  // roundoff)
  __shared__ double tL[kLesRowMax], tD[kLesRowMax], tC[kLesRowMax], tV[kLesRowMax];
  __shared__ double rUp[kLesRowMax], rDn[kLesRowMax], rC[kLesRowMax], kLo[kLesRowMax], kUp[kLesRowMax];
  __shared__ int tBc[kLesRowMax];
  const Geo &g = a.g;
  const int f = a.f0 + blockIdx.y;
  const long long pp = g.pp, col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = col < pp;
  const int NY = g.ny, nrow = NY + 2;
  const double temp1 = a.in.nu * a.clk->h_bar[a.rk] / 2.0;
  const double temp2 = a.clk->h_bar[a.rk] / 2.0;
  const double *DY = a.in.dy, *DYF = a.in.dyf;
  const double tv = (f < 3 ? temp1 : temp1 / a.in.pr[f - 3]) * a.in.nu_v_scale;
  for (int J = threadIdx.x; J < nrow; J += blockDim.x) {
    double L, D, C, V;
    int bc;
    real_row(a.in, f, temp1, J, L, D, C, V, bc);
    tL[J] = L;
    tD[J] = D;
    tC[J] = C;
    tV[J] = V;
    tBc[J] = bc;
    if (!bc) {  // interior rows: 1 <= J <= NY
      if (f == 1) {
        rUp[J] = 1.0 / DYF[J];
        rDn[J] = 1.0 / DYF[J - 1];
        rC[J] = 1.0 / DY[J];
        kLo[J] = temp2 * 0.5 / (DYF[J - 1] * DY[J]);
        kUp[J] = temp2 * 0.5 / (DYF[J] * DY[J]);
      } else {
        rUp[J] = 1.0 / DY[J + 1];
        rDn[J] = 1.0 / DY[J];
        rC[J] = 1.0 / DYF[J];
        kLo[J] = temp2 / (DY[J] * DYF[J]);
        kUp[J] = temp2 / (DY[J + 1] * DYF[J]);
      }
    }
  }
  __syncthreads();
  const bool cour = a.courant && f < 3;
  const double dxf = (f == 0) ? a.dx : a.dz;
  double dtmin = __longlong_as_double(0x7FF0000000000000ll);
  if (live) {
    const double *U = a.uphys + f * a.fs + col, *R = a.rphys + f * a.fs + col, *NT = a.nut + col;
    const long long ws = (long long)a.in.nfields * nrow * pp;
    double *Bw = a.wk + (long long)f * nrow * pp + col, *Gw = Bw + ws, *Cw = Gw + ws;
    // row J from its neighbours' U and NU_T and its R (loads issued by the
    // caller, eight rows at a time, ahead of the rows' arithmetic: with one
    // column per thread there are too few waves to hide memory latency):
    // the explicit Crank-Nicolson half, molecular (channel.f:606-626,
    // 739-746) then the eddy viscosity or diffusivity (NU_T interpolated to
    // GYF for U2), and the NU_T parts of the rows
    auto row = [&](int J, double um, double u0, double up, double nm, double n0, double np, double r, double &L,
                   double &D, double &C, double &V) {
      L = tL[J];
      D = tD[J];
      C = tC[J];
      V = tV[J];
      if (tBc[J]) return;
      const double gu = (up - u0) * rUp[J], gd = (u0 - um) * rDn[J];
      V = r + tv * ((gu - gd) * rC[J]);
      const double lo = (f == 1) ? (n0 + nm) : n0, hi = (f == 1) ? (n0 + np) : np;
      const double h = (f == 1) ? 0.5 : 1.0;
      V = V + temp2 * ((h * hi * gu - h * lo * gd) * rC[J]);
      L = L - kLo[J] * lo;
      D = D + kUp[J] * hi + kLo[J] * lo;
      C = C - kUp[J] * hi;
    };
    constexpr int kB = 8;
    double L, D, C, V;
    row(0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, L, D, C, V);  // a boundary row of every field
    double bp = D, gp = V, cp = C;
    Bw[0] = bp;
    Gw[0] = gp;
    Cw[0] = cp;
    double um = U[0], u0 = U[pp], nm = NT[0], n0 = NT[pp];
    for (int J0 = 1; J0 <= NY + 1; J0 += kB) {
      double rr[kB], uu[kB], nn[kB];
#pragma unroll
      for (int q = 0; q < kB; ++q) {
        const int J = J0 + q;
        rr[q] = (J <= NY + 1) ? R[J * pp] : 0.0;
        uu[q] = (J <= NY) ? U[(J + 1) * pp] : 0.0;
        nn[q] = (J <= NY) ? NT[(J + 1) * pp] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < kB; ++q) {
        const int J = J0 + q;
        if (J > NY + 1) break;
        row(J, um, u0, uu[q], nm, n0, nn[q], rr[q], L, D, C, V);
        const double am = -L / bp;
        D = D + am * cp;
        V = V + am * gp;
        Bw[J * pp] = D;
        Gw[J * pp] = V;
        Cw[J * pp] = C;
        bp = D;
        gp = V;
        cp = C;
        um = u0;
        u0 = uu[q];
        nm = n0;
        n0 = nn[q];
      }
    }
    double *Ub = a.dst[f] + col;
    const int wlo = a.wlo[f], whi = a.whi[f];
    auto put = [&](int J, double x) {
      if (J >= wlo && J <= whi) {
        Ub[J * pp] = x;
        if (cour && J >= 1 && J <= NY) {
          const double dl = (f == 1) ? g.dy[J] : dxf;
          dtmin = fmin(dtmin, a.cfl * dl / fabs(x));
        }
      }
    };
    double xn = gp / bp;  // row NY+1
    put(NY + 1, xn);
    for (int J1 = NY; J1 >= 0; J1 -= kB) {
      double bb[kB], gg[kB], cc[kB];
#pragma unroll
      for (int q = 0; q < kB; ++q) {
        const int J = J1 - q;
        bb[q] = (J >= 0) ? Bw[J * pp] : 1.0;
        gg[q] = (J >= 0) ? Gw[J * pp] : 0.0;
        cc[q] = (J >= 0) ? Cw[J * pp] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < kB; ++q) {
        const int J = J1 - q;
        if (J < 0) break;
        const double x = (bb[q] != 0.0) ? (gg[q] - cc[q] * xn) / bb[q] : 0.0;
        xn = x;
        put(J, x);
      }
    }
  }
  if (a.courant && f < 3) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int o = 32; o > 0; o >>= 1) dtmin = fmin(dtmin, __shfl_down(dtmin, o));
    if (lane == 0) red[w] = dtmin;
    __syncthreads();
    if (threadIdx.x == 0) {
      double m = red[0];
      for (int q = 1; q < (int)(blockDim.x >> 6); ++q) m = fmin(m, red[q]);
      if (m >= 0.0) atomicMin(&a.clk->cfl_min_bits, (unsigned long long)__double_as_longlong(m));
    }
  }
}

// The same systems solved with each column split over 16 lanes (4 columns
// per wave, lane = 4 p + s): lane p holds rows [R p, R p + R) of its column
// (rows beyond NY+1 are identity rows), forms them as k_solve_les does and
// eliminates them locally into x_r = y_r - v_r X_L - w_r X_R, X_L and X_R the
// solution just outside the segment (the partition or SPIKE method).  The
// first and last rows of the 16 segments give a 2 x 2 block-tridiagonal
// system over the 15 interfaces, solved by one lane per column; every lane
// then sweeps its rows again with the boundary terms moved to the
// right-hand side.  Sixteen times the waves of one column per thread; the
// sums differ from the sequential sweep's by rounding (as the segmented
// constant-matrix solve's do).
constexpr int kLesPartThreads = 256;
template <int R>
__global__ void __launch_bounds__(kLesPartThreads) k_solve_les_part(LesSolveArgs a) {
#pragma clang fp contract(off)
  __shared__ double red[kLesPartThreads / 64];
  __shared__ double xch[kLesPartThreads / 64][4][16 * 8];  // per wave, column, segment: 8 values
  __shared__ double tL[kLesRowMax], tD[kLesRowMax], tC[kLesRowMax], tV[kLesRowMax];
  __shared__ double rUp[kLesRowMax], rDn[kLesRowMax], rC[kLesRowMax], kLo[kLesRowMax], kUp[kLesRowMax];
  __shared__ int tBc[kLesRowMax];
  const Geo &g = a.g;
  const int f = a.f0 + blockIdx.y;
  const int NY = g.ny, nrow = NY + 2;
  const double temp1 = a.in.nu * a.clk->h_bar[a.rk] / 2.0;
  const double temp2 = a.clk->h_bar[a.rk] / 2.0;
  const double *DY = a.in.dy, *DYF = a.in.dyf;
  const double tv = (f < 3 ? temp1 : temp1 / a.in.pr[f - 3]) * a.in.nu_v_scale;
  for (int J = threadIdx.x; J < nrow; J += blockDim.x) {
    double L, D, C, V;
    int bc;
    real_row(a.in, f, temp1, J, L, D, C, V, bc);
    tL[J] = L;
    tD[J] = D;
    tC[J] = C;
    tV[J] = V;
    tBc[J] = bc;
    if (!bc) {
      if (f == 1) {
        rUp[J] = 1.0 / DYF[J];
        rDn[J] = 1.0 / DYF[J - 1];
        rC[J] = 1.0 / DY[J];
        kLo[J] = temp2 * 0.5 / (DYF[J - 1] * DY[J]);
        kUp[J] = temp2 * 0.5 / (DYF[J] * DY[J]);
      } else {
        rUp[J] = 1.0 / DY[J + 1];
        rDn[J] = 1.0 / DY[J];
        rC[J] = 1.0 / DYF[J];
        kLo[J] = temp2 / (DY[J] * DYF[J]);
        kUp[J] = temp2 / (DY[J + 1] * DYF[J]);
      }
    }
  }
  __syncthreads();
  const long long pp = g.pp;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, p = lane >> 2, sw = lane & 3;
  const long long col = ((long long)blockIdx.x * (kLesPartThreads / 64) + w) * 4 + sw;
  const bool live = col < pp;
  const long long colc = live ? col : pp - 1;  // idle lanes load in bounds
  const int r0 = p * R;
  const double *U = a.uphys + f * a.fs + colc, *Rp = a.rphys + f * a.fs + colc, *NT = a.nut + colc;
  // the segment's rows and their neighbours' U and NU_T
  double uu[R + 2], nn[R + 2], rr[R];
#pragma unroll
  for (int q = 0; q < R + 2; ++q) {
    const int J = r0 - 1 + q;
    const bool in = J >= 0 && J < nrow;
    uu[q] = in ? U[J * pp] : 0.0;
    nn[q] = in ? NT[J * pp] : 0.0;
  }
#pragma unroll
  for (int q = 0; q < R; ++q) rr[q] = (r0 + q < nrow) ? Rp[(r0 + q) * pp] : 0.0;
  double L[R], D[R], C[R], V[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int J = r0 + q;
    if (J >= nrow) {
      L[q] = 0.0; D[q] = 1.0; C[q] = 0.0; V[q] = 0.0;
      continue;
    }
    L[q] = tL[J]; D[q] = tD[J]; C[q] = tC[J]; V[q] = tV[J];
    if (!tBc[J]) {
      const double um = uu[q], u0 = uu[q + 1], up = uu[q + 2], nm = nn[q], n0 = nn[q + 1], np = nn[q + 2];
      const double gu = (up - u0) * rUp[J], gd = (u0 - um) * rDn[J];
      double v = rr[q] + tv * ((gu - gd) * rC[J]);
      const double lo = (f == 1) ? (n0 + nm) : n0, hi = (f == 1) ? (n0 + np) : np;
      const double h = (f == 1) ? 0.5 : 1.0;
      V[q] = v + temp2 * ((h * hi * gu - h * lo * gd) * rC[J]);
      L[q] = L[q] - kLo[J] * lo;
      D[q] = D[q] + kUp[J] * hi + kLo[J] * lo;
      C[q] = C[q] - kUp[J] * hi;
    }
    if (J == 0) L[q] = 0.0;          // no row below 0
    if (J == nrow - 1) C[q] = 0.0;   // nor above NY+1
  }
  // local elimination: cp = C/piv, ip = 1/piv, V -> forward-swept; the left
  // spike's last forward value
  double cp[R], ip[R];
  double vlast = 0.0;
  {
    double cprev = 0.0, vprev = 0.0;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const double piv = (q == 0) ? D[q] : D[q] - L[q] * cprev;
      const double inv = 1.0 / piv;
      ip[q] = inv;
      cp[q] = C[q] * inv;
      if (q > 0) V[q] = V[q] - L[q] * V[q - 1];
      V[q] = V[q] * inv;
      vprev = ((q == 0) ? L[q] : -L[q] * vprev) * inv;
      cprev = cp[q];
    }
    vlast = vprev;
  }
  // first-row values: y_0 and w_0 by the backward recurrences, v_0 = L_0 /
  // q_0 with q the pivots of an elimination from the bottom
  double yf, wf, vf;
  {
    double yy = V[R - 1], ww = cp[R - 1];
#pragma unroll
    for (int q = R - 2; q >= 0; --q) {
      yy = V[q] - cp[q] * yy;
      ww = -cp[q] * ww;
    }
    yf = yy;
    wf = ww;
    double qq = D[R - 1];
#pragma unroll
    for (int q = R - 2; q >= 0; --q) qq = D[q] - C[q] * L[q + 1] / qq;
    vf = L[0] / qq;
  }
  double *rw = xch[w][sw];
  {
    double *q8 = rw + p * 8;
    q8[0] = yf; q8[1] = V[R - 1]; q8[2] = vf; q8[3] = vlast; q8[4] = wf; q8[5] = cp[R - 1];
  }
  wave_lds_sync();
  // interfaces i: P_i = x(last row of segment i), Q_i = x(first row of
  // segment i+1):  P_i + vl_i P_{i-1} + wl_i Q_i = yl_i,
  //                vf_{i+1} P_i + Q_i + wf_{i+1} Q_{i+1} = yf_{i+1}
  if (p == 0) {
    double pc0 = 0.0, prP = 0.0;
    // forward: r0 -> [i][1], c0 -> [i][6], r1 -> [i+1][0], c1 -> [i+1][7]
#pragma unroll 1
    for (int i = 0; i < 15; ++i) {
      double *qi = rw + i * 8, *qn = rw + (i + 1) * 8;
      const double yl = qi[1], vl = qi[3], wl = qi[5], yfn = qn[0], vfn = qn[2], wfn = qn[4];
      const double m01 = wl - vl * pc0;
      const double idet = 1.0 / (1.0 - m01 * vfn);
      const double t0 = yl - vl * prP;
      const double c0 = -m01 * wfn * idet, c1 = wfn * idet;
      const double x0 = (t0 - m01 * yfn) * idet, x1 = (yfn - vfn * t0) * idet;
      qi[1] = x0; qi[6] = c0; qn[0] = x1; qn[7] = c1;
      pc0 = c0;
      prP = x0;
    }
    double qn = 0.0;
#pragma unroll 1
    for (int i = 14; i >= 0; --i) {
      double *qi = rw + i * 8, *qx = rw + (i + 1) * 8;
      const double P = qi[1] - qi[6] * qn, Q = qx[0] - qx[7] * qn;
      qn = Q;
      qx[0] = P;  // X_L of segment i+1
      qi[1] = Q;  // X_R of segment i
    }
    rw[0] = 0.0;
    rw[15 * 8 + 1] = 0.0;
  }
  wave_lds_sync();
  const double XL = rw[p * 8 + 0], XR = rw[p * 8 + 1];
  // the rows with the boundary terms moved right: forward-swept
  // V - X_L v' - X_R w', then the backward sweep
  {
    double vprev = 0.0;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      vprev = ((q == 0) ? L[q] : -L[q] * vprev) * ip[q];
      V[q] = V[q] - vprev * XL;
    }
    V[R - 1] = V[R - 1] - cp[R - 1] * XR;
  }
#pragma unroll
  for (int q = R - 2; q >= 0; --q) V[q] = V[q] - cp[q] * V[q + 1];
  const bool cour = a.courant && f < 3;
  const double dxf = (f == 0) ? a.dx : a.dz;
  double dtmin = __longlong_as_double(0x7FF0000000000000ll);
  if (live) {
    double *Ub = a.dst[f] + col;
    const int wlo = a.wlo[f], whi = a.whi[f];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int J = r0 + q;
      if (J < nrow && J >= wlo && J <= whi) {
        Ub[J * pp] = V[q];
        if (cour && J >= 1 && J <= NY) {
          const double dl = (f == 1) ? g.dy[J] : dxf;
          dtmin = fmin(dtmin, a.cfl * dl / fabs(V[q]));
        }
      }
    }
  }
  if (a.courant && f < 3) {
    for (int o = 32; o > 0; o >>= 1) dtmin = fmin(dtmin, __shfl_down(dtmin, o));
    if (lane == 0) red[w] = dtmin;
    __syncthreads();
    if (threadIdx.x == 0) {
      double m = red[0];
      for (int q = 1; q < (int)(blockDim.x >> 6); ++q) m = fmin(m, red[q]);
      if (m >= 0.0) atomicMin(&a.clk->cfl_min_bits, (unsigned long long)__double_as_longlong(m));
    }
  }
}

hipError_t launch_solve_les_range(Ctx &c, int rk, int courant, int f0, int f1, hipStream_t st) {
  if (f1 <= f0) return hipSuccess;
  LesSolveArgs a{};
  a.g = c.g;
  a.clk = c.clock;
  a.rk = rk;
  a.f0 = f0;
  a.in = make_factor_in(c);
  a.rphys = c.rphys;
  a.uphys = c.uphys;
  a.fs = (long long)(c.g.ny + 2) * c.g.pp;
  for (int f = 0; f < c.nfields; ++f) {
    a.dst[f] = (f < 3 ? c.rphys : c.uphys) + f * a.fs;
    a.wlo[f] = c.wlo[f];
    a.whi[f] = c.whi[f];
  }
  a.nut = c.nut;
  a.wk = c.les_wk;
  a.courant = courant;
  a.cfl = c.p.cfl;
  a.dx = c.p.lx / c.g.nx;
  a.dz = c.p.lz / c.g.nz;
  const int nrow = c.g.ny + 2;
  if (nrow <= 16 * 17) {
    // columns split over 16 lanes: 16 columns per block of 4 waves
    const dim3 grid((unsigned)((c.g.pp + 15) / 16), f1 - f0);
    const void *k = (nrow <= 16 * 9) ? (const void *)k_solve_les_part<9> : (const void *)k_solve_les_part<17>;
    return klaunch(c, k, nullptr, grid, dim3(kLesPartThreads), 0, st, a);
  }
  const dim3 grid((unsigned)((c.g.pp + kLesSolveThreads - 1) / kLesSolveThreads), f1 - f0);
  return klaunch(c, (const void *)k_solve_les, nullptr, grid, dim3(kLesSolveThreads), 0, st, a);
}

static size_t rem_div_thomas_lds(const Ctx &c) {
  const size_t ntot = c.g.ny + 1 + 2 * kRdtPad;
  return ntot * kRdtModes * 5 * sizeof(double) + ntot * sizeof(double);
}

static RemDivArgs make_rem_div_args(Ctx &c, int rk) {
  RemDivArgs a{};
  a.g = c.g;
  a.clk = c.clock;
  a.rk = rk;
  a.cp_old = (rk == 2 && c.p.variable_dt) ? 1 : 0;
  for (int i = 0; i < 3; ++i) {
    a.cin[i] = (rk >= 0) ? c.cun[i] : c.cu[i];  // a substep projects the transformed velocity
    a.cu[i] = c.cu[i];
  }
  a.cp = c.cp;
  a.phi = c.phi;
  a.rl = c.rd_l;
  a.ru = c.rd_u;
  a.ta = c.rd_ta;
  a.tb = c.rd_tb;
  a.tr = c.rd_tr;
  a.seg = c.rd_seg;
  return a;
}

hipError_t launch_rem_div_factors(Ctx &c) {
  RemDivArgs a = make_rem_div_args(c, -1);
  hipLaunchKernelGGL(k_rem_div_factors, dim3((c.g.pf + 255) / 256), dim3(256), 0, c.stream, a);
  return hipGetLastError();
}

hipError_t launch_rem_div(Ctx &c, int rk, int bnd_end, int bnd_begin) {
  RemDivArgs a = make_rem_div_args(c, rk);
  a.bnd_end = bnd_end;
  a.bnd_begin = bnd_begin;
  const int nb = (c.g.pf + kRdtModes - 1) / kRdtModes;
  return klaunch(c, (const void *)k_rem_div_thomas, (const void *)k_rem_div_thomas_b, dim3(nb), dim3(256),
                 (unsigned)rem_div_thomas_lds(c), c.stream, a);
}

// REM_DIV and the CN tables for the DELTA_T COURANT just set, in one launch
// (not in recording mode, which has no batched form of it).
static hipError_t launch_rem_div_cn(Ctx &c, int rk, int bnd_end, int bnd_begin) {
  if (c.recording) return hipErrorNotSupported;
  RemDivArgs a = make_rem_div_args(c, rk);
  a.bnd_end = bnd_end;
  a.bnd_begin = bnd_begin;
  const CnArgs cn{c.clock, make_factor_in(c), c.fac};
  const int nb = (c.g.pf + kRdtModes - 1) / kRdtModes, ncn = 3 * c.nfields;
  const size_t lrd = rem_div_thomas_lds(c), lcn = (size_t)8 * (c.g.ny + 2) * sizeof(double);
  const size_t lds = lrd > lcn ? lrd : lcn;
  hipLaunchKernelGGL(k_rem_div_thomas_cn, dim3(nb + ncn), dim3(256), (unsigned)lds, c.stream, a, cn, ncn);
  return hipGetLastError();
}

// COURANT, then the solve tables for the new DELTA_T on the side stream: the
// next solve is a substep away, so the rebuild runs beside this substep's
// transform and pressure solve and is joined at the end of the substep.
static hipError_t launch_courant_dt(Ctx &c) {
  const CourantArgs ca{c.clock, c.dt_cap, c.p.variable_dt, c.p.update_dt};
  return klaunch(c, (const void *)k_courant_dt, (const void *)k_courant_dt_b, dim3(1), dim3(64), 0, c.stream, ca);
}
static hipError_t launch_courant(Ctx &c) {
  hipError_t e = launch_courant_dt(c);
  if (e != hipSuccess) return e;
  if ((e = krecord(c, 2, c.stream)) != hipSuccess) return e;  // ev_fork
  if ((e = kwait(c, c.side, 2)) != hipSuccess) return e;
  if ((e = launch_cn_tables(c, c.side)) != hipSuccess) return e;
  if ((e = krecord(c, 3, c.side)) != hipSuccess) return e;  // ev_join
  c.tables_pending = true;
  return hipSuccess;
}

// Kernels of one substep in launch order (channel.f:386-1065).
const char *const kPhaseNames[kNumPhases] = {"to_physical", "nonlinear", "cn_solve", "courant", "to_fourier",
                                             "rem_div"};

hipError_t launch_phase(Ctx &c, int rk, int phase) {
  const int courant = (c.p.variable_dt && rk == 2) ? 1 : 0;
  switch (phase) {
    case 0: {
      if (!c.les) return launch_state_to_phys(c);
      hipError_t e = launch_les_strain_and_state(c, c.stream);
      if (e == hipSuccess) e = launch_les_model_rest(c, c.stream);
      return e;
    }
    case 1: return launch_nonlinear(c, rk);
    case 2: return c.les ? launch_solve_les_range(c, rk, courant, 0, c.nfields, c.stream) : launch_solve(c, rk, courant);
    case 3: return courant ? launch_courant(c) : hipSuccess;
    case 4: return launch_step_to_four(c, 0, c.nfields, c.stream);
    case 5: return launch_rem_div(c, rk);
  }
  return hipErrorInvalidValue;
}

hipError_t join_side(Ctx &c) {
  if (!c.tables_pending) return hipSuccess;
  c.tables_pending = false;
  return kwait(c, c.stream, 3);  // ev_join
}

// One substep.  With planes in LDS each phase is one launch over all fields
// on the context stream (flat_substep).  On the large-plane path the
// nonlinear pass of U1 and U2 runs on the context stream and the last
// velocity component's nonlinear pass, solve and transform back (then the
// scalars' nonlinear pass and solve) as a branch beside it and U1's and
// U2's solve and transform; the scalars' transform back, COURANT and
// REM_DIV follow on the context stream once the branch has joined.  The
// kernels and their arithmetic are those of launch_phase.
hipError_t launch_substep(Ctx &c, int rk, int bnd_end, int bnd_begin) {
  const int courant = (c.p.variable_dt && rk == 2) ? 1 : 0;
  // after COURANT the CN tables of the new DELTA_T are built in REM_DIV's
  // launch (a fork to a side stream and its join cost more on the device
  // than the tables take)
  const bool fuse_tables = courant && !c.recording;
  const int nv = c.nfields < 3 ? c.nfields : 3;
  hipError_t e;
#define SCHK(expr) \
  if ((e = (expr)) != hipSuccess) return e;
  if (c.les) {
    // les_chan's BCs and the model's strain in one launch, then its inverse
    // transforms and U's in one launch (both read the BC'd Fourier copy)
    SCHK(launch_les_strain_and_state(c, c.stream));
    SCHK(launch_les_model_rest(c, c.stream));
  } else {
    SCHK(launch_state_to_phys(c));
  }
  if (flat_substep(c)) {
    // one launch per phase, in launch_phase's order (the solve tables of a
    // new DELTA_T in REM_DIV's launch)
    SCHK(launch_nonlinear_range(c, rk, 0, c.nfields, c.stream));
    SCHK(c.les ? launch_solve_les_range(c, rk, courant, 0, c.nfields, c.stream)
               : launch_solve_range(c, rk, courant, 0, c.nfields, c.stream));
    if (courant) SCHK(fuse_tables ? launch_courant_dt(c) : launch_courant(c));
    SCHK(launch_step_to_four(c, 0, c.nfields, c.stream));
    SCHK(fuse_tables ? launch_rem_div_cn(c, rk, bnd_end, bnd_begin) : launch_rem_div(c, rk, bnd_end, bnd_begin));
    return join_side(c);
  }
  // planes beyond LDS (FP64, no LES model): U3 and the scalars run on the
  // branch, which forks as soon as the fields are in physical space (all
  // its nonlinear pass reads), each pass on its own scratch half
  SCHK(krecord(c, 0, c.stream));  // ev_nl01
  SCHK(kwait(c, c.branch, 0));
  SCHK(launch_nonlinear_range(c, rk, 0, 2, c.stream));
  SCHK(launch_nonlinear_range(c, rk, 2, c.nfields, c.branch));
  SCHK(launch_solve_range(c, rk, courant, 2, c.nfields, c.branch));
  SCHK(launch_step_to_four(c, 2, nv, c.branch));
  SCHK(krecord(c, 1, c.branch));  // ev_branch_done
  SCHK(launch_solve_range(c, rk, courant, 0, 2, c.stream));
  SCHK(launch_step_to_four(c, 0, 2, c.stream));
  SCHK(kwait(c, c.stream, 1));
  if (courant) SCHK(fuse_tables ? launch_courant_dt(c) : launch_courant(c));
  SCHK(launch_step_to_four(c, 3, c.nfields, c.stream));
  SCHK(fuse_tables ? launch_rem_div_cn(c, rk, bnd_end, bnd_begin) : launch_rem_div(c, rk, bnd_end, bnd_begin));
#undef SCHK
  return join_side(c);
}

// Step bookkeeping of diablo.f:66-110 between substeps: TIME = TIME +
// DELTA_T ending a step (with the DELTA_T COURANT set in it), then
// TIME_STEP = TIME_STEP + 1 starting the next; one launch serves both.
struct BoundaryArgs {
  Clock *clk;
  int end_prev, begin_next;
};
__device__ __forceinline__ void step_boundary_body(const BoundaryArgs &a) {
  if (threadIdx.x != 0) return;
  Clock *clk = a.clk;
  if (a.end_prev) clk->time = clk->time + clk->dt;
  if (a.begin_next) clk->time_step += 1;
}
__global__ void k_step_boundary(BoundaryArgs a) { step_boundary_body(a); }
__global__ void k_step_boundary_b(const BoundaryArgs *__restrict__ av) { step_boundary_body(av[blockIdx.z]); }

hipError_t launch_step_boundary(Ctx &c, bool end_prev, bool begin_next) {
  const BoundaryArgs ba{c.clock, end_prev ? 1 : 0, begin_next ? 1 : 0};
  return klaunch(c, (const void *)k_step_boundary, (const void *)k_step_boundary_b, dim3(1), dim3(64), 0, c.stream,
                 ba);
}

hipError_t prepare_step_kernels(Ctx &c) {
  if ((c.lds_bytes > kLdsMax && !c.gp) || (long long)rem_div_thomas_lds(c) > kLdsMax) return hipErrorInvalidValue;
  const void *fns[] = {(const void *)k_nonlinear<true, kNlRegQ>,    (const void *)k_nonlinear<false, kMaxQ>,
                       (const void *)k_nonlinear<true, kNlRegQ, true>, (const void *)k_nonlinear<false, kMaxQ, true>,
                       (const void *)k_nonlinear<false, kMaxQ, false, 512>,
                       (const void *)k_nonlinear<false, kMaxQ, true, 512>,
                       (const void *)k_nonlinear_b<false, kMaxQ, 512>,
                       (const void *)k_nonlinear32<true, kNlRegQ>,  (const void *)k_nonlinear32<false, kMaxQ>,
                       (const void *)k_nonlinear_b<true, kNlRegQ>,  (const void *)k_nonlinear_b<false, kMaxQ>,
                       (const void *)k_nonlinear32_b<true, kNlRegQ>, (const void *)k_nonlinear32_b<false, kMaxQ>,
                       (const void *)k_rem_div_thomas,              (const void *)k_rem_div_thomas_b,
                       (const void *)k_rem_div_thomas_cn};
  for (const void *fn : fns) {
    const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, kLdsMax);
    if (e != hipSuccess) return e;
  }
  return hipFuncSetAttribute((const void *)k_nonlinear_gp, hipFuncAttributeMaxDynamicSharedMemorySize, kGpLds);
}

}  // namespace dh
