// Batched kernels outside the channel context, on caller-owned device
// arrays:
//   k_cn_solve   complex tridiagonal systems of the Crank-Nicolson form
//                (channel.f:914-963) with per-system coefficients, the work
//                of THOMAS_FORWARD/BACKWARD_COMPLEX_MPI (mpi.f:708-865) on
//                many systems at once;
//   k_fft512_*   2-D real transforms of 512 x 512 planes in the directions
//                and normalisation of FFT_XZ_TO_FOURIER/PHYSICAL
//                (mpi.f:1477-1606): a real-to-complex x pass and a complex z
//                pass, full spectra (no de-aliasing truncation).
// FP64 (double2) or FP32 (float2) storage and arithmetic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>

#include "dh_fft_dev.h"

namespace dh {

// ---------------------------------------------------------------------------
// Batched Crank-Nicolson solve.  System s, row j (0..ny-1):
//   a_j x_{j-1} + b_j x_j + c_j x_{j+1} = d_j,
//   a_j = -alpha_s sl_j, c_j = -alpha_s su_j, b_j = 1 + alpha_s (sl_j + su_j),
// the rows of channel.f:914-963 with alpha_s = NU T1 / h^2 and sl, su the
// metric factors h^2/(DY(J) DYF(J)), h^2/(DY(J+1) DYF(J)); sl = su = 0 gives
// an identity (Dirichlet) row.  The couplings a_0 and c_{ny-1} reach outside
// the system and are dropped, as the reference's sweeps never read them.
// Layout [row][system], systems fastest.
//
// A wave solves 64/SEG systems, SEG lanes each (lane = (64/SEG) p + s).
// Lane p holds rows [R p, R p + R) of its system in registers (rows at or beyond ny are
// identity rows with zero right-hand side, which leave the others alone) and
// eliminates them locally, Thomas on the segment with its outer couplings
// dropped, so that
//   x_r = y_r - v_r X_L - w_r X_R,
// X_L, X_R the solution just outside the segment (the partition or SPIKE
// method): y = T^-1 d, v = T^-1 a_0 e_0, w = T^-1 c_{R-1} e_{R-1}.  The first
// and last rows of every segment give a 2 x 2 block-tridiagonal system over
// the SEG-1 interfaces, which one lane per system solves by block elimination.
// Each lane then sweeps its rows again with the boundary terms moved to the
// right-hand side.  The forward-swept right-hand side, c/pivot and 1/pivot
// stay in registers between the sweeps (4 reals per row); the spike's
// forward values are formed again when needed.  Loads and
// stores are the right-hand side and the solution only (the diagonals come
// from alpha and the two row tables): 32 bytes per row in FP64.
// ---------------------------------------------------------------------------
constexpr int kCnThreads = 256;
constexpr int kCnRows = 9;       // rows per lane
constexpr int kCnMaxNy = 32 * kCnRows;

// SEG segments (lanes) per system, 64 / SEG systems per wave: 16 for ny <=
// 144, 32 for ny <= 288 (9 rows per lane either way, which keeps the FP64
// solve within 168 registers).
template <int R, int SEG, class C>
__device__ __forceinline__ void cn_solve_wave(const double *__restrict__ alpha, const C *__restrict__ rhs,
                                              C *__restrict__ x, int ny, long long nsys, Sc<C> *red,
                                              const Sc<C> *sl, const Sc<C> *su) {
  using S = Sc<C>;
  constexpr int kCnSeg = SEG, kCnSys = 64 / SEG;
  const int lane = threadIdx.x & 63;
  const int p = lane / kCnSys, sw = lane % kCnSys;
  const long long wave = (long long)blockIdx.x * (kCnThreads / 64) + (threadIdx.x >> 6);
  const long long sys = wave * kCnSys + sw;
  const bool live = sys < nsys;
  const long long sysc = live ? sys : nsys - 1;  // loads of idle lanes stay in bounds
  const int r0 = p * R;
  const S al = S(alpha[sysc]);
  // row coefficients (rows beyond ny: identity; the outer couplings dropped)
  auto coef = [&](int r, S &a, S &b, S &c) {
    const int j = r0 + r;
    if (j < ny) {
      a = (j == 0) ? S(0) : -al * sl[j];
      c = (j == ny - 1) ? S(0) : -al * su[j];
      b = S(1) + al * (sl[j] + su[j]);
    } else {
      a = S(0);
      b = S(1);
      c = S(0);
    }
  };

  C d[R];
  {
    // rows at stride nsys from the segment's first: the per-row offsets are
    // multiples of a scalar, so no per-row address stays live in registers
    const C *rp = rhs + (long long)r0 * nsys + sysc;
#pragma unroll
    for (int r = 0; r < R; ++r) d[r] = (r0 + r < ny) ? rp[(long long)r * nsys] : mkc<C>(S(0), S(0));
  }
  // forward: cp_r = c_r / piv_r, d'_r; the left spike v'_r = -a_r v'_{r-1} / piv_r
  // (v'_0 = a_0 / piv_0) is carried only to its last value
  S cp[R], ip[R];  // c_r / piv_r and 1 / piv_r
  S a0 = S(0), vlast = S(0);
  {
    S cprev = S(0), vprev = S(0);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      S a, b, c;
      coef(r, a, b, c);
      if (r == 0) a0 = a;
      const S piv = (r == 0) ? b : b - a * cprev;
      const S inv = S(1) / piv;
      ip[r] = inv;
      cp[r] = c * inv;
      if (r > 0) d[r] = mkc<C>(d[r].x - a * d[r - 1].x, d[r].y - a * d[r - 1].y);
      d[r] = mkc<C>(d[r].x * inv, d[r].y * inv);
      vprev = ((r == 0) ? a : -a * vprev) * inv;
      cprev = cp[r];
    }
    vlast = vprev;
  }
  // first-row values: y_0 and w_0 by the backward recurrences (no storage),
  // v_0 = a_0 (T^-1)_00 = a_0 / q_0 with q the pivots of an elimination from
  // the bottom: q_{R-1} = b_{R-1}, q_r = b_r - c_r a_{r+1} / q_{r+1}
  // (the empty asm makes every sweep read its row tables from LDS again
  // rather than keep the previous sweep's coefficients in registers)
  asm volatile("" ::: "memory");
  C yf, yl = d[R - 1];
  S wf, wl = cp[R - 1], vf, vl = vlast;
  {
    C yy = d[R - 1];
    S ww = cp[R - 1];
#pragma unroll
    for (int r = R - 2; r >= 0; --r) {
      yy = mkc<C>(d[r].x - cp[r] * yy.x, d[r].y - cp[r] * yy.y);
      ww = -cp[r] * ww;
    }
    yf = yy;
    wf = ww;
    asm volatile("" ::: "memory");
    S an, bn, cn;
    coef(R - 1, an, bn, cn);
    S q = bn, anext = an;
#pragma unroll
    for (int r = R - 2; r >= 0; --r) {
      S a, b, c;
      coef(r, a, b, c);
      q = b - c * anext / q;
      anext = a;
    }
    vf = a0 / q;
  }
  // the segment's first and last rows to the wave's exchange area:
  // [sys][p][8] = y_first (2), y_last (2), v_first, v_last, w_first, w_last
  S *rw = red + ((threadIdx.x >> 6) * kCnSys + sw) * (kCnSeg * 8);
  {
    S *q = rw + p * 8;
    q[0] = yf.x; q[1] = yf.y; q[2] = yl.x; q[3] = yl.y;
    q[4] = vf; q[5] = vl; q[6] = wf; q[7] = wl;
  }
  wave_lds_sync();
  // interfaces i = 0..SEG-2: P_i = x(last row of segment i), Q_i = x(first
  // row of segment i+1):
  //   P_i + vl_i P_{i-1} + wl_i Q_i         = yl_i
  //   vf_{i+1} P_i + Q_i + wf_{i+1} Q_{i+1} = yf_{i+1}
  // with P_{-1} = Q_{SEG-1} = 0 (no rows outside the system).  Block elimination
  // keeps U-hat = [[0, c0], [0, c1]] (only Q_{i+1} couples forward).
  if (p == 0) {
    // forward elimination; the eliminated coefficients of interface i go
    // to slots the later steps no longer read: r0 -> [i][2,3], c0 -> [i][5],
    // r1 -> [i+1][0,1], c1 -> [i+1][4]
    S pc0 = S(0);
    C prP = mkc<C>(S(0), S(0));
#pragma unroll 1
    for (int i = 0; i < kCnSeg - 1; ++i) {
      S *qi = rw + i * 8, *qn = rw + (i + 1) * 8;
      const C ylv = mkc<C>(qi[2], qi[3]), yfv = mkc<C>(qn[0], qn[1]);
      const S vli = qi[5], wli = qi[7], vfn = qn[4], wfn = qn[6];
      // M = [[1, wl - vl c0_{i-1}], [vf, 1]], rhs = (yl - vl rP_{i-1}, yf),
      // M^{-1} = [[1, -m01], [-vf, 1]] / det
      const S m01 = wli - vli * pc0;
      const S idet = S(1) / (S(1) - m01 * vfn);
      const C t0 = mkc<C>(ylv.x - vli * prP.x, ylv.y - vli * prP.y);
      const S c0 = -m01 * wfn * idet, c1 = wfn * idet;
      const C rr0 = mkc<C>((t0.x - m01 * yfv.x) * idet, (t0.y - m01 * yfv.y) * idet);
      const C rr1 = mkc<C>((yfv.x - vfn * t0.x) * idet, (yfv.y - vfn * t0.y) * idet);
      qi[2] = rr0.x; qi[3] = rr0.y; qi[5] = c0;
      qn[0] = rr1.x; qn[1] = rr1.y; qn[4] = c1;
      pc0 = c0;
      prP = rr0;
    }
    // back substitution u_i = r_i - (c0_i, c1_i) Q_{i+1}; X_L of segment i+1
    // (= P_i) to [i+1][0,1], X_R of segment i (= Q_i) to [i][2,3]
    C qn = mkc<C>(S(0), S(0));
#pragma unroll 1
    for (int i = kCnSeg - 2; i >= 0; --i) {
      S *qi = rw + i * 8, *qx = rw + (i + 1) * 8;
      const C rr0 = mkc<C>(qi[2], qi[3]), rr1 = mkc<C>(qx[0], qx[1]);
      const S c0 = qi[5], c1 = qx[4];
      const C P = mkc<C>(rr0.x - c0 * qn.x, rr0.y - c0 * qn.y);
      const C Q = mkc<C>(rr1.x - c1 * qn.x, rr1.y - c1 * qn.y);
      qn = Q;
      qx[0] = P.x;
      qx[1] = P.y;
      qi[2] = Q.x;
      qi[3] = Q.y;
    }
    rw[0] = S(0);
    rw[1] = S(0);
    rw[(kCnSeg - 1) * 8 + 2] = S(0);
    rw[(kCnSeg - 1) * 8 + 3] = S(0);
  }
  wave_lds_sync();
  const C XL = mkc<C>(rw[p * 8 + 0], rw[p * 8 + 1]);
  const C XR = mkc<C>(rw[p * 8 + 2], rw[p * 8 + 3]);
  // the segment's system with the boundary terms on the right: forward-swept
  // d''_r = d'_r - X_L v'_r - X_R w'_r (w'_r = 0 below the last row, cp_{R-1}
  // there), then the backward sweep x_r = d''_r - cp_r x_{r+1}
  {
    S vprev = S(0);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      S a, b, c;
      coef(r, a, b, c);
      vprev = ((r == 0) ? a : -a * vprev) * ip[r];
      d[r] = mkc<C>(d[r].x - vprev * XL.x, d[r].y - vprev * XL.y);
    }
    d[R - 1] = mkc<C>(d[R - 1].x - cp[R - 1] * XR.x, d[R - 1].y - cp[R - 1] * XR.y);
  }
#pragma unroll
  for (int r = R - 2; r >= 0; --r) d[r] = mkc<C>(d[r].x - cp[r] * d[r + 1].x, d[r].y - cp[r] * d[r + 1].y);
  if (!live) return;
  C *xp = x + (long long)r0 * nsys + sys;
  asm volatile("" : "+v"(xp));  // a fresh base: the loads' addresses are not kept for the stores
#pragma unroll
  for (int r = 0; r < R; ++r)
    if (r0 + r < ny) xp[(long long)r * nsys] = d[r];
}

template <int SEG, class C>
__global__ void __launch_bounds__(kCnThreads) k_cn_solve(const double *__restrict__ sl, const double *__restrict__ su,
                                                       const double *__restrict__ alpha, const C *__restrict__ rhs,
                                                       C *__restrict__ x, int ny, long long nsys) {
  using S = Sc<C>;
  __shared__ S red[(kCnThreads / 64) * 64 * 8];
  __shared__ S tab[2 * kCnMaxNy];
  for (int j = threadIdx.x; j < ny; j += blockDim.x) {
    tab[j] = S(sl[j]);
    tab[ny + j] = S(su[j]);
  }
  __syncthreads();
  cn_solve_wave<kCnRows, SEG, C>(alpha, rhs, x, ny, nsys, red, tab, tab + ny);
}

// ---------------------------------------------------------------------------
// 512-point complex DFT in one wave: lane l holds x_{l + 64 j} in v[j]
// (j < 8) and ends holding X_{l + 64 j}.  512 = 8 x 64 and 64 = 8 x 8:
//   an 8-point DFT over j in registers, twiddle W512^{l k1};
//   transpose (lane l1 + 8 k1 takes the values of l = l1 + 8 l2), an
//   8-point DFT over l2, twiddle W64^{l1 k2a};
//   transpose (lane k1 + 8 k2a takes the values of l1), an 8-point DFT
//   over l1, giving X_{k1 + 8 k2a + 64 k2b} in slot k2b.
// The exchanges go through 512 slots of LDS at s[idx * stride] that only
// this wave touches.  tw512 holds exp(-2 pi i m/512), m < 512.
// ---------------------------------------------------------------------------
template <bool INV, class C>
__device__ __forceinline__ C tw_of(const C *tw, int m) {
  const C w = tw[m & 511];
  return INV ? cconj(w) : w;
}

template <bool INV, class C>
__device__ __forceinline__ void fft512_wave(C (&v)[8], C *s, int stride, const C *tw) {
  const int lane = threadIdx.x & 63;
  dft_small<8, INV>(v);
#pragma unroll
  for (int k1 = 1; k1 < 8; ++k1) v[k1] = cmul(v[k1], tw_of<INV>(tw, lane * k1));
  // lane l = l1 + 8 l2 writes Y[l][k1] to slot 64 k1 + l
#pragma unroll
  for (int k1 = 0; k1 < 8; ++k1) s[(64 * k1 + lane) * stride] = v[k1];
  wave_lds_sync();
  {
    const int l1 = lane & 7, k1 = lane >> 3;
#pragma unroll
    for (int l2 = 0; l2 < 8; ++l2) v[l2] = s[(64 * k1 + l1 + 8 * l2) * stride];
    dft_small<8, INV>(v);  // v[k2a]
#pragma unroll
    for (int k2a = 1; k2a < 8; ++k2a) v[k2a] = cmul(v[k2a], tw_of<INV>(tw, 8 * l1 * k2a));
    wave_lds_sync();
    // lane (l1, k1) writes G[l1][k2a] to slot 64 k1 + 8 k2a + l1
#pragma unroll
    for (int k2a = 0; k2a < 8; ++k2a) s[(64 * k1 + 8 * k2a + l1) * stride] = v[k2a];
  }
  wave_lds_sync();
  {
    const int k1 = lane & 7, k2a = lane >> 3;
#pragma unroll
    for (int l1 = 0; l1 < 8; ++l1) v[l1] = s[(64 * k1 + 8 * k2a + l1) * stride];
    dft_small<8, INV>(v);  // v[k2b] = X_{k1 + 8 k2a + 64 k2b} = X_{lane + 64 k2b}
  }
  wave_lds_sync();
}

// x pass, forward: two rows z, z+1 of a plane per wave as one complex
// transform of x_n + i y_n; X_k = (Z_k + conj Z_{512-k})/2, Y_k = (Z_k -
// conj Z_{512-k})/(2i), k = 0..256, each divided by NX = 512 (mpi.f:1490-1500
// scales by 1/NX).  phys [plane][512][512], four [plane][512][257].
constexpr int kFftN = 512, kFftK = 257;

template <class C>
__global__ void __launch_bounds__(256) k_fft512_rows_fwd(const Sc<C> *__restrict__ phys, C *__restrict__ four,
                                                       const C *__restrict__ tw, long long p0, long long np) {
  using S = Sc<C>;
  __shared__ C sm[4][kFftN];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const long long pair = (long long)blockIdx.x * 4 + w;  // row pair in the chunk
  if (pair >= np * (kFftN / 2)) return;                  // whole waves exit together
  const long long plane = p0 + pair / (kFftN / 2);
  const int z = (int)(pair % (kFftN / 2)) * 2;
  const S *ra = phys + (plane * kFftN + z) * kFftN, *rb = ra + kFftN;
  C v[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) v[j] = mkc<C>(ra[lane + 64 * j], rb[lane + 64 * j]);
  fft512_wave<false>(v, sm[w], 1, tw);
#pragma unroll
  for (int j = 0; j < 8; ++j) sm[w][lane + 64 * j] = v[j];
  wave_lds_sync();
  C *oa = four + (plane * kFftN + z) * kFftK, *ob = oa + kFftK;
  const S half = S(0.5), nxd = S(kFftN), rnx = S(1.0 / kFftN);
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    const int k = lane + 64 * j;
    if (k <= kFftN / 2) {
      const C zk = sm[w][k], zm = sm[w][(kFftN - k) & (kFftN - 1)];
      const C X = mkc<C>(half * (zk.x + zm.x), half * (zk.y - zm.y));
      const C Y = mkc<C>(half * (zk.y + zm.y), half * (zm.x - zk.x));
      oa[k] = mkc<C>(div_by_recip(X.x, nxd, rnx), div_by_recip(X.y, nxd, rnx));
      ob[k] = mkc<C>(div_by_recip(Y.x, nxd, rnx), div_by_recip(Y.y, nxd, rnx));
    }
  }
}

// x pass, inverse (the c2r of mpi.f:1600-1604, unnormalised): rows z, z+1
// of the half spectra X, Y to x_n + i y_n = sum_k (X_k + i Y_k) e^{+2 pi i
// nk/512}, with X_{512-k} = conj X_k and the imaginary parts of X_0, X_256
// ignored as a c2r ignores them.
template <class C>
__global__ void __launch_bounds__(256) k_fft512_rows_inv(const C *__restrict__ four, Sc<C> *__restrict__ phys,
                                                       const C *__restrict__ tw, long long p0, long long np,
                                                       long long pout0) {
  using S = Sc<C>;
  __shared__ C sm[4][kFftN];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const long long pair = (long long)blockIdx.x * 4 + w;
  if (pair >= np * (kFftN / 2)) return;
  const long long plane = pair / (kFftN / 2);
  const int z = (int)(pair % (kFftN / 2)) * 2;
  const C *ia = four + ((p0 + plane) * kFftN + z) * kFftK, *ib = ia + kFftK;
  C v[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int k = lane + 64 * j;
    const int kk = (k <= kFftN / 2) ? k : kFftN - k;
    C X = ia[kk], Y = ib[kk];
    if (kk == 0 || kk == kFftN / 2) {
      X.y = S(0);
      Y.y = S(0);
    }
    if (k > kFftN / 2) {
      X.y = -X.y;
      Y.y = -Y.y;
    }
    v[j] = mkc<C>(X.x - Y.y, X.y + Y.x);  // X + i Y
  }
  fft512_wave<true>(v, sm[w], 1, tw);
  S *oa = phys + ((pout0 + plane) * kFftN + z) * kFftN, *ob = oa + kFftN;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    oa[lane + 64 * j] = v[j].x;
    ob[lane + 64 * j] = v[j].y;
  }
}

// z pass: columns kx of [plane][512][257], 8 columns per block of 8 waves,
// staged through an LDS tile [512][9] (one column of padding); wave c
// transforms column c in place in the tile.  Forward divides by NZ = 512
// (mpi.f:1529-1531); inverse is unnormalised.  src/dst may be equal.
constexpr int kColTile = 8, kColPad = kColTile + 1;

template <bool INV, class C>
__global__ void __launch_bounds__(512) k_fft512_cols(const C *__restrict__ src, C *__restrict__ dst,
                                                   const C *__restrict__ tw, long long p0, long long np,
                                                   long long pdst0) {
  using S = Sc<C>;
  extern __shared__ double2 smraw[];
  C *tile = reinterpret_cast<C *>(smraw);
  constexpr int nblk = (kFftK + kColTile - 1) / kColTile;
  const long long plane = blockIdx.x / nblk;
  if (plane >= np) return;
  const int kx0 = (int)(blockIdx.x % nblk) * kColTile;
  const C *in = src + (p0 + plane) * (long long)kFftN * kFftK;
  C *out = dst + (pdst0 + plane) * (long long)kFftN * kFftK;
  const int c = threadIdx.x & 7, zr = threadIdx.x >> 3;  // 64 rows per sweep
  const bool cin = kx0 + c < kFftK;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int z = zr + 64 * j;
    tile[z * kColPad + c] = cin ? in[(long long)z * kFftK + kx0 + c] : mkc<C>(S(0), S(0));
  }
  __syncthreads();
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  {
    C v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = tile[(lane + 64 * j) * kColPad + w];
    fft512_wave<INV>(v, tile + w, kColPad, tw);
    const S nzd = S(kFftN), rnz = S(1.0 / kFftN);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      C o = v[j];
      if (!INV) o = mkc<C>(div_by_recip(o.x, nzd, rnz), div_by_recip(o.y, nzd, rnz));
      tile[(lane + 64 * j) * kColPad + w] = o;
    }
  }
  __syncthreads();
  if (!cin) return;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int z = zr + 64 * j;
    out[(long long)z * kFftK + kx0 + c] = tile[z * kColPad + c];
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {
// Twiddles exp(-2 pi i m/512), m < 512, in file-scope device arrays (one
// copy per device), filled from the host at the first call on each device.
__device__ double2 g_tw64_dev[kFftN];
__device__ float2 g_tw32_dev[kFftN];

std::mutex g_mu;
const void *g_tw64[64] = {nullptr}, *g_tw32[64] = {nullptr};  // their addresses, per device
// staging of the inverse transform, allocated on first use and grown
void *g_stage[64] = {nullptr};
size_t g_stage_bytes[64] = {0};

// the twiddle tables of the current device (computed in long double and
// rounded once; no device memory is allocated)
hipError_t twiddles(bool fp32, const void **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_tw64[dev]) {
    double2 h64[kFftN];
    float2 h32[kFftN];
    for (int m = 0; m < kFftN; ++m) {
      const long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)m / (long double)kFftN;
      const long double cr = cosl(a), ci = sinl(a);
      h64[m] = make_double2((double)cr, (double)ci);
      h32[m] = make_float2((float)cr, (float)ci);
    }
    void *p64 = nullptr, *p32 = nullptr;
    if ((e = hipGetSymbolAddress(&p64, HIP_SYMBOL(g_tw64_dev))) != hipSuccess) return e;
    if ((e = hipGetSymbolAddress(&p32, HIP_SYMBOL(g_tw32_dev))) != hipSuccess) return e;
    if ((e = hipMemcpy(p64, h64, sizeof(h64), hipMemcpyHostToDevice)) != hipSuccess) return e;
    if ((e = hipMemcpy(p32, h32, sizeof(h32), hipMemcpyHostToDevice)) != hipSuccess) return e;
    g_tw32[dev] = p32;
    g_tw64[dev] = p64;
  }
  *out = fp32 ? g_tw32[dev] : g_tw64[dev];
  return hipSuccess;
}

thread_local std::string g_err;
int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}
}  // namespace

const char *batched_set_error(int code, const char *msg) {
  (void)code;
  g_err = msg;
  return g_err.c_str();
}

// the FP64 column tile (73.7 KB) is beyond the default dynamic-LDS limit
hipError_t fft512_lds_limits() {
  static hipError_t once = [] {
    const void *k[4] = {(const void *)k_fft512_cols<false, double2>, (const void *)k_fft512_cols<true, double2>,
                        (const void *)k_fft512_cols<false, float2>, (const void *)k_fft512_cols<true, float2>};
    for (const void *f : k) {
      const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, kLdsMax);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }();
  return once;
}

hipError_t fft512_forward(const void *phys, void *four, long long nplanes, bool fp32, hipStream_t st,
                          long long chunk) {
  const void *tw = nullptr;
  hipError_t e = fft512_lds_limits();
  if (e != hipSuccess) return e;
  e = twiddles(fp32, &tw);
  if (e != hipSuccess) return e;
  constexpr int nblk = (kFftK + kColTile - 1) / kColTile;
  const size_t lds = (size_t)kFftN * kColPad * (fp32 ? sizeof(float2) : sizeof(double2));
  for (long long p0 = 0; p0 < nplanes; p0 += chunk) {
    const long long np = std::min(chunk, nplanes - p0);
    const unsigned rb = (unsigned)((np * (kFftN / 2) + 3) / 4);
    if (fp32) {
      hipLaunchKernelGGL(k_fft512_rows_fwd<float2>, dim3(rb), dim3(256), 0, st, (const float *)phys, (float2 *)four,
                         (const float2 *)tw, p0, np);
      hipLaunchKernelGGL((k_fft512_cols<false, float2>), dim3((unsigned)(np * nblk)), dim3(512), lds, st,
                         (const float2 *)four, (float2 *)four, (const float2 *)tw, p0, np, p0);
    } else {
      hipLaunchKernelGGL(k_fft512_rows_fwd<double2>, dim3(rb), dim3(256), 0, st, (const double *)phys,
                         (double2 *)four, (const double2 *)tw, p0, np);
      hipLaunchKernelGGL((k_fft512_cols<false, double2>), dim3((unsigned)(np * nblk)), dim3(512), lds, st,
                         (const double2 *)four, (double2 *)four, (const double2 *)tw, p0, np, p0);
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

// stage: chunk planes of [512][257] values
hipError_t fft512_inverse(const void *four, void *phys, long long nplanes, bool fp32, hipStream_t st, void *stage,
                          long long chunk) {
  const void *tw = nullptr;
  hipError_t e = fft512_lds_limits();
  if (e != hipSuccess) return e;
  e = twiddles(fp32, &tw);
  if (e != hipSuccess) return e;
  constexpr int nblk = (kFftK + kColTile - 1) / kColTile;
  const size_t lds = (size_t)kFftN * kColPad * (fp32 ? sizeof(float2) : sizeof(double2));
  for (long long p0 = 0; p0 < nplanes; p0 += chunk) {
    const long long np = std::min(chunk, nplanes - p0);
    const unsigned rb = (unsigned)((np * (kFftN / 2) + 3) / 4);
    if (fp32) {
      hipLaunchKernelGGL((k_fft512_cols<true, float2>), dim3((unsigned)(np * nblk)), dim3(512), lds, st,
                         (const float2 *)four, (float2 *)stage, (const float2 *)tw, p0, np, 0LL);
      hipLaunchKernelGGL(k_fft512_rows_inv<float2>, dim3(rb), dim3(256), 0, st, (const float2 *)stage,
                         (float *)phys, (const float2 *)tw, 0LL, np, p0);
    } else {
      hipLaunchKernelGGL((k_fft512_cols<true, double2>), dim3((unsigned)(np * nblk)), dim3(512), lds, st,
                         (const double2 *)four, (double2 *)stage, (const double2 *)tw, p0, np, 0LL);
      hipLaunchKernelGGL(k_fft512_rows_inv<double2>, dim3(rb), dim3(256), 0, st, (const double2 *)stage,
                         (double *)phys, (const double2 *)tw, 0LL, np, p0);
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace dh

using namespace dh;

namespace {
constexpr long long kFftChunk = 16;  // planes per x/z pass pair: the pair's data stays in the last-level cache
size_t plane_four_bytes(bool fp32) { return (size_t)kFftN * kFftK * (fp32 ? sizeof(float2) : sizeof(double2)); }
}  // namespace

extern "C" {

const char *dh_batched_last_error(void) { return g_err.c_str(); }

int dh_cn_solve_batched_dev(const double *sl, const double *su, const double *alpha, const void *rhs, void *x,
                            int32_t ny, int64_t nsys, int32_t fp32, void *stream) {
  if (!sl || !su || !alpha || !rhs || !x) return fail(DH_EARG, "null pointer");
  if (ny < 2 || ny > kCnMaxNy) return fail(DH_EARG, "ny must lie in 2..288");
  if (nsys <= 0) return fail(DH_EARG, "nsys must be positive");
  if ((nsys + 7) / 8 > 0x7fffffffLL || (long long)ny * nsys > (1LL << 40)) return fail(DH_EARG, "too many systems");
  hipStream_t st = static_cast<hipStream_t>(stream);
  // 16 lanes of 9 rows per system (ny <= 144) or 32 (ny <= 288)
  const dim3 b(kCnThreads);
  if (ny <= 16 * kCnRows) {
    const dim3 g((unsigned)((nsys + 15) / 16));  // 4 waves of 4 systems
    if (fp32)
      hipLaunchKernelGGL((k_cn_solve<16, float2>), g, b, 0, st, sl, su, alpha, (const float2 *)rhs, (float2 *)x,
                         (int)ny, (long long)nsys);
    else
      hipLaunchKernelGGL((k_cn_solve<16, double2>), g, b, 0, st, sl, su, alpha, (const double2 *)rhs, (double2 *)x,
                         (int)ny, (long long)nsys);
  } else {
    const dim3 g((unsigned)((nsys + 7) / 8));  // 4 waves of 2 systems
    if (fp32)
      hipLaunchKernelGGL((k_cn_solve<32, float2>), g, b, 0, st, sl, su, alpha, (const float2 *)rhs, (float2 *)x,
                         (int)ny, (long long)nsys);
    else
      hipLaunchKernelGGL((k_cn_solve<32, double2>), g, b, 0, st, sl, su, alpha, (const double2 *)rhs, (double2 *)x,
                         (int)ny, (long long)nsys);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(DH_EHIP, hipGetErrorString(e));
  return DH_OK;
}

int dh_fft512_xz_to_fourier_dev(const void *phys, void *four, int64_t nplanes, int32_t fp32, void *stream) {
  if (!phys || !four) return fail(DH_EARG, "null pointer");
  if (nplanes <= 0) return fail(DH_EARG, "nplanes must be positive");
  const hipError_t e = fft512_forward(phys, four, nplanes, fp32 != 0, static_cast<hipStream_t>(stream), kFftChunk);
  if (e != hipSuccess) return fail(DH_EHIP, hipGetErrorString(e));
  return DH_OK;
}

int dh_fft512_xz_to_physical_dev(const void *four, void *phys, int64_t nplanes, int32_t fp32, void *stream) {
  if (!phys || !four) return fail(DH_EARG, "null pointer");
  if (nplanes <= 0) return fail(DH_EARG, "nplanes must be positive");
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(DH_EHIP, hipGetErrorString(e));
  if (dev < 0 || dev >= 64) return fail(DH_EARG, "device index beyond 63");
  // one staging buffer of kFftChunk FP64 planes per device, made at the first call
  const long long chunk = kFftChunk;
  const size_t need = (size_t)chunk * plane_four_bytes(false);
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_stage[dev]) {
      if ((e = hipMalloc(&g_stage[dev], need)) != hipSuccess) return fail(DH_EHIP, hipGetErrorString(e));
      g_stage_bytes[dev] = need;
    }
  }
  e = fft512_inverse(four, phys, nplanes, fp32 != 0, static_cast<hipStream_t>(stream), g_stage[dev], chunk);
  if (e != hipSuccess) return fail(DH_EHIP, hipGetErrorString(e));
  return DH_OK;
}

}  // extern "C"
