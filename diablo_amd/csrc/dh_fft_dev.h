// Device-side FFT engine: batched in-place radix-2/3/4/8 transforms of one
// x-z plane held in LDS, shared by the plane-transform kernels and the fused nonlinear
// kernel.
//
// Plane layout in LDS: A[z][0..LW-1] complex, LW = NX/2.
//   physical view: A[z][m] = (x_{2m}, x_{2m+1}) of row z (a real row of NX
//                  doubles is exactly LW complex values);
//   spectral view: A[z][kx] for kx <= NKX after the x transform.
//
// The real-to-complex x transform of length NX is done as a complex
// transform of length NX/2 plus the usual even/odd split; FFTW's r2c/c2r
// (fft.f:73-76) compute the same unnormalised DFTs.
#pragma once

#include "dh_internal.h"

namespace dh {

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cscale(double s, double2 a) { return make_double2(s * a.x, s * a.y); }
// (i*k) * z with the component order of a COMPLEX*16 product (0,k)*(x,y)
__device__ __forceinline__ double2 cmul_ik(double k, double2 z) { return make_double2(-(k * z.y), k * z.x); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ double2 cmul_mi(double2 a) {
  return INV ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

template <bool INV>
__device__ __forceinline__ double2 tw_get(const double2 *tw, int idx) {
  double2 w = tw[idx];
  return INV ? cconj(w) : w;
}

template <int R, bool INV>
__device__ __forceinline__ void dft_small(double2 *v) {
  if constexpr (R == 2) {
    double2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 3) {
    const double c1 = -0.5;
    const double s1 = INV ? 0.86602540378443864676 : -0.86602540378443864676;
    double2 t1 = cadd(v[1], v[2]);
    double2 t2 = csub(v[1], v[2]);
    double2 m = make_double2(v[0].x + c1 * t1.x, v[0].y + c1 * t1.y);
    // i*s1*t2
    double2 it = make_double2(-s1 * t2.y, s1 * t2.x);
    v[0] = cadd(v[0], t1);
    v[1] = cadd(m, it);
    v[2] = csub(m, it);
  } else if constexpr (R == 4) {
    double2 a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
    double2 c = cadd(v[1], v[3]), d = cmul_mi<INV>(csub(v[1], v[3]));
    v[0] = cadd(a, c);
    v[2] = csub(a, c);
    v[1] = cadd(b, d);
    v[3] = csub(b, d);
  } else if constexpr (R == 8) {
    double2 e[4] = {v[0], v[2], v[4], v[6]};
    double2 o[4] = {v[1], v[3], v[5], v[7]};
    dft_small<4, INV>(e);
    dft_small<4, INV>(o);
    const double h = 0.70710678118654752440;
    // w^1 = h(1 -/+ i), w^2 = -/+ i, w^3 = h(-1 -/+ i)
    double2 o1 = INV ? make_double2(h * (o[1].x - o[1].y), h * (o[1].x + o[1].y))
                     : make_double2(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
    double2 o2 = cmul_mi<INV>(o[2]);
    double2 o3 = INV ? make_double2(-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y))
                     : make_double2(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
  }
}

// One in-place decimation-in-frequency stage (radix R, block length L,
// span S = L/R) over `ntrans` transforms of length N.  Each butterfly writes
// its outputs back to the addresses it read, so a stage needs no read/write
// barrier; after all stages position p holds output index digrev(p) (the
// pos_* tables of Geo give the position of each index).
//   rows (COLS=false): transform t is LDS row t, element i at A[t*lw + i]
//   cols (COLS=true) : transform t is column colpos[t] (or t), element i at
//                      A[i*lw + col]; consecutive threads take consecutive
//                      columns so LDS accesses stay on distinct banks.
template <int R, bool INV, bool COLS>
__device__ __forceinline__ void dif_stage(double2 *A, int N, int L, int ntrans, int lw, const int *colpos,
                                          const double2 *tw) {
  const int S = L / R;
  const int nb = N / R;
  const int total = ntrans * nb;
  const int twm = N / L;
  for (int gi = threadIdx.x; gi < total; gi += blockDim.x) {
    int t, b;
    if (COLS) {
      t = gi % ntrans;
      b = gi / ntrans;
    } else {
      t = gi / nb;
      b = gi - t * nb;
    }
    const int g = b / S, j = b - g * S;
    const int i0 = g * L + j;
    double2 *base;
    int es;
    if (COLS) {
      base = A + (colpos ? colpos[t] : t) + i0 * lw;
      es = S * lw;
    } else {
      base = A + t * lw + i0;
      es = S;
    }
    double2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = base[r * es];
    dft_small<R, INV>(v);
    if (j > 0) {
#pragma unroll
      for (int q = 1; q < R; ++q) v[q] = cmul(v[q], tw_get<INV>(tw, q * j * twm));
    }
#pragma unroll
    for (int q = 0; q < R; ++q) base[q * es] = v[q];
  }
  __syncthreads();
}

template <bool INV, bool COLS>
__device__ void lds_fft(double2 *A, const FftPlan &pl, int ntrans, int lw, const int *colpos, const double2 *tw) {
  int L = pl.n;
  for (int s = 0; s < pl.nst; ++s) {
    const int R = pl.rad[s];
    switch (R) {
      case 8: dif_stage<8, INV, COLS>(A, pl.n, L, ntrans, lw, colpos, tw); break;
      case 4: dif_stage<4, INV, COLS>(A, pl.n, L, ntrans, lw, colpos, tw); break;
      case 3: dif_stage<3, INV, COLS>(A, pl.n, L, ntrans, lw, colpos, tw); break;
      default: dif_stage<2, INV, COLS>(A, pl.n, L, ntrans, lw, colpos, tw); break;
    }
    L /= R;
  }
}

// Forward x transform of every row (FFT_XZ_TO_FOURIER's r2c, mpi.f:1490-1506):
// A holds the physical plane (natural order); on exit the LDS slot
// posr[k] of each row holds X_k / NX for k <= NKX.
__device__ void plane_rows_fwd(double2 *A, const Geo &g) {
  const int h = g.lw;  // NX/2
  lds_fft<false, false>(A, g.prow, g.nz, g.lw, nullptr, g.tw_row);
  // even/odd split: X_k = (Z_k + conj(Z_{h-k}))/2 - i w^-k (Z_k - conj(Z_{h-k}))/2
  const int npair = h / 2 + 1;  // k = 0..h/2 paired with h-k
  const int total = g.nz * npair;
  const int *pos = g.posr;
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    const int z = e / npair, k = e - z * npair;
    double2 *row = A + z * g.lw;
    const int k2 = (k == 0) ? 0 : h - k;
    const int pk = pos[k], pk2 = pos[k2];
    const double2 zk = row[pk];
    const double2 zm = row[pk2];
    {
      double2 a = make_double2(0.5 * (zk.x + zm.x), 0.5 * (zk.y - zm.y));  // (Z_k + conj Z_{h-k})/2
      double2 b = make_double2(0.5 * (zk.x - zm.x), 0.5 * (zk.y + zm.y));  // (Z_k - conj Z_{h-k})/2
      double2 wb = cmul(g.tw_real[k], b);                                 // w = e^{-2 pi i k/NX}
      double2 X = make_double2(a.x + wb.y, a.y - wb.x);                  // a - i*w*b
      if (k <= g.nkx) row[pk] = make_double2(X.x / g.nx, X.y / g.nx);
    }
    if (k != 0 && k2 != k && k2 <= g.nkx) {
      double2 a = make_double2(0.5 * (zm.x + zk.x), 0.5 * (zm.y - zk.y));
      double2 b = make_double2(0.5 * (zm.x - zk.x), 0.5 * (zm.y + zk.y));
      double2 wb = cmul(g.tw_real[k2], b);
      double2 X = make_double2(a.x + wb.y, a.y - wb.x);
      row[pk2] = make_double2(X.x / g.nx, X.y / g.nx);
    }
  }
  __syncthreads();
}

// Inverse x transform of every row (FFT_XZ_TO_PHYSICAL's c2r, mpi.f:1600-1604):
// on entry A[z][k], k <= NKX in natural order (higher modes and the Nyquist
// term are zero, the imaginary part of the k=0 term is ignored); on exit
// z_m = (x_{2m}, x_{2m+1}) sits in slot posr[m] of each row.
__device__ void plane_rows_inv(double2 *A, const Geo &g) {
  const int h = g.lw;
  const int npair = h / 2 + 1;
  const int total = g.nz * npair;
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    const int z = e / npair, k = e - z * npair;
    double2 *row = A + z * g.lw;
    const int k2 = h - k;  // partner (k=0 pairs with the Nyquist term)
    double2 xk = (k <= g.nkx) ? row[k] : make_double2(0.0, 0.0);
    double2 xm = (k2 <= g.nkx && k2 < h) ? row[k2] : make_double2(0.0, 0.0);
    if (k == 0) xk.y = 0.0;
    // Z_k = (X_k + conj X_{h-k}) + i w^k (X_k - conj X_{h-k}),  w = e^{+2 pi i/NX}
    {
      double2 a = make_double2(xk.x + xm.x, xk.y - xm.y);
      double2 b = make_double2(xk.x - xm.x, xk.y + xm.y);
      double2 wb = cmul(cconj(g.tw_real[k]), b);
      if (k < h) row[k] = make_double2(a.x - wb.y, a.y + wb.x);
    }
    if (k != 0 && k2 != k) {
      double2 a = make_double2(xm.x + xk.x, xm.y - xk.y);
      double2 b = make_double2(xm.x - xk.x, xm.y + xk.y);
      double2 wb = cmul(cconj(g.tw_real[k2]), b);
      row[k2] = make_double2(a.x - wb.y, a.y + wb.x);
    }
  }
  __syncthreads();
  lds_fft<true, false>(A, g.prow, g.nz, g.lw, nullptr, g.tw_row);
}

// z transforms over the de-aliased columns kx = 0..NKX.  Forward: the
// columns sit at the row-transform output slots posr[kx]; inverse: natural.
// Output row position p holds z (or kz) index digrev(p); posz maps back.
__device__ __forceinline__ void plane_cols_fwd(double2 *A, const Geo &g) {
  lds_fft<false, true>(A, g.pcol, g.nkx1, g.lw, g.posr, g.tw_col);
}
__device__ __forceinline__ void plane_cols_inv(double2 *A, const Geo &g) {
  lds_fft<true, true>(A, g.pcol, g.nkx1, g.lw, nullptr, g.tw_col);
}

// LDS slot of the forward-transformed mode (kx, packed kz index kzp).
__device__ __forceinline__ int fwd_slot(const Geo &g, int kx, int kzp);

// Row of the unpacked z spectrum that holds packed index kzp (mpi.f:1564-1568).
__device__ __forceinline__ int kz_row(const Geo &g, int kzp) {
  return (kzp <= g.nkz) ? kzp : (g.nz - 1 + (kzp - g.nkz) - g.nkz);
}

__device__ __forceinline__ int fwd_slot(const Geo &g, int kx, int kzp) {
  return g.posz[kz_row(g, kzp)] * g.lw + g.posr[kx];
}

// LDS slot of physical pair m (x = 2m, 2m+1) of row z after the inverse.
__device__ __forceinline__ int inv_slot(const Geo &g, int z, int m) { return g.posz[z] * g.lw + g.posr[m]; }

// Scatter a de-aliased Fourier plane into A (unpack + zero fill of the
// columns 0..NKX, mpi.f:1562-1574).  `src` may be null (all zero).
__device__ void plane_unpack(double2 *A, const Geo &g, const double2 *src) {
  const int total = g.nz * g.nkx1;
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    const int zr = e / g.nkx1, kx = e - zr * g.nkx1;
    int kzp = -1;
    if (zr <= g.nkz)
      kzp = zr;
    else if (zr >= g.nz - g.nkz)
      kzp = zr - (g.nz - 1 - g.nkz) + g.nkz;
    A[zr * g.lw + kx] = (kzp >= 0 && src) ? src[kzp * g.nkx1 + kx] : make_double2(0.0, 0.0);
  }
  __syncthreads();
}

}  // namespace dh
