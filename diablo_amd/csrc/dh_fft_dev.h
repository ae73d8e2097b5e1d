// Device-side FFT engine: batched in-place radix-2/3/4/8 transforms of one
// x-z plane held in LDS, shared by the plane-transform kernels and the fused
// nonlinear kernel.
//
// Plane layout in LDS: A[z][0..LW-1] complex, LW = NX/2, row stride LS =
// LW + 1 complex (16 B of padding per row, so that the same slot of
// consecutive rows falls on different LDS banks).
//   physical view: row z holds (x_{2m}, x_{2m+1}) in natural slot m;
//   spectral view: A[.][kx] for kx <= NKX after the x transform.
// Every transform runs between natural order and a digit-reversed order
// that is absorbed by a table where the plane meets HBM:
//   forward: physical pairs are loaded into digit-reversed row slots (posr),
//            rows by decimation in time (natural output), even/odd split in
//            place, columns by decimation in frequency (rows come out in
//            the order posz, read back by the pack);
//   inverse: the unpack writes rows in the order posc, columns by
//            decimation in time (natural rows), even/odd split in place,
//            rows by decimation in frequency (pairs come out in the order
//            posrd, read back by the store).
// All stages run in place: every butterfly writes back to the slots it
// read, so a stage needs one barrier and only R complex values per thread.
// Work is spread across the batch dimension first (consecutive threads take
// consecutive rows for a row pass and consecutive columns for a column
// pass), so the lanes of a wave touch distinct banks and read the same
// twiddle (an LDS broadcast).
//
// The real x transform of length NX is a complex transform of length NX/2
// plus the even/odd split; FFTW's r2c/c2r (fft.f:73-76) compute the same
// unnormalised DFTs.
#pragma once

#include "dh_internal.h"

namespace dh {

// Complex element of a plane: double2 (FP64 transforms) or float2 (the
// mixed form's FP32 transforms).  Sc<C> is its real type and mkc builds one.
template <class C>
struct CxOf;
template <>
struct CxOf<double2> {
  using S = double;
};
template <>
struct CxOf<float2> {
  using S = float;
};
template <class C>
using Sc = typename CxOf<C>::S;
template <class C>
__device__ __forceinline__ C mkc(Sc<C> x, Sc<C> y) {
  C r;
  r.x = x;
  r.y = y;
  return r;
}
// an FP64 complex value as C (rounded to nearest for float2)
template <class C>
__device__ __forceinline__ C cvt(double2 v) {
  return mkc<C>(static_cast<Sc<C>>(v.x), static_cast<Sc<C>>(v.y));
}

template <class C>
__device__ __forceinline__ C cadd(C a, C b) {
  return mkc<C>(a.x + b.x, a.y + b.y);
}
template <class C>
__device__ __forceinline__ C csub(C a, C b) {
  return mkc<C>(a.x - b.x, a.y - b.y);
}
template <class C>
__device__ __forceinline__ C cmul(C a, C b) {
  return mkc<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <class C>
__device__ __forceinline__ C cconj(C a) {
  return mkc<C>(a.x, -a.y);
}
// (i*k) * z with the component order of a COMPLEX*16 product (0,k)*(x,y)
template <class C>
__device__ __forceinline__ C cmul_ik(Sc<C> k, C z) {
  return mkc<C>(-(k * z.y), k * z.x);
}
// multiply by -i (forward) or +i (inverse)
template <bool INV, class C>
__device__ __forceinline__ C cmul_mi(C a) {
  return INV ? mkc<C>(-a.y, a.x) : mkc<C>(a.y, -a.x);
}

// Workgroup barrier that orders LDS only: global loads already issued stay
// in flight across it (a plain __syncthreads also waits for them).  None of
// these kernels exchanges data between the threads of a block through
// global memory.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// n / b for a divisor b whose correctly rounded reciprocal r = 1/b is known:
// q0 = n r, e = n - b q0 (exact by fma), q = q0 + e r.  With r correctly
// rounded this is the correctly rounded quotient (Markstein's theorem) away
// from overflow and underflow, i.e. the same double as n / b, in three
// dependent operations instead of a full division.
template <class S>
__device__ __forceinline__ S div_by_recip(S n, S b, S r) {
  const S q0 = n * r;
  const S e = fma(-b, q0, n);
  return fma(e, r, q0);
}

// n / d for n, d < 2^16 by multiply-high (m = ceil(2^32 / d)).
__device__ __forceinline__ unsigned fdiv(unsigned n, unsigned long long m) {
  return (unsigned)(((unsigned long long)n * m) >> 32);
}

template <int R, bool INV, class C>
__device__ __forceinline__ void dft_small(C *v) {
  using S = Sc<C>;
  if constexpr (R == 2) {
    C a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 3) {
    const S c1 = -0.5;
    const S s1 = INV ? 0.86602540378443864676 : -0.86602540378443864676;
    C t1 = cadd(v[1], v[2]);
    C t2 = csub(v[1], v[2]);
    C m = mkc<C>(v[0].x + c1 * t1.x, v[0].y + c1 * t1.y);
    C it = mkc<C>(-s1 * t2.y, s1 * t2.x);  // i*s1*t2
    v[0] = cadd(v[0], t1);
    v[1] = cadd(m, it);
    v[2] = csub(m, it);
  } else if constexpr (R == 4) {
    C a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
    C c = cadd(v[1], v[3]), d = cmul_mi<INV>(csub(v[1], v[3]));
    v[0] = cadd(a, c);
    v[2] = csub(a, c);
    v[1] = cadd(b, d);
    v[3] = csub(b, d);
  } else if constexpr (R == 8) {
    C e[4] = {v[0], v[2], v[4], v[6]};
    C o[4] = {v[1], v[3], v[5], v[7]};
    dft_small<4, INV>(e);
    dft_small<4, INV>(o);
    const S h = 0.70710678118654752440;
    // w^1 = h(1 -/+ i), w^2 = -/+ i, w^3 = h(-1 -/+ i)
    C o1 = INV ? mkc<C>(h * (o[1].x - o[1].y), h * (o[1].x + o[1].y))
                     : mkc<C>(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
    C o2 = cmul_mi<INV>(o[2]);
    C o3 = INV ? mkc<C>(-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y))
                     : mkc<C>(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
  } else if constexpr (R == 16) {
    // 16 = 4 x 4: DFT4 over n1 of x[4 n1 + n2], twiddle W16^{n2 k1}, DFT4
    // over n2; output X[k1 + 4 k2] in natural order
    C y[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
      C t[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
      dft_small<4, INV>(t);
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) y[n2][k1] = t[k1];
    }
    const S c1 = 0.92387953251128675613, s1 = 0.38268343236508977173, h = 0.70710678118654752440;
    // W16^j = cos(2 pi j/16) -/+ i sin(2 pi j/16) for j = 1, 2, 3, 4, 6, 9
    const S sg = INV ? S(1) : S(-1);
    auto tw = [&](C a, S c, S sn) {  // a * (c + i sg sn)
      return mkc<C>(a.x * c - a.y * (sg * sn), a.x * (sg * sn) + a.y * c);
    };
    y[1][1] = tw(y[1][1], c1, s1);
    y[1][2] = tw(y[1][2], h, h);
    y[1][3] = tw(y[1][3], s1, c1);
    y[2][1] = tw(y[2][1], h, h);
    y[2][2] = cmul_mi<INV>(y[2][2]);
    y[2][3] = tw(y[2][3], -h, h);
    y[3][1] = tw(y[3][1], s1, c1);
    y[3][2] = tw(y[3][2], -h, h);
    y[3][3] = tw(y[3][3], -c1, -s1);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      C t[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
      dft_small<4, INV>(t);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
    }
  }
}

// Transform tables staged in LDS behind the plane (see load_tables).
template <class C>
struct TablesT {
  C *twr;        // exp(-2 pi i m/LW), m < LW
  C *twc;        // exp(-2 pi i m/NZ), m < NZ
  C *twx;        // exp(-2 pi i k/NX), k <= LW
  int *posr;     // DIT input slot of row index m
  int *posrd;    // slot of row output index k after the DIF
  int *posz;     // slot of column output index k after the DIF
  int *posc;     // DIT input slot of column index k
};
using Tables = TablesT<double2>;

// Copies the tables into LDS at p; returns pointers into it.
template <class C>
__device__ TablesT<C> load_tables_at(C *p, const Geo &g) {
  TablesT<C> t;
  t.twr = p;
  t.twc = p + g.lw;
  t.twx = p + g.lw + g.nz;
  int *q = reinterpret_cast<int *>(p + g.lw + g.nz + g.lw + 1);
  t.posr = q;
  t.posrd = q + g.lw;
  t.posz = q + 2 * g.lw;
  t.posc = q + 2 * g.lw + g.nz;
  for (int i = threadIdx.x; i < g.lw; i += blockDim.x) {
    t.twr[i] = cvt<C>(g.tw_row[i]);
    t.posr[i] = g.posr[i];
    t.posrd[i] = g.posrd[i];
  }
  for (int i = threadIdx.x; i < g.nz; i += blockDim.x) {
    t.twc[i] = cvt<C>(g.tw_col[i]);
    t.posz[i] = g.posz[i];
    t.posc[i] = g.posc[i];
  }
  for (int i = threadIdx.x; i <= g.lw; i += blockDim.x) t.twx[i] = cvt<C>(g.tw_real[i]);
  lds_barrier();
  return t;
}
// The same behind an LDS plane that starts at smem.
template <class C>
__device__ TablesT<C> load_tables(C *smem, const Geo &g) {
  return load_tables_at(smem + g.nz * g.ls, g);
}

// One in-place stage of radix R over `ntrans` transforms of length N.
//   DIT: twiddles W_L^{r j} on the inputs, then the butterfly (L grows);
//   DIF: butterfly, then twiddles W_L^{q j} on the outputs (L shrinks).
// rows (COLS=false): transform t is LDS row t, element i at A[t*ls + i];
// cols (COLS=true):  transform t is column t, element i at A[i*ls + t].
// Consecutive threads take consecutive transforms t (fnt = reciprocal of
// ntrans) with the same butterfly.
template <int R, bool INV, bool COLS, bool DIT, class C>
__device__ __forceinline__ void fft_stage(C *A, const FftStage &st, int N, int ntrans, unsigned long long fnt, int ls,
                                          const C *tw) {
  const int S = st.S;
  const int nb = N / R;
  const int total = ntrans * nb;
  for (int gi = threadIdx.x; gi < total; gi += blockDim.x) {
    const int b = (int)fdiv((unsigned)gi, fnt);
    const int t = gi - b * ntrans;
    const int g = (int)fdiv((unsigned)b, st.fS), j = b - g * S;
    const int i0 = g * st.L + j;
    C *base;
    int es;
    if (COLS) {
      base = A + t + i0 * ls;
      es = S * ls;
    } else {
      base = A + t * ls + i0;
      es = S;
    }
    C v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = base[r * es];
    if (DIT && j > 0) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        C w = tw[r * j * st.twm];
        v[r] = cmul(v[r], INV ? cconj(w) : w);
      }
    }
    dft_small<R, INV>(v);
    if (!DIT && j > 0) {
#pragma unroll
      for (int q = 1; q < R; ++q) {
        C w = tw[q * j * st.twm];
        v[q] = cmul(v[q], INV ? cconj(w) : w);
      }
    }
#pragma unroll
    for (int q = 0; q < R; ++q) base[q * es] = v[q];
  }
  lds_barrier();
}

template <bool INV, bool COLS, bool DIT, class C>
__device__ void lds_fft(C *A, const FftPlan &pl, int ntrans, unsigned long long fnt, int ls, const C *tw) {
#pragma unroll 1
  for (int s = 0; s < pl.nst; ++s) {
    const FftStage &st = DIT ? pl.dit[s] : pl.dif[s];
    switch (st.R) {
      case 16: fft_stage<16, INV, COLS, DIT>(A, st, pl.n, ntrans, fnt, ls, tw); break;
      case 8: fft_stage<8, INV, COLS, DIT>(A, st, pl.n, ntrans, fnt, ls, tw); break;
      case 4: fft_stage<4, INV, COLS, DIT>(A, st, pl.n, ntrans, fnt, ls, tw); break;
      case 3: fft_stage<3, INV, COLS, DIT>(A, st, pl.n, ntrans, fnt, ls, tw); break;
      default: fft_stage<2, INV, COLS, DIT>(A, st, pl.n, ntrans, fnt, ls, tw); break;
    }
  }
}

// LDS slot through which physical pair m of row z enters the forward transform.
template <class TT>
__device__ __forceinline__ int phys_slot(const TT &T, const Geo &g, int z, int m) {
  return z * g.ls + T.posr[m];
}

// Work split for filling the digit-reversed input slots from natural-order
// HBM rows: a wave covers a tile of 8 rows x 8 consecutive pairs, lane =
// 8 * pair + row.  Each row segment is one 128 B read, and for radix-8 row
// plans the 16 lanes of an LDS cycle land on 16 distinct banks.
struct LoadTile {
  int ntm, ntiles;
  unsigned long long fntm;
  __device__ LoadTile(const Geo &g) {
    ntm = (g.lw + 7) >> 3;
    ntiles = ((g.nz + 7) >> 3) * ntm;
    fntm = ((1ull << 32) + ntm - 1) / ntm;
  }
  // (z, m) of this lane in tile t; false when outside the plane
  __device__ __forceinline__ bool at(const Geo &g, int t, int &z, int &m) const {
    const int lane = threadIdx.x & 63;
    const int tz = (int)fdiv((unsigned)t, fntm), tm = t - tz * ntm;
    z = (tz << 3) + (lane & 7);
    m = (tm << 3) + (lane >> 3);
    return z < g.nz && m < g.lw;
  }
};

// Forward x transform of every row (FFT_XZ_TO_FOURIER's r2c, mpi.f:1490-1506).
// On entry row z holds (x_{2m}, x_{2m+1}) in slot posr[m]; on exit slot k
// holds X_k / NX for k <= NKX (natural order).
// nz rows (fnz: the reciprocal for fdiv) instead of the plane's NZ: a chunk
// of a plane's rows.
template <class C>
__device__ void plane_rows_fwd(C *A, const Geo &g, const TablesT<C> &T, int nz, unsigned long long fnz) {
  using S = Sc<C>;
  const int h = g.lw;  // NX/2
  lds_fft<false, false, true>(A, g.prow, nz, fnz, g.ls, T.twr);
  // even/odd split: X_k = (Z_k + conj(Z_{h-k}))/2 - i w^-k (Z_k - conj(Z_{h-k}))/2
  const int npair = h / 2 + 1;  // k = 0..h/2 paired with h-k
  const int total = nz * npair;
  const S nxd = S(g.dnx), rnx = S(g.rnx);  // X / NX as the correctly rounded quotient
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    const int k = (int)fdiv((unsigned)e, fnz), z = e - k * nz;
    C *row = A + z * g.ls;
    const int k2 = (k == 0) ? 0 : h - k;
    const C zk = row[k];
    const C zm = row[k2];
    {
      C a = mkc<C>(S(0.5) * (zk.x + zm.x), S(0.5) * (zk.y - zm.y));  // (Z_k + conj Z_{h-k})/2
      C b = mkc<C>(S(0.5) * (zk.x - zm.x), S(0.5) * (zk.y + zm.y));  // (Z_k - conj Z_{h-k})/2
      C wb = cmul(T.twx[k], b);                                     // w = e^{-2 pi i k/NX}
      C X = mkc<C>(a.x + wb.y, a.y - wb.x);                  // a - i*w*b
      if (k <= g.nkx) row[k] = mkc<C>(div_by_recip(X.x, nxd, rnx), div_by_recip(X.y, nxd, rnx));
    }
    if (k != 0 && k2 != k && k2 <= g.nkx) {
      C a = mkc<C>(S(0.5) * (zm.x + zk.x), S(0.5) * (zm.y - zk.y));
      C b = mkc<C>(S(0.5) * (zm.x - zk.x), S(0.5) * (zm.y + zk.y));
      C wb = cmul(T.twx[k2], b);
      C X = mkc<C>(a.x + wb.y, a.y - wb.x);
      row[k2] = mkc<C>(div_by_recip(X.x, nxd, rnx), div_by_recip(X.y, nxd, rnx));
    }
  }
  lds_barrier();
}

// Inverse x transform of every row (FFT_XZ_TO_PHYSICAL's c2r, mpi.f:1600-1604).
// On entry slot k of each row holds X_k for k <= NKX (higher modes and the
// Nyquist term are zero; the imaginary part of the k=0 term is ignored); on
// exit slot posrd[m] holds (x_{2m}, x_{2m+1}).
template <class C>
__device__ void plane_rows_inv(C *A, const Geo &g, const TablesT<C> &T, int nz, unsigned long long fnz) {
  const int h = g.lw;
  const int npair = h / 2 + 1;
  const int total = nz * npair;
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    const int k = (int)fdiv((unsigned)e, fnz), z = e - k * nz;
    C *row = A + z * g.ls;
    const int k2 = h - k;  // partner (k=0 pairs with the Nyquist term)
    C xk = (k <= g.nkx) ? row[k] : mkc<C>(0.0, 0.0);
    C xm = (k2 <= g.nkx && k2 < h) ? row[k2] : mkc<C>(0.0, 0.0);
    if (k == 0) xk.y = 0.0;
    // Z_k = (X_k + conj X_{h-k}) + i w^k (X_k - conj X_{h-k}),  w = e^{+2 pi i/NX}
    {
      C a = mkc<C>(xk.x + xm.x, xk.y - xm.y);
      C b = mkc<C>(xk.x - xm.x, xk.y + xm.y);
      C wb = cmul(cconj(T.twx[k]), b);
      if (k < h) row[k] = mkc<C>(a.x - wb.y, a.y + wb.x);
    }
    if (k != 0 && k2 != k) {
      C a = mkc<C>(xm.x + xk.x, xm.y - xk.y);
      C b = mkc<C>(xm.x - xk.x, xm.y + xk.y);
      C wb = cmul(cconj(T.twx[k2]), b);
      row[k2] = mkc<C>(a.x - wb.y, a.y + wb.x);
    }
  }
  lds_barrier();
  lds_fft<true, false, false>(A, g.prow, nz, fnz, g.ls, T.twr);
}
template <class C>
__device__ __forceinline__ void plane_rows_fwd(C *A, const Geo &g, const TablesT<C> &T) {
  plane_rows_fwd(A, g, T, g.nz, g.f_nz);
}
template <class C>
__device__ __forceinline__ void plane_rows_inv(C *A, const Geo &g, const TablesT<C> &T) {
  plane_rows_inv(A, g, T, g.nz, g.f_nz);
}

// z transforms over the de-aliased columns kx = 0..NKX: forward by DIF
// (natural rows in, row of index k at posz[k] out), inverse by DIT (row of
// index k at posc[k] in, natural rows out).
template <class C>
__device__ __forceinline__ void plane_cols_fwd(C *A, const Geo &g, const TablesT<C> &T) {
  lds_fft<false, true, false>(A, g.pcol, g.nkx1, g.f_nkx1, g.ls, T.twc);
}
template <class C>
__device__ __forceinline__ void plane_cols_inv(C *A, const Geo &g, const TablesT<C> &T) {
  lds_fft<true, true, true>(A, g.pcol, g.nkx1, g.f_nkx1, g.ls, T.twc);
}

// Row of the unpacked z spectrum that holds packed index kzp (mpi.f:1564-1568).
__device__ __forceinline__ int kz_row(const Geo &g, int kzp) {
  return (kzp <= g.nkz) ? kzp : (g.nz - 1 + (kzp - g.nkz) - g.nkz);
}

// LDS slot of the forward-transformed mode (kx, packed kz index kzp).
template <class TT>
__device__ __forceinline__ int fwd_slot(const Geo &g, const TT &T, int kx, int kzp) {
  return T.posz[kz_row(g, kzp)] * g.ls + kx;
}

// LDS slot of physical pair m (x = 2m, 2m+1) of row z after the inverse.
template <class TT>
__device__ __forceinline__ int inv_slot(const Geo &g, const TT &T, int z, int m) {
  return z * g.ls + T.posrd[m];
}

// LDS row of the unpacked z spectrum entering the inverse column transform
// at natural index zr.
template <class TT>
__device__ __forceinline__ int inv_row(const TT &T, int zr) {
  return T.posc[zr];
}

// Scatter a de-aliased Fourier plane into A (unpack + zero fill of the
// columns 0..NKX, mpi.f:1562-1574).  `src` may be null (all zero);
// drop_mean leaves out the (0,0) coefficient (the plane mean).  Loads are
// issued in batches of 8 per thread before any LDS store, so their
// latencies overlap.
template <class C>
__device__ void plane_unpack(C *A, const Geo &g, const TablesT<C> &T, const double2 *src, bool drop_mean = false) {
  constexpr int B = 8;
  const int total = g.nz * g.nkx1;
  for (int base = threadIdx.x; base < total; base += B * blockDim.x) {
    C v[B];
    int slot[B];
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int e = base + b * blockDim.x;
      slot[b] = -1;
      if (e < total) {
        const int zr = (int)fdiv((unsigned)e, g.f_nkx1), kx = e - zr * g.nkx1;
        int kzp = -1;
        if (zr <= g.nkz)
          kzp = zr;
        else if (zr >= g.nz - g.nkz)
          kzp = zr - (g.nz - 1 - g.nkz) + g.nkz;
        slot[b] = inv_row(T, zr) * g.ls + kx;
        v[b] = (kzp >= 0 && src && !(drop_mean && kzp == 0 && kx == 0)) ? cvt<C>(src[kzp * g.nkx1 + kx]) : mkc<C>(0, 0);
      }
    }
#pragma unroll
    for (int b = 0; b < B; ++b)
      if (slot[b] >= 0) A[slot[b]] = v[b];
  }
  lds_barrier();
}

// Physical plane (natural order in HBM) into the digit-reversed input slots.
template <class C>
__device__ void plane_load_phys(C *A, const Geo &g, const TablesT<C> &T, const C *src) {
  constexpr int B = 4;
  const LoadTile lt(g);
  const int nw = blockDim.x >> 6;
  for (int t0 = threadIdx.x >> 6; t0 < lt.ntiles; t0 += B * nw) {
    C v[B];
    int sl[B];
#pragma unroll
    for (int b = 0; b < B; ++b) {
      int z, m;
      sl[b] = -1;
      const int t = t0 + b * nw;
      v[b] = mkc<C>(0.0, 0.0);
      if (t < lt.ntiles && lt.at(g, t, z, m)) {
        sl[b] = phys_slot(T, g, z, m);
        const C *q = src + z * g.lw + m;
        v[b].x = q->x;
        v[b].y = q->y;
      }
    }
#pragma unroll
    for (int b = 0; b < B; ++b)
      if (sl[b] >= 0) A[sl[b]] = v[b];
  }
  lds_barrier();
}

// ---------------------------------------------------------------------------
// Register-resident row transforms for NX = 128 (64 complex pairs per row).
// A wave owns a slab of 8 rows; lane = 8 * row + t.  The 64-point complex
// DFT is split 64 = 8 x 8 (index m = t + 8r, frequency k = k1 + 8 k2): an
// 8-point DFT over r in each lane's registers, the twiddle W64^{t k1}, an
// 8 x 8 transpose through the wave's own rows of the LDS plane, and an
// 8-point DFT over t.  The even/odd split of the real transform exchanges
// Z_k and Z_{64-k} through the same rows.  Only lanes of one wave touch a
// slab, and LDS operations of a wave complete in order, so the exchanges
// need no workgroup barrier.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// Forward x transform of every row, physical pairs from load(z, m) (global
// memory or computed products) to X_k / NX for k <= NKX in natural slots of
// the LDS rows (FFT_XZ_TO_FOURIER's r2c, mpi.f:1490-1506).  MERGE: the rows
// already hold the transform R_k of another product; the result stored is
// i KX(k) R_k + X_k / NX (kxv = KX), the sum whose z transform is that of
// both products' contributions.
template <bool MERGE = false, class C, class Fetch, class Eval>
__device__ void rows_fwd_reg64(C *A, const Geo &g, const TablesT<C> &T, Fetch fetch, Eval eval,
                               const double *kxv = nullptr) {
  using S = Sc<C>;
  const int lane = threadIdx.x & 63, rho = lane >> 3, t = lane & 7;
  const int nw = blockDim.x >> 6;
  const S nxd = S(g.dnx), rnx = S(g.rnx);  // X / NX as the correctly rounded quotient
  for (int slab = threadIdx.x >> 6; slab * 8 < g.nz; slab += nw) {
    const int z = slab * 8 + rho;  // < NZ: register rows require NZ % 8 == 0
    C *row = A + z * g.ls;
    // the loads of a group of G samples are all issued before any arithmetic
    // on them; left to itself the scheduler interleaves loads and arithmetic
    // sample by sample and the memory latency is exposed eight times per
    // slab.  Samples wider than three complex values (two complex pairs in
    // FP32, whose kernels run with half the registers) go in two groups: the
    // register file holds the raw values of one group.
    using Raw = decltype(fetch(0, 0));
    constexpr int G = (sizeof(C) == sizeof(double2)) ? (sizeof(Raw) > 48 ? 4 : 8) : (sizeof(Raw) > 16 ? 4 : 8);
    C v[8];
#pragma unroll
    for (int r0 = 0; r0 < 8; r0 += G) {
      Raw raw[G];
#pragma unroll
      for (int r = 0; r < G; ++r) raw[r] = fetch(z, t + 8 * (r0 + r));
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int r = 0; r < G; ++r) v[r0 + r] = eval(raw[r]);
    }
    // the other product's outputs of this lane, read before the row is
    // reused (after the samples are consumed, to spare registers)
    C prev[kRowK2];
    S kxp[kRowK2];
    if constexpr (MERGE) {
#pragma unroll
      for (int k2 = 0; k2 < kRowK2; ++k2) {
        const int k = min(t + 8 * k2, g.nkx);
        prev[k2] = row[k];
        kxp[k2] = S(kxv[k]);
      }
    }
    dft_small<8, false>(v);
#pragma unroll
    for (int k1 = 1; k1 < 8; ++k1) v[k1] = cmul(v[k1], T.twr[t * k1]);
    // transpose: lane t's Y_t[k1] to slot 8 k1 + t; lane k1 reads slots 8 k1 + t'
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1) row[8 * k1 + t] = v[k1];
    wave_lds_sync();
#pragma unroll
    for (int tt = 0; tt < 8; ++tt) v[tt] = row[8 * t + tt];  // lane plays k1 = t
    dft_small<8, false>(v);  // v[k2] = Z_{t + 8 k2}
    wave_lds_sync();
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) row[t + 8 * k2] = v[k2];
    wave_lds_sync();
    C zm[kRowK2];
#pragma unroll
    for (int k2 = 0; k2 < kRowK2; ++k2) zm[k2] = row[(64 - t - 8 * k2) & 63];
    wave_lds_sync();
    // X_k = (Z_k + conj Z_{64-k})/2 - i w^k (Z_k - conj Z_{64-k})/2, w = e^{-2 pi i/NX}
#pragma unroll
    for (int k2 = 0; k2 < kRowK2; ++k2) {
      const int k = t + 8 * k2;
      const C zk = v[k2], zmk = zm[k2];
      C a = mkc<C>(S(0.5) * (zk.x + zmk.x), S(0.5) * (zk.y - zmk.y));
      C b = mkc<C>(S(0.5) * (zk.x - zmk.x), S(0.5) * (zk.y + zmk.y));
      C wb = cmul(T.twx[k], b);
      C x = mkc<C>(div_by_recip(a.x + wb.y, nxd, rnx), div_by_recip(a.y - wb.x, nxd, rnx));
      if constexpr (MERGE) {
        const C ik = cmul_ik(kxp[k2], prev[k2]);
        x = mkc<C>(ik.x + x.x, ik.y + x.y);
      }
      if (k <= g.nkx) row[k] = x;
    }
  }
  lds_barrier();
}

template <class C, class Load>
__device__ __forceinline__ void rows_fwd_reg64(C *A, const Geo &g, const TablesT<C> &T, Load load) {
  rows_fwd_reg64<false>(A, g, T, load, [](C v) { return v; });
}

// Forward z transform of the de-aliased columns kx <= NKX for NZ = 128, in
// registers: 16 lanes of a wave own a column, lane t holding x_{t+16r},
// r < 8.  X_{k1+8 k2} = sum_t W16^{t k2} W128^{t k1} sum_r x_{t+16r} W8^{r k1}
// (k1 < 8, k2 < 16): an 8-point DFT over r in each lane and the twiddle
// W128^{t k1}, giving Y_t[k1], then a transpose through the column's own
// LDS slots (only this wave touches them, so no workgroup barrier).  The
// 16-point sums over t are split by the parity p of k2 = 2m + p:
// X_{k1+8p+16m} = sum_{s<8} W8^{s m} W16^{s p} (Y_s[k1] + (-1)^p Y_{s+8}[k1]),
// so lane (k1, p) folds 16 values into 8 as it reads them and ends with an
// 8-point DFT: 8 live values per lane throughout.  Each result goes to
// sink(kx, z index, X).  The rows must be complete in LDS (a workgroup
// barrier before the call).
template <class C, class Sink>
__device__ void cols_fwd_reg128(C *A, const Geo &g, const TablesT<C> &T, Sink sink) {
  const int lane = threadIdx.x & 63, grp = lane >> 4, t = lane & 15;
  const int k1 = t & 7, p = t >> 3;
  const int nw = blockDim.x >> 6;
  const Sc<C> sgn = p ? Sc<C>(-1) : Sc<C>(1);
  for (int kx = (int)(threadIdx.x >> 6) * 4 + grp; kx < g.nkx1; kx += nw * 4) {
    C v[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = A[(t + 16 * r) * g.ls + kx];
    dft_small<8, false>(v);
#pragma unroll
    for (int j = 1; j < 8; ++j) v[j] = cmul(v[j], T.twc[t * j]);
    wave_lds_sync();
#pragma unroll
    for (int j = 0; j < 8; ++j) A[(16 * j + t) * g.ls + kx] = v[j];
    wave_lds_sync();
#pragma unroll
    for (int s2 = 0; s2 < 8; ++s2) {
      const C y0 = A[(16 * k1 + s2) * g.ls + kx], y1 = A[(16 * k1 + s2 + 8) * g.ls + kx];
      const C z = mkc<C>(y0.x + sgn * y1.x, y0.y + sgn * y1.y);
      v[s2] = s2 ? cmul(z, T.twc[8 * s2 * p]) : z;
    }
    dft_small<8, false>(v);
#pragma unroll
    for (int m = 0; m < 8; ++m) sink(kx, k1 + 8 * p + 16 * m, v[m]);
  }
}

// Inverse x transform of every row (FFT_XZ_TO_PHYSICAL's c2r,
// mpi.f:1600-1604): X_k (k <= NKX) in natural slots of the LDS rows to the
// physical pairs, handed to store(z, m, value).
template <class C, class Store>
__device__ void rows_inv_reg64(C *A, const Geo &g, const TablesT<C> &T, Store store) {
  const int lane = threadIdx.x & 63, rho = lane >> 3, t = lane & 7;
  const int nw = blockDim.x >> 6;
  for (int slab = threadIdx.x >> 6; slab * 8 < g.nz; slab += nw) {
    const int z = slab * 8 + rho;
    const bool zin = z < g.nz;
    C *row = A + z * g.ls;
    // Z_k = (X_k + conj X_{64-k}) + i w^k (X_k - conj X_{64-k}), w = e^{+2 pi i/NX},
    // for k = t + 8 k2 (X_k = 0 above NKX; Im X_0 ignored)
    C v[8];
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) {
      const int k = t + 8 * k2, km = (64 - k) & 63;
      C xk = mkc<C>(0.0, 0.0), xm = mkc<C>(0.0, 0.0);
      if (k <= g.nkx) xk = row[k];
      if (k != 0 && km <= g.nkx) xm = row[km];
      if (k == 0) xk.y = 0.0;
      C a = mkc<C>(xk.x + xm.x, xk.y - xm.y);
      C b = mkc<C>(xk.x - xm.x, xk.y + xm.y);
      C wb = cmul(cconj(T.twx[k]), b);
      v[k2] = mkc<C>(a.x - wb.y, a.y + wb.x);
    }
    wave_lds_sync();
    // c_{t' + 8r} = sum_k1 W8^{-r k1} W64^{-t' k1} sum_k2 W8^{-t' k2} Z_{k1 + 8 k2}
    dft_small<8, true>(v);  // lane plays k1 = t: v[t'] = sum_k2 Z W8^{-t' k2}
#pragma unroll
    for (int tt = 1; tt < 8; ++tt) v[tt] = cmul(v[tt], cconj(T.twr[tt * t]));
#pragma unroll
    for (int tt = 0; tt < 8; ++tt) row[8 * tt + t] = v[tt];
    wave_lds_sync();
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1) v[k1] = row[8 * t + k1];  // lane plays t' = t
    dft_small<8, true>(v);  // v[r] = c_{t + 8r}
    wave_lds_sync();
    if (zin) {
#pragma unroll
      for (int r = 0; r < 8; ++r) store(z, t + 8 * r, v[r]);
    }
  }
}

// ---------------------------------------------------------------------------
// Register-resident row transforms for NX = 64 (32 complex pairs per row).
// A wave owns a slab of 16 rows; lane = 4 * row + t (t < 4), eight values
// per lane.  The 32-point complex DFT is split 32 = 4 x 8 with an 8-point
// DFT in each lane's registers, a twiddle, a 4 x 8 transpose through the
// row's own LDS slots (lane t then takes the two columns 2t and 2t + 1),
// and 4-point DFTs.  As for NX = 128, only lanes of one wave touch a slab,
// so the exchanges need no workgroup barrier.
// ---------------------------------------------------------------------------

// Forward x transform of every row, physical pair m of row z from load(z, m),
// to X_k / NX for k <= NKX in natural slots of the LDS rows
// (FFT_XZ_TO_FOURIER's r2c, mpi.f:1490-1506).  With m = t + 4r (r < 8) and
// k = k1 + 8 k2 (k1 < 8, k2 < 4):
//   Z_{k1+8k2} = sum_t W4^{t k2} W32^{t k1} sum_r W8^{r k1} x_{t+4r}.
template <class C, class Load>
__device__ void rows_fwd_reg32(C *A, const Geo &g, const TablesT<C> &T, Load load) {
  using S = Sc<C>;
  const int lane = threadIdx.x & 63, rho = lane >> 2, t = lane & 3;
  const int nw = blockDim.x >> 6;
  const S nxd = S(g.dnx), rnx = S(g.rnx);  // X / NX as the correctly rounded quotient
  for (int slab = threadIdx.x >> 6; slab * 16 < g.nz; slab += nw) {
    const int z = slab * 16 + rho;  // < NZ: these rows require NZ % 16 == 0
    C *row = A + z * g.ls;
    C v[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = load(z, t + 4 * r);
    dft_small<8, false>(v);  // v[k1] = sum_r W8^{r k1} x_{t+4r}
#pragma unroll
    for (int k1 = 1; k1 < 8; ++k1) v[k1] = cmul(v[k1], T.twr[t * k1]);  // W32^{t k1}
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1) row[4 * k1 + t] = v[k1];
    wave_lds_sync();
    C w[8];  // w[4h + k2] = Z_{k1 + 8 k2}, k1 = 2t + h
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      C q[4];
#pragma unroll
      for (int tt = 0; tt < 4; ++tt) q[tt] = row[4 * (2 * t + h) + tt];
      dft_small<4, false>(q);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) w[4 * h + k2] = q[k2];
    }
    wave_lds_sync();
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) row[2 * t + h + 8 * k2] = w[4 * h + k2];
    wave_lds_sync();
    C zm[8];  // Z_{32-k}
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) zm[4 * h + k2] = row[(32 - (2 * t + h) - 8 * k2) & 31];
    wave_lds_sync();
    // X_k = (Z_k + conj Z_{32-k})/2 - i w^k (Z_k - conj Z_{32-k})/2, w = e^{-2 pi i/NX}
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) {
        const int k = 2 * t + h + 8 * k2;
        const C zk = w[4 * h + k2], zmk = zm[4 * h + k2];
        C a = mkc<C>(S(0.5) * (zk.x + zmk.x), S(0.5) * (zk.y - zmk.y));
        C b = mkc<C>(S(0.5) * (zk.x - zmk.x), S(0.5) * (zk.y + zmk.y));
        C wb = cmul(T.twx[k], b);
        if (k <= g.nkx) row[k] = mkc<C>(div_by_recip(a.x + wb.y, nxd, rnx), div_by_recip(a.y - wb.x, nxd, rnx));
      }
  }
  lds_barrier();
}

// Inverse x transform of every row for NX = 64 (FFT_XZ_TO_PHYSICAL's c2r,
// mpi.f:1600-1604): X_k (k <= NKX) in natural slots of the LDS rows to the
// physical pairs, handed to store(z, m, value).  Lane t forms Z_k for
// k = t + 4 k2 (k2 < 8); with m = m1 + 8 m2 (m1 < 8, m2 < 4):
//   c_{m1+8m2} = sum_{k1} W4^{-m2 k1} W32^{-m1 k1} sum_{k2} W8^{-m1 k2} Z_{k1+4k2}.
template <class C, class Store>
__device__ void rows_inv_reg32(C *A, const Geo &g, const TablesT<C> &T, Store store) {
  const int lane = threadIdx.x & 63, rho = lane >> 2, t = lane & 3;
  const int nw = blockDim.x >> 6;
  for (int slab = threadIdx.x >> 6; slab * 16 < g.nz; slab += nw) {
    const int z = slab * 16 + rho;
    C *row = A + z * g.ls;
    // Z_k = (X_k + conj X_{32-k}) + i w^k (X_k - conj X_{32-k}), w = e^{+2 pi i/NX}
    // (X_k = 0 above NKX; Im X_0 ignored)
    C v[8];
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) {
      const int k = t + 4 * k2, km = (32 - k) & 31;
      C xk = mkc<C>(0.0, 0.0), xm = mkc<C>(0.0, 0.0);
      if (k <= g.nkx) xk = row[k];
      if (k != 0 && km <= g.nkx) xm = row[km];
      if (k == 0) xk.y = 0.0;
      C a = mkc<C>(xk.x + xm.x, xk.y - xm.y);
      C b = mkc<C>(xk.x - xm.x, xk.y + xm.y);
      C wb = cmul(cconj(T.twx[k]), b);
      v[k2] = mkc<C>(a.x - wb.y, a.y + wb.x);
    }
    wave_lds_sync();
    dft_small<8, true>(v);  // v[m1] = sum_k2 W8^{-m1 k2} Z_{t+4k2}
#pragma unroll
    for (int m1 = 1; m1 < 8; ++m1) v[m1] = cmul(v[m1], cconj(T.twr[m1 * t]));  // W32^{-m1 t}
#pragma unroll
    for (int m1 = 0; m1 < 8; ++m1) row[4 * m1 + t] = v[m1];
    wave_lds_sync();
    C w[8];  // w[4h + m2] = c_{m1 + 8 m2}, m1 = 2t + h
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      C q[4];
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) q[kk] = row[4 * (2 * t + h) + kk];
      dft_small<4, true>(q);
#pragma unroll
      for (int m2 = 0; m2 < 4; ++m2) w[4 * h + m2] = q[m2];
    }
    wave_lds_sync();
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int m2 = 0; m2 < 4; ++m2) store(z, 2 * t + h + 8 * m2, w[4 * h + m2]);
  }
}

// Physical plane from the inverse-transform slots to natural order in HBM.
template <class C>
__device__ void plane_store_phys(const C *A, const Geo &g, const TablesT<C> &T, C *dst) {
  const LoadTile lt(g);
  const int nw = blockDim.x >> 6;
  for (int t = threadIdx.x >> 6; t < lt.ntiles; t += nw) {
    int z, m;
    if (lt.at(g, t, z, m)) dst[z * g.lw + m] = A[inv_slot(g, T, z, m)];
  }
}

// ---------------------------------------------------------------------------
// Planes too large for LDS.  A transform runs as a row pass and a column
// pass, each staged through an LDS buffer `buf` of `stage` complex values:
// a chunk of whole rows (row stride LS, as in an LDS plane) or of whole
// columns (stride K + 1) is filled, transformed by the routines above over
// the chunk's rows or columns, and drained.  Between the
// passes the plane lives in device memory as Q: NZ rows of the NKX + 1 kept
// x modes (row stride NKX + 1), written once and read once per transform.
// Rows and columns are transformed independently of each other, so the
// results are those of the whole-plane routines, bit for bit.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long recip_dev(unsigned d) { return ((1ull << 32) + d - 1) / d; }

// For e = threadIdx.x, + blockDim.x, ... < n: store(e, load(e)), the loads of
// B iterations issued before their stores so that their latencies overlap.
template <int B, class C, class Load, class Store>
__device__ __forceinline__ void gp_move(int n, Load load, Store store) {
  for (int base = threadIdx.x; base < n; base += B * blockDim.x) {
    C v[B];
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int e = base + b * blockDim.x;
      if (e < n) v[b] = load(e);
    }
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int e = base + b * blockDim.x;
      if (e < n) store(e, v[b]);
    }
  }
}

// Forward x transform of every row: phys(z, m) gives physical pair m
// (x = 2m, 2m+1) of row z; X_k / NX for k <= NKX goes to Q[z (NKX+1) + k].
// B: the loads issued together (phys may read several values per pair).
template <int B = 4, class C, class Phys>
__device__ void gp_rows_fwd(C *Q, const Geo &g, const TablesT<C> &T, C *buf, int stage, Phys phys) {
  const int R = stage / g.ls;  // rows per chunk
  for (int r0 = 0; r0 < g.nz; r0 += R) {
    const int nr = min(R, g.nz - r0);
    gp_move<B, C>(
        nr * g.lw,
        [&](int e) {
          const int z = (int)fdiv((unsigned)e, g.f_lw);
          return phys(r0 + z, e - z * g.lw);
        },
        [&](int e, C v) {
          const int z = (int)fdiv((unsigned)e, g.f_lw);
          buf[z * g.ls + T.posr[e - z * g.lw]] = v;
        });
    lds_barrier();
    plane_rows_fwd(buf, g, T, nr, recip_dev((unsigned)nr));  // ends on a barrier
    C *q = Q + (long long)r0 * g.nkx1;
    for (int e = threadIdx.x; e < nr * g.nkx1; e += blockDim.x) {
      const int z = (int)fdiv((unsigned)e, g.f_nkx1);
      q[e] = buf[z * g.ls + (e - z * g.nkx1)];
    }
    lds_barrier();  // buf is refilled
  }
}

// Forward z transform of Q's columns: out(kx, kzp, V) takes each kept mode
// (packed kz index kzp) before the division by NZ.
template <class C, class Out>
__device__ void gp_cols_fwd(const C *Q, const Geo &g, const TablesT<C> &T, C *buf, int stage, Out out) {
  const int K = stage / g.nz - 1, lc = K + 1;  // columns per chunk, their stride in buf
  for (int k0 = 0; k0 < g.nkx1; k0 += K) {
    const int nk = min(K, g.nkx1 - k0);
    const unsigned long long fk = recip_dev((unsigned)nk);
    gp_move<4, C>(
        g.nz * nk,
        [&](int e) {
          const int z = (int)fdiv((unsigned)e, fk);
          return Q[z * g.nkx1 + k0 + (e - z * nk)];
        },
        [&](int e, C v) {
          const int z = (int)fdiv((unsigned)e, fk);
          buf[z * lc + (e - z * nk)] = v;
        });
    lds_barrier();
    lds_fft<false, true, false>(buf, g.pcol, nk, fk, lc, T.twc);  // plane_cols_fwd over the chunk
    for (int e = threadIdx.x; e < (g.tnkz + 1) * nk; e += blockDim.x) {
      const int kzp = (int)fdiv((unsigned)e, fk), j = e - kzp * nk;
      out(k0 + j, kzp, buf[T.posz[kz_row(g, kzp)] * lc + j]);
    }
    lds_barrier();
  }
}

// Inverse z transform of the columns kx <= NKX: spec(kx, zr) gives row zr
// (natural order, zeros in the de-aliased gap) of the unpacked z spectrum;
// the natural rows go to Q.
template <class C, class Spec>
__device__ void gp_cols_inv(C *Q, const Geo &g, const TablesT<C> &T, C *buf, int stage, Spec spec) {
  const int K = stage / g.nz - 1, lc = K + 1;
  for (int k0 = 0; k0 < g.nkx1; k0 += K) {
    const int nk = min(K, g.nkx1 - k0);
    const unsigned long long fk = recip_dev((unsigned)nk);
    gp_move<4, C>(
        g.nz * nk,
        [&](int e) {
          const int zr = (int)fdiv((unsigned)e, fk);
          return spec(k0 + (e - zr * nk), zr);
        },
        [&](int e, C v) {
          const int zr = (int)fdiv((unsigned)e, fk);
          buf[inv_row(T, zr) * lc + (e - zr * nk)] = v;
        });
    lds_barrier();
    lds_fft<true, true, true>(buf, g.pcol, nk, fk, lc, T.twc);  // plane_cols_inv over the chunk
    for (int e = threadIdx.x; e < g.nz * nk; e += blockDim.x) {
      const int z = (int)fdiv((unsigned)e, fk), j = e - z * nk;
      Q[z * g.nkx1 + k0 + j] = buf[z * lc + j];
    }
    lds_barrier();
  }
}

// Inverse x transform of Q's rows: store(z, m, v) takes physical pair m of
// row z.
template <class C, class Store>
__device__ void gp_rows_inv(const C *Q, const Geo &g, const TablesT<C> &T, C *buf, int stage, Store store) {
  const int R = stage / g.ls;
  for (int r0 = 0; r0 < g.nz; r0 += R) {
    const int nr = min(R, g.nz - r0);
    const C *q = Q + (long long)r0 * g.nkx1;
    gp_move<4, C>(
        nr * g.nkx1, [&](int e) { return q[e]; },
        [&](int e, C v) {
          const int z = (int)fdiv((unsigned)e, g.f_nkx1);
          buf[z * g.ls + (e - z * g.nkx1)] = v;
        });
    lds_barrier();
    plane_rows_inv(buf, g, T, nr, recip_dev((unsigned)nr));  // reads modes k <= NKX only; ends on a barrier
    for (int e = threadIdx.x; e < nr * g.lw; e += blockDim.x) {
      const int z = (int)fdiv((unsigned)e, g.f_lw), m = e - z * g.lw;
      store(r0 + z, m, buf[z * g.ls + T.posrd[m]]);
    }
    lds_barrier();
  }
}

}  // namespace dh
