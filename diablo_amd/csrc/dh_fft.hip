// Plane transforms between the de-aliased Fourier layout and physical
// planes: FFT_XZ_TO_PHYSICAL (mpi.f:1547-1606) and FFT_XZ_TO_FOURIER
// (mpi.f:1477-1542) for a batch of y-planes and fields.  One workgroup owns
// one x-z plane of one field; the whole plane lives in LDS between the z
// and x passes, so each plane is read from and written to HBM once.
#include "dh_fft_dev.h"

namespace dh {

// Fields per launch: the state (U1..3, TH) and the LES model's strain
// slots fit one launch.
constexpr int kBatchFields = 16;

// Physical planes are double (FP64 transforms, C = double2) or float (the
// mixed form, C = float2); the Fourier side is always FP64.
struct PlaneBatch {
  const double2 *four[kBatchFields];  // Fourier source/destination per field
  void *phys[kBatchFields];           // physical destination/source per field (Sc<C>)
  int j0;                             // first plane
  // forward only: planes alo..ahi of field f come from alt[f] when set
  const void *alt[kBatchFields];
  int alo[kBatchFields], ahi[kBatchFields];
  // forward only, the mixed form: when base[f] is set, planes alo..ahi of
  // alt[f] hold increments and the destination becomes base + transform;
  // the other planes are base itself (base may be the destination)
  const double2 *base[kBatchFields];
  // inverse only: planes of field f below jlo[f] are left alone
  int jlo[kBatchFields];
  // inverse only, the mixed form: transform the fluctuations (no plane mean)
  int drop_mean;
};

struct PlaneArgs {
  Geo g;
  PlaneBatch b;
};

// Fourier -> physical for plane j of field f.  A: the plane in LDS, T: the
// transform tables in LDS.
template <bool REG, class C, class TT>
__device__ __forceinline__ void plane_inv_body(const Geo &g, const PlaneBatch &b, C *A, const TT &T, int f,
                                               long long j) {
  if (j < b.jlo[f]) return;
  plane_unpack(A, g, T, b.four[f] + j * g.pf, b.drop_mean != 0);
  plane_cols_inv(A, g, T);
  C *dst = reinterpret_cast<C *>(static_cast<Sc<C> *>(b.phys[f]) + j * g.pp);
  if constexpr (REG) {
    if (g.lw == 64)
      rows_inv_reg64(A, g, T, [&](int z, int m, C v) { dst[z * 64 + m] = v; });
    else
      rows_inv_reg32(A, g, T, [&](int z, int m, C v) { dst[z * 32 + m] = v; });
  } else {
    plane_rows_inv(A, g, T);
    plane_store_phys(A, g, T, dst);
  }
}

// physical -> Fourier (x r2c with 1/NX and truncation, z c2c with 1/NZ and
// packing of kz = -NKZ..-1 after 0..NKZ).
// COLS: the register column transform when NZ = 128 (left out of the
// batched kernel, where its registers do not fit beside the batch indexing).
template <bool REG, class C, bool COLS = REG, class TT>
__device__ __forceinline__ void plane_fwd_body(const Geo &g, const PlaneBatch &b, C *A, const TT &T, int f,
                                               long long j) {
  const double2 *base = b.base[f] ? b.base[f] + j * g.pf : nullptr;
  double2 *dst = const_cast<double2 *>(b.four[f]) + j * g.pf;
  if (base && (j < b.alo[f] || j > b.ahi[f])) {  // no increment on this plane
    if (base != dst)
      for (int m = threadIdx.x; m < g.pf; m += blockDim.x) dst[m] = base[m];
    return;
  }
  const void *ph = (b.alt[f] && j >= b.alo[f] && j <= b.ahi[f]) ? b.alt[f] : b.phys[f];
  const C *src = reinterpret_cast<const C *>(static_cast<const Sc<C> *>(ph) + j * g.pp);
  if constexpr (REG) {
    if (g.lw == 64)
      rows_fwd_reg64(A, g, T, [&](int z, int m) { return src[z * 64 + m]; });
    else
      rows_fwd_reg32(A, g, T, [&](int z, int m) { return src[z * 32 + m]; });
  } else {
    plane_load_phys(A, g, T, src);
    plane_rows_fwd(A, g, T);
  }
  if (COLS && g.nz == 128) {
    // z transform in registers, each mode stored where it is computed
    cols_fwd_reg128(A, g, T, [&](int kx, int zr, C v) {
      const double nzd = g.dnz, rnz = g.rnz;  // V / NZ as the correctly rounded quotient
      int kzp = -1;
      if (zr <= g.nkz)
        kzp = zr;
      else if (zr >= g.nz - g.nkz)
        kzp = zr - (g.nz - 1 - 2 * g.nkz);
      if (kzp < 0) return;
      const int m = kzp * g.nkx1 + kx;
      double2 r = make_double2(div_by_recip((double)v.x, nzd, rnz), div_by_recip((double)v.y, nzd, rnz));
      if (base) r = make_double2(base[m].x + r.x, base[m].y + r.y);
      dst[m] = r;
    });
    return;
  }
  plane_cols_fwd(A, g, T);
  const double nzd = g.dnz, rnz = g.rnz;  // V / NZ as the correctly rounded quotient
  for (int m = threadIdx.x; m < g.pf; m += blockDim.x) {
    const int kzp = (int)fdiv((unsigned)m, g.f_nkx1), kx = m - kzp * g.nkx1;
    const C v = A[fwd_slot(g, T, kx, kzp)];
    double2 r = make_double2(div_by_recip((double)v.x, nzd, rnz), div_by_recip((double)v.y, nzd, rnz));
    if (base) r = make_double2(base[m].x + r.x, base[m].y + r.y);
    dst[m] = r;
  }
}

// the kernels: one context (arguments by value), or a batch of contexts of
// one shape (argument structs in device memory, member blockIdx.z)
// plane j0 + blockIdx.x of field blockIdx.y
template <bool REG, class C>
__global__ void __launch_bounds__(1024) k_plane_inv(PlaneArgs pa) {
  extern __shared__ double2 smem[];
  C *A = reinterpret_cast<C *>(smem);
  plane_inv_body<REG, C>(pa.g, pa.b, A, load_tables(A, pa.g), blockIdx.y, pa.b.j0 + blockIdx.x);
}
template <bool REG, class C>
__global__ void __launch_bounds__(1024) k_plane_inv_b(const PlaneArgs *__restrict__ av) {
  extern __shared__ double2 smem[];
  const PlaneArgs &pa = av[blockIdx.z];
  C *A = reinterpret_cast<C *>(smem);
  plane_inv_body<REG, C>(pa.g, pa.b, A, load_tables(A, pa.g), blockIdx.y, pa.b.j0 + blockIdx.x);
}
template <bool REG, class C>
__global__ void __launch_bounds__(1024) k_plane_fwd(PlaneArgs pa) {
  extern __shared__ double2 smem[];
  C *A = reinterpret_cast<C *>(smem);
  plane_fwd_body<REG, C>(pa.g, pa.b, A, load_tables(A, pa.g), blockIdx.y, pa.b.j0 + blockIdx.x);
}
template <bool REG, class C>
__global__ void __launch_bounds__(1024) k_plane_fwd_b(const PlaneArgs *__restrict__ av) {
  extern __shared__ double2 smem[];
  const PlaneArgs &pa = av[blockIdx.z];
  C *A = reinterpret_cast<C *>(smem);
  plane_fwd_body<REG, C, false>(pa.g, pa.b, A, load_tables(A, pa.g), blockIdx.y, pa.b.j0 + blockIdx.x);
}

// Planes too large for LDS (FP64): block b keeps the plane between its
// row and column pass in its own scratch slot Q = gp + b gp_slot (device
// memory) and takes the (plane, field) items b, b + gridDim.x, ...: nplanes
// planes of each field.  The tables and the staging buffer are in LDS.
// Fourier -> physical as plane_inv_body; physical -> Fourier as
// plane_fwd_body.
__device__ __forceinline__ void plane_inv_gp_body(const Geo &g, const PlaneBatch &b, double2 *Q, const Tables &T,
                                                  double2 *buf, int stage, int f, long long j) {
  if (j < b.jlo[f]) return;
  const double2 *src = b.four[f] ? b.four[f] + j * g.pf : nullptr;
  const bool drop_mean = b.drop_mean != 0;
  gp_cols_inv(Q, g, T, buf, stage, [&](int kx, int zr) {
    int kzp = -1;  // unpack of mpi.f:1562-1574
    if (zr <= g.nkz)
      kzp = zr;
    else if (zr >= g.nz - g.nkz)
      kzp = zr - (g.nz - 1 - g.nkz) + g.nkz;
    return (kzp >= 0 && src && !(drop_mean && kzp == 0 && kx == 0)) ? src[kzp * g.nkx1 + kx] : make_double2(0.0, 0.0);
  });
  __syncthreads();  // Q complete before the row pass reads it
  double2 *dst = reinterpret_cast<double2 *>(static_cast<double *>(b.phys[f]) + j * g.pp);
  gp_rows_inv(Q, g, T, buf, stage, [&](int z, int m, double2 v) { dst[z * g.lw + m] = v; });
}

__device__ __forceinline__ void plane_fwd_gp_body(const Geo &g, const PlaneBatch &b, double2 *Q, const Tables &T,
                                                  double2 *buf, int stage, int f, long long j) {
  const double2 *base = b.base[f] ? b.base[f] + j * g.pf : nullptr;
  double2 *dst = const_cast<double2 *>(b.four[f]) + j * g.pf;
  if (base && (j < b.alo[f] || j > b.ahi[f])) {  // no increment on this plane
    if (base != dst)
      for (int m = threadIdx.x; m < g.pf; m += blockDim.x) dst[m] = base[m];
    return;
  }
  const void *ph = (b.alt[f] && j >= b.alo[f] && j <= b.ahi[f]) ? b.alt[f] : b.phys[f];
  const double2 *src = reinterpret_cast<const double2 *>(static_cast<const double *>(ph) + j * g.pp);
  gp_rows_fwd(Q, g, T, buf, stage, [&](int z, int m) { return src[z * g.lw + m]; });
  __syncthreads();
  const double nzd = g.dnz, rnz = g.rnz;  // V / NZ as the correctly rounded quotient
  gp_cols_fwd(Q, g, T, buf, stage, [&](int kx, int kzp, double2 v) {
    const int m = kzp * g.nkx1 + kx;
    double2 r = make_double2(div_by_recip(v.x, nzd, rnz), div_by_recip(v.y, nzd, rnz));
    if (base) r = make_double2(base[m].x + r.x, base[m].y + r.y);
    dst[m] = r;
  });
}

__global__ void __launch_bounds__(kGpThreads, 4) k_plane_inv_gp(PlaneArgs pa, double2 *gp, long long gp_slot,
                                                                int nplanes, int nitems) {
  extern __shared__ double2 smem[];
  const Tables T = load_tables_at(smem, pa.g);
  double2 *buf = smem + gp_table_elems(pa.g);
  const int stage = gp_stage_elems(pa.g);
  double2 *Q = gp + (long long)blockIdx.x * gp_slot;
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const int f = it / nplanes;
    plane_inv_gp_body(pa.g, pa.b, Q, T, buf, stage, f, pa.b.j0 + (it - f * nplanes));
    __syncthreads();  // the next item writes over Q
  }
}
__global__ void __launch_bounds__(kGpThreads, 4) k_plane_fwd_gp(PlaneArgs pa, double2 *gp, long long gp_slot,
                                                                int nplanes, int nitems) {
  extern __shared__ double2 smem[];
  const Tables T = load_tables_at(smem, pa.g);
  double2 *buf = smem + gp_table_elems(pa.g);
  const int stage = gp_stage_elems(pa.g);
  double2 *Q = gp + (long long)blockIdx.x * gp_slot;
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const int f = it / nplanes;
    plane_fwd_gp_body(pa.g, pa.b, Q, T, buf, stage, f, pa.b.j0 + (it - f * nplanes));
    __syncthreads();
  }
}

static hipError_t set_lds(const void *fn, int bytes) {
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

// Launches of the plane kernels in precision C (register rows when the
// plane allows them).
// Large planes: the persistent kernels above over grid.x planes of grid.y
// fields, on the scratch half of the stream (the branch stream has its own).
template <bool INV>
static hipError_t plane_gp(Ctx &c, dim3 grid, hipStream_t st, const PlaneBatch &b) {
  if (c.recording) return hipErrorNotSupported;  // no batched forms
  PlaneArgs pa{c.g, b};
  double2 *gp = gp_base(c, st);
  long long slot = c.gp_slot;
  int nplanes = (int)grid.x, nitems = (int)(grid.x * grid.y);
  void *args[] = {&pa, &gp, &slot, &nplanes, &nitems};
  const unsigned blocks = (unsigned)(nitems < kGpSlots ? nitems : kGpSlots);
  return hipLaunchKernel(INV ? (const void *)k_plane_inv_gp : (const void *)k_plane_fwd_gp, dim3(blocks),
                         dim3(kGpThreads), args, (unsigned)kGpLds, st);
}

template <class C>
static hipError_t plane_inv(Ctx &c, dim3 grid, hipStream_t st, const PlaneBatch &b) {
  if constexpr (sizeof(C) == sizeof(double2))
    if (c.gp) return plane_gp<true>(c, grid, st, b);
  const PlaneArgs pa{c.g, b};
  const unsigned lds = lds_bytes_of<C>(c);
  if (c.reg_rows)
    return klaunch(c, (const void *)k_plane_inv<true, C>, (const void *)k_plane_inv_b<true, C>, grid,
                   dim3(c.plane_threads), lds, st, pa);
  return klaunch(c, (const void *)k_plane_inv<false, C>, (const void *)k_plane_inv_b<false, C>, grid,
                 dim3(c.plane_threads), lds, st, pa);
}
template <class C>
static hipError_t plane_fwd(Ctx &c, dim3 grid, hipStream_t st, const PlaneBatch &b) {
  if constexpr (sizeof(C) == sizeof(double2))
    if (c.gp) return plane_gp<false>(c, grid, st, b);
  const PlaneArgs pa{c.g, b};
  const unsigned lds = lds_bytes_of<C>(c);
  if (c.reg_rows)
    return klaunch(c, (const void *)k_plane_fwd<true, C>, (const void *)k_plane_fwd_b<true, C>, grid,
                   dim3(c.plane_threads), lds, st, pa);
  return klaunch(c, (const void *)k_plane_fwd<false, C>, (const void *)k_plane_fwd_b<false, C>, grid,
                 dim3(c.plane_threads), lds, st, pa);
}

hipError_t launch_planes_inv(Ctx &c, const double2 *four, double *phys, int nf, int j0, int nplanes, hipStream_t st) {
  const long long fs = (long long)(c.g.ny + 2) * c.g.pp, cfs = (long long)(c.g.ny + 2) * c.g.pf;
  for (int f0 = 0; f0 < nf; f0 += kBatchFields) {
    const int n = nf - f0 < kBatchFields ? nf - f0 : kBatchFields;
    PlaneBatch b{};
    for (int i = 0; i < n; ++i) {
      b.four[i] = four + (f0 + i) * cfs;
      b.phys[i] = phys + (f0 + i) * fs;
    }
    b.j0 = j0;
    const hipError_t e = plane_inv<double2>(c, dim3(nplanes, n), st, b);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_planes_fwd(Ctx &c, const double *phys, double2 *four, int nf, int j0, int nplanes, hipStream_t st) {
  const long long fs = (long long)(c.g.ny + 2) * c.g.pp, cfs = (long long)(c.g.ny + 2) * c.g.pf;
  for (int f0 = 0; f0 < nf; f0 += kBatchFields) {
    const int n = nf - f0 < kBatchFields ? nf - f0 : kBatchFields;
    PlaneBatch b{};
    for (int i = 0; i < n; ++i) {
      b.four[i] = four + (f0 + i) * cfs;
      b.phys[i] = const_cast<double *>(phys) + (f0 + i) * fs;
    }
    b.j0 = j0;
    const hipError_t e = plane_fwd<double2>(c, dim3(nplanes, n), st, b);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_fft_to_phys(Ctx &c, const double2 *four, double *phys, int nplanes) {
  PlaneBatch b{};
  b.four[0] = four;
  b.phys[0] = phys;
  b.j0 = 0;
  return plane_inv<double2>(c, dim3(nplanes, 1), c.stream, b);
}

hipError_t launch_fft_to_four(Ctx &c, const double *phys, double2 *four, int nplanes) {
  PlaneBatch b{};
  b.four[0] = four;
  b.phys[0] = const_cast<double *>(phys);
  b.j0 = 0;
  return plane_fwd<double2>(c, dim3(nplanes, 1), c.stream, b);
}

// All state fields (U1..3, TH) to the substep's physical arrays, planes
// 0..NY+1 (channel.f:386-399).
hipError_t launch_state_to_phys(Ctx &c) {
  PlaneBatch b{};
  const long long fs = (long long)(c.g.ny + 2) * c.g.pp;
  const long long cfs = (long long)(c.g.ny + 2) * c.g.pf;
  for (int f = 0; f < c.nfields; ++f) {
    // LES: velocity from the copy of CU that carries les_chan's BCs
    b.four[f] = (f < 3) ? (c.les ? c.cul + f * cfs : c.cu[f]) : c.cth + (f - 3) * cfs;
    b.phys[f] = c.mixed ? (void *)(c.uphys32 + f * fs) : (void *)(c.uphys + f * fs);
  }
  b.j0 = 0;
  b.drop_mean = c.mixed;  // the mixed form keeps the plane means in FP64
  const dim3 grid(c.g.ny + 2, c.nfields);
  return c.mixed ? plane_inv<float2>(c, grid, c.stream, b) : plane_inv<double2>(c, grid, c.stream, b);
}

// The LES substep's inverse transforms in one launch: the state fields as
// launch_state_to_phys (FP64, U from the copy with les_chan's BCs), planes
// 0..NY+1, and n more fields four[i] -> phys[i] (the model's strain and
// gradients), planes 1..NY+1.  Two launches where they exceed a batch.
hipError_t launch_state_and_fields_to_phys(Ctx &c, const double2 *const *four, double *const *phys, int n,
                                           hipStream_t st) {
  if (c.mixed || !c.les || c.nfields + n > kBatchFields) {
    hipError_t e = launch_state_to_phys(c);
    for (int i = 0; i < n && e == hipSuccess; ++i) e = launch_planes_inv(c, four[i], phys[i], 1, 1, c.g.ny + 1, st);
    return e;
  }
  PlaneBatch b{};
  const long long fs = (long long)(c.g.ny + 2) * c.g.pp;
  const long long cfs = (long long)(c.g.ny + 2) * c.g.pf;
  for (int f = 0; f < c.nfields; ++f) {
    b.four[f] = (f < 3) ? c.cul + f * cfs : c.cth + (f - 3) * cfs;
    b.phys[f] = c.uphys + f * fs;
  }
  for (int i = 0; i < n; ++i) {
    b.four[c.nfields + i] = four[i];
    b.phys[c.nfields + i] = phys[i];
    b.jlo[c.nfields + i] = 1;
  }
  b.j0 = 0;
  return plane_inv<double2>(c, dim3(c.g.ny + 2, c.nfields + n), st, b);
}

// FP64 fields f0..f1-1 (U1,U2,U3,TH...) back to Fourier space, planes
// 0..NY+1 (channel.f:1043-1050).
hipError_t launch_fields_to_four(Ctx &c, int f0, int f1) {
  PlaneBatch b{};
  const long long fs = (long long)(c.g.ny + 2) * c.g.pp;
  const long long cfs = (long long)(c.g.ny + 2) * c.g.pf;
  for (int f = f0; f < f1; ++f) {
    b.four[f - f0] = (f < 3) ? c.cu[f] : c.cth + (f - 3) * cfs;
    b.phys[f - f0] = c.uphys + f * fs;
  }
  b.j0 = 0;
  if (f1 <= f0) return hipSuccess;
  return plane_fwd<double2>(c, dim3(c.g.ny + 2, f1 - f0), c.stream, b);
}

hipError_t launch_state_to_four(Ctx &c) { return launch_fields_to_four(c, 0, c.nfields); }

// Fields f0..f1-1 back to Fourier space at the end of a substep
// (channel.f:1043-1050): the solve left planes wlo..whi of a velocity field
// in its R array (the other planes keep the values transformed at the start
// of the substep); velocity goes to CUN, for REM_DIV to project into CU, and
// scalars straight to CTH.
hipError_t launch_step_to_four(Ctx &c, int f0, int f1, hipStream_t st) {
  if (f1 <= f0) return hipSuccess;
  PlaneBatch b{};
  const long long fs = (long long)(c.g.ny + 2) * c.g.pp;
  const long long cfs = (long long)(c.g.ny + 2) * c.g.pf;
  for (int f = f0; f < f1; ++f) {
    const int i = f - f0;
    const double2 *four = (f < 3) ? c.cun[f] : c.cth + (f - 3) * cfs;
    b.four[i] = four;
    if (c.mixed) {
      // increments over the state at the start of the substep
      b.phys[i] = c.rphys32 + f * fs;
      b.alt[i] = c.rphys32 + f * fs;
      b.base[i] = (f < 3) ? c.cu[f] : four;
      b.alo[i] = c.wlo[f];
      b.ahi[i] = c.whi[f];
    } else {
      b.phys[i] = c.uphys + f * fs;
      if (f < 3) {
        b.alt[i] = c.rphys + f * fs;
        b.alo[i] = c.wlo[f];
        b.ahi[i] = c.whi[f];
      }
    }
  }
  b.j0 = 0;
  const dim3 grid(c.g.ny + 2, f1 - f0);
  return c.mixed ? plane_fwd<float2>(c, grid, st, b) : plane_fwd<double2>(c, grid, st, b);
}

template <bool REG, class C>
static hipError_t set_lds_plane() {
  const void *fns[] = {(const void *)k_plane_inv<REG, C>, (const void *)k_plane_inv_b<REG, C>,
                       (const void *)k_plane_fwd<REG, C>, (const void *)k_plane_fwd_b<REG, C>};
  for (const void *fn : fns) {
    const hipError_t e = set_lds(fn, kLdsMax);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t prepare_fft_kernels(Ctx &c) {
  if (c.lds_bytes > kLdsMax && !c.gp) return hipErrorInvalidValue;
  hipError_t e;
  if ((e = set_lds((const void *)k_plane_inv_gp, kGpLds)) != hipSuccess) return e;
  if ((e = set_lds((const void *)k_plane_fwd_gp, kGpLds)) != hipSuccess) return e;
  if ((e = set_lds_plane<true, double2>()) != hipSuccess) return e;
  if ((e = set_lds_plane<false, double2>()) != hipSuccess) return e;
  if ((e = set_lds_plane<true, float2>()) != hipSuccess) return e;
  return set_lds_plane<false, float2>();
}

}  // namespace dh
