// Plane transforms between the de-aliased Fourier layout and physical
// planes: FFT_XZ_TO_PHYSICAL (mpi.f:1547-1606) and FFT_XZ_TO_FOURIER
// (mpi.f:1477-1542) for a batch of y-planes and fields.  One workgroup owns
// one x-z plane of one field; the whole plane lives in LDS between the z
// and x passes, so each plane is read from and written to HBM once.
#include "dh_fft_dev.h"

namespace dh {

struct PlaneBatch {
  const double2 *four[kMaxFields];  // Fourier source/destination per field
  double *phys[kMaxFields];         // physical destination/source per field
  int j0;                           // first plane
  // forward only: planes alo..ahi of field f come from alt[f] when set
  const double *alt[kMaxFields];
  int alo[kMaxFields], ahi[kMaxFields];
};

// Fourier -> physical for planes j0 + blockIdx.x of field blockIdx.y.
template <bool REG>
__global__ void __launch_bounds__(1024) k_plane_inv(Geo g, PlaneBatch b) {
  extern __shared__ double2 A[];
  const Tables T = load_tables(A, g);
  const int f = blockIdx.y;
  const long long j = b.j0 + blockIdx.x;
  plane_unpack(A, g, T, b.four[f] + j * g.pf);
  plane_cols_inv(A, g, T);
  double2 *dst = reinterpret_cast<double2 *>(b.phys[f] + j * g.pp);
  if constexpr (REG) {
    rows_inv_reg64(A, g, T, [&](int z, int m, double2 v) { dst[z * 64 + m] = v; });
  } else {
    plane_rows_inv(A, g, T);
    plane_store_phys(A, g, T, dst);
  }
}

// physical -> Fourier (x r2c with 1/NX and truncation, z c2c with 1/NZ and
// packing of kz = -NKZ..-1 after 0..NKZ).
template <bool REG>
__global__ void __launch_bounds__(1024) k_plane_fwd(Geo g, PlaneBatch b) {
  extern __shared__ double2 A[];
  const Tables T = load_tables(A, g);
  const int f = blockIdx.y;
  const long long j = b.j0 + blockIdx.x;
  const double *ph = (b.alt[f] && j >= b.alo[f] && j <= b.ahi[f]) ? b.alt[f] : b.phys[f];
  const double2 *src = reinterpret_cast<const double2 *>(ph + j * g.pp);
  if constexpr (REG) {
    rows_fwd_reg64(A, g, T, [&](int z, int m) { return src[z * 64 + m]; });
  } else {
    plane_load_phys(A, g, T, src);
    plane_rows_fwd(A, g, T);
  }
  plane_cols_fwd(A, g, T);
  double2 *dst = const_cast<double2 *>(b.four[f]) + j * g.pf;
  const double nzd = g.dnz, rnz = g.rnz;  // V / NZ as the correctly rounded quotient
  for (int m = threadIdx.x; m < g.pf; m += blockDim.x) {
    const int kzp = (int)fdiv((unsigned)m, g.f_nkx1), kx = m - kzp * g.nkx1;
    const double2 v = A[fwd_slot(g, T, kx, kzp)];
    dst[m] = make_double2(div_by_recip(v.x, nzd, rnz), div_by_recip(v.y, nzd, rnz));
  }
}

static hipError_t set_lds(const void *fn, int bytes) {
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

hipError_t launch_fft_to_phys(Ctx &c, const double2 *four, double *phys, int nplanes) {
  PlaneBatch b{};
  b.four[0] = four;
  b.phys[0] = phys;
  b.j0 = 0;
  if (c.reg_rows)
    hipLaunchKernelGGL(k_plane_inv<true>, dim3(nplanes, 1), dim3(c.plane_threads), c.lds_bytes, c.stream, c.g, b);
  else
    hipLaunchKernelGGL(k_plane_inv<false>, dim3(nplanes, 1), dim3(c.plane_threads), c.lds_bytes, c.stream, c.g, b);
  return hipGetLastError();
}

hipError_t launch_fft_to_four(Ctx &c, const double *phys, double2 *four, int nplanes) {
  PlaneBatch b{};
  b.four[0] = four;
  b.phys[0] = const_cast<double *>(phys);
  b.j0 = 0;
  if (c.reg_rows)
    hipLaunchKernelGGL(k_plane_fwd<true>, dim3(nplanes, 1), dim3(c.plane_threads), c.lds_bytes, c.stream, c.g, b);
  else
    hipLaunchKernelGGL(k_plane_fwd<false>, dim3(nplanes, 1), dim3(c.plane_threads), c.lds_bytes, c.stream, c.g, b);
  return hipGetLastError();
}

// All state fields (U1..3, TH) to physical space, planes 0..NY+1
// (channel.f:386-399).
hipError_t launch_state_to_phys(Ctx &c) {
  PlaneBatch b{};
  const long long fs = (long long)(c.g.ny + 2) * c.g.pp;
  const long long cfs = (long long)(c.g.ny + 2) * c.g.pf;
  for (int f = 0; f < c.nfields; ++f) {
    b.four[f] = (f < 3) ? c.cu[f] : c.cth + (f - 3) * cfs;
    b.phys[f] = c.uphys + f * fs;
  }
  b.j0 = 0;
  if (c.reg_rows)
    hipLaunchKernelGGL(k_plane_inv<true>, dim3(c.g.ny + 2, c.nfields), dim3(c.plane_threads), c.lds_bytes, c.stream,
                       c.g, b);
  else
    hipLaunchKernelGGL(k_plane_inv<false>, dim3(c.g.ny + 2, c.nfields), dim3(c.plane_threads), c.lds_bytes, c.stream,
                       c.g, b);
  return hipGetLastError();
}

// Fields f0..f1-1 (U1,U2,U3,TH...) back to Fourier space, planes 0..NY+1
// (channel.f:1043-1050).
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
  if (c.reg_rows)
    hipLaunchKernelGGL(k_plane_fwd<true>, dim3(c.g.ny + 2, f1 - f0), dim3(c.plane_threads), c.lds_bytes, c.stream,
                       c.g, b);
  else
    hipLaunchKernelGGL(k_plane_fwd<false>, dim3(c.g.ny + 2, f1 - f0), dim3(c.plane_threads), c.lds_bytes, c.stream,
                       c.g, b);
  return hipGetLastError();
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
    b.phys[i] = c.uphys + f * fs;
    if (f < 3) {
      b.four[i] = c.cun[f];
      b.alt[i] = c.rphys + f * fs;
      b.alo[i] = c.wlo[f];
      b.ahi[i] = c.whi[f];
    } else {
      b.four[i] = c.cth + (f - 3) * cfs;
    }
  }
  b.j0 = 0;
  if (c.reg_rows)
    hipLaunchKernelGGL(k_plane_fwd<true>, dim3(c.g.ny + 2, f1 - f0), dim3(c.plane_threads), c.lds_bytes, st, c.g, b);
  else
    hipLaunchKernelGGL(k_plane_fwd<false>, dim3(c.g.ny + 2, f1 - f0), dim3(c.plane_threads), c.lds_bytes, st, c.g, b);
  return hipGetLastError();
}

hipError_t prepare_fft_kernels(Ctx &c) {
  const void *fns[] = {(const void *)k_plane_inv<true>, (const void *)k_plane_inv<false>,
                       (const void *)k_plane_fwd<true>, (const void *)k_plane_fwd<false>};
  for (const void *fn : fns) {
    hipError_t e = set_lds(fn, c.lds_bytes);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace dh
