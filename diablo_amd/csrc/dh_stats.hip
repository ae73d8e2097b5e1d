// Plane statistics of SAVE_STATS_CHAN (save_stats.f:136-262) on the device.
// The velocities are transformed to physical space by the plane kernels;
// k_stats then reduces each plane j = 1..NY in FP64 (one block per plane,
// a fixed per-thread order and a fixed tree, so the result is
// deterministic) into the profiles of dh_save_stats.
#include "dh_fft_dev.h"

namespace dh {

constexpr int kStatsThreads = 256;

struct StatsArgs {
  Geo g;
  const double2 *cu[3];
  const double *u[3];  // physical U1..U3, planes 0..NY+1
  const double *gyf;
  double *out;         // [DH_NSTATS][NY+2]
};

__device__ __forceinline__ double mean_mode(const double2 *cu, long long pf, int j) { return cu[pf * j].x; }

__global__ void __launch_bounds__(kStatsThreads) k_stats(StatsArgs a) {
#pragma clang fp contract(off)
  __shared__ double red[7][kStatsThreads / 64];
  const Geo &g = a.g;
  const int j = blockIdx.x + 1;
  const long long pf = g.pf, pp = g.pp;
  const int ny2 = g.ny + 2;
  const double ume = mean_mode(a.cu[0], pf, j), vme = mean_mode(a.cu[1], pf, j), wme = mean_mode(a.cu[2], pf, j);
  const double umm = mean_mode(a.cu[0], pf, j - 1), vmp = mean_mode(a.cu[1], pf, j + 1);
  const double wmm = mean_mode(a.cu[2], pf, j - 1);
  const double *dy = g.dy, *dyf = g.dyf;
  const double *U1 = a.u[0] + pp * j, *U2 = a.u[1] + pp * j, *U3 = a.u[2] + pp * j;
  double su = 0., sv = 0., sw = 0., suv = 0., suw = 0., swv = 0., ssh = 0.;
  for (long long e = threadIdx.x; e < pp; e += blockDim.x) {
    const double u = U1[e], w = U3[e], v0 = U2[e], vp = U2[e + pp];
    const double um = U1[e - pp], wm = U3[e - pp], up = U1[e + pp], wp = U3[e + pp];
    su = su + (u - ume) * (u - ume);
    sv = sv + 0.5 * ((v0 - vme) * (v0 - vme) + (vp - vmp) * (vp - vmp));
    sw = sw + (w - wme) * (w - wme);
    suv = suv + (dyf[j - 1] * (u - ume) + dyf[j] * (um - umm)) / (2.0 * dy[j]) * (v0 - vme);
    swv = swv + (dyf[j - 1] * (w - wme) + dyf[j] * (wm - wmm)) / (2.0 * dy[j]) * (v0 - vme);
    suw = suw + (u - ume) * (w - wme);
    const double da = (up - um) / (2.0 * dyf[j]), db = (wp - wm) / (2.0 * dyf[j]);
    ssh = ssh + da * da + db * db;
  }
  double v[7] = {su, sv, sw, suv, suw, swv, ssh};
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    double x = v[q];
    for (int o = 32; o > 0; o >>= 1) x = x + __shfl_down(x, o);
    if (lane == 0) red[q][w] = x;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    double x = red[q][0];
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) x = x + red[q][i];
    v[q] = x;
  }
  const double n = (double)g.nz * (double)g.nx;
  double *o = a.out;
  o[DH_ST_URMS * ny2 + j] = sqrt(v[0] / n);
  o[DH_ST_VRMS * ny2 + j] = sqrt(v[1] / n);
  o[DH_ST_WRMS * ny2 + j] = sqrt(v[2] / n);
  o[DH_ST_UV * ny2 + j] = v[3] / n;
  o[DH_ST_UW * ny2 + j] = v[4] / n;
  o[DH_ST_WV * ny2 + j] = v[5] / n;
  o[DH_ST_SHEAR * ny2 + j] = v[6] / (double)(g.nx * g.nz);
  o[DH_ST_DUDY * ny2 + j] = (ume - umm) / (a.gyf[j] - a.gyf[j - 1]);
  o[DH_ST_DWDY * ny2 + j] = (wme - wmm) / (a.gyf[j] - a.gyf[j - 1]);
  o[DH_ST_UME * ny2 + j] = ume;
  o[DH_ST_VME * ny2 + j] = vme;
  o[DH_ST_WME * ny2 + j] = wme;
  if (j == 1) {  // the ghost rows of the mean profiles
    for (int jj : {0, g.ny + 1}) {
      o[DH_ST_UME * ny2 + jj] = mean_mode(a.cu[0], pf, jj);
      o[DH_ST_VME * ny2 + jj] = mean_mode(a.cu[1], pf, jj);
      o[DH_ST_WME * ny2 + jj] = mean_mode(a.cu[2], pf, jj);
    }
  }
}

hipError_t launch_stats(Ctx &c, double *out_dev) {
  const long long fs = (long long)(c.g.ny + 2) * c.g.pp;
  for (int f = 0; f < 3; ++f) {
    hipError_t e = launch_fft_to_phys(c, c.cu[f], c.uphys + f * fs, c.g.ny + 2);
    if (e != hipSuccess) return e;
  }
  hipError_t e = hipMemsetAsync(out_dev, 0, sizeof(double) * DH_NSTATS * (c.g.ny + 2), c.stream);
  if (e != hipSuccess) return e;
  StatsArgs a{};
  a.g = c.g;
  for (int f = 0; f < 3; ++f) {
    a.cu[f] = c.cu[f];
    a.u[f] = c.uphys + f * fs;
  }
  a.gyf = c.d_gyf;
  a.out = out_dev;
  hipLaunchKernelGGL(k_stats, dim3(c.g.ny), dim3(kStatsThreads), 0, c.stream, a);
  return hipGetLastError();
}

}  // namespace dh
