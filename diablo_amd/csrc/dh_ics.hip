// Initial-condition pipeline on the device: CREATE_FLOW_CHAN
// (set_ics.f:2-200), CREATE_TH_CHAN (set_ics.f:204-273) and
// APPLY_BC_VEL_LOWER/UPPER (channel.f:2285-2492).  The compiler's
// RANDOM_NUMBER is replaced by a counter-based generator (SplitMix64 +
// Box-Muller), so every seed gives the same field on every device.
#include "dh_fft_dev.h"

namespace dh {

__device__ __forceinline__ unsigned long long splitmix64(unsigned long long x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__device__ double gauss(unsigned long long seed, unsigned long long counter) {
  const unsigned long long a = splitmix64(seed * 0x100000001B3ull ^ splitmix64(2 * counter));
  const unsigned long long b = splitmix64(seed * 0x100000001B3ull ^ splitmix64(2 * counter + 1));
  const double u1 = ((double)(a >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  const double u2 = ((double)(b >> 11)) * (1.0 / 9007199254740992.0);
  return sqrt(-2.0 * log(u1)) * cos(2.0 * 3.141592653589793 * u2);
}

struct IcArgs {
  Geo g;
  double *uphys;
  long long fs;
  const double *gyf;
  int ic_type;
  double ubulk0, amp, ymask, lx;
  unsigned long long seed;
  int n_th;
  double th_min_c1[DH_MAX_TH], th_max_c1[DH_MAX_TH];
  int th_min[DH_MAX_TH], th_max[DH_MAX_TH];
};

// Physical U1..U3 of set_ics.f:26-108 on planes 0..NY+1 (profile on
// 0..NY, noise on 0..NY+1 where |GYF| < ymask).
__global__ void k_ic_velocity(IcArgs a) {
  const Geo &g = a.g;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = (long long)(g.ny + 2) * g.pp;
  if (e >= n) return;
  const int J = (int)(e / g.pp);
  const long long r = e - (long long)J * g.pp;
  const int K = (int)(r / g.nx), i = (int)(r - (long long)K * g.nx);
  const double y = a.gyf[J];
  double u1 = 0.0;
  if (J <= g.ny) {
    if (a.ic_type == 0)
      u1 = (3. / 2.) * a.ubulk0 * (1.0 - y * y);
    else if (a.ic_type == 2)
      u1 = y;
    else if (a.ic_type == 3)
      u1 = tanh(y);
  }
  double u2 = 0.0, u3 = 0.0;
  if (a.ymask <= 0.0 || fabs(y) < a.ymask) {
    const unsigned long long c =
        3ull * ((unsigned long long)i + (unsigned long long)g.nx * ((unsigned long long)K + (unsigned long long)g.nz * J));
    u1 += a.amp * gauss(a.seed, c);
    u2 += a.amp * gauss(a.seed, c + 1);
    u3 += a.amp * gauss(a.seed, c + 2);
  }
  a.uphys[e] = u1;
  a.uphys[a.fs + e] = u2;
  a.uphys[2 * a.fs + e] = u3;
}

// Physical TH of set_ics.f:214-262 on planes 1..NY (0 and NY+1 zero).
__global__ void k_ic_scalar(IcArgs a) {
  const Geo &g = a.g;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = (long long)(g.ny + 2) * g.pp;
  if (e >= n) return;
  const int J = (int)(e / g.pp);
  const long long r = e - (long long)J * g.pp;
  const int K = (int)(r / g.nx), i = (int)(r - (long long)K * g.nx);
  (void)K;
  const double y = a.gyf[J];
  for (int nn = 0; nn < a.n_th; ++nn) {
    double v = 0.0;
    if (J >= 1 && J <= g.ny) {
      if (a.ic_type == 0) {
        const double PI = 3.14159265358979323846;
        const double gx = (i * a.lx) / g.nx;
        v = sin(2.0 * PI * gx / a.lx) / (4.0 * PI * PI);
      } else if (a.ic_type == 1 || a.ic_type == 2) {
        if (a.th_min[nn] == 0 && a.th_max[nn] == 0)
          v = (y <= 2.0) ? (a.th_max_c1[nn] - a.th_min_c1[nn]) * (y + 1.) / 2.0 + a.th_min_c1[nn] : a.th_max_c1[nn];
        else if (a.th_min[nn] == 1 && a.th_max[nn] == 1)
          v = a.th_max_c1[nn] * y;
      } else {
        v = tanh(y);
      }
    }
    a.uphys[(3 + nn) * a.fs + e] = v;
  }
}

struct BcArgs {
  Geo g;
  double2 *cu1, *cu2, *cu3;
  const double *dy, *dyf;
  int ubmin, vbmin, wbmin, ubmax, vbmax, wbmax;
  double uc1min, vc1min, wc1min, uc1max, vc1max, wc1max;
};

// APPLY_BC_VEL_LOWER/UPPER in Fourier space for one mode (mean: the (0,0)
// mode), on its edge planes through at(f, e): f = 0..2 for U1..U3, e = 0..5
// for planes 0, 1, 2, NY-1, NY, NY+1 (distinct: NY >= 4), in the
// reference's statement order.
template <class At>
__device__ __forceinline__ void bc_vel_edges(const BcArgs &a, bool mean, At at) {
  const int NY = a.g.ny;
  const double2 zero = make_double2(0.0, 0.0);
  // lower wall
  if (a.ubmin == 0) {
    at(0, 0) = zero;
    at(0, 1) = zero;
    if (mean) at(0, 1) = at(0, 0) = make_double2(a.uc1min, 0.0);
  } else {
    at(0, 1) = at(0, 2);
    at(0, 0) = at(0, 1);
    if (mean) {
      at(0, 1) = make_double2(at(0, 2).x - a.dy[2] * a.uc1min, at(0, 2).y);
      at(0, 0) = make_double2(at(0, 1).x - a.dy[1] * a.uc1min, at(0, 1).y);
    }
  }
  if (a.wbmin == 0) {
    at(2, 0) = zero;
    at(2, 1) = zero;
    if (mean) at(2, 1) = at(2, 0) = make_double2(a.wc1min, 0.0);
  } else {
    at(2, 1) = at(2, 2);
    at(2, 0) = at(2, 1);
    if (mean) {
      at(2, 1) = make_double2(at(2, 2).x - a.dy[2] * a.wc1min, at(2, 2).y);
      at(2, 0) = make_double2(at(2, 1).x - a.dy[1] * a.wc1min, at(2, 1).y);
    }
  }
  if (a.vbmin == 0) {
    const double2 v2 = at(1, 2);
    at(1, 1) = make_double2(2.0 * a.vc1min - v2.x, -v2.y);
    at(1, 0) = at(1, 1);
  } else {
    at(1, 1) = at(1, 2);
    at(1, 0) = at(1, 1);
    if (mean) {
      at(1, 1) = make_double2(at(1, 2).x - a.dyf[1] * a.vc1min, at(1, 2).y);
      at(1, 0) = make_double2(at(1, 1).x - a.dyf[1] * a.vc1min, at(1, 1).y);
    }
  }
  // upper wall
  if (a.ubmax == 0) {
    at(0, 4) = zero;
    at(0, 5) = zero;
    if (mean) at(0, 4) = at(0, 5) = make_double2(a.uc1max, 0.0);
  } else {
    at(0, 4) = at(0, 3);
    at(0, 5) = at(0, 4);
    if (mean) {
      at(0, 4) = make_double2(at(0, 3).x + a.dy[NY] * a.uc1max, at(0, 3).y);
      at(0, 5) = make_double2(at(0, 4).x + a.dy[NY] * a.uc1max, at(0, 4).y);
    }
  }
  if (a.wbmax == 0) {
    at(2, 4) = zero;
    at(2, 5) = zero;
    if (mean) {
      at(2, 4) = make_double2(a.wc1max, 0.0);
      at(2, 5) = zero;  // channel.f:2445 overrides the ghost value
    }
  } else {
    at(2, 4) = at(2, 3);
    at(2, 5) = at(2, 4);
    if (mean) {
      at(2, 4) = make_double2(at(2, 3).x + a.dy[NY] * a.wc1max, at(2, 3).y);
      at(2, 5) = make_double2(at(2, 4).x + a.dy[NY] * a.wc1max, at(2, 4).y);
    }
  }
  if (a.vbmax == 0) {
    if (mean) at(1, 5) = make_double2(2.0 * a.vc1max - at(1, 4).x, -at(1, 4).y);
  } else {
    at(1, 4) = at(1, 3);
    at(1, 5) = at(1, 4);
    if (mean) at(1, 5) = make_double2(at(1, 4).x + a.dy[NY] * a.vc1max, at(1, 4).y);
  }
}

// edge index e of bc_vel_edges -> plane
__device__ __forceinline__ int edge_plane(int NY, int e) { return e < 3 ? e : NY - 4 + e; }

// APPLY_BC_VEL_LOWER/UPPER in Fourier space, one thread per mode (RANKZ=0).
__device__ void apply_bc_vel_mode(const BcArgs &a, int m) {
  const long long pf = a.g.pf;
  double2 *F[3] = {a.cu1 + m, a.cu2 + m, a.cu3 + m};
  bc_vel_edges(a, m == 0, [&](int f, int e) -> double2 & { return F[f][pf * edge_plane(a.g.ny, e)]; });
}

__global__ void k_apply_bc_vel(BcArgs a) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < a.g.pf) apply_bc_vel_mode(a, m);
}

// The opening of les_chan (les.f:62-73) when the model runs (channel.f:303:
// not a new flow, or past step 100): APPLY_BC_VEL_MPI, then APPLY_BC_LES
// (les.f:805-862), whose Neumann walls copy the first interior plane into
// the ghost planes of U1 and U3 so the model sees no wall gradient.  They
// act on a copy of CU (b.cu1..3), from which the substep's physical U and
// the model are formed: the explicit Fourier terms formed before les_chan
// (channel.f:88-290) see CU without them, and at the end of the substep CU
// is made anew from the physical fields.  While the gate is closed the copy
// is CU itself.
struct LesBcArgs {
  BcArgs b;
  const double2 *cu[3];
  const Clock *clk;
  int cnf;
};
// APPLY_BC_LES (les.f:805-862) on the edge planes of bc_vel_edges: Neumann
// walls copy the first interior plane outwards.
template <class At>
__device__ __forceinline__ void bc_les_edges(const BcArgs &a, At at) {
  if (a.ubmax == 1) {
    at(0, 4) = at(0, 3);
    at(0, 5) = at(0, 4);
  }
  if (a.wbmax == 1) {
    at(2, 4) = at(2, 3);
    at(2, 5) = at(2, 4);
  }
  if (a.ubmin == 1) {
    at(0, 1) = at(0, 2);
    at(0, 0) = at(0, 1);
  }
  if (a.wbmin == 1) {
    at(2, 1) = at(2, 2);
    at(2, 0) = at(2, 1);
  }
}

// Mode m of plane j = blockIdx.y: planes 3..NY-2 are copied by their own
// threads; the thread of plane 0 loads the planes the BCs read or write
// (0..2, NY-1..NY+1), applies the BCs to them in registers and stores
// them, so no other thread writes those planes after it (planes 2 and
// NY-1, which the BCs only read, are also copied by their own threads, with
// the same values).
__global__ void k_les_bc_four(LesBcArgs la) {
  const BcArgs &a = la.b;
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= a.g.pf) return;
  const long long pf = a.g.pf;
  const int j = blockIdx.y, NYp = a.g.ny;
  double2 *dst[3] = {a.cu1, a.cu2, a.cu3};
  if (j != 0) {
    if (j >= 2 && j <= NYp - 1)
      for (int f = 0; f < 3; ++f) dst[f][pf * j + m] = la.cu[f][pf * j + m];
    return;
  }
  double2 r[3][6];  // edge planes 0, 1, 2, NY-1, NY, NY+1 of U1..U3
#pragma unroll
  for (int f = 0; f < 3; ++f)
#pragma unroll
    for (int e = 0; e < 6; ++e) r[f][e] = la.cu[f][pf * edge_plane(NYp, e) + m];
  if (!(la.cnf && la.clk->time_step <= 100)) {
    auto at = [&](int f, int e) -> double2 & { return r[f][e]; };
    bc_vel_edges(a, m == 0, at);
    bc_les_edges(a, at);
  }
#pragma unroll
  for (int f = 0; f < 3; ++f)
#pragma unroll
    for (int e = 0; e < 6; ++e) dst[f][pf * edge_plane(NYp, e) + m] = r[f][e];
}

// compute_strain_chan (les.f:536-556) of mode m at plane J from U1..U3 at
// J-1, J, J+1 (u[f][d], plane J-1+d) with les_chan's BCs: S11, S22, S33,
// S5 on J <= NY (zero at NY+1), S4, S6 on J = 1..NY+1, dU1/dz with the
// closure; ikx TH and ikz TH per scalar (channel.f:328-358).
__device__ __forceinline__ void les_strain_mode(const Geo &g, const LesStrainOut &so, int m, int J,
                                                const double2 (&u)[3][3]) {
#pragma clang fp contract(off)
  const int kzp = m / g.nkx1, kx = m - kzp * g.nkx1;
  const double KX = g.kx[kx], KZ = g.kz[kzp];
  const long long o = (long long)J * g.pf + m;
  auto F = [&](int s) -> double2 & { return so.four[s * so.cfs + o]; };
  const double2 U1 = u[0][1], U2 = u[1][1], U3 = u[2][1], U1m = u[0][0], U3m = u[2][0];
  const double2 zero = make_double2(0.0, 0.0);
  if (J <= g.ny) {
    const double2 U2p = u[1][2];
    const double dyf = g.dyf[J];
    F(0) = cmul_ik(KX, U1);
    F(1) = make_double2((U2p.x - U2.x) / dyf, (U2p.y - U2.y) / dyf);
    F(2) = cmul_ik(KZ, U3);
    const double2 t = cadd(cmul_ik(KZ, U1), cmul_ik(KX, U3));
    F(4) = make_double2(0.5 * t.x, 0.5 * t.y);
    if (so.mlp) F(6) = cmul_ik(KZ, U1);  // dU1/dz, for the rotation the closure reads
  } else {
    F(0) = F(1) = F(2) = F(4) = zero;
    if (so.mlp) F(6) = zero;
  }
  const double dy = g.dy[J];
  const double2 ik2x = cmul_ik(KX, U2), ik2z = cmul_ik(KZ, U2);
  F(3) = make_double2(0.5 * ((U1.x - U1m.x) / dy + ik2x.x), 0.5 * ((U1.y - U1m.y) / dy + ik2x.y));
  F(5) = make_double2(0.5 * (ik2z.x + (U3.x - U3m.x) / dy), 0.5 * (ik2z.y + (U3.y - U3m.y) / dy));
  for (int n = 0; n < so.nth; ++n) {
    const double2 th = so.cth[n * so.cfs + o];
    F(kLesFour + 2 * n) = cmul_ik(KX, th);
    F(kLesFour + 2 * n + 1) = cmul_ik(KZ, th);
  }
}

// k_les_bc_four and the strain in one launch: mode m of plane j =
// blockIdx.y (0..NY+1) takes U1..U3 at j-1, j, j+1; within three planes of
// a wall it also loads the mode's six edge planes and applies les_chan's
// BCs to them in registers (the same values k_les_bc_four stores), then
// writes plane j of the copy and, with the spin-up gate open and j >= 1,
// the strain at j.  No thread reads what another writes.
__global__ void k_les_bc_strain(LesBcArgs la, LesStrainOut so) {
  const BcArgs &a = la.b;
  const Geo &g = a.g;
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= g.pf) return;
  const long long pf = g.pf;
  const int j = blockIdx.y, NY = g.ny;
  const bool open = !(la.cnf && la.clk->time_step <= 100);
  const double2 zero = make_double2(0.0, 0.0);
  double2 u[3][3];
#pragma unroll
  for (int f = 0; f < 3; ++f)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int p = j - 1 + d;
      double2 v = zero;
      if (p >= 0 && p <= NY + 1) v = la.cu[f][pf * p + m];
      u[f][d].x = v.x;  // component-wise: a select of double2 values goes through scratch
      u[f][d].y = v.y;
    }
  if (open && (j <= 3 || j >= NY - 2)) {
    double2 r[3][6];
#pragma unroll
    for (int f = 0; f < 3; ++f)
#pragma unroll
      for (int e = 0; e < 6; ++e) r[f][e] = la.cu[f][pf * edge_plane(NY, e) + m];
    auto at = [&](int f, int e) -> double2 & { return r[f][e]; };
    bc_vel_edges(a, m == 0, at);
    bc_les_edges(a, at);
#pragma unroll
    for (int e = 0; e < 6; ++e)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const bool hit = j - 1 + d == edge_plane(NY, e);
#pragma unroll
        for (int f = 0; f < 3; ++f) {
          u[f][d].x = hit ? r[f][e].x : u[f][d].x;
          u[f][d].y = hit ? r[f][e].y : u[f][d].y;
        }
      }
  }
  double2 *dst[3] = {a.cu1, a.cu2, a.cu3};
#pragma unroll
  for (int f = 0; f < 3; ++f) dst[f][pf * j + m] = u[f][1];
  if (open && j >= 1) les_strain_mode(g, so, m, j, u);
}

__global__ void k_fill_zero(double2 *p, long long n) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) p[e] = make_double2(0.0, 0.0);
}

static IcArgs ic_args(Ctx &c) {
  IcArgs a{};
  a.g = c.g;
  a.uphys = c.uphys;
  a.fs = (long long)(c.g.ny + 2) * c.g.pp;
  a.gyf = c.d_gyf;
  a.lx = c.p.lx;
  a.n_th = c.p.n_th;
  for (int n = 0; n < c.p.n_th; ++n) {
    a.th_min[n] = c.p.th_bc_ymin[n];
    a.th_max[n] = c.p.th_bc_ymax[n];
    a.th_min_c1[n] = c.p.th_bc_ymin_c1[n];
    a.th_max_c1[n] = c.p.th_bc_ymax_c1[n];
  }
  return a;
}

static BcArgs bc_args(Ctx &c) {
  BcArgs b{};
  b.g = c.g;
  b.cu1 = c.cu[0];
  b.cu2 = c.cu[1];
  b.cu3 = c.cu[2];
  b.dy = c.d_dy;
  b.dyf = c.d_dyf;
  const dh_params &p = c.p;
  b.ubmin = p.u_bc_ymin; b.vbmin = p.v_bc_ymin; b.wbmin = p.w_bc_ymin;
  b.ubmax = p.u_bc_ymax; b.vbmax = p.v_bc_ymax; b.wbmax = p.w_bc_ymax;
  b.uc1min = p.u_bc_ymin_c1; b.vc1min = p.v_bc_ymin_c1; b.wc1min = p.w_bc_ymin_c1;
  b.uc1max = p.u_bc_ymax_c1; b.vc1max = p.v_bc_ymax_c1; b.wc1max = p.w_bc_ymax_c1;
  return b;
}

hipError_t launch_apply_bc_vel(Ctx &c) {
  const BcArgs b = bc_args(c);
  hipLaunchKernelGGL(k_apply_bc_vel, dim3((c.g.pf + 63) / 64), dim3(64), 0, c.stream, b);
  return hipGetLastError();
}

hipError_t launch_les_bc_four(Ctx &c, hipStream_t st) {
  LesBcArgs la{bc_args(c), {c.cu[0], c.cu[1], c.cu[2]}, c.clock, c.p.create_new_flow};
  const long long cfs = (long long)(c.g.ny + 2) * c.g.pf;
  la.b.cu1 = c.cul;
  la.b.cu2 = c.cul + cfs;
  la.b.cu3 = c.cul + 2 * cfs;
  return klaunch(c, (const void *)k_les_bc_four, nullptr, dim3((c.g.pf + 255) / 256, c.g.ny + 2), dim3(256), 0, st,
                 la);
}

hipError_t launch_les_bc_strain(Ctx &c, hipStream_t st, const LesStrainOut &so) {
  LesBcArgs la{bc_args(c), {c.cu[0], c.cu[1], c.cu[2]}, c.clock, c.p.create_new_flow};
  const long long cfs = (long long)(c.g.ny + 2) * c.g.pf;
  la.b.cu1 = c.cul;
  la.b.cu2 = c.cul + cfs;
  la.b.cu3 = c.cul + 2 * cfs;
  hipLaunchKernelGGL(k_les_bc_strain, dim3((c.g.pf + 255) / 256, c.g.ny + 2), dim3(256), 0, st, la, so);
  return hipGetLastError();
}

hipError_t launch_create_flow(Ctx &c, int ic_type, double ubulk0, unsigned long long seed, double amp, double ymask) {
  IcArgs a = ic_args(c);
  a.ic_type = ic_type;
  a.ubulk0 = ubulk0;
  a.seed = seed;
  a.amp = amp;
  a.ymask = ymask;
  const long long n = (long long)(c.g.ny + 2) * c.g.pp;
  hipLaunchKernelGGL(k_ic_velocity, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if ((e = launch_fields_to_four(c, 0, 3)) != hipSuccess) return e;
  const long long cf = (long long)(c.g.ny + 2) * c.g.pf;
  hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((cf + 255) / 256)), dim3(256), 0, c.stream, c.cp, cf);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = launch_apply_bc_vel(c)) != hipSuccess) return e;
  return launch_rem_div(c, -1);
}

hipError_t launch_create_th(Ctx &c, int ic_type) {
  if (c.p.n_th == 0) return hipSuccess;
  IcArgs a = ic_args(c);
  a.ic_type = ic_type;
  const long long n = (long long)(c.g.ny + 2) * c.g.pp;
  hipLaunchKernelGGL(k_ic_scalar, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_fields_to_four(c, 3, c.nfields);
}

}  // namespace dh
