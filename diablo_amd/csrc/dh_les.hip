// Constant-Smagorinsky LES on the device: les_chan (les.f:1-336,
// LES_MODEL_TYPE=1) with compute_strain_chan (les.f:526-569), the one-rank
// ghost rule of GHOST_LES_MPI (mpi.f:493-509), APPLY_BC_VEL_PHYS_MPI
// (mpi.f:1366-1378, channel.f:2495-2650) and the scalars' eddy diffusivity
// KAPPA_T = NU_T (channel.f:313-375).  The opening Fourier BCs are
// k_les_bc_four (dh_ics.hip) and the NU_T terms of the implicit solve are
// k_solve_les (dh_step.hip).
//
// The model runs from the Fourier velocity at the start of a substep with
// les_chan's BCs applied (a copy of CU, k_les_bc_four) and the physical
// U1..3 the substep's inverse transform made of that copy:
//   k_les_strain   S_ij (and ikx TH, ikz TH) in Fourier space, planes 1..NY+1
//   inverse transforms of those fields
//   k_les_nut      |S| at GYF and GY, NU_T (with its zero ghost rows),
//                  |S| S_ij, the cross terms NU_T dU/dy and KAPPA_T dTH/dx
//   k_les_bc_phys  the physical BCs of U1..3
//   forward transforms of the products, planes 1..NY
//   k_les_cf       the subgrid divergence as increments of CF1..3 and CFTH,
//                  which the nonlinear pass adds to its explicit terms
// Every kernel checks the spin-up gate of channel.f:303 on the device clock
// (the model waits for step 101 of a new flow), so one captured graph
// serves both sides of it; while the gate is closed the increments are zero
// and NU_T keeps its last value (zero before the first call), as the
// reference's implicit NU_T terms, which are not gated, then see.
#include "dh_fft_dev.h"
#include "dh_internal.h"

namespace dh {

struct LesArgs {
  Geo g;
  const Clock *clk;
  int cnf, nth;
  const double2 *cu[3];
  const double2 *cth;
  double2 *four;       // Fourier work fields (slots as kLesFour)
  double *phys;        // physical work fields (slots as kLesPhys)
  double *uphys;       // physical U1..3 of the substep
  double *nut;         // NU_T [NY+2][NZ][NX]
  double2 *sgs;        // subgrid increments of CF1..3, CFTH
  const double *dyw, *dyfw;
  long long fs, cfs;
  int jstart, jend, jstart_th[DH_MAX_TH], jend_th[DH_MAX_TH];
  int ubmin, vbmin, wbmin, ubmax, vbmax, wbmax;
  double uc1min, vc1min, wc1min, uc1max, vc1max, wc1max;
  // learned closure: features out of k_les_nut, NU_T from k_les_mlp
  int mlp;
  double *feat;
  const short *mlp_w;
  const float *mlp_b;
};

__device__ __forceinline__ bool les_gate_open(const LesArgs &a) { return !a.cnf || a.clk->time_step > 100; }

// compute_strain_chan (les.f:536-556) at plane J = 1 + blockIdx.y: S11,
// S22, S33, S5 on J <= NY (zero at NY+1, where the model reads none), S4, S6
// on J = 1..NY+1; for each scalar ikx TH and ikz TH (channel.f:328-358).
__global__ void k_les_strain(LesArgs a) {
#pragma clang fp contract(off)
  if (!les_gate_open(a)) return;
  const Geo &g = a.g;
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= g.pf) return;
  const int J = 1 + blockIdx.y;
  const int kzp = m / g.nkx1, kx = m - kzp * g.nkx1;
  const double KX = g.kx[kx], KZ = g.kz[kzp];
  const long long pf = g.pf, o = (long long)J * pf + m;
  auto F = [&](int s) -> double2 & { return a.four[s * a.cfs + o]; };
  const double2 U1 = a.cu[0][o], U2 = a.cu[1][o], U3 = a.cu[2][o];
  const double2 U1m = a.cu[0][o - pf], U3m = a.cu[2][o - pf];
  const double2 zero = make_double2(0.0, 0.0);
  if (J <= g.ny) {
    const double2 U2p = a.cu[1][o + pf];
    const double dyf = g.dyf[J];
    F(0) = cmul_ik(KX, U1);
    F(1) = make_double2((U2p.x - U2.x) / dyf, (U2p.y - U2.y) / dyf);
    F(2) = cmul_ik(KZ, U3);
    const double2 t = cadd(cmul_ik(KZ, U1), cmul_ik(KX, U3));
    F(4) = make_double2(0.5 * t.x, 0.5 * t.y);
  } else {
    F(0) = F(1) = F(2) = F(4) = zero;
  }
  const double dy = g.dy[J];
  const double2 ik2x = cmul_ik(KX, U2), ik2z = cmul_ik(KZ, U2);
  F(3) = make_double2(0.5 * ((U1.x - U1m.x) / dy + ik2x.x), 0.5 * ((U1.y - U1m.y) / dy + ik2x.y));
  F(5) = make_double2(0.5 * (ik2z.x + (U3.x - U3m.x) / dy), 0.5 * (ik2z.y + (U3.y - U3m.y) / dy));
  for (int n = 0; n < a.nth; ++n) {
    const double2 th = a.cth[n * a.cfs + o];
    F(kLesFour + 2 * n) = cmul_ik(KX, th);
    F(kLesFour + 2 * n + 1) = cmul_ik(KZ, th);
  }
}

// Point (x, z) of plane J = 1 + blockIdx.y: |S| at GYF (les.f:113-134) and
// at GY (137-160) with the mean shear of the plane averages removed, the
// products |S| S_ij (164-190), NU_T = -DYW/2 |S| (221-227) zeroed on rows
// 1, 2, NY, NY+1 (one rank), the cross terms NU_T dU1/dy, NU_T dU3/dy
// (307-328) and KAPPA_T times the scalars' horizontal gradients.  Inputs
// and outputs are separate fields: neighbours in y read the inputs.
__global__ void k_les_nut(LesArgs a) {
#pragma clang fp contract(off)
  if (!les_gate_open(a)) return;
  const Geo &g = a.g;
  const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= g.pp) return;
  const int J = 1 + blockIdx.y, NY = g.ny;
  const long long pp = g.pp, fs = a.fs;
  auto S = [&](int s, int j) { return a.phys[s * fs + (long long)j * pp + col]; };
  const int out = kLesPhys - kLesFour + 2 * a.nth;  // first product slot (7 + 2 N_TH)
  auto P = [&](int s) -> double & { return a.phys[(out + s) * fs + (long long)J * pp + col]; };
  const double *U1 = a.uphys, *U3 = a.uphys + 2 * fs;
  auto u = [&](const double *U, int j) { return U[(long long)j * pp + col]; };
  auto mean = [&](int c, int j) { return a.cu[c][(long long)j * g.pf].x; };
  const double *DY = g.dy, *DYF = g.dyf;
  // |S| at GYF
  double sf = 0.0;
  if (J >= a.jstart && J <= a.jend) {
    const double a4 = 0.5 * (S(3, J + 1) + S(3, J) - 0.5 * (mean(0, J) - mean(0, J - 1)) / DY[J] -
                             0.5 * (mean(0, J + 1) - mean(0, J)) / DY[J + 1]);
    const double a6 = 0.5 * (S(5, J + 1) + S(5, J) - 0.5 * (mean(2, J) - mean(2, J - 1)) / DY[J] -
                             0.5 * (mean(2, J + 1) - mean(2, J)) / DY[J + 1]);
    const double s11 = S(0, J), s22 = S(1, J), s33 = S(2, J), s5 = S(4, J);
    sf = sqrt(2.0 * (s11 * s11) + 4.0 * (a4 * a4) + 4.0 * (s5 * s5) + 2.0 * (s22 * s22) + 4.0 * (a6 * a6) +
              2.0 * (s33 * s33));
  }
  // |S| at GY
  double sy = 0.0;
  if (J >= 2) {
    auto gyi = [&](int s) { return (S(s, J) * DYF[J - 1] + S(s, J - 1) * DYF[J]) / (2.0 * DY[J]); };
    const double i11 = gyi(0), i5 = gyi(4), i22 = gyi(1), i33 = gyi(2);
    const double a4 = S(3, J) - 0.5 * (mean(0, J) - mean(0, J - 1)) / DY[J];
    const double a6 = S(5, J) - 0.5 * (mean(2, J) - mean(2, J - 1)) / DY[J];
    sy = sqrt(2.0 * (i11 * i11) + 4.0 * (a4 * a4) + 4.0 * (i5 * i5) + 2.0 * (i22 * i22) + 4.0 * (a6 * a6) +
              2.0 * (i33 * i33));
  }
  const bool gyf_row = J >= a.jstart && J <= a.jend;
  P(0) = gyf_row ? sf * S(0, J) : 0.0;  // |S| S11
  P(1) = gyf_row ? sf * S(2, J) : 0.0;  // |S| S33
  P(3) = gyf_row ? sf * S(4, J) : 0.0;  // |S| S5
  const double du1 = u(U1, J) - u(U1, J - 1), du3 = u(U3, J) - u(U3, J - 1);
  P(2) = (J >= 2) ? sy * (S(3, J) - 0.5 * du1 / DY[J]) : 0.0;  // |S| S4 less the implicit dU1/dy
  P(4) = (J >= 2) ? sy * (S(5, J) - 0.5 * du3 / DY[J]) : 0.0;  // |S| S6 less the implicit dU3/dy
  if (a.mlp) {
    // the closure's inputs at GY (zero on row 1); NU_T and the terms that
    // use it follow in k_les_mlp and k_les_post
    double f[kLesFeat] = {};
    if (J >= 2) {
      auto gyi = [&](int s) { return (S(s, J) * DYF[J - 1] + S(s, J - 1) * DYF[J]) / (2.0 * DY[J]); };
      f[0] = sy;
      f[1] = S(3, J) - 0.5 * (mean(0, J) - mean(0, J - 1)) / DY[J];
      f[2] = S(5, J) - 0.5 * (mean(2, J) - mean(2, J - 1)) / DY[J];
      f[3] = gyi(0);
      f[4] = gyi(1);
      f[5] = gyi(2);
      f[6] = gyi(4);
    }
    for (int q = 0; q < kLesFeat; ++q) a.feat[q * fs + (long long)J * pp + col] = f[q];
    return;
  }
  const double nut = (J >= 3 && J <= NY - 1) ? -0.5 * a.dyw[J] * sy : 0.0;
  a.nut[(long long)J * pp + col] = nut;
  const bool cross = J >= 2 && J <= NY;
  P(5) = cross ? nut * du1 / DY[J] : 0.0;
  P(6) = cross ? nut * du3 / DY[J] : 0.0;
  for (int n = 0; n < a.nth; ++n) {
    double *tx = a.phys + (kLesFour + 2 * n) * fs + (long long)J * pp + col;
    double *tz = tx + fs;
    *tx = nut * *tx;
    *tz = nut * *tz;
  }
}

// APPLY_BC_VEL_PHYS_LOWER/UPPER (channel.f:2495-2650) on the physical
// U1..3 after the model has read them.
__global__ void k_les_bc_phys(LesArgs a) {
  if (!les_gate_open(a)) return;
  const Geo &g = a.g;
  const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= g.pp) return;
  const long long pp = g.pp;
  const int NY = g.ny;
  const double *DY = g.dy, *DYF = g.dyf;
  double *U1 = a.uphys + col, *U2 = a.uphys + a.fs + col, *U3 = a.uphys + 2 * a.fs + col;
  auto at = [&](double *U, int j) -> double & { return U[(long long)j * pp]; };
  if (a.ubmin == 0) { at(U1, 0) = a.uc1min; at(U1, 1) = a.uc1min; }
  else { at(U1, 1) = at(U1, 2) - DY[2] * a.uc1min; at(U1, 0) = at(U1, 1) - DY[1] * a.uc1min; }
  if (a.wbmin == 0) { at(U3, 0) = a.wc1min; at(U3, 1) = a.wc1min; }
  else { at(U3, 1) = at(U3, 2) - DY[2] * a.wc1min; at(U3, 0) = at(U3, 1) - DY[1] * a.wc1min; }
  if (a.vbmin == 0) { at(U2, 1) = 2.0 * a.vc1min - at(U2, 2); at(U2, 0) = at(U2, 1); }
  else { at(U2, 1) = at(U2, 2) - DYF[1] * a.vc1min; at(U2, 0) = at(U2, 1) - DYF[1] * a.vc1min; }
  if (a.ubmax == 0) { at(U1, NY) = a.uc1max; at(U1, NY + 1) = a.uc1max; }
  else { at(U1, NY) = at(U1, NY - 1) + DY[NY] * a.uc1max; at(U1, NY + 1) = at(U1, NY) + DY[NY] * a.uc1max; }
  if (a.wbmax == 0) { at(U3, NY) = a.wc1max; at(U3, NY + 1) = a.wc1max; }
  else { at(U3, NY) = at(U3, NY - 1) + DY[NY] * a.wc1max; at(U3, NY + 1) = at(U3, NY) + DY[NY] * a.wc1max; }
  if (a.vbmax == 0) at(U2, NY + 1) = 2.0 * a.vc1max - at(U2, NY);
  else at(U2, NY + 1) = at(U2, NY) + DY[NY] * a.vc1max;
}

// The subgrid terms of plane J = blockIdx.y (les.f:246-336, channel.f:
// 346-372) as increments: tau = DELTA F[|S| S_ij] (DYFW at GYF, DYW at GY),
// CF1 and CF3 on J = 2..NY-1 (the upper-wall rank's J2 = NY-1), CF2 on
// J = 2..NY with the cross terms, CFTH on JSTART_TH..JEND_TH; zero
// elsewhere and while the gate is closed.
__global__ void k_les_cf(LesArgs a) {
#pragma clang fp contract(off)
  const Geo &g = a.g;
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= g.pf) return;
  const int J = blockIdx.y, NY = g.ny;
  const long long pf = g.pf, o = (long long)J * pf + m;
  const double2 zero = make_double2(0.0, 0.0);
  auto G = [&](int c) -> double2 & { return a.sgs[c * a.cfs + o]; };
  if (!les_gate_open(a)) {
    for (int c = 0; c < 3 + a.nth; ++c) G(c) = zero;
    return;
  }
  const int kzp = m / g.nkx1, kx = m - kzp * g.nkx1;
  const double KX = g.kx[kx], KZ = g.kz[kzp];
  auto F = [&](int s, int j) { return a.four[s * a.cfs + (long long)j * pf + m]; };
  auto sc = [](double w, double2 z) { return make_double2(w * z.x, w * z.y); };
  if (J >= 2 && J <= NY - 1) {
    const double wf = a.dyfw[J], dyfJ = g.dyf[J];
    const double2 t11 = sc(wf, F(0, J)), t33 = sc(wf, F(1, J)), t5 = sc(wf, F(3, J));
    const double2 t4 = sc(a.dyw[J], F(2, J)), t4p = sc(a.dyw[J + 1], F(2, J + 1));
    const double2 t6 = sc(a.dyw[J], F(4, J)), t6p = sc(a.dyw[J + 1], F(4, J + 1));
    const double2 a1 = cmul_ik(KX, t11), c1 = cmul_ik(KZ, t5);
    G(0) = make_double2(-a1.x - (t4p.x - t4.x) / dyfJ - c1.x, -a1.y - (t4p.y - t4.y) / dyfJ - c1.y);
    const double2 a3 = cmul_ik(KX, t5), c3 = cmul_ik(KZ, t33);
    G(2) = make_double2(-a3.x - (t6p.x - t6.x) / dyfJ - c3.x, -a3.y - (t6p.y - t6.y) / dyfJ - c3.y);
  } else {
    G(0) = G(2) = zero;
  }
  if (J >= 2 && J <= NY) {
    const double2 t4 = sc(a.dyw[J], F(2, J)), t6 = sc(a.dyw[J], F(4, J));
    const double2 a2 = cmul_ik(KX, t4), c2 = cmul_ik(KZ, t6);
    const double2 x1 = cmul_ik(KX, F(5, J)), x3 = cmul_ik(KZ, F(6, J));
    G(1) = make_double2(((-a2.x - c2.x) + x1.x) + x3.x, ((-a2.y - c2.y) + x1.y) + x3.y);
  } else {
    G(1) = zero;
  }
  for (int n = 0; n < a.nth; ++n) {
    if (J >= a.jstart_th[n] && J <= a.jend_th[n]) {
      const double2 x = cmul_ik(KX, F(kLesFour + 2 * n, J)), z = cmul_ik(KZ, F(kLesFour + 2 * n + 1, J));
      G(3 + n) = make_double2(x.x + z.x, x.y + z.y);
    } else {
      G(3 + n) = zero;
    }
  }
}

// The learned closure (no counterpart in the reference): NU_T at the GY
// point of plane J = -DYW(J)/2 max(0, y), y the output of an MLP
// 8 -> 32 -> 32 -> 1 (ReLU hidden layers) of the point's strain features
// (|S| at GY, S4 and S6 less their mean shear, S11, S22, S33 and S5
// interpolated to GY, one spare zero), on J = 3..NY-1 as the model's own
// NU_T; a network that returns |S| is the constant-Smagorinsky model.
// BF16 MFMA (16x16x16, FP32 accumulation): a wave takes 16 points at a time.
// Layer 1 runs transposed, H1^T = W1^T X^T (A = W1^T from the table, B = the
// points' features, lane l holding features 4(l/16)..+3 of point l%16), so
// that its result, hidden units 16t + 4(l/16) + i of point l%16 in lane l,
// is already the A operand of layer 2 (points x hidden).  Layer 2 leaves
// hidden unit 16nt + l%16 of points 4(l/16) + i in lane l; layer 3 is a sum
// over the 16 lanes of each group.
typedef short mlp_s4 __attribute__((ext_vector_type(4)));
typedef float mlp_f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short to_bf16(float x) {  // round to nearest even
  unsigned u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

__global__ void __launch_bounds__(256) k_les_mlp(LesArgs a) {
  if (!les_gate_open(a)) return;
  const Geo &g = a.g;
  const int NY = g.ny;
  const long long pp = g.pp, npts = (long long)(NY + 1) * pp;  // planes 1..NY+1
  const int lane = threadIdx.x & 63, grp = lane >> 4, col16 = lane & 15;
  const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
  const mlp_s4 *W = reinterpret_cast<const mlp_s4 *>(a.mlp_w);
  const mlp_s4 w1a[2] = {W[lane], W[64 + lane]};
  const mlp_s4 w2b[2][2] = {{W[128 + lane], W[192 + lane]}, {W[256 + lane], W[320 + lane]}};
  const float *b1 = a.mlp_b, *b2 = a.mlp_b + kMlpHidden, *w3 = a.mlp_b + 2 * kMlpHidden;
  const float b3 = a.mlp_b[3 * kMlpHidden];
  for (long long base = wave * 16; base < npts; base += nwaves * 16) {
    // B of layer 1: features 4 grp..4 grp+3 of point base + col16 (groups 2, 3: zero)
    const long long q = base + col16;
    mlp_s4 x = {0, 0, 0, 0};
    if (grp < 2 && q < npts) {
      const double *fp = a.feat + pp + q;  // point q is plane 1 + q / pp
      x[0] = to_bf16((float)fp[(4 * grp + 0) * a.fs]);
      x[1] = to_bf16((float)fp[(4 * grp + 1) * a.fs]);
      x[2] = to_bf16((float)fp[(4 * grp + 2) * a.fs]);
      x[3] = to_bf16((float)fp[(4 * grp + 3) * a.fs]);
    }
    mlp_s4 h1[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      mlp_f4 d = {0.f, 0.f, 0.f, 0.f};
      d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w1a[t], x, d, 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 4; ++i) h1[t][i] = to_bf16(fmaxf(d[i] + b1[16 * t + 4 * grp + i], 0.f));
    }
    float part[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
      mlp_f4 d = {0.f, 0.f, 0.f, 0.f};
      d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(h1[0], w2b[0][nt], d, 0, 0, 0);
      d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(h1[1], w2b[1][nt], d, 0, 0, 0);
      const float bb = b2[16 * nt + col16], ww = w3[16 * nt + col16];
#pragma unroll
      for (int i = 0; i < 4; ++i) part[i] += fmaxf(d[i] + bb, 0.f) * ww;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
      for (int o = 8; o > 0; o >>= 1) part[i] += __shfl_xor(part[i], o, 16);
    if (col16 == 0) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const long long p = base + 4 * grp + i;
        if (p < npts) {
          const int J = 1 + (int)(p / pp);
          const double y = (double)(part[i] + b3);
          a.nut[pp + p] = (J >= 3 && J <= NY - 1) ? -0.5 * a.dyw[J] * fmax(y, 0.0) : 0.0;
        }
      }
    }
  }
}

// With the closure: the terms k_les_nut leaves to after NU_T, the cross
// terms NU_T dU1/dy, NU_T dU3/dy and KAPPA_T times the scalar gradients.
__global__ void k_les_post(LesArgs a) {
#pragma clang fp contract(off)
  if (!les_gate_open(a)) return;
  const Geo &g = a.g;
  const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= g.pp) return;
  const int J = 1 + blockIdx.y, NY = g.ny;
  const long long pp = g.pp, fs = a.fs, o = (long long)J * pp + col;
  const int out = kLesPhys - kLesFour + 2 * a.nth;
  const double *U1 = a.uphys, *U3 = a.uphys + 2 * fs, *DY = g.dy;
  const double nut = a.nut[o];
  const bool cross = J >= 2 && J <= NY;
  a.phys[(out + 5) * fs + o] = cross ? nut * (U1[o] - U1[o - pp]) / DY[J] : 0.0;
  a.phys[(out + 6) * fs + o] = cross ? nut * (U3[o] - U3[o - pp]) / DY[J] : 0.0;
  for (int n = 0; n < a.nth; ++n) {
    double *tx = a.phys + (kLesFour + 2 * n) * fs + o, *tz = tx + fs;
    *tx = nut * *tx;
    *tz = nut * *tz;
  }
}

static LesArgs les_args(Ctx &c) {
  LesArgs a{};
  a.g = c.g;
  a.clk = c.clock;
  a.cnf = c.p.create_new_flow;
  a.nth = c.p.n_th;
  for (int i = 0; i < 3; ++i) a.cu[i] = c.cul + i * (long long)(c.g.ny + 2) * c.g.pf;  // with les_chan's BCs
  a.cth = c.cth;
  a.four = c.les_four;
  a.phys = c.les_phys;
  a.uphys = c.uphys;
  a.nut = c.nut;
  a.sgs = c.sgs;
  a.dyw = c.d_lesw;
  a.dyfw = c.d_lesw + (c.g.ny + 2);
  a.fs = (long long)(c.g.ny + 2) * c.g.pp;
  a.cfs = (long long)(c.g.ny + 2) * c.g.pf;
  a.jstart = c.jstart;
  a.jend = c.jend;
  for (int n = 0; n < DH_MAX_TH; ++n) {
    a.jstart_th[n] = c.jstart_th[n];
    a.jend_th[n] = c.jend_th[n];
  }
  const dh_params &p = c.p;
  a.ubmin = p.u_bc_ymin; a.vbmin = p.v_bc_ymin; a.wbmin = p.w_bc_ymin;
  a.ubmax = p.u_bc_ymax; a.vbmax = p.v_bc_ymax; a.wbmax = p.w_bc_ymax;
  a.uc1min = p.u_bc_ymin_c1; a.vc1min = p.v_bc_ymin_c1; a.wc1min = p.w_bc_ymin_c1;
  a.uc1max = p.u_bc_ymax_c1; a.vc1max = p.v_bc_ymax_c1; a.wc1max = p.w_bc_ymax_c1;
  a.mlp = c.les_mlp ? 1 : 0;
  a.feat = c.les_feat;
  a.mlp_w = c.d_mlp_w;
  a.mlp_b = c.d_mlp_b;
  return a;
}

hipError_t launch_les_model(Ctx &c, hipStream_t st) {
  const LesArgs a = les_args(c);
  const int NY = c.g.ny, nth = c.p.n_th;
  const unsigned mb = (unsigned)((c.g.pf + 255) / 256), pb = (unsigned)((c.g.pp + 255) / 256);
  hipError_t e;
  if ((e = klaunch(c, (const void *)k_les_strain, nullptr, dim3(mb, NY + 1), dim3(256), 0, st, a)) != hipSuccess)
    return e;
  // the strain (slots 0..5) and the scalars' gradients (7 + 2n, 8 + 2n), planes 1..NY+1
  const long long fs = a.fs, cfs = a.cfs;
  if ((e = launch_planes_inv(c, c.les_four, c.les_phys, 6, 1, NY + 1, st)) != hipSuccess) return e;
  if (nth > 0 && (e = launch_planes_inv(c, c.les_four + kLesFour * cfs, c.les_phys + kLesFour * fs, 2 * nth, 1,
                                        NY + 1, st)) != hipSuccess)
    return e;
  if ((e = klaunch(c, (const void *)k_les_nut, nullptr, dim3(pb, NY + 1), dim3(256), 0, st, a)) != hipSuccess)
    return e;
  if (c.les_mlp) {
    // 16 points per wave and pass; a grid of about four waves per SIMD
    if ((e = klaunch(c, (const void *)k_les_mlp, nullptr, dim3(1024), dim3(256), 0, st, a)) != hipSuccess) return e;
    if ((e = klaunch(c, (const void *)k_les_post, nullptr, dim3(pb, NY + 1), dim3(256), 0, st, a)) != hipSuccess)
      return e;
  }
  if ((e = klaunch(c, (const void *)k_les_bc_phys, nullptr, dim3(pb), dim3(256), 0, st, a)) != hipSuccess) return e;
  // the products (Fourier slots 0..6) and the scalars' fluxes, planes 1..NY
  const int out = kLesPhys - kLesFour + 2 * nth;
  if ((e = launch_planes_fwd(c, c.les_phys + out * fs, c.les_four, 7, 1, NY, st)) != hipSuccess) return e;
  if (nth > 0 && (e = launch_planes_fwd(c, c.les_phys + kLesFour * fs, c.les_four + kLesFour * cfs, 2 * nth, 1, NY,
                                        st)) != hipSuccess)
    return e;
  return klaunch(c, (const void *)k_les_cf, nullptr, dim3(mb, NY + 2), dim3(256), 0, st, a);
}

}  // namespace dh
