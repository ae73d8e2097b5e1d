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
//   forward transforms of the products, planes 1..NY
//   k_les_cf       the subgrid divergence as increments of CF1..3 and CFTH,
//                  which the nonlinear pass adds to its explicit terms, and
//                  in its last row of blocks the physical BCs of U1..3
// Every kernel checks the spin-up gate of channel.f:303 on the device clock
// (the model waits for step 101 of a new flow), so one captured graph
// serves both sides of it; while the gate is closed the increments are zero
// and NU_T keeps its last value (zero before the first call), as the
// reference's implicit NU_T terms, which are not gated, then see.
#include "dh_fft_dev.h"
#include "dh_internal.h"
#include "dh_mlp_dev.h"

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
  // learned closure: inputs out of k_les_nut, NU_T from k_les_mlp
  int mlp;
  double *feat;
  const bf16x8 *mlp_pk;
  const float *mlp_p;
  double mlp_ts, mlp_ls, mlp_nus;
  const double *gyf;  // GYF (0:NY+1)
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
    if (a.mlp) F(6) = cmul_ik(KZ, U1);  // dU1/dz, for the rotation the closure reads
  } else {
    F(0) = F(1) = F(2) = F(4) = zero;
    if (a.mlp) F(6) = zero;
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

// The learned closure's inputs at the GY point (x, z) of plane J (zero on
// row 1): the five invariants of the resolved strain S and rotation O
// (Pope 1975): tr S^2, tr O^2, tr S^3, tr O^2 S, tr O^2 S^2, in units of
// the time scale ts, and the wall distance (1 - |y|) in units of the length
// scale ls, each through asinh; then S11, S22, S33, S12, S13, S23 and |S|
// (the full strain, with the mean shear) for the training targets.  S11,
// S22, S33, S13 and dU1/dz live at GYF and are interpolated to GY as
// les_chan interpolates (les.f:137-160); S12, S23 and the wall-normal
// derivatives are at GY.  The products the subgrid stress needs are left as
// the strain (S_ij, with the implicit dU/dy parts taken out of S12 and S23);
// k_les_post multiplies them by -2 NU_T over the filter widths once NU_T is
// known.
__device__ void les_closure_inputs(const LesArgs &a, int J, long long col, double sy) {
  const Geo &g = a.g;
  const long long pp = g.pp, fs = a.fs;
  const int NY = g.ny;
  const double *DY = g.dy, *DYF = g.dyf;
  auto S = [&](int s, int j) { return a.phys[s * fs + (long long)j * pp + col]; };
  const int out = kLesPhys - kLesFour + 2 * a.nth;
  auto P = [&](int s) -> double & { return a.phys[(out + s) * fs + (long long)J * pp + col]; };
  const double *U1 = a.uphys, *U3 = a.uphys + 2 * fs;
  auto u = [&](const double *U, int j) { return U[(long long)j * pp + col]; };
  const bool gyf_row = J >= a.jstart && J <= a.jend;
  P(0) = gyf_row ? S(0, J) : 0.0;
  P(1) = gyf_row ? S(2, J) : 0.0;
  P(3) = gyf_row ? S(4, J) : 0.0;
  double f[kLesFeat] = {};
  if (J >= 2) {
    const double du1 = (u(U1, J) - u(U1, J - 1)) / DY[J], du3 = (u(U3, J) - u(U3, J - 1)) / DY[J];
    P(2) = S(3, J) - 0.5 * du1;
    P(4) = S(5, J) - 0.5 * du3;
    auto gyi = [&](int s) { return (S(s, J) * DYF[J - 1] + S(s, J - 1) * DYF[J]) / (2.0 * DY[J]); };
    const double s11 = gyi(0), s22 = gyi(1), s33 = gyi(2), s13 = gyi(4), s12 = S(3, J), s23 = S(5, J);
    const double o12 = du1 - s12, o23 = s23 - du3, o13 = gyi(6) - s13;  // O_ij = (du_i/dx_j - du_j/dx_i)/2
    // S and O as 3 x 3 matrices (O antisymmetric)
    const double Sm[3][3] = {{s11, s12, s13}, {s12, s22, s23}, {s13, s23, s33}};
    const double Om[3][3] = {{0.0, o12, o13}, {-o12, 0.0, o23}, {-o13, -o23, 0.0}};
    double S2[3][3], O2[3][3];
    for (int i = 0; i < 3; ++i)
      for (int k = 0; k < 3; ++k) {
        S2[i][k] = Sm[i][0] * Sm[0][k] + Sm[i][1] * Sm[1][k] + Sm[i][2] * Sm[2][k];
        O2[i][k] = Om[i][0] * Om[0][k] + Om[i][1] * Om[1][k] + Om[i][2] * Om[2][k];
      }
    double l1 = 0, l2 = 0, l3 = 0, l4 = 0, l5 = 0;
    for (int i = 0; i < 3; ++i)
      for (int k = 0; k < 3; ++k) {
        l1 += Sm[i][k] * Sm[k][i];
        l2 += Om[i][k] * Om[k][i];
        l3 += S2[i][k] * Sm[k][i];
        l4 += O2[i][k] * Sm[k][i];
        l5 += O2[i][k] * S2[k][i];
      }
    const double t1 = a.mlp_ts, t2 = t1 * t1, t3 = t2 * t1;
    f[0] = asinh(l1 * t2);
    f[1] = asinh(l2 * t2);
    f[2] = asinh(l3 * t3);
    f[3] = asinh(l4 * t3);
    f[4] = asinh(l5 * t2 * t2);
    const double gyJ = a.gyf[J] - 0.5 * DYF[J];  // GY(J)
    f[5] = asinh(fmax(0.0, 1.0 - fabs(gyJ)) / a.mlp_ls);
    f[6] = s11; f[7] = s22; f[8] = s33; f[9] = s12; f[10] = s13; f[11] = s23;
    f[12] = sqrt(2.0 * l1);
  } else {
    P(2) = P(4) = 0.0;
  }
  for (int q = 0; q < kLesFeat; ++q) a.feat[q * fs + (long long)J * pp + col] = f[q];
  (void)sy;
  (void)NY;
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
    les_closure_inputs(a, J, col, sy);
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
// U1..3 after the model has read them: the last row of k_les_cf's blocks
// (nothing between the model's last read of U and the nonlinear pass reads
// it, and k_les_cf works on Fourier fields only).
__device__ __forceinline__ void les_bc_phys_body(const LesArgs &a) {
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
  if (blockIdx.y == (unsigned)g.ny + 2) {  // the physical BCs of U1..3
    les_bc_phys_body(a);
    return;
  }
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
// point of plane J = NUS softplus(y) on J = 3..NY-1, the model's own rows
// (zero elsewhere, as GHOST_LES_MPI leaves the Smagorinsky NU_T), y the
// network of dh_mlp_dev.h on the point's inputs (feature slots 0..5), 16
// points per wave pass on BF16 MFMA.
__global__ void __launch_bounds__(256) k_les_mlp(LesArgs a) {
  if (!les_gate_open(a)) return;
  const Geo &g = a.g;
  const int NY = g.ny;
  const long long pp = g.pp, npts = (long long)(NY + 1) * pp;  // planes 1..NY+1
  const int lane = threadIdx.x & 63, c = lane & 15, grp = lane >> 4;
  const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
  const float b4 = a.mlp_p[kOB4];
  const int ldi = (int)a.fs;
  MlpRegs r;
  mlp_load_regs(a.mlp_pk, r);
  // the features of the wave's next 16 points are requested before the
  // current 16 go through the network, so their latency overlaps it
  double f[kMlpIn];
  auto fetch = [&](long long q) {
#pragma unroll
    for (int j = 0; j < kMlpIn; ++j) f[j] = 0.0;
    if (grp == 0 && q < npts) {
      const double *fp = a.feat + pp + q;  // point q is plane 1 + q / pp
#pragma unroll
      for (int j = 0; j < kMlpIn; ++j) f[j] = fp[j * ldi];
    }
  };
  fetch(wave * 16 + c);
  for (long long base = wave * 16; base < npts; base += nwaves * 16) {
    const long long q = base + c;
    bf16x8 x;
#pragma unroll
    for (int j = 0; j < 8; ++j) x[j] = (__bf16)0.0f;
#pragma unroll
    for (int j = 0; j < kMlpIn; ++j) x[j] = (__bf16)(float)f[j];
    fetch(base + nwaves * 16 + c);
    const float y = mlp_forward_regs(r, b4, x);
    if (lane < 16 && q < npts) {
      const int J = 1 + (int)(q / pp);
      a.nut[pp + q] = (J >= 3 && J <= NY - 1) ? a.mlp_nus * (double)softplus(y) : 0.0;
    }
  }
}

// With the closure: the terms k_les_nut leaves to after NU_T: the subgrid
// stress -2 NU_T S_ij (over the filter width), the cross terms NU_T dU1/dy,
// NU_T dU3/dy and KAPPA_T times the scalar gradients.
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
  // the stress -2 NU_T S_ij, over the width k_les_cf multiplies by: NU_T at
  // GY for S12, S23; at GYF the mean of the two GY points around it
  if (J >= 2) {
    const double fy = -2.0 * nut / a.dyw[J];
    a.phys[(out + 2) * fs + o] *= fy;
    a.phys[(out + 4) * fs + o] *= fy;
  }
  if (J >= a.jstart && J <= a.jend) {
    const double ff = -2.0 * (0.5 * (nut + a.nut[o + pp])) / a.dyfw[J];
    a.phys[(out + 0) * fs + o] *= ff;
    a.phys[(out + 1) * fs + o] *= ff;
    a.phys[(out + 3) * fs + o] *= ff;
  }
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
  a.mlp_pk = static_cast<const bf16x8 *>(c.d_mlp_pk);
  a.mlp_p = c.d_mlp_p;
  a.mlp_ts = c.mlp_ts;
  a.mlp_ls = c.mlp_ls;
  a.mlp_nus = c.mlp_nus;
  a.gyf = c.d_gyf;
  return a;
}

// The rest, once the strain and the physical U are both there: NU_T (or
// the closure's), the products' forward transforms and the increments.
hipError_t launch_les_model_rest(Ctx &c, hipStream_t st) {
  const LesArgs a = les_args(c);
  const int NY = c.g.ny, nth = c.p.n_th;
  const unsigned mb = (unsigned)((c.g.pf + 255) / 256), pb = (unsigned)((c.g.pp + 255) / 256);
  const long long fs = a.fs, cfs = a.cfs;
  hipError_t e;
  if ((e = klaunch(c, (const void *)k_les_nut, nullptr, dim3(pb, NY + 1), dim3(256), 0, st, a)) != hipSuccess)
    return e;
  if (c.les_mlp) {
    // 16 points per wave and pass; a grid of two waves per SIMD (512 blocks
    // of 4 waves over 256 CUs of 4 SIMDs), as many as its ~150 VGPRs allow
    if ((e = klaunch(c, (const void *)k_les_mlp, nullptr, dim3(512), dim3(256), 0, st, a)) != hipSuccess) return e;
    if ((e = klaunch(c, (const void *)k_les_post, nullptr, dim3(pb, NY + 1), dim3(256), 0, st, a)) != hipSuccess)
      return e;
  }
  // the products (Fourier slots 0..6) and the scalars' fluxes, planes 1..NY
  const int out = kLesPhys - kLesFour + 2 * nth;
  if ((e = launch_planes_fwd(c, c.les_phys + out * fs, c.les_four, 7, 1, NY, st)) != hipSuccess) return e;
  if (nth > 0 && (e = launch_planes_fwd(c, c.les_phys + kLesFour * fs, c.les_four + kLesFour * cfs, 2 * nth, 1, NY,
                                        st)) != hipSuccess)
    return e;
  // rows 0..NY+1: the increments of mode blocks 0..mb-1; row NY+2: the
  // physical BCs of column blocks 0..pb-1
  return klaunch(c, (const void *)k_les_cf, nullptr, dim3(mb > pb ? mb : pb, NY + 3), dim3(256), 0, st, a);
}

// The model's first part with the substep's inverse transform of the state:
// les_chan's Fourier BCs onto the copy of CU and, from that copy, the strain
// and the scalars' gradients in Fourier space (k_les_bc_strain), then one
// inverse transform launch of the state fields (planes 0..NY+1, U from the
// copy), the strain's slots (0..5; 6: dU1/dz with the closure) and the
// gradients (7 + 2n, 8 + 2n), planes 1..NY+1 of the last two groups.
hipError_t launch_les_strain_and_state(Ctx &c, hipStream_t st) {
  const LesArgs a = les_args(c);
  const int nth = c.p.n_th;
  hipError_t e;
  if (c.recording) return hipErrorNotSupported;  // LES contexts are never recorded
  const LesStrainOut so{c.les_four, c.cth, a.cfs, nth, c.les_mlp ? 1 : 0};
  if ((e = launch_les_bc_strain(c, st, so)) != hipSuccess) return e;
  const double2 *four[kLesFour + 2 * DH_MAX_TH];
  double *phys[kLesFour + 2 * DH_MAX_TH];
  int n = 0;
  for (int s = 0; s < (c.les_mlp ? 7 : 6); ++s, ++n) {
    four[n] = c.les_four + s * a.cfs;
    phys[n] = c.les_phys + s * a.fs;
  }
  for (int s = 0; s < 2 * nth; ++s, ++n) {
    four[n] = c.les_four + (kLesFour + s) * a.cfs;
    phys[n] = c.les_phys + (kLesFour + s) * a.fs;
  }
  return launch_state_and_fields_to_phys(c, four, phys, n, st);
}

// The closure's inputs from the current physical U and the Fourier copy
// with les_chan's BCs (both made by the caller), as the model's first
// kernels form them in a substep; nothing else is written.
hipError_t launch_les_features(Ctx &c, hipStream_t st) {
  const LesArgs a = les_args(c);
  const int NY = c.g.ny;
  const unsigned mb = (unsigned)((c.g.pf + 255) / 256), pb = (unsigned)((c.g.pp + 255) / 256);
  hipError_t e;
  if ((e = klaunch(c, (const void *)k_les_strain, nullptr, dim3(mb, NY + 1), dim3(256), 0, st, a)) != hipSuccess)
    return e;
  if ((e = launch_planes_inv(c, c.les_four, c.les_phys, 7, 1, NY + 1, st)) != hipSuccess) return e;
  return klaunch(c, (const void *)k_les_nut, nullptr, dim3(pb, NY + 1), dim3(256), 0, st, a);
}

}  // namespace dh
