// Internal definitions shared by the HIP sources of libdiablo_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/diablo_hip.h"

namespace dh {

constexpr int kMaxStages = 8;
constexpr int kMaxFields = 3 + DH_MAX_TH;  // U1,U2,U3,TH(1..N_TH)

// One in-place radix-R stage of a 1-D transform of length n: block length
// L, span S = L/R, twiddle stride n/L; fS/fnb are multiply-high
// reciprocals of S and n/R.
struct FftStage {
  int R, L, S, twm;
  unsigned long long fS, fnb;
};

// Radix plan of one 1-D transform length (radices 16,8,4,2,3), as decimation
// in time (L grows) and decimation in frequency (L shrinks).
struct FftPlan {
  int n;
  int nst;
  FftStage dit[kMaxStages];
  FftStage dif[kMaxStages];
};

// Geometry and transform tables of one grid, passed by value to kernels.
struct Geo {
  int nx, ny, nz;       // grid_def
  int nkx, nkz, tnkz;   // header:21-28
  int nkx1, nkz1;       // de-aliased extents NKX+1, TNKZ+1
  int pf;               // Fourier plane elements nkx1*nkz1
  int lw;               // complex elements per row = NX/2
  int ls;               // LDS row stride in complex elements (lw + 1: rows start on distinct banks)
  long long pp;         // physical plane elements NX*NZ
  double dnx, dnz, rnx, rnz;  // NX, NZ as doubles and their correctly rounded reciprocals
  FftPlan prow, pcol;   // row c2c (length NX/2), column c2c (length NZ)
  const double2 *tw_row;   // exp(-2 pi i m/(NX/2)), m < NX/2
  const double2 *tw_col;   // exp(-2 pi i m/NZ),     m < NZ
  const double2 *tw_real;  // exp(-2 pi i k/NX),      k <= NX/2
  const double *kx, *kz, *kx2, *kz2;  // wavenumbers (fft.f:77-100), de-aliased
  const double *dy, *dyf;             // grid spacing (0:NY+1)
  const double *rdy, *rdyf;           // their correctly rounded reciprocals (0 where the spacing is 0)
  const int *posr;   // DIT input slot of row index m (length NX/2)
  const int *posrd;  // DIF output slot of row index k (length NX/2)
  const int *posz;   // DIF output slot of column index k (length NZ)
  const int *posc;   // DIT input slot of column index k (length NZ)
  unsigned long long f_npair, f_nkx1, f_lw, f_nz;  // reciprocals of NX/4+1, NKX+1, NX/2, NZ
};

// Device-resident clock (/RUNTIME_VARS/ + /RKW3_PARAMS/).
struct Clock {
  double time;
  double dt;
  double h_bar[3];
  double beta_bar[3];
  double zeta_bar[3];
  double h_cp;                      // H_BAR(3) of the substep in which COURANT ran (channel.f:80)
  unsigned long long cfl_min_bits;  // running CFL minimum (non-negative double bits)
  int time_step;
  int pad;
};

// Parallel-cyclic-reduction tables of the y-implicit (Crank-Nicolson)
// matrices, one set per RK substep and field; with LES off each matrix
// depends only on j, so every (x,z) column of a field shares it.
// Rows of one column held by one lane of the Crank-Nicolson solve; a column
// is split into 4 (NY+2 <= 144) or 8 such segments.
constexpr int kSegRows = 36;
constexpr int kStepsPerGraph = 8;  // time steps captured in one graph for long runs
// Segments of one (kx,kz) mode's pressure-Poisson sweep over rows 2..NY, one
// lane each, and the most rows a segment holds in registers.
constexpr int kRdsSeg = 8;
constexpr int kRdsRows = 17;
// Output k = t + 8 k2 of a register row transform (NX = 128) is stored only
// for k <= NKX = NX/3 = 42, so k2 < kRowK2 covers every stored output.
constexpr int kRowK2 = 6;
// Largest dynamic LDS of a workgroup on gfx950 (160 KB per CU).  The
// dynamic-LDS limit is an attribute of a kernel function, shared by every
// context in the process, so kernels are given the hardware maximum rather
// than one context's need (a smaller context created later would otherwise
// lower it under a larger one).
constexpr int kLdsMax = 160 * 1024;
// Planes too large for LDS (dh_create sets Ctx::gp) are transformed in
// chunks staged through LDS, the plane between the row and the column pass
// held in a scratch slot in device memory: kGpSlots slots per stream half,
// one per block of the persistent large-plane kernels (kGpThreads threads).
constexpr int kGpSlots = 256, kGpThreads = 1024;

// Tables of the y-implicit (Crank-Nicolson) matrices, one set per RK
// substep and field; with LES off each matrix depends only on j, so every
// (x,z) column of a field shares them.
struct RealFactors {
  double *tl, *tb, *tc, *tr;  // [rk][field][row] Thomas multipliers, pivots, upper diagonal, 1/pivots
  double *pf, *pb;            // [rk][field][row] segment carry factors of the two sweeps
  int *isbc;                  // [field][row] 1: row takes a constant right-hand side
  double *vconst;             // [field][row]
  double *bcl, *bcd, *bcu;    // [field][row] coefficients of the boundary rows (isbc = 1)
  int seg_rows;               // rows per register segment of a column
  int nseg;                   // segments per column (4 or 8)
};

// A launch, event record or stream wait of the step sequence, as the
// launchers append it in recording mode (Ctx::recording) instead of issuing
// it: ensembles record one step of each member and replay the launches of
// members of one shape as single launches of the kernels' batched variants,
// which read their argument struct from a device array at blockIdx.z.
struct RecOp {
  int kind;        // 0: kernel, 1: event record, 2: stream waits on an event
  const void *fnb; // batched kernel
  dim3 grid, block;
  unsigned lds;
  int role;        // stream: 0 the context's main stream, 1 branch, 2 side
  int ev;          // event: 0 ev_nl01, 1 ev_branch_done, 2 ev_fork, 3 ev_join
  std::vector<unsigned char> args;  // the kernel's argument struct
};

struct Ctx {
  dh_params p;
  Geo g;
  int nfields;  // 3 + n_th
  int jstart, jend, jstart_th[DH_MAX_TH], jend_th[DH_MAX_TH];
  int wlo[kMaxFields], whi[kMaxFields];  // solution write-back rows per field
  int lds_bytes;     // dynamic LDS of the FP64 plane kernels
  int lds_bytes32;   // the same for the FP32 transforms of the mixed form
  int plane_threads;
  bool reg_rows = false;     // plane kernels: row transforms in registers (NX = 128, NZ a multiple of 8; NX = 64, NZ of 16)
  bool reg_rows_nl = false;  // the same in k_nonlinear
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;          // table rebuild after COURANT, overlapped with the substep's tail
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool tables_pending = false;
  // branch of a substep: the last velocity component's (and the scalars')
  // nonlinear pass, solve and forward transform, beside the solve and
  // transform of components 0 and 1
  hipStream_t branch = nullptr;
  hipEvent_t ev_nl01 = nullptr, ev_branch_done = nullptr;
  std::string err;
  bool grid_set = false;
  // host copies
  std::vector<double> gy, gyf, dy, dyf;
  double dt_cap;  // courant.f:15-31 limits that do not depend on U
  // device buffers
  double2 *tw_row = nullptr, *tw_col = nullptr, *tw_real = nullptr;
  int *posr = nullptr, *posrd = nullptr, *posz = nullptr, *posc = nullptr;
  double *kx = nullptr, *kz = nullptr, *kx2 = nullptr, *kz2 = nullptr;
  double *d_dy = nullptr, *d_dyf = nullptr, *d_gyf = nullptr, *d_rdy = nullptr, *d_rdyf = nullptr;
  double2 *cu[3] = {}, *cp = nullptr, *cf[3] = {}, *cth = nullptr, *cfth = nullptr;
  double2 *phi = nullptr;       // CR1 after REM_DIV
  double2 *cun[3] = {};         // velocity back in Fourier space before REM_DIV projects it into CU
  double *uphys = nullptr;      // [nfields][NY+2][NZ][NX]: U1,U2,U3,TH
  double *rphys = nullptr;      // [nfields][NY+2][NZ][NX]: R1,R2,R3,RTH
  // mixed form (dh_set_mixed_precision): the substep's transforms and
  // physical arrays in FP32, allocated when it is first switched on
  bool mixed = false;
  float *uphys32 = nullptr, *rphys32 = nullptr;
  bool th_phys32 = false;  // the physical TH of the last substep is in uphys32
  double *rd_ta = nullptr;      // REM_DIV sweep multipliers [NY+2][pf]
  double *rd_tb = nullptr;      // REM_DIV sweep pivots      [NY+2][pf]
  double *rd_tr = nullptr;      // REM_DIV 1/pivots          [NY+2][pf]
  int rd_seg = 0;               // rows per REM_DIV sweep segment (0: one sequential sweep per mode)
  double *stats = nullptr;      // dh_save_stats profiles [DH_NSTATS][NY+2]
  Clock *clock = nullptr;
  RealFactors fac{};
  double *rd_l = nullptr, *rd_u = nullptr;  // 1/(DY DYF) tables for REM_DIV
  // graphs of one and of kStepsPerGraph full time steps
  hipGraphExec_t step_graph = nullptr, steps_graph = nullptr;
  bool graph_ok = false, steps_graph_ok = false;
  // recording mode (see RecOp)
  bool recording = false;
  std::vector<RecOp> rec;
  // constant-Smagorinsky LES (les.f:1-336): NU_T at GY points [NY+2][NZ][NX]
  // (also the scalars' KAPPA_T), the model's Fourier and physical work
  // fields, the subgrid increments of CF1..3 and CFTH the nonlinear pass
  // adds, the per-column solve's sweep storage, and the filter widths
  bool les = false;
  double *nut = nullptr;
  double2 *cul = nullptr;       // CU1..3 with les_chan's BCs, [3][NY+2][pf]
  double2 *les_four = nullptr;  // [kLesFour + 2 N_TH][NY+2][pf]
  double *les_phys = nullptr;   // [kLesPhys + 2 N_TH][NY+2][NZ][NX]
  double2 *sgs = nullptr;       // [3 + N_TH][NY+2][pf]
  double *les_wk = nullptr;     // [2][nfields][NY+2][NZ][NX]
  double *d_lesw = nullptr;     // DYW (0:NY+1) then DYFW (0:NY+1)
  // learned closure in place of NU_T = C_s^2 Delta^2 |S| (no counterpart in
  // the reference): the network's FP32 parameters, their BF16 operand
  // packing (dh_mlp_dev.h), the scales of its inputs and output, and the
  // inputs and strain it reads and writes [kLesFeat][NY+2][NZ][NX]
  bool les_mlp = false;
  float *d_mlp_p = nullptr;     // parameters (kMlpNP floats)
  void *d_mlp_pk = nullptr;     // packed operands (kPAll x 16 bytes)
  double mlp_ts = 1.0, mlp_ls = 1.0, mlp_nus = 1.0;  // time scale, length scale, NU_T scale
  double *les_feat = nullptr;
  // planes too large for LDS: kGpSlots scratch slots for the main stream's
  // kernels and kGpSlots for the branch's, gp_slot complex values each (a
  // plane between the row and the column pass, NZ rows of the NKX+1 kept x
  // modes, then the nonlinear pass's accumulators of the PF modes)
  bool gp = false;
  double2 *gp_scratch = nullptr;
  long long gp_slot = 0;
};
inline double2 *gp_base(const Ctx &c, hipStream_t st) {
  return c.gp_scratch + (st == c.branch ? (long long)kGpSlots * c.gp_slot : 0LL);
}
// The large-plane kernels' LDS (all of a CU's: one block of 16 waves per
// CU, 4 per SIMD, which their launch bounds keep to 128 VGPRs; against two
// blocks of half the size, the larger chunks halve the barriers per plane
// and the step at 256 x 193 x 256 takes 7.4 instead of 7.9 ms): the
// transform tables (load_tables_at), in double2 units rounded up, then the
// staging buffer of the row and column passes.
constexpr int kGpLds = kLdsMax;
__host__ __device__ inline int gp_table_elems(const Geo &g) {
  return (2 * g.lw + g.nz + 1) + ((2 * g.lw + 2 * g.nz) * (int)sizeof(int) + 15) / 16;
}
__host__ __device__ inline int gp_stage_elems(const Geo &g) { return kGpLds / 16 - gp_table_elems(g); }
// feature slots of the learned closure: 0..5 the network's inputs, 6..11
// S11, S22, S33, S12, S13, S23 at GY, 12 |S| at GY
constexpr int kLesFeat = 13;

// Work fields of the LES pass: Fourier slots 0..6 (S11, S22, S33, S4, S5,
// S6 in, then |S| S products and cross terms out), 7 + 2n and 8 + 2n for
// scalar n; physical slots 0..5 the strain, 7 + 2n, 8 + 2n the scalar
// gradients, 7 + 2 N_TH + 0..6 the products.
constexpr int kLesFour = 7, kLesPhys = 14;

inline int stream_role(const Ctx &c, hipStream_t st) { return st == c.branch ? 1 : (st == c.side ? 2 : 0); }
inline hipEvent_t event_of(const Ctx &c, int ev) {
  return ev == 0 ? c.ev_nl01 : ev == 1 ? c.ev_branch_done : ev == 2 ? c.ev_fork : c.ev_join;
}
// Launch of kernel fn (one argument struct A by value) on st, or its record
// with the batched variant fnb.
template <class A>
inline hipError_t klaunch(Ctx &c, const void *fn, const void *fnb, dim3 grid, dim3 block, unsigned lds, hipStream_t st,
                          const A &a) {
  if (c.recording) {
    RecOp o{};
    o.kind = 0;
    o.fnb = fnb;
    o.grid = grid;
    o.block = block;
    o.lds = lds;
    o.role = stream_role(c, st);
    const unsigned char *b = reinterpret_cast<const unsigned char *>(&a);
    o.args.assign(b, b + sizeof(A));
    c.rec.push_back(std::move(o));
    return hipSuccess;
  }
  void *args[] = {const_cast<A *>(&a)};
  return hipLaunchKernel(fn, grid, block, args, lds, st);
}
inline hipError_t krecord(Ctx &c, int ev, hipStream_t st) {
  if (c.recording) {
    RecOp o{};
    o.kind = 1;
    o.ev = ev;
    o.role = stream_role(c, st);
    c.rec.push_back(std::move(o));
    return hipSuccess;
  }
  return hipEventRecord(event_of(c, ev), st);
}
inline hipError_t kwait(Ctx &c, hipStream_t st, int ev) {
  if (c.recording) {
    RecOp o{};
    o.kind = 2;
    o.ev = ev;
    o.role = stream_role(c, st);
    c.rec.push_back(std::move(o));
    return hipSuccess;
  }
  return hipStreamWaitEvent(st, event_of(c, ev), 0);
}

// ---- launchers (dh_fft.hip) ----
hipError_t prepare_fft_kernels(Ctx &c);
// dynamic LDS of a plane kernel whose plane holds complex values of type C
template <class C>
inline int lds_bytes_of(const Ctx &c) {
  return sizeof(C) == sizeof(double2) ? c.lds_bytes : c.lds_bytes32;
}
hipError_t launch_fft_to_phys(Ctx &c, const double2 *four, double *phys, int nplanes);
hipError_t launch_stats(Ctx &c, double *out_dev);  // profiles of dh_save_stats into out_dev
hipError_t launch_fft_to_four(Ctx &c, const double *phys, double2 *four, int nplanes);
hipError_t launch_state_to_phys(Ctx &c);
// LES: the state (planes 0..NY+1) and n strain fields (planes 1..NY+1) in one launch
hipError_t launch_state_and_fields_to_phys(Ctx &c, const double2 *const *four, double *const *phys, int n,
                                           hipStream_t st);
hipError_t launch_state_to_four(Ctx &c);
hipError_t launch_fields_to_four(Ctx &c, int f0, int f1);
hipError_t launch_step_to_four(Ctx &c, int f0, int f1, hipStream_t st);  // end of a substep (into CUN / CTH)
// ---- launchers (dh_ics.hip) ----
hipError_t launch_apply_bc_vel(Ctx &c);
// LES (les.f): the gated Fourier BCs that open les_chan, before the state
// goes to physical space; the model itself (strain, NU_T, subgrid
// increments, physical BCs) after it; the per-column solve with NU_T
hipError_t launch_les_bc_four(Ctx &c, hipStream_t st);
hipError_t launch_les_model_rest(Ctx &c, hipStream_t st);  // NU_T, the products' transforms, the increments
// k_les_strain, then the state's and the strain's inverse transforms in one launch
hipError_t launch_les_strain_and_state(Ctx &c, hipStream_t st);
// where k_les_bc_strain writes the strain and the scalars' gradients
struct LesStrainOut {
  double2 *four;        // slots [s][NY+2][pf], stride cfs
  const double2 *cth;   // the scalars' Fourier fields
  long long cfs;
  int nth, mlp;
};
// les_chan's Fourier BCs onto the copy of CU and compute_strain_chan from it, one launch
hipError_t launch_les_bc_strain(Ctx &c, hipStream_t st, const LesStrainOut &so);
hipError_t launch_les_features(Ctx &c, hipStream_t st);  // the closure's inputs only (no model step)
hipError_t launch_mlp_pack(const float *params, void *packed, hipStream_t st);
hipError_t launch_solve_les_range(Ctx &c, int rk, int courant, int f0, int f1, hipStream_t st);
// fields f0..f0+nf-1 of strided work arrays (Fourier stride (NY+2) pf,
// physical (NY+2) NX NZ), planes j0..j0+nplanes-1, in batches of kMaxFields
hipError_t launch_planes_inv(Ctx &c, const double2 *four, double *phys, int nf, int j0, int nplanes, hipStream_t st);
hipError_t launch_planes_fwd(Ctx &c, const double *phys, double2 *four, int nf, int j0, int nplanes, hipStream_t st);
hipError_t launch_create_flow(Ctx &c, int ic_type, double ubulk0, unsigned long long seed, double amp, double ymask);
hipError_t launch_create_th(Ctx &c, int ic_type);
// ---- launchers (dh_step.hip) ----
hipError_t prepare_step_kernels(Ctx &c);
// rk 0..2; bnd_end/bnd_begin: the step boundary after it (TIME += DELTA_T,
// TIME_STEP += 1), done within REM_DIV's launch
hipError_t launch_substep(Ctx &c, int rk, int bnd_end = 0, int bnd_begin = 0);
constexpr int kNumPhases = 6;
extern const char *const kPhaseNames[kNumPhases];
hipError_t launch_phase(Ctx &c, int rk, int phase);
hipError_t launch_rem_div(Ctx &c, int rk_for_cp, int bnd_end = 0, int bnd_begin = 0);  // rk_for_cp<0: no CP update
hipError_t launch_step_boundary(Ctx &c, bool end_prev, bool begin_next);  // TIME += DELTA_T, TIME_STEP += 1
hipError_t launch_cn_tables(Ctx &c, hipStream_t st);
hipError_t join_side(Ctx &c);  // main stream waits for a pending table rebuild
hipError_t launch_rem_div_factors(Ctx &c);
// ---- launchers (dh_poisson.hip) ----
hipError_t launch_poisson_p(Ctx &c);  // POISSON_P_CHAN into CP  // once per grid  // for the DELTA_T in the device clock

}  // namespace dh
