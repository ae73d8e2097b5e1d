// C ABI of libdiablo_hip.so (declarations and reference citations in
// include/diablo_hip.h): context set-up, host<->device layout conversion,
// substep/step dispatch and the captured step graph.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>

#include "dh_internal.h"
#include "dh_mlp_dev.h"

using namespace dh;

struct dh_ctx {
  Ctx c;
};

// Members of one shape: their recorded steps have the same launches (kernel,
// grid, block, LDS, stream role) and events, so each launch runs once for all
// of them with the kernel's batched variant over their argument structs.
struct EnsGroup {
  std::vector<dh_ctx *> m;
  // recorded sequences of member 0: the start of the first step (pro), one
  // step ending where the next begins (unit), the last step (epi)
  std::vector<RecOp> seq[3];
  std::vector<void *> dargs[3];  // per launch: device array of the members' argument structs
  hipStream_t st[3] = {};        // main, branch, side
  hipEvent_t ev[4] = {};         // ev_nl01, ev_branch_done, ev_fork, ev_join
  hipEvent_t done = nullptr;
};

// Ensemble over existing contexts (not owned).
struct dh_ens {
  std::vector<dh_ctx *> m;
  std::vector<EnsGroup> g;
  bool batched = false;   // groups built
  bool fallback = false;  // grouping failed: each member's own step graph, member by member
  hipStream_t root = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  hipGraphExec_t graph[2] = {};  // 1 and kStepsPerGraph steps
};

namespace {

const double kPi = 3.14159265358979323846;

int fail(Ctx &c, int code, const std::string &msg) {
  c.err = msg;
  return code;
}

#define HIPCHK(expr)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) return fail(c, DH_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

unsigned long long recip(unsigned d) { return ((1ull << 32) + d - 1) / d; }

// Slot of each output index after the in-place DIF with radices rad[0..nst):
// slot s holds index digrev(s), digrev(s) = s/S + R*digrev'(s%S).
void digrev(int n, const int *rad, int nst, std::vector<int> &idx_of_slot) {
  idx_of_slot.assign(n, 0);
  if (n == 1 || nst == 0) return;
  const int R = rad[0], S = n / R;
  std::vector<int> sub;
  digrev(S, rad + 1, nst - 1, sub);
  for (int s = 0; s < n; ++s) idx_of_slot[s] = s / S + R * sub[s % S];
}

bool make_plan(int n, FftPlan &p, std::vector<int> &dit_in_slot, std::vector<int> &dif_out_slot, bool radix16) {
  int rad[kMaxStages];
  int nst = 0, m = n;
  for (int R : {16, 8, 4, 2, 3})
    while (m % R == 0 && nst < kMaxStages && (R != 16 || radix16)) {
      rad[nst++] = R;
      m /= R;
    }
  if (m != 1) return false;
  p.n = n;
  p.nst = nst;
  int L = n;
  for (int s = 0; s < nst; ++s) {  // decimation in frequency
    FftStage &st = p.dif[s];
    st.R = rad[s];
    st.L = L;
    st.S = L / st.R;
    st.twm = n / L;
    st.fS = recip(st.S);
    st.fnb = recip(n / st.R);
    L /= st.R;
  }
  L = 1;
  for (int s = 0; s < nst; ++s) {  // decimation in time
    FftStage &st = p.dit[s];
    st.R = rad[s];
    L *= st.R;
    st.L = L;
    st.S = L / st.R;
    st.twm = n / L;
    st.fS = recip(st.S);
    st.fnb = recip(n / st.R);
  }
  std::vector<int> ios;
  digrev(n, rad, nst, ios);  // DIF: slot s holds index ios[s]
  dif_out_slot.assign(n, 0);
  for (int s = 0; s < n; ++s) dif_out_slot[ios[s]] = s;
  int rrev[kMaxStages];
  for (int s = 0; s < nst; ++s) rrev[s] = rad[nst - 1 - s];
  digrev(n, rrev, nst, ios);  // DIT input: slot s must hold index ios[s]
  dit_in_slot.assign(n, 0);
  for (int s = 0; s < n; ++s) dit_in_slot[ios[s]] = s;
  return true;
}

std::vector<double2> twiddles(int n, int count) {
  std::vector<double2> t(count);
  for (int m = 0; m < count; ++m) {
    long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)m / (long double)n;
    t[m] = make_double2((double)cosl(a), (double)sinl(a));
  }
  return t;
}

template <class T>
hipError_t dalloc(T **p, size_t count) {
  hipError_t e = hipMalloc((void **)p, sizeof(T) * (count ? count : 1));
  if (e == hipSuccess) e = hipMemset(*p, 0, sizeof(T) * (count ? count : 1));
  return e;
}

void free_all(Ctx &c) {
  void *ptrs[] = {c.tw_row, c.tw_col, c.tw_real, c.posr, c.posrd, c.posz, c.posc, c.kx, c.kz, c.kx2, c.kz2, c.d_dy, c.d_dyf, c.d_gyf, c.d_rdy, c.d_rdyf, c.cu[0], c.cu[1],
                  c.cu[2], c.cp, c.cf[0], c.cf[1], c.cf[2], c.cth, c.cfth, c.phi, c.uphys, c.rphys, c.rd_ta,
                  c.rd_tb, c.clock, c.fac.pf, c.fac.pb, c.fac.tl, c.fac.tb, c.fac.tc, c.fac.tr, c.fac.isbc, c.rd_tr, c.fac.vconst, c.rd_l, c.rd_u, c.stats, c.cun[0], c.cun[1], c.cun[2], c.uphys32, c.rphys32, c.fac.bcl, c.fac.bcd, c.fac.bcu, c.nut, c.cul, c.les_four, c.les_phys, c.sgs, c.les_wk, c.d_lesw, c.d_mlp_p, c.d_mlp_pk, c.les_feat, c.gp_scratch};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (c.step_graph) (void)hipGraphExecDestroy(c.step_graph);
  if (c.steps_graph) (void)hipGraphExecDestroy(c.steps_graph);
  if (c.stream) (void)hipStreamDestroy(c.stream);
  if (c.side) (void)hipStreamDestroy(c.side);
  if (c.ev_fork) (void)hipEventDestroy(c.ev_fork);
  if (c.ev_join) (void)hipEventDestroy(c.ev_join);
  if (c.branch) (void)hipStreamDestroy(c.branch);
  if (c.ev_nl01) (void)hipEventDestroy(c.ev_nl01);
  if (c.ev_branch_done) (void)hipEventDestroy(c.ev_branch_done);
}

// courant.f:15-31: the part of the time-step bound independent of U.
double dt_cap_of(const Ctx &c) {
  const dh_params &p = c.p;
  double dt = 999.0;
  const double dx1 = p.lx / p.nx, dy1 = c.dy[1];
  const double m = dx1 < dy1 ? dx1 : dy1;
  const double v = 0.5 * (m * m) / p.nu;
  if (v < dt) dt = v;
  for (int n = 0; n < p.n_th; ++n) {
    const double w = dt * p.nu / (p.nu / p.pr[n]);
    if (w < dt) dt = w;
  }
  if (p.i_ro != 0.0) {
    const double w = 2.0 * kPi / std::fabs(p.i_ro) / 20.0;
    if (w < dt) dt = w;
  }
  for (int n = 0; n < p.n_th; ++n) {
    if (p.ri[n] != 0.0) {
      const double Nmax = std::sqrt(std::fabs(p.ri[n] * p.th_bc_ymin_c1[n]));
      const double w = (double)0.1f * 2.0 * kPi / Nmax;  // 0.1 is a default-real literal
      if (w < dt) dt = w;
    }
  }
  return dt;
}

int rebuild_tables(Ctx &c) {
  HIPCHK(launch_cn_tables(c, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  return DH_OK;
}

}  // namespace

extern "C" {

int dh_create(const dh_params *p, dh_ctx **out) {
  if (!p || !out) return DH_EARG;
  *out = nullptr;
  dh_ctx *h = new (std::nothrow) dh_ctx();
  if (!h) return DH_EARG;
  Ctx &c = h->c;
  c.p = *p;
  auto reject = [&](const char *msg) {
    c.err = msg;
    std::fprintf(stderr, "dh_create: %s\n", msg);
    delete h;
    return DH_EARG;
  };
  if (p->nx < 8 || p->nx % 4 != 0) return reject("NX must be a multiple of 4, >= 8");
  if (p->nz < 1 || p->ny < 4) return reject("NZ >= 1 and NY >= 4 required");
  if (p->n_th < 0 || p->n_th > DH_MAX_TH) return reject("N_TH out of range");
  const int bcs[] = {p->u_bc_ymin, p->v_bc_ymin, p->w_bc_ymin, p->u_bc_ymax, p->v_bc_ymax, p->w_bc_ymax};
  for (int b : bcs)
    if (b != 0 && b != 1) return reject("velocity BC types must be 0 (Dirichlet) or 1 (Neumann)");
  for (int n = 0; n < p->n_th; ++n)
    if ((p->th_bc_ymin[n] != 0 && p->th_bc_ymin[n] != 1) || (p->th_bc_ymax[n] != 0 && p->th_bc_ymax[n] != 1))
      return reject("scalar BC types must be 0 or 1");
  if (p->variable_dt && p->update_dt < 1) return reject("UPDATE_DT must be >= 1");
  Geo &g = c.g;
  g.nx = p->nx;
  g.ny = p->ny;
  g.nz = p->nz;
  g.nkx = p->nx / 3;
  g.nkz = p->nz / 3;
  g.tnkz = 2 * g.nkz;
  g.nkx1 = g.nkx + 1;
  g.nkz1 = g.tnkz + 1;
  g.pf = g.nkx1 * g.nkz1;
  g.lw = p->nx / 2;
  g.pp = (long long)p->nx * p->nz;
  g.ls = g.lw + 1;
  g.dnx = p->nx;
  g.dnz = p->nz;
  g.rnx = 1.0 / g.dnx;
  g.rnz = 1.0 / g.dnz;
  std::vector<int> posr, posrd, posc, posz;
  // radix 16 only along z: its 16-point butterflies cost registers that the
  // row transforms of the 1024-thread kernels do not have
  if (!make_plan(g.lw, g.prow, posr, posrd, false) || !make_plan(g.nz, g.pcol, posc, posz, true))
    return reject("NX/2 and NZ must factor into 2 and 3");
  c.nfields = 3 + p->n_th;
  c.lds_bytes = g.nz * g.ls * (int)sizeof(double2) + (2 * g.lw + g.nz + 1) * (int)sizeof(double2) +
                (2 * g.lw + 2 * g.nz) * (int)sizeof(int);
  c.lds_bytes32 = g.nz * g.ls * (int)sizeof(float2) + (2 * g.lw + g.nz + 1) * (int)sizeof(float2) +
                  (2 * g.lw + 2 * g.nz) * (int)sizeof(int);
  // planes beyond LDS go through the staged large-plane kernels (FP64, no LES)
  c.gp = c.lds_bytes > kLdsMax;
  if (c.gp && gp_stage_elems(g) < std::max(g.ls, 2 * g.nz))
    return reject("x-z plane too large (a row or a column of it does not fit the staging buffer)");
  if (g.ny + 2 > 8 * kSegRows) return reject("NY+2 is beyond the column solver's register segments (8 x 36 rows)");
  int thr = (g.nz * g.lw + 7) / 8;
  thr = ((thr + 63) / 64) * 64;
  if (thr < 64) thr = 64;
  if (thr > 1024) thr = 1024;
  while (thr * 4 < g.pf && thr < 1024) thr += 64;  // k_nonlinear keeps <= 4 modes per thread
  if (!c.gp && thr * 4 < g.pf) return reject("plane has too many Fourier modes");
  c.plane_threads = thr;
  // register rows: NX = 128 (slabs of 8 rows per wave) or NX = 64 (slabs of 16)
  c.reg_rows = !c.gp && thr % 64 == 0 &&
               ((g.lw == 64 && g.nz % 8 == 0 && g.nkx < 8 * kRowK2) || (g.lw == 32 && g.nz % 16 == 0));
  // k_nonlinear's register variant (NX = 128): 512 threads x 8 modes
  c.reg_rows_nl = c.reg_rows && g.lw == 64 && g.pf <= 512 * 8;
  // NX = 64 with at most 2048 modes: the nonlinear pass's register rows are
  // in its <= 512-thread instantiation, and every offered size runs them
  if (c.reg_rows && g.lw == 32 && g.pf <= 4 * 512 && c.plane_threads > 512) c.plane_threads = 512;
  {
  }
  g.f_npair = recip(g.lw / 2 + 1);
  g.f_nkx1 = recip(g.nkx1);
  g.f_lw = recip(g.lw);
  g.f_nz = recip(g.nz);
  // mpi.f:1022-1046 (RANKY=0 branch on one rank)
  c.jstart = 2;
  c.jend = g.ny;
  for (int n = 0; n < DH_MAX_TH; ++n) {
    c.jend_th[n] = g.ny;
    c.jstart_th[n] = (n < p->n_th && p->th_bc_ymin[n] == 1) ? 1 : 2;
  }
  // write-back rows (channel.f:826-830, 895-899, 959-963, 1023-1027)
  c.wlo[0] = c.jstart - 1; c.whi[0] = c.jend + 1;
  c.wlo[1] = 1; c.whi[1] = g.ny + 1;
  c.wlo[2] = c.jstart - 1; c.whi[2] = c.jend + 1;
  for (int n = 0; n < p->n_th; ++n) {
    c.wlo[3 + n] = c.jstart_th[n] - 1;
    c.whi[3 + n] = c.jend_th[n] + 1;
  }
  c.gy.assign(g.ny + 2, 0.0);
  c.gyf.assign(g.ny + 2, 0.0);
  c.dy.assign(g.ny + 2, 0.0);
  c.dyf.assign(g.ny + 2, 0.0);

  hipError_t e;
#define ACHK(expr)                          \
  if ((e = (expr)) != hipSuccess) {         \
    c.err = hipGetErrorString(e);           \
    std::fprintf(stderr, "dh_create: %s: %s\n", #expr, c.err.c_str()); \
    free_all(c);                            \
    delete h;                               \
    return DH_EHIP;                         \
  }
  ACHK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  ACHK(hipStreamCreateWithFlags(&c.side, hipStreamNonBlocking));
  ACHK(hipEventCreateWithFlags(&c.ev_fork, hipEventDisableTiming));
  ACHK(hipEventCreateWithFlags(&c.ev_join, hipEventDisableTiming));
  ACHK(hipStreamCreateWithFlags(&c.branch, hipStreamNonBlocking));
  ACHK(hipEventCreateWithFlags(&c.ev_nl01, hipEventDisableTiming));
  ACHK(hipEventCreateWithFlags(&c.ev_branch_done, hipEventDisableTiming));
  const size_t nyy = (size_t)g.ny + 2;
  const size_t cf_elems = nyy * g.pf;
  const size_t ph_elems = nyy * (size_t)g.pp;
  ACHK(dalloc(&c.tw_row, g.lw));
  ACHK(dalloc(&c.tw_col, g.nz));
  ACHK(dalloc(&c.tw_real, g.lw + 1));
  ACHK(dalloc(&c.posr, g.lw));
  ACHK(dalloc(&c.posrd, g.lw));
  ACHK(dalloc(&c.posz, g.nz));
  ACHK(dalloc(&c.posc, g.nz));
  ACHK(dalloc(&c.kx, g.nkx1));
  ACHK(dalloc(&c.kx2, g.nkx1));
  ACHK(dalloc(&c.kz, g.nkz1));
  ACHK(dalloc(&c.kz2, g.nkz1));
  ACHK(dalloc(&c.d_dy, nyy));
  ACHK(dalloc(&c.d_dyf, nyy));
  ACHK(dalloc(&c.d_gyf, nyy));
  ACHK(dalloc(&c.d_rdy, nyy));
  ACHK(dalloc(&c.d_rdyf, nyy));
  for (int i = 0; i < 3; ++i) {
    ACHK(dalloc(&c.cu[i], cf_elems));
    ACHK(dalloc(&c.cf[i], cf_elems));
  }
  ACHK(dalloc(&c.cp, cf_elems));
  ACHK(dalloc(&c.phi, cf_elems));
  for (int i = 0; i < 3; ++i) ACHK(dalloc(&c.cun[i], cf_elems));
  ACHK(dalloc(&c.cth, cf_elems * (p->n_th > 0 ? p->n_th : 1)));
  ACHK(dalloc(&c.cfth, cf_elems * (p->n_th > 0 ? p->n_th : 1)));
  ACHK(dalloc(&c.uphys, ph_elems * c.nfields));
  ACHK(dalloc(&c.rphys, ph_elems * c.nfields));
  if (c.gp) {
    c.gp_slot = (long long)g.nz * g.nkx1 + g.pf;  // the plane between passes, then the accumulators
    ACHK(dalloc(&c.gp_scratch, 2 * (size_t)kGpSlots * (size_t)c.gp_slot));
  }
  ACHK(dalloc(&c.rd_ta, cf_elems));
  ACHK(dalloc(&c.rd_tb, cf_elems));
  ACHK(dalloc(&c.rd_tr, cf_elems));
  ACHK(dalloc(&c.stats, (size_t)DH_NSTATS * nyy));
  ACHK(dalloc(&c.rd_l, nyy));
  ACHK(dalloc(&c.rd_u, nyy));
  ACHK(dalloc(&c.clock, 1));
  ACHK(dalloc(&c.fac.pf, 3 * (size_t)c.nfields * nyy));
  ACHK(dalloc(&c.fac.pb, 3 * (size_t)c.nfields * nyy));
  ACHK(dalloc(&c.fac.tl, 3 * (size_t)c.nfields * nyy));
  ACHK(dalloc(&c.fac.tb, 3 * (size_t)c.nfields * nyy));
  ACHK(dalloc(&c.fac.tc, 3 * (size_t)c.nfields * nyy));
  ACHK(dalloc(&c.fac.tr, 3 * (size_t)c.nfields * nyy));
  {
    const int rows = g.ny - 1;  // Poisson rows 2..NY swept per mode
    const int sl = (rows + kRdsSeg - 1) / kRdsSeg;
    c.rd_seg = (sl <= kRdsRows) ? sl : 0;
  }
  c.fac.nseg = (nyy <= 4 * (size_t)kSegRows) ? 4 : 8;
  c.fac.seg_rows = (int)((nyy + c.fac.nseg - 1) / c.fac.nseg);
  ACHK(dalloc(&c.fac.isbc, (size_t)c.nfields * nyy));
  ACHK(dalloc(&c.fac.vconst, (size_t)c.nfields * nyy));
  ACHK(dalloc(&c.fac.bcl, (size_t)c.nfields * nyy));
  ACHK(dalloc(&c.fac.bcd, (size_t)c.nfields * nyy));
  ACHK(dalloc(&c.fac.bcu, (size_t)c.nfields * nyy));
  {
    auto tr = twiddles(g.lw, g.lw), tc = twiddles(g.nz, g.nz), tx = twiddles(g.nx, g.lw + 1);
    ACHK(hipMemcpy(c.tw_row, tr.data(), tr.size() * sizeof(double2), hipMemcpyHostToDevice));
    ACHK(hipMemcpy(c.tw_col, tc.data(), tc.size() * sizeof(double2), hipMemcpyHostToDevice));
    ACHK(hipMemcpy(c.tw_real, tx.data(), tx.size() * sizeof(double2), hipMemcpyHostToDevice));
    ACHK(hipMemcpy(c.posr, posr.data(), posr.size() * sizeof(int), hipMemcpyHostToDevice));
    ACHK(hipMemcpy(c.posrd, posrd.data(), posrd.size() * sizeof(int), hipMemcpyHostToDevice));
    ACHK(hipMemcpy(c.posz, posz.data(), posz.size() * sizeof(int), hipMemcpyHostToDevice));
    ACHK(hipMemcpy(c.posc, posc.data(), posc.size() * sizeof(int), hipMemcpyHostToDevice));
    // fft.f:77-100
    const double PI = 4.0 * std::atan(1.0);
    std::vector<double> kx(g.nkx1), kx2(g.nkx1), kz(g.nkz1), kz2(g.nkz1);
    for (int i = 0; i < g.nkx1; ++i) {
      kx[i] = (double)i * (2. * PI) / p->lx;
      kx2[i] = kx[i] * kx[i];
    }
    for (int k = 0; k <= g.nkz; ++k) kz[k] = (double)k * (2. * PI) / p->lz;
    for (int k = 1; k <= g.nkz; ++k) kz[g.tnkz + 1 - k] = -(double)k * (2. * PI) / p->lz;
    for (int k = 0; k < g.nkz1; ++k) kz2[k] = kz[k] * kz[k];
    ACHK(hipMemcpy(c.kx, kx.data(), g.nkx1 * 8, hipMemcpyHostToDevice));
    ACHK(hipMemcpy(c.kx2, kx2.data(), g.nkx1 * 8, hipMemcpyHostToDevice));
    ACHK(hipMemcpy(c.kz, kz.data(), g.nkz1 * 8, hipMemcpyHostToDevice));
    ACHK(hipMemcpy(c.kz2, kz2.data(), g.nkz1 * 8, hipMemcpyHostToDevice));
  }
  g.tw_row = c.tw_row;
  g.tw_col = c.tw_col;
  g.tw_real = c.tw_real;
  g.posr = c.posr;
  g.posrd = c.posrd;
  g.posz = c.posz;
  g.posc = c.posc;
  g.kx = c.kx;
  g.kz = c.kz;
  g.kx2 = c.kx2;
  g.kz2 = c.kz2;
  g.dy = c.d_dy;
  g.dyf = c.d_dyf;
  g.rdy = c.d_rdy;
  g.rdyf = c.d_rdyf;
  {
    Clock ck{};
    ck.time = 0.0;
    ck.dt = p->delta_t;
    // diablo_io.f:116-124
    ck.h_bar[0] = p->delta_t * (8.0 / 15.0);
    ck.h_bar[1] = p->delta_t * (2.0 / 15.0);
    ck.h_bar[2] = p->delta_t * (5.0 / 15.0);
    ck.beta_bar[0] = 1.0;
    ck.beta_bar[1] = 25.0 / 8.0;
    ck.beta_bar[2] = 9.0 / 4.0;
    ck.zeta_bar[0] = 0.0;
    ck.zeta_bar[1] = -17.0 / 8.0;
    ck.zeta_bar[2] = -5.0 / 4.0;
    ck.cfl_min_bits = 0x7FF0000000000000ull;
    ck.time_step = 0;
    ACHK(hipMemcpy(c.clock, &ck, sizeof(Clock), hipMemcpyHostToDevice));
  }
  ACHK(prepare_fft_kernels(c));
  ACHK(prepare_step_kernels(c));
#undef ACHK
  *out = h;
  return DH_OK;
}

int dh_destroy(dh_ctx *h) {
  if (!h) return DH_EARG;
  (void)hipStreamSynchronize(h->c.stream);
  free_all(h->c);
  delete h;
  return DH_OK;
}

const char *dh_last_error(dh_ctx *h) { return h ? h->c.err.c_str() : "null context"; }

static hipError_t upload_les_widths(Ctx &c);

int dh_set_grid(dh_ctx *h, const double *gy, const double *gyf, const double *dy, const double *dyf) {
  if (!h || !gy || !gyf || !dy || !dyf) return DH_EARG;
  Ctx &c = h->c;
  const int nyy = c.g.ny + 2;
  c.gy.assign(gy, gy + nyy);
  c.gyf.assign(gyf, gyf + nyy);
  c.dy.assign(dy, dy + nyy);
  c.dyf.assign(dyf, dyf + nyy);
  for (int j = 1; j <= c.g.ny; ++j)
    if (!(c.dy[j] > 0.0) || !(c.dyf[j] > 0.0)) return fail(c, DH_EARG, "DY/DYF must be positive on rows 1..NY");
  HIPCHK(hipMemcpy(c.d_dy, dy, nyy * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c.d_dyf, dyf, nyy * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c.d_gyf, gyf, nyy * 8, hipMemcpyHostToDevice));
  if (c.d_lesw) HIPCHK(upload_les_widths(c));
  {
    std::vector<double> r(nyy), rf(nyy);
    for (int j = 0; j < nyy; ++j) {
      r[j] = (c.dy[j] != 0.0) ? 1.0 / c.dy[j] : 0.0;
      rf[j] = (c.dyf[j] != 0.0) ? 1.0 / c.dyf[j] : 0.0;
    }
    HIPCHK(hipMemcpy(c.d_rdy, r.data(), nyy * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c.d_rdyf, rf.data(), nyy * 8, hipMemcpyHostToDevice));
  }
  // channel.f:1126-1134 coefficient tables
  std::vector<double> rl(nyy, 0.0), ru(nyy, 0.0);
  for (int j = 1; j <= c.g.ny; ++j) {
    rl[j] = 1. / (c.dy[j] * c.dyf[j]);
    ru[j] = 1. / (c.dy[j + 1] * c.dyf[j]);
  }
  HIPCHK(hipMemcpy(c.rd_l, rl.data(), nyy * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c.rd_u, ru.data(), nyy * 8, hipMemcpyHostToDevice));
  c.grid_set = true;
  c.dt_cap = dt_cap_of(c);
  HIPCHK(launch_rem_div_factors(c));
  return rebuild_tables(c);
}

int dh_set_grid_h5(dh_ctx *h, const char *path) {
  if (!h || !path) return DH_EARG;
  Ctx &c = h->c;
  const int NY = c.g.ny;
  std::vector<double> y(NY + 2, 0.0);
  int64_t n = 0;
  char err[256];
  int r = dh_h5_read_f64(path, "/grids/y", y.data(), NY + 2, &n, err, (int32_t)sizeof(err));
  if (r != DH_OK) return fail(c, r, err);
  if (n != NY + 1) return fail(c, DH_EARG, "Grid file and program do not match: /grids/y must hold NY+1 values");
  // hdf5.f:690-705 and channel.f:1646-1652 on one rank
  std::vector<double> GY(NY + 2, 0.0), GYF(NY + 2, 0.0), DY(NY + 2, 0.0), DYF(NY + 2, 0.0);
  for (int j = 1; j <= NY + 1; ++j) GY[j] = y[j - 1];
  for (int j = 1; j <= NY; ++j) GYF[j] = 0.5 * (GY[j] + GY[j + 1]);
  GYF[0] = 2.0 * GYF[1] - GYF[2];
  GYF[NY + 1] = 2.0 * GYF[NY] - GYF[NY - 1];
  for (int j = 1; j <= NY + 1; ++j) DY[j] = GYF[j] - GYF[j - 1];
  for (int j = 1; j <= NY; ++j) DYF[j] = GY[j + 1] - GY[j];
  DYF[NY + 1] = DYF[NY];
  return dh_set_grid(h, GY.data(), GYF.data(), DY.data(), DYF.data());
}

int dh_set_clock(dh_ctx *h, double time, double delta_t, int32_t time_step) {
  if (!h) return DH_EARG;
  Ctx &c = h->c;
  if (!c.grid_set) return fail(c, DH_ESTATE, "dh_set_grid first");
  HIPCHK(hipStreamSynchronize(c.stream));
  Clock ck;
  HIPCHK(hipMemcpy(&ck, c.clock, sizeof(Clock), hipMemcpyDeviceToHost));
  ck.time = time;
  ck.time_step = time_step;
  const bool newdt = (delta_t != ck.dt);
  if (newdt) {
    ck.dt = delta_t;
    ck.h_bar[0] = delta_t * (8.0 / 15.0);
    ck.h_bar[1] = delta_t * (2.0 / 15.0);
    ck.h_bar[2] = delta_t * (5.0 / 15.0);
  }
  HIPCHK(hipMemcpy(c.clock, &ck, sizeof(Clock), hipMemcpyHostToDevice));
  if (newdt) return rebuild_tables(c);
  return DH_OK;
}

int dh_get_clock(dh_ctx *h, double *time, double *delta_t, double *h_bar3, int32_t *time_step) {
  if (!h) return DH_EARG;
  Ctx &c = h->c;
  HIPCHK(hipStreamSynchronize(c.stream));
  Clock ck;
  HIPCHK(hipMemcpy(&ck, c.clock, sizeof(Clock), hipMemcpyDeviceToHost));
  if (time) *time = ck.time;
  if (delta_t) *delta_t = ck.dt;
  if (h_bar3) std::memcpy(h_bar3, ck.h_bar, 3 * sizeof(double));
  if (time_step) *time_step = ck.time_step;
  return DH_OK;
}

}  // extern "C"

namespace {

// Fortran complex (0:NXP,0:NZ+1,0:NY+1) <-> de-aliased [J][K][I].
int pack_fourier(Ctx &c, const double *src, double2 *dst_dev, int nfield, const char *name) {
  const Geo &g = c.g;
  const int nxp = g.nx / 2, nzz = g.nz + 2, nyy = g.ny + 2;
  std::vector<double2> buf((size_t)nyy * g.pf);
  for (int f = 0; f < nfield; ++f) {
    const double *s = src + (size_t)f * 2 * (nxp + 1) * nzz * nyy;
    for (int j = 0; j < nyy; ++j)
      for (int k = 0; k < g.nkz1; ++k)
        for (int i = 0; i <= nxp; ++i) {
          const size_t si = 2 * ((size_t)i + (size_t)(nxp + 1) * ((size_t)k + (size_t)nzz * j));
          if (i < g.nkx1) {
            buf[(size_t)j * g.pf + (size_t)k * g.nkx1 + i] = make_double2(s[si], s[si + 1]);
          } else if (s[si] != 0.0 || s[si + 1] != 0.0) {
            char msg[160];
            std::snprintf(msg, sizeof msg, "%s(%d,%d,%d) is non-zero beyond the de-aliased modes (I>NKX=%d)", name, i,
                          k, j, g.nkx);
            return fail(c, DH_ELAYOUT, msg);
          }
        }
    HIPCHK(hipMemcpy(dst_dev + (size_t)f * nyy * g.pf, buf.data(), buf.size() * sizeof(double2),
                     hipMemcpyHostToDevice));
  }
  return DH_OK;
}

int unpack_fourier(Ctx &c, const double2 *src_dev, double *dst, int nfield) {
  const Geo &g = c.g;
  const int nxp = g.nx / 2, nzz = g.nz + 2, nyy = g.ny + 2;
  std::vector<double2> buf((size_t)nyy * g.pf);
  for (int f = 0; f < nfield; ++f) {
    HIPCHK(hipMemcpy(buf.data(), src_dev + (size_t)f * nyy * g.pf, buf.size() * sizeof(double2),
                     hipMemcpyDeviceToHost));
    double *d = dst + (size_t)f * 2 * (nxp + 1) * nzz * nyy;
    std::memset(d, 0, sizeof(double) * 2 * (size_t)(nxp + 1) * nzz * nyy);
    for (int j = 0; j < nyy; ++j)
      for (int k = 0; k < g.nkz1; ++k)
        for (int i = 0; i < g.nkx1; ++i) {
          const size_t di = 2 * ((size_t)i + (size_t)(nxp + 1) * ((size_t)k + (size_t)nzz * j));
          const double2 v = buf[(size_t)j * g.pf + (size_t)k * g.nkx1 + i];
          d[di] = v.x;
          d[di + 1] = v.y;
        }
  }
  return DH_OK;
}

// Fortran real (0:NX+1,0:NZ+1,0:NY+1) <-> [J][z][x]
int pack_phys(Ctx &c, const double *src, double *dst_dev, int nfield) {
  const Geo &g = c.g;
  const int nxx = g.nx + 2, nzz = g.nz + 2, nyy = g.ny + 2;
  std::vector<double> buf((size_t)nyy * g.pp);
  for (int f = 0; f < nfield; ++f) {
    const double *s = src + (size_t)f * nxx * nzz * nyy;
    for (int j = 0; j < nyy; ++j)
      for (int k = 0; k < g.nz; ++k)
        for (int i = 0; i < g.nx; ++i)
          buf[(size_t)j * g.pp + (size_t)k * g.nx + i] = s[(size_t)i + (size_t)nxx * ((size_t)k + (size_t)nzz * j)];
    HIPCHK(hipMemcpy(dst_dev + (size_t)f * nyy * g.pp, buf.data(), buf.size() * 8, hipMemcpyHostToDevice));
  }
  return DH_OK;
}

int unpack_phys(Ctx &c, const double *src_dev, double *dst, int nfield) {
  const Geo &g = c.g;
  const int nxx = g.nx + 2, nzz = g.nz + 2, nyy = g.ny + 2;
  std::vector<double> buf((size_t)nyy * g.pp);
  for (int f = 0; f < nfield; ++f) {
    HIPCHK(hipMemcpy(buf.data(), src_dev + (size_t)f * nyy * g.pp, buf.size() * 8, hipMemcpyDeviceToHost));
    double *d = dst + (size_t)f * nxx * nzz * nyy;
    std::memset(d, 0, sizeof(double) * (size_t)nxx * nzz * nyy);
    for (int j = 0; j < nyy; ++j)
      for (int k = 0; k < g.nz; ++k)
        for (int i = 0; i < g.nx; ++i)
          d[(size_t)i + (size_t)nxx * ((size_t)k + (size_t)nzz * j)] = buf[(size_t)j * g.pp + (size_t)k * g.nx + i];
  }
  return DH_OK;
}

}  // namespace

extern "C" {

int dh_upload(dh_ctx *h, const dh_host_state *st) {
  if (!h || !st) return DH_EARG;
  Ctx &c = h->c;
  HIPCHK(hipStreamSynchronize(c.stream));
  const long long fs = (long long)(c.g.ny + 2) * c.g.pp;
  int r;
  const double *fl[7] = {st->cu1, st->cu2, st->cu3, st->cp, st->cf1, st->cf2, st->cf3};
  double2 *dv[7] = {c.cu[0], c.cu[1], c.cu[2], c.cp, c.cf[0], c.cf[1], c.cf[2]};
  const char *nm[7] = {"CU1", "CU2", "CU3", "CP", "CF1", "CF2", "CF3"};
  for (int i = 0; i < 7; ++i)
    if (fl[i] && (r = pack_fourier(c, fl[i], dv[i], 1, nm[i])) != DH_OK) return r;
  const int nth = c.p.n_th;
  if (nth > 0) {
    if (st->cth && (r = pack_fourier(c, st->cth, c.cth, nth, "CTH")) != DH_OK) return r;
    if (st->cfth && (r = pack_fourier(c, st->cfth, c.cfth, nth, "CFTH")) != DH_OK) return r;
    if (st->th && (r = pack_phys(c, st->th, c.uphys + 3 * fs, nth)) != DH_OK) return r;
    if (st->th) c.th_phys32 = false;
  }
  return DH_OK;
}

// FP32 physical fields widened to FP64 (the mixed form's physical TH).
__global__ void k_widen(const float *src, double *dst, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

int dh_download(dh_ctx *h, dh_host_state *st) {
  if (!h || !st) return DH_EARG;
  Ctx &c = h->c;
  HIPCHK(hipStreamSynchronize(c.stream));
  const long long fs = (long long)(c.g.ny + 2) * c.g.pp;
  int r;
  double *fl[7] = {st->cu1, st->cu2, st->cu3, st->cp, st->cf1, st->cf2, st->cf3};
  double2 *dv[7] = {c.cu[0], c.cu[1], c.cu[2], c.cp, c.cf[0], c.cf[1], c.cf[2]};
  for (int i = 0; i < 7; ++i)
    if (fl[i] && (r = unpack_fourier(c, dv[i], fl[i], 1)) != DH_OK) return r;
  const int nth = c.p.n_th;
  if (nth > 0) {
    if (st->cth && (r = unpack_fourier(c, c.cth, st->cth, nth)) != DH_OK) return r;
    if (st->cfth && (r = unpack_fourier(c, c.cfth, st->cfth, nth)) != DH_OK) return r;
    if (st->th && c.th_phys32) {
      // the last substep ran in the mixed form: its TH is in the FP32 array
      hipLaunchKernelGGL(k_widen, dim3(1024), dim3(256), 0, c.stream, c.uphys32 + 3 * fs, c.uphys + 3 * fs,
                         (long long)nth * fs);
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(c.stream));
      c.th_phys32 = false;
    }
    if (st->th && (r = unpack_phys(c, c.uphys + 3 * fs, st->th, nth)) != DH_OK) return r;
  }
  return DH_OK;
}

int dh_rk_substep(dh_ctx *h, int32_t rk_step) {
  if (!h) return DH_EARG;
  Ctx &c = h->c;
  if (rk_step < 1 || rk_step > 3) return fail(c, DH_EARG, "rk_step must be 1..3");
  if (!c.grid_set) return fail(c, DH_ESTATE, "dh_set_grid first");
  HIPCHK(launch_substep(c, rk_step - 1));
  c.th_phys32 = c.mixed;
  return DH_OK;
}

static void drop_graphs(Ctx &c) {
  if (c.step_graph) (void)hipGraphExecDestroy(c.step_graph);
  if (c.steps_graph) (void)hipGraphExecDestroy(c.steps_graph);
  c.step_graph = c.steps_graph = nullptr;
  c.graph_ok = c.steps_graph_ok = false;
}

int dh_set_mixed_precision(dh_ctx *h, int32_t on) {
  if (!h || (on != 0 && on != 1)) return DH_EARG;
  Ctx &c = h->c;
  if ((on != 0) == c.mixed) return DH_OK;
  if (on && c.les) return fail(c, DH_EARG, "the mixed form does not carry the LES model (FP64 only)");
  if (on && c.gp) return fail(c, DH_EARG, "planes beyond LDS run in FP64 only");
  HIPCHK(hipStreamSynchronize(c.stream));
  if (on && !c.uphys32) {
    const size_t n = (size_t)(c.g.ny + 2) * (size_t)c.g.pp * (size_t)c.nfields;
    HIPCHK(dalloc(&c.uphys32, n));
    HIPCHK(dalloc(&c.rphys32, n));
  }
  c.mixed = (on != 0);
  drop_graphs(c);  // the captured steps launch the other form's kernels
  return DH_OK;
}

int dh_mixed_precision(dh_ctx *h) { return h ? (h->c.mixed ? 1 : 0) : DH_EARG; }

// LES filter widths (les.f:196-217) with -2 C_SMAG^2 absorbed, C_SMAG =
// 0.13, beta_sgs = 1: DYW(J) at GY, DYFW(J) at GYF for J = 1..NY+1, from the
// host's DY, DYF so they match the reference's real powers.
static hipError_t upload_les_widths(Ctx &c) {
  const int nyy = c.g.ny + 2;
  std::vector<double> w(2 * (size_t)nyy, 0.0);
  const double C_SMAG = 0.13, beta_sgs = 1.0, dx = c.p.lx / c.g.nx, dz = c.p.lz / c.g.nz;
  for (int J = 1; J <= c.g.ny + 1; ++J) {
    w[J] = -2.0 * pow(C_SMAG, 2.0) * pow(dx * beta_sgs * c.dy[J] * 2.0 * dz * beta_sgs, 2.0 / 3.0);
    w[nyy + J] = -2.0 * pow(C_SMAG, 2.0) * pow(dx * beta_sgs * c.dyf[J] * 2.0 * dz * beta_sgs, 2.0 / 3.0);
  }
  return hipMemcpy(c.d_lesw, w.data(), w.size() * 8, hipMemcpyHostToDevice);
}

int dh_set_les(dh_ctx *h, int32_t on) {
  if (!h || (on != 0 && on != 1)) return DH_EARG;
  Ctx &c = h->c;
  if ((on != 0) == c.les) return DH_OK;
  if (on && c.mixed) return fail(c, DH_EARG, "LES runs in FP64: switch the mixed form off first");
  if (on && c.gp) return fail(c, DH_EARG, "LES needs planes that fit in LDS");
  HIPCHK(hipStreamSynchronize(c.stream));
  if (on && !c.nut) {
    const size_t nyy = (size_t)c.g.ny + 2, pp = (size_t)c.g.pp, pf = (size_t)c.g.pf, nth = (size_t)c.p.n_th;
    HIPCHK(dalloc(&c.nut, nyy * pp));
    HIPCHK(dalloc(&c.cul, 3 * nyy * pf));
    HIPCHK(dalloc(&c.les_four, (kLesFour + 2 * nth) * nyy * pf));
    HIPCHK(dalloc(&c.les_phys, (kLesPhys + 2 * nth) * nyy * pp));
    HIPCHK(dalloc(&c.sgs, (3 + nth) * nyy * pf));
    HIPCHK(dalloc(&c.les_wk, 3 * (size_t)c.nfields * nyy * pp));
    HIPCHK(dalloc(&c.d_lesw, 2 * nyy));
  }
  if (on && c.grid_set) HIPCHK(upload_les_widths(c));
  c.les = (on != 0);
  drop_graphs(c);  // the captured steps launch the other path's kernels
  return DH_OK;
}

int dh_les(dh_ctx *h) { return h ? (h->c.les ? 1 : 0) : DH_EARG; }

int dh_set_les_closure(dh_ctx *h, const float *w, int32_t n) {
  if (!h) return DH_EARG;
  Ctx &c = h->c;
  if (!w) {
    if (c.les_mlp) {
      HIPCHK(hipStreamSynchronize(c.stream));
      c.les_mlp = false;
      drop_graphs(c);
    }
    return DH_OK;
  }
  if (n != kMlpNP + 3)
    return fail(c, DH_EARG, "the closure takes the network's 10497 parameters then its time, length and NU_T scales");
  for (int i = 0; i < n; ++i)
    if (!std::isfinite(w[i])) return fail(c, DH_EARG, "closure parameters must be finite");
  if (!(w[kMlpNP] > 0.f && w[kMlpNP + 1] > 0.f && w[kMlpNP + 2] > 0.f))
    return fail(c, DH_EARG, "the closure's scales must be positive");
  HIPCHK(hipStreamSynchronize(c.stream));
  if (!c.d_mlp_p) {
    HIPCHK(dalloc(&c.d_mlp_p, (size_t)kMlpNP));
    HIPCHK(hipMalloc(&c.d_mlp_pk, 16 * (size_t)kPAll));
    HIPCHK(dalloc(&c.les_feat, (size_t)kLesFeat * (c.g.ny + 2) * c.g.pp));
  }
  HIPCHK(hipMemcpy(c.d_mlp_p, w, sizeof(float) * kMlpNP, hipMemcpyHostToDevice));
  HIPCHK(launch_mlp_pack(c.d_mlp_p, c.d_mlp_pk, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  c.mlp_ts = w[kMlpNP];
  c.mlp_ls = w[kMlpNP + 1];
  c.mlp_nus = w[kMlpNP + 2];
  if (!c.les_mlp) c.les_mlp = true;
  drop_graphs(c);  // the captured steps hold the old scales
  return DH_OK;
}

int dh_les_features(dh_ctx *h) {
  if (!h) return DH_EARG;
  Ctx &c = h->c;
  if (!c.les || !c.les_mlp) return fail(c, DH_ESTATE, "the closure's inputs need LES and the closure on");
  if (!c.grid_set) return fail(c, DH_ESTATE, "no grid");
  HIPCHK(launch_les_bc_four(c, c.stream));
  HIPCHK(launch_state_to_phys(c));
  HIPCHK(launch_les_features(c, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  return DH_OK;
}

static int enqueue_steps(Ctx &c, int nsteps);

// Whether the nonlinear pass's LDS variant of <= 512 threads serves the
// plane with register x rows (NX = 64): then only blocks of up to 512
// threads are offered, so that every valid size runs the same rows.
static bool rows32_caps_threads(const Ctx &c) {
  return c.reg_rows && c.g.lw == 32 && c.g.pf <= 4 * 512;
}

// Block sizes of the plane kernels and of the nonlinear pass's LDS variant:
// any multiple of 64 up to 1024 (512 where rows32_caps_threads) that leaves
// the nonlinear pass at most 4 modes per thread.  The size only decides
// which thread computes which butterfly or mode, so results do not depend
// on it.
static bool plane_threads_ok(const Ctx &c, int t) {
  return t >= 64 && t <= (rows32_caps_threads(c) ? 512 : 1024) && t % 64 == 0 &&
         (c.reg_rows_nl || t * 4 >= c.g.pf);
}

int dh_set_plane_threads(dh_ctx *h, int32_t threads) {
  if (!h) return DH_EARG;
  Ctx &c = h->c;
  if (!plane_threads_ok(c, threads))
    return fail(c, DH_EARG, "plane threads must be a multiple of 64 in 64..1024 (64..512 for NX = 64 planes "
                            "of at most 2048 Fourier modes) with 4 x threads >= the plane's Fourier modes");
  if (threads == c.plane_threads) return DH_OK;
  HIPCHK(hipStreamSynchronize(c.stream));
  c.plane_threads = threads;
  drop_graphs(c);  // the captured steps hold the old block size
  return DH_OK;
}

int dh_plane_threads(dh_ctx *h) { return h ? h->c.plane_threads : DH_EARG; }

// A copy of a context's time-stepped state (Fourier fields, CF, CP, the
// scalars, NU_T with LES, and the clock) on the device, for trial steps that
// must leave it unchanged; restore() also rebuilds the CN tables of the
// restored DELTA_T.
struct StateCopy {
  Ctx *c = nullptr;
  int n = 0;  // fields saved: the nine of the state, and NU_T with LES
  double2 *src[10] = {}, *save[10] = {};
  size_t cnt[10] = {};
  Clock *ck = nullptr;
  hipError_t init(Ctx &cx) {
    c = &cx;
    const size_t cf = (size_t)(cx.g.ny + 2) * cx.g.pf, nth = (size_t)(cx.p.n_th > 0 ? cx.p.n_th : 1);
    double2 *s0[10] = {cx.cu[0], cx.cu[1], cx.cu[2], cx.cp, cx.cf[0], cx.cf[1], cx.cf[2], cx.cth, cx.cfth,
                       reinterpret_cast<double2 *>(cx.nut)};
    // NU_T: (NY+2) NX NZ doubles, an even count (NX is even)
    const size_t n0[10] = {cf, cf, cf, cf, cf, cf, cf, cf * nth, cf * nth, (size_t)(cx.g.ny + 2) * cx.g.pp / 2};
    n = cx.nut ? 10 : 9;
    hipError_t e = hipSuccess;
    for (int i = 0; i < n && e == hipSuccess; ++i) {
      src[i] = s0[i];
      cnt[i] = n0[i];
      e = hipMalloc(&save[i], cnt[i] * sizeof(double2));
    }
    if (e == hipSuccess) e = hipMalloc(&ck, sizeof(Clock));
    if (e == hipSuccess) e = copy(false);
    return e;
  }
  hipError_t copy(bool restore) {
    hipError_t e = hipSuccess;
    for (int i = 0; i < n && e == hipSuccess; ++i)
      e = hipMemcpyAsync(restore ? src[i] : save[i], restore ? save[i] : src[i], cnt[i] * sizeof(double2),
                         hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(restore ? c->clock : ck, restore ? ck : c->clock, sizeof(Clock), hipMemcpyDeviceToDevice,
                         c->stream);
    return e;
  }
  hipError_t restore() {
    hipError_t e = copy(true);
    if (e == hipSuccess) e = launch_cn_tables(*c, c->stream);
    c->tables_pending = false;
    return e;
  }
  ~StateCopy() {
    for (double2 *q : save)
      if (q) (void)hipFree(q);
    if (ck) (void)hipFree(ck);
  }
};

// Times nsteps uncaptured steps (the best of two runs) with each valid block
// size from the context's current state, restoring the state before each
// run and at the end, and keeps the fastest.
int dh_autotune(dh_ctx *h, int32_t nsteps, int32_t *chosen) {
  if (!h || nsteps < 1) return DH_EARG;
  Ctx &c = h->c;
  if (!c.grid_set) return fail(c, DH_ESTATE, "dh_set_grid first");
  HIPCHK(hipStreamSynchronize(c.stream));
  StateCopy sc;
  hipError_t e = sc.init(c);
  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (e == hipSuccess) e = hipEventCreate(&t0);
  if (e == hipSuccess) e = hipEventCreate(&t1);
  const int keep = c.plane_threads;
  int best = keep;
  float best_ms = 0.f;
  for (int t = 64; t <= 1024 && e == hipSuccess; t *= 2) {
    if (!plane_threads_ok(c, t)) continue;
    c.plane_threads = t;
    float run_best = 0.f;
    for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {
      e = sc.restore();
      if (e == hipSuccess) e = hipEventRecord(t0, c.stream);
      if (e == hipSuccess && enqueue_steps(c, nsteps) != DH_OK) e = hipErrorLaunchFailure;
      if (e == hipSuccess) e = hipEventRecord(t1, c.stream);
      if (e == hipSuccess) e = hipEventSynchronize(t1);
      float ms = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
      if (e == hipSuccess && (rep == 0 || ms < run_best)) run_best = ms;
    }
    if (e == hipSuccess && (best_ms == 0.f || run_best < best_ms)) {
      best_ms = run_best;
      best = t;
    }
  }
  c.plane_threads = (e == hipSuccess) ? best : keep;
  if (e == hipSuccess) e = sc.restore();
  if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
  if (t0) (void)hipEventDestroy(t0);
  if (t1) (void)hipEventDestroy(t1);
  drop_graphs(c);
  if (e != hipSuccess) return fail(c, DH_EHIP, std::string("dh_autotune: ") + hipGetErrorString(e));
  if (chosen) *chosen = best;
  return DH_OK;
}

// nsteps iterations of diablo.f:66-110 without I/O: TIME_STEP += 1, the
// three substeps, TIME += DELTA_T; the end of one step and the start of the
// next share one launch.
static int enqueue_steps(Ctx &c, int nsteps) {
  HIPCHK(launch_step_boundary(c, false, true));
  for (int s = 0; s < nsteps; ++s) {
    // the boundary into the next step (or the end of the last) rides in the
    // last substep's REM_DIV launch
    for (int rk = 0; rk < 2; ++rk) HIPCHK(launch_substep(c, rk));
    HIPCHK(launch_substep(c, 2, 1, s + 1 < nsteps ? 1 : 0));
  }
  return DH_OK;
}

static int enqueue_step(Ctx &c) { return enqueue_steps(c, 1); }

// Captures nsteps steps into *exec; false (and a cleared error) if the
// capture or instantiation fails, in which case the steps run uncaptured.
static bool capture_steps(Ctx &c, int nsteps, hipGraphExec_t *exec) {
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal);
  int r = DH_OK;
  if (e == hipSuccess) {
    r = enqueue_steps(c, nsteps);
    e = hipStreamEndCapture(c.stream, &graph);
  }
  bool ok = false;
  if (e == hipSuccess && r == DH_OK && graph) {
    ok = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0) == hipSuccess;
  }
  if (graph) (void)hipGraphDestroy(graph);
  if (!ok) (void)hipGetLastError();
  return ok;
}

int dh_time_step(dh_ctx *h, int32_t nsteps) {
  if (!h || nsteps < 0) return DH_EARG;
  Ctx &c = h->c;
  if (!c.grid_set) return fail(c, DH_ESTATE, "dh_set_grid first");
  if (nsteps == 0) return DH_OK;
  if (nsteps >= kStepsPerGraph && !c.steps_graph_ok) c.steps_graph_ok = capture_steps(c, kStepsPerGraph, &c.steps_graph);
  if (!c.graph_ok) c.graph_ok = capture_steps(c, 1, &c.step_graph);
  int left = nsteps;
  while (c.steps_graph_ok && left >= kStepsPerGraph) {
    HIPCHK(hipGraphLaunch(c.steps_graph, c.stream));
    left -= kStepsPerGraph;
  }
  for (; left > 0; --left) {
    if (c.graph_ok) {
      HIPCHK(hipGraphLaunch(c.step_graph, c.stream));
    } else {
      const int r = enqueue_step(c);
      if (r != DH_OK) return r;
    }
  }
  c.th_phys32 = c.mixed;
  return DH_OK;
}

int dh_profile_substep(dh_ctx *h, int32_t rk_step, int32_t nrep, double *ms, int32_t nms) {
  if (!h || !ms || nrep < 1 || nms < kNumPhases) return DH_EARG;
  Ctx &c = h->c;
  if (rk_step < 1 || rk_step > 3) return fail(c, DH_EARG, "rk_step must be 1..3");
  if (!c.grid_set) return fail(c, DH_ESTATE, "dh_set_grid first");
  hipEvent_t ev[kNumPhases + 1];
  for (auto &e : ev) HIPCHK(hipEventCreate(&e));
  for (int i = 0; i < nms; ++i) ms[i] = 0.0;
  for (int r = 0; r < nrep; ++r) {
    HIPCHK(hipEventRecord(ev[0], c.stream));
    for (int ph = 0; ph < kNumPhases; ++ph) {
      HIPCHK(launch_phase(c, rk_step - 1, ph));
      HIPCHK(hipEventRecord(ev[ph + 1], c.stream));
    }
    HIPCHK(join_side(c));
    HIPCHK(hipEventSynchronize(ev[kNumPhases]));
    for (int ph = 0; ph < kNumPhases; ++ph) {
      float t = 0.f;
      HIPCHK(hipEventElapsedTime(&t, ev[ph], ev[ph + 1]));
      ms[ph] += t / nrep;
    }
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  return DH_OK;
}

const char *dh_phase_name(int32_t phase) {
  return (phase >= 0 && phase < kNumPhases) ? kPhaseNames[phase] : "";
}

namespace {

bool same_op(const RecOp &x, const RecOp &y) {
  return x.kind == y.kind && x.fnb == y.fnb && x.grid.x == y.grid.x && x.grid.y == y.grid.y && x.grid.z == y.grid.z &&
         x.block.x == y.block.x && x.block.y == y.block.y && x.block.z == y.block.z && x.lds == y.lds &&
         x.role == y.role && x.ev == y.ev && x.args.size() == y.args.size();
}
bool same_seq(const std::vector<RecOp> &x, const std::vector<RecOp> &y) {
  if (x.size() != y.size()) return false;
  for (size_t i = 0; i < x.size(); ++i)
    if (!same_op(x[i], y[i])) return false;
  return true;
}

// One step of context c as recorded launches: pro = the start of a first
// step, unit = three substeps and the boundary into the next step, epi =
// three substeps and the end of the last step.
struct StepRec {
  std::vector<RecOp> q[3];
};

hipError_t record_step(Ctx &c, StepRec &sr) {
  std::vector<RecOp> *seq = sr.q;
  c.recording = true;
  c.rec.clear();
  hipError_t e = launch_step_boundary(c, false, true);
  seq[0] = std::move(c.rec);
  c.rec.clear();
  for (int rk = 0; rk < 3 && e == hipSuccess; ++rk) e = launch_substep(c, rk);
  if (e == hipSuccess) e = launch_step_boundary(c, true, true);
  seq[1] = std::move(c.rec);
  c.rec.clear();
  for (int rk = 0; rk < 3 && e == hipSuccess; ++rk) e = launch_substep(c, rk);
  if (e == hipSuccess) e = launch_step_boundary(c, true, false);
  seq[2] = std::move(c.rec);
  c.rec.clear();
  c.recording = false;
  c.tables_pending = false;
  for (int k = 0; k < 3; ++k)
    for (const RecOp &o : seq[k])
      if (o.kind == 0 && o.grid.z != 1) return hipErrorInvalidValue;
  return e;
}

void free_groups(dh_ens &e) {
  for (int k = 0; k < 2; ++k)
    if (e.graph[k]) (void)hipGraphExecDestroy(e.graph[k]);
  e.graph[0] = e.graph[1] = nullptr;
  for (EnsGroup &g : e.g) {
    for (auto &v : g.dargs)
      for (void *p : v)
        if (p) (void)hipFree(p);
    for (hipStream_t s : g.st)
      if (s) (void)hipStreamDestroy(s);
    for (hipEvent_t v : g.ev)
      if (v) (void)hipEventDestroy(v);
    if (g.done) (void)hipEventDestroy(g.done);
  }
  e.g.clear();
  if (e.root) (void)hipStreamDestroy(e.root);
  if (e.ev_in) (void)hipEventDestroy(e.ev_in);
  if (e.ev_out) (void)hipEventDestroy(e.ev_out);
  e.root = nullptr;
  e.ev_in = e.ev_out = nullptr;
  e.batched = false;
}

// Groups the members by recorded step and uploads their argument structs.
hipError_t build_groups(dh_ens &e) {
  for (dh_ctx *h : e.m)
    if (h->c.les) return hipErrorNotSupported;  // the LES kernels have no batched forms
  std::vector<StepRec> recs(e.m.size());
  for (size_t i = 0; i < e.m.size(); ++i) {
    const hipError_t r = record_step(e.m[i]->c, recs[i]);
    if (r != hipSuccess) return r;
  }
  std::vector<int> gi(e.m.size(), -1);
  std::vector<size_t> first;  // member whose records define each group
  for (size_t i = 0; i < e.m.size(); ++i) {
    for (size_t k = 0; k < first.size() && gi[i] < 0; ++k) {
      const StepRec &r0 = recs[first[k]];
      if (same_seq(r0.q[0], recs[i].q[0]) && same_seq(r0.q[1], recs[i].q[1]) && same_seq(r0.q[2], recs[i].q[2]))
        gi[i] = (int)k;
    }
    if (gi[i] < 0) {
      gi[i] = (int)first.size();
      first.push_back(i);
    }
  }
  e.g.resize(first.size());
  hipError_t r;
#define GCHK(x) \
  if ((r = (x)) != hipSuccess) return r;
  GCHK(hipStreamCreateWithFlags(&e.root, hipStreamNonBlocking));
  GCHK(hipEventCreateWithFlags(&e.ev_in, hipEventDisableTiming));
  GCHK(hipEventCreateWithFlags(&e.ev_out, hipEventDisableTiming));
  for (size_t k = 0; k < first.size(); ++k) {
    EnsGroup &g = e.g[k];
    std::vector<size_t> idx;
    for (size_t i = 0; i < e.m.size(); ++i)
      if (gi[i] == (int)k) {
        g.m.push_back(e.m[i]);
        idx.push_back(i);
      }
    for (int q = 0; q < 3; ++q) {
      g.seq[q] = recs[first[k]].q[q];
      g.dargs[q].assign(g.seq[q].size(), nullptr);
      for (size_t o = 0; o < g.seq[q].size(); ++o) {
        if (g.seq[q][o].kind != 0) continue;
        const size_t sz = g.seq[q][o].args.size();
        std::vector<unsigned char> host(sz * idx.size());
        for (size_t j = 0; j < idx.size(); ++j) std::memcpy(&host[j * sz], recs[idx[j]].q[q][o].args.data(), sz);
        GCHK(hipMalloc(&g.dargs[q][o], host.size()));
        GCHK(hipMemcpy(g.dargs[q][o], host.data(), host.size(), hipMemcpyHostToDevice));
      }
    }
    for (hipStream_t &st : g.st) GCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    for (hipEvent_t &v : g.ev) GCHK(hipEventCreateWithFlags(&v, hipEventDisableTiming));
    GCHK(hipEventCreateWithFlags(&g.done, hipEventDisableTiming));
  }
#undef GCHK
  return hipSuccess;
}

hipError_t replay(EnsGroup &g, int q) {
  const unsigned n = (unsigned)g.m.size();
  for (size_t i = 0; i < g.seq[q].size(); ++i) {
    const RecOp &o = g.seq[q][i];
    hipStream_t st = g.st[o.role];
    hipError_t r;
    if (o.kind == 0) {
      void *args[] = {&g.dargs[q][i]};
      r = hipLaunchKernel(o.fnb, dim3(o.grid.x, o.grid.y, n), o.block, args, o.lds, st);
    } else if (o.kind == 1) {
      r = hipEventRecord(g.ev[o.ev], st);
    } else {
      r = hipStreamWaitEvent(st, g.ev[o.ev], 0);
    }
    if (r != hipSuccess) return r;
  }
  return hipSuccess;
}

// nsteps steps of every group, forked from and joined into the root stream.
hipError_t enqueue_ensemble(dh_ens &e, int nsteps) {
  hipError_t r;
  if ((r = hipEventRecord(e.ev_in, e.root)) != hipSuccess) return r;
  for (EnsGroup &g : e.g) {
    if ((r = hipStreamWaitEvent(g.st[0], e.ev_in, 0)) != hipSuccess) return r;
    if ((r = replay(g, 0)) != hipSuccess) return r;
    for (int s = 0; s + 1 < nsteps; ++s)
      if ((r = replay(g, 1)) != hipSuccess) return r;
    if ((r = replay(g, 2)) != hipSuccess) return r;
    if ((r = hipEventRecord(g.done, g.st[0])) != hipSuccess) return r;
    if ((r = hipStreamWaitEvent(e.root, g.done, 0)) != hipSuccess) return r;
  }
  return hipSuccess;
}

bool capture_ensemble(dh_ens &e, int nsteps, hipGraphExec_t *exec) {
  hipGraph_t graph = nullptr;
  hipError_t r = hipStreamBeginCapture(e.root, hipStreamCaptureModeThreadLocal);
  if (r == hipSuccess) {
    const hipError_t q = enqueue_ensemble(e, nsteps);
    r = hipStreamEndCapture(e.root, &graph);
    if (q != hipSuccess) r = q;
  }
  bool ok = r == hipSuccess && graph && hipGraphInstantiate(exec, graph, nullptr, nullptr, 0) == hipSuccess;
  if (graph) (void)hipGraphDestroy(graph);
  if (!ok) {
    (void)hipGetLastError();
    *exec = nullptr;  // callers take a non-null handle as a usable graph
  }
  return ok;
}

// Batched stepping (build_groups, capture_ensemble) is switched off: the
// first batched step of an ensemble of KH instances (NX 64 and 96) ended the
// process with a host segmentation fault on the MI355X, and no cause has been
// found in the host logic (run against a stand-in runtime under
// AddressSanitizer) or in the batched kernels' indexing.  Ensembles step
// their members one by one, each with its own captured step graph.
constexpr bool kBatchedEnsembles = false;

}  // namespace

int dh_ensemble_create(int32_t n, dh_ctx *const *members, dh_ens **out) {
  if (!out) return DH_EARG;
  *out = nullptr;
  if (n < 1 || !members) return DH_EARG;
  for (int i = 0; i < n; ++i) {
    if (!members[i]) return DH_EARG;
    for (int k = 0; k < i; ++k)
      if (members[k] == members[i]) return DH_EARG;  // a context appears once
  }
  dh_ens *e = new (std::nothrow) dh_ens();
  if (!e) return DH_EARG;
  e->m.assign(members, members + n);
  *out = e;
  return DH_OK;
}

// Members of one shape advance together: every launch of their step runs
// once for all of them (the kernels' batched variants, member blockIdx.z),
// groups of different shapes on streams of their own, the whole in one
// captured graph of 1 or kStepsPerGraph steps.  The ensemble's streams wait
// for work already queued on the members and the members' streams wait for
// the ensemble, so the members' own entry points stay ordered with it.  The
// groups are formed at the first call from the members' state then (grid,
// sizes, precision form); dh_ensemble_time_step falls back to stepping the
// members one by one if that is not possible.
int dh_ensemble_time_step(dh_ens *e, int32_t nsteps) {
  if (!e || nsteps < 0) return DH_EARG;
  for (dh_ctx *h : e->m)
    if (!h->c.grid_set) return fail(h->c, DH_ESTATE, "dh_set_grid first");
  if (nsteps == 0) return DH_OK;
  Ctx &c = e->m[0]->c;  // errors are reported on the first member
  if (!e->batched && !e->fallback) {
    if (kBatchedEnsembles && build_groups(*e) == hipSuccess) {
      e->batched = true;
    } else {
      (void)hipGetLastError();
      free_groups(*e);
      e->fallback = true;
    }
  }
  if (!e->batched) {
    for (int s = 0; s < nsteps; ++s)
      for (dh_ctx *h : e->m) {
        const int r = dh_time_step(h, 1);
        if (r != DH_OK) return r;
      }
    return DH_OK;
  }
  for (dh_ctx *h : e->m) {
    HIPCHK(hipEventRecord(h->c.ev_fork, h->c.stream));
    HIPCHK(hipStreamWaitEvent(e->root, h->c.ev_fork, 0));
  }
  int left = nsteps;
  if (left >= kStepsPerGraph && !e->graph[1]) (void)capture_ensemble(*e, kStepsPerGraph, &e->graph[1]);
  if (!e->graph[0]) (void)capture_ensemble(*e, 1, &e->graph[0]);
  while (e->graph[1] && left >= kStepsPerGraph) {
    HIPCHK(hipGraphLaunch(e->graph[1], e->root));
    left -= kStepsPerGraph;
  }
  for (; left > 0; --left) {
    if (e->graph[0])
      HIPCHK(hipGraphLaunch(e->graph[0], e->root));
    else
      HIPCHK(enqueue_ensemble(*e, 1));
  }
  HIPCHK(hipEventRecord(e->ev_out, e->root));
  for (dh_ctx *h : e->m) {
    HIPCHK(hipStreamWaitEvent(h->c.stream, e->ev_out, 0));
    h->c.th_phys32 = h->c.mixed;
  }
  return DH_OK;
}

// Tunes the block size of each shape (sizes, scalars, precision form) as
// the ensemble runs it: the members of the shape stepped as one batch (a
// temporary ensemble of them), nsteps steps (best of two) with each
// power-of-two size, the members' states restored before each run and at
// the end.  The batches are formed again at the next dh_ensemble_time_step.
int dh_ensemble_autotune(dh_ens *e, int32_t nsteps) {
  if (!e || nsteps < 1) return DH_EARG;
  Ctx &c0 = e->m[0]->c;
  {
    Ctx &c = c0;
    if (e->root) HIPCHK(hipStreamSynchronize(e->root));
    for (dh_ctx *h : e->m)
      if (!h->c.grid_set) return fail(h->c, DH_ESTATE, "dh_set_grid first");
  }
  free_groups(*e);
  e->fallback = false;
  std::vector<bool> done(e->m.size(), false);
  for (size_t i = 0; i < e->m.size(); ++i) {
    if (done[i]) continue;
    const Ctx &ci = e->m[i]->c;
    std::vector<dh_ctx *> shape;
    for (size_t j = i; j < e->m.size(); ++j) {
      const Ctx &cj = e->m[j]->c;
      if (!done[j] && cj.g.nx == ci.g.nx && cj.g.ny == ci.g.ny && cj.g.nz == ci.g.nz && cj.nfields == ci.nfields &&
          cj.mixed == ci.mixed && cj.p.variable_dt == ci.p.variable_dt) {
        shape.push_back(e->m[j]);
        done[j] = true;
      }
    }
    std::vector<StateCopy> saved(shape.size());
    hipError_t r = hipSuccess;
    for (size_t k = 0; k < shape.size() && r == hipSuccess; ++k) {
      r = hipStreamSynchronize(shape[k]->c.stream);
      if (r == hipSuccess) r = saved[k].init(shape[k]->c);
    }
    if (r != hipSuccess) return fail(c0, DH_EHIP, std::string("dh_ensemble_autotune: ") + hipGetErrorString(r));
    const int keep = shape[0]->c.plane_threads;
    int best = keep;
    double best_s = -1.0;
    for (int t = 64; t <= 1024; t *= 2) {
      if (!plane_threads_ok(shape[0]->c, t)) continue;
      int rc = DH_OK;
      for (dh_ctx *h : shape)
        if ((rc = dh_set_plane_threads(h, t)) != DH_OK) return rc;
      dh_ens tmp;
      tmp.m = shape;
      double run_best = -1.0;
      for (int rep = 0; rep < 3 && rc == DH_OK; ++rep) {
        for (StateCopy &sc : saved)
          if ((r = sc.restore()) != hipSuccess) break;
        for (dh_ctx *h : shape)
          if (r == hipSuccess) r = hipStreamSynchronize(h->c.stream);
        if (r != hipSuccess) break;
        const auto t0 = std::chrono::steady_clock::now();
        rc = dh_ensemble_time_step(&tmp, nsteps);
        if (rc == DH_OK) rc = dh_ensemble_synchronize(&tmp);
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (rep > 0 && (run_best < 0 || sec < run_best)) run_best = sec;  // the first run builds and captures
      }
      free_groups(tmp);
      if (r != hipSuccess) return fail(c0, DH_EHIP, std::string("dh_ensemble_autotune: ") + hipGetErrorString(r));
      if (rc != DH_OK) return rc;
      if (best_s < 0 || run_best < best_s) {
        best_s = run_best;
        best = t;
      }
    }
    for (dh_ctx *h : shape) {
      const int rc = dh_set_plane_threads(h, best);
      if (rc != DH_OK) return rc;
    }
    for (StateCopy &sc : saved)
      if ((r = sc.restore()) != hipSuccess) break;
    for (dh_ctx *h : shape)
      if (r == hipSuccess) r = hipStreamSynchronize(h->c.stream);
    if (r != hipSuccess) return fail(c0, DH_EHIP, std::string("dh_ensemble_autotune: ") + hipGetErrorString(r));
  }
  return DH_OK;
}

int dh_ensemble_synchronize(dh_ens *e) {
  if (!e) return DH_EARG;
  Ctx &c = e->m[0]->c;
  if (e->root) HIPCHK(hipStreamSynchronize(e->root));
  for (dh_ctx *h : e->m) {
    const int r = dh_synchronize(h);
    if (r != DH_OK) return r;
  }
  return DH_OK;
}

int dh_ensemble_destroy(dh_ens *e) {
  if (!e) return DH_EARG;
  (void)dh_ensemble_synchronize(e);
  free_groups(*e);
  delete e;
  return DH_OK;
}

int dh_synchronize(dh_ctx *h) {
  if (!h) return DH_EARG;
  Ctx &c = h->c;
  HIPCHK(hipStreamSynchronize(c.stream));
  return DH_OK;
}

int dh_create_flow_chan(dh_ctx *h, int32_t ic_type, double ubulk0, uint64_t seed, double amp, double ymask) {
  if (!h) return DH_EARG;
  Ctx &c = h->c;
  if (!c.grid_set) return fail(c, DH_ESTATE, "dh_set_grid first");
  if (ic_type < 0 || ic_type > 3) return fail(c, DH_EARG, "IC_TYPE must be 0..3");
  HIPCHK(launch_create_flow(c, ic_type, ubulk0, seed, amp, ymask));
  HIPCHK(hipStreamSynchronize(c.stream));
  return DH_OK;
}

int dh_create_th_chan(dh_ctx *h, int32_t ic_type) {
  if (!h) return DH_EARG;
  Ctx &c = h->c;
  if (!c.grid_set) return fail(c, DH_ESTATE, "dh_set_grid first");
  HIPCHK(launch_create_th(c, ic_type));
  HIPCHK(hipStreamSynchronize(c.stream));
  return DH_OK;
}

int dh_apply_bc_vel(dh_ctx *h) {
  if (!h) return DH_EARG;
  Ctx &c = h->c;
  if (!c.grid_set) return fail(c, DH_ESTATE, "dh_set_grid first");
  HIPCHK(launch_apply_bc_vel(c));
  return DH_OK;
}

int dh_rem_div(dh_ctx *h) {
  if (!h) return DH_EARG;
  Ctx &c = h->c;
  if (!c.grid_set) return fail(c, DH_ESTATE, "dh_set_grid first");
  HIPCHK(launch_rem_div(c, -1));
  return DH_OK;
}

int dh_poisson_p_chan(dh_ctx *h) {
  if (!h) return DH_EARG;
  Ctx &c = h->c;
  if (!c.grid_set) return fail(c, DH_ESTATE, "dh_set_grid first");
  HIPCHK(launch_poisson_p(c));
  return DH_OK;
}

int dh_save_stats(dh_ctx *h, double ly, double *out, double *ubulk) {
  if (!h || !out || !ubulk || !(ly > 0.0)) return DH_EARG;
  Ctx &c = h->c;
  if (!c.grid_set) return fail(c, DH_ESTATE, "dh_set_grid first");
  const int ny2 = c.g.ny + 2;
  HIPCHK(launch_stats(c, c.stats));
  HIPCHK(hipMemcpyAsync(out, c.stats, sizeof(double) * DH_NSTATS * ny2, hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  // INTEGRATE_Y_VAR (mpi.f:1610-1632), in its order
  const double *ume = out + DH_ST_UME * ny2;
  double res = 0.;
  for (int j = 2; j <= c.g.ny; ++j) res = res + 0.5 * (ume[j] + ume[j - 1]) * c.dy[j];
  *ubulk = res / ly;
  return DH_OK;
}

int dh_fft_xz_to_fourier_dev(dh_ctx *h, const double *phys, double *four, int32_t nplanes) {
  if (!h || !phys || !four || nplanes < 0) return DH_EARG;
  Ctx &c = h->c;
  if (nplanes == 0) return DH_OK;
  HIPCHK(launch_fft_to_four(c, phys, reinterpret_cast<double2 *>(four), nplanes));
  return DH_OK;
}

int dh_fft_xz_to_physical_dev(dh_ctx *h, const double *four, double *phys, int32_t nplanes) {
  if (!h || !phys || !four || nplanes < 0) return DH_EARG;
  Ctx &c = h->c;
  if (nplanes == 0) return DH_OK;
  HIPCHK(launch_fft_to_phys(c, reinterpret_cast<const double2 *>(four), phys, nplanes));
  return DH_OK;
}

double *dh_field_dev(dh_ctx *h, int32_t which) {
  if (!h) return nullptr;
  Ctx &c = h->c;
  const int nth = c.p.n_th;
  const long long cfs = (long long)(c.g.ny + 2) * c.g.pf;
  const long long fs = (long long)(c.g.ny + 2) * c.g.pp;
  if (which >= 0 && which < 3) return reinterpret_cast<double *>(c.cu[which]);
  if (which == 3) return reinterpret_cast<double *>(c.cp);
  if (which >= 4 && which < 7) return reinterpret_cast<double *>(c.cf[which - 4]);
  if (which == 7) return reinterpret_cast<double *>(c.phi);
  if (which >= 8 && which < 8 + nth) return reinterpret_cast<double *>(c.cth + (which - 8) * cfs);
  if (which >= 8 + nth && which < 8 + 2 * nth) return reinterpret_cast<double *>(c.cfth + (which - 8 - nth) * cfs);
  if (which >= 100 && which < 100 + c.nfields) return c.uphys + (which - 100) * fs;
  if (which >= 200 && which < 200 + c.nfields && c.uphys32)
    return reinterpret_cast<double *>(c.uphys32 + (which - 200) * fs);
  if (which == 300) return c.nut;
  if (which == 301) return c.les_feat;
  return nullptr;
}

void *dh_stream(dh_ctx *h) { return h ? (void *)h->c.stream : nullptr; }

}  // extern "C"
