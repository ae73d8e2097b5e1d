// Training and inference kernels of the learned eddy-viscosity closure's
// network (dh_mlp_dev.h): operand packing, batched inference, the forward
// and backward pass of a batch, the weight gradients as BF16 MFMA GEMMs
// over the batch, the reduction of per-block partial gradients, and Adam.
// No counterpart in the reference (its LES has no learned model).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "dh_internal.h"
#include "dh_mlp_dev.h"

namespace dh {

// Packs the FP32 parameters into the BF16 operand order of dh_mlp_dev.h.
// One thread per (array, lane) writes the lane's 8 values.
__global__ void k_mlp_pack(const float *__restrict__ P, bf16x8 *__restrict__ pk) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= kPAll) return;
  const int lane = t & 63, g = lane >> 4, c = lane & 15;
  if (t >= kPBias) {  // biases and w4 in the result-tile layout
    const int r = (t - kPBias) >> 6, layer = r >> 2, mt = r & 3;  // layer 3: w4
    const int off = (layer == 0) ? kOB1 : (layer == 1) ? kOB2 : (layer == 2) ? kOB3 : kOW4;
    f32x4 b;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float v = P[off + 16 * mt + 4 * g + i];
      b[i] = (layer == 3) ? (float)(__bf16)v : v;
    }
    reinterpret_cast<f32x4 *>(pk)[t] = b;
    return;
  }
  bf16x8 v;
  if (t < kPA2) {  // A1[mt][lane]: W1[8 g + j][16 mt + c]
    const int mt = t >> 6;
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = (__bf16)P[kOW1 + (8 * g + j) * kMlpH + 16 * mt + c];
  } else if (t < kPA4) {  // A2, A3 [mt][kb][lane]: W[u(kb,g,j)][16 mt + c]
    const int base = (t < kPA3) ? kOW2 : kOW3, r = (t - ((t < kPA3) ? kPA2 : kPA3)) >> 6;
    const int mt = r >> 1, kb = r & 1;
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = (__bf16)P[base + mlp_unit(kb, g, j) * kMlpH + 16 * mt + c];
  } else if (t < kPB2) {  // A4[kb][lane]: row 0 is w4
    const int kb = (t - kPA4) >> 6;
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = (__bf16)(c == 0 ? P[kOW4 + mlp_unit(kb, g, j)] : 0.0f);
  } else {  // B2, B3 [mt][kb][lane]: W[16 mt + c][u(kb,g,j)] (untransposed)
    const int base = (t < kPB3) ? kOW2 : kOW3, r = (t - ((t < kPB3) ? kPB2 : kPB3)) >> 6;
    const int mt = r >> 1, kb = r & 1;
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = (__bf16)P[base + (16 * mt + c) * kMlpH + mlp_unit(kb, g, j)];
  }
  pk[t] = v;
}

// layer-1 operand of this lane: features 8 g + j (< nin) of sample n, from a
// feature-major array x[f * ld + n] (zero for n >= N)
// (only lane group 0 holds inputs; ld * 8 < 2^31, checked by the host)
__device__ __forceinline__ bf16x8 load_x(const float *__restrict__ x, long long ld, long long n, long long N, int nin) {
  const int g = (threadIdx.x & 63) >> 4;
  bf16x8 v;
#pragma unroll
  for (int j = 0; j < 8; ++j) v[j] = (__bf16)0.0f;
  if (g == 0 && n < N) {
    const float *xn = x + n;
    const int ldi = (int)ld;
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (j < nin) v[j] = (__bf16)xn[j * ldi];
  }
  return v;
}

// out[n] = scale * softplus(y(x_n)), x feature-major [6][ld]
__global__ void __launch_bounds__(256) k_mlp_infer(const bf16x8 *__restrict__ pk, const float *__restrict__ P,
                                                 const float *__restrict__ x, long long ld, long long N,
                                                 float scale, float *__restrict__ out) {
  const int lane = threadIdx.x & 63;
  const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long long nw = ((long long)gridDim.x * blockDim.x) >> 6;
  MlpRegs r;
  mlp_load_regs(pk, r);
  const float b4 = P[kOB4];
  for (long long n0 = wave * 16; n0 < N; n0 += nw * 16) {
    const long long n = n0 + (lane & 15);
    const bf16x8 xb = load_x(x, ld, n, N, kMlpIn);
    const float y = mlp_forward_regs(r, b4, xb);
    if (lane < 16 && n < N) out[n] = scale * softplus(y);
  }
}

// Forward and backward pass of a batch.  Loss: sum_n (softplus(y_n) -
// t_n)^2 / N.  Writes, as BF16 rows of N samples, the layer inputs H0
// (features, 8 rows), H1, H2, H3, the output gradient dL/dy (1 row) and the
// pre-activation gradients dZ1, dZ2, dZ3, for the weight-gradient GEMMs,
// and per wave the partial sum of the loss.
__global__ void __launch_bounds__(64) k_mlp_fwd_bwd(const bf16x8 *pk, const float *P, const float *__restrict__ x,
                                                  const float *__restrict__ t, long long N, float inv_n,
                                                  __bf16 *__restrict__ act, float *__restrict__ part) {
  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  // rows: H0 [8][N], H1..H3 [64][N], DY [8][N] (row 0 used), dZ1..dZ3 [64][N]
  __bf16 *H0 = act, *H1 = H0 + 8 * N, *H2 = H1 + 64 * N, *H3 = H2 + 64 * N;
  __bf16 *DY = H3 + 64 * N, *D1 = DY + 8 * N, *D2 = D1 + 64 * N, *D3 = D2 + 64 * N;
  float sloss = 0.f;
  const long long nt = (N + 15) / 16;
  for (long long tile = blockIdx.x; tile < nt; tile += gridDim.x) {
    asm volatile("" ::: "memory");  // operands re-read each pass, not held across it
    const long long n = tile * 16 + c;
    const bool live = n < N;
    const bf16x8 xb = load_x(x, N, n, N, kMlpIn);
    f32x4 z1[4], z2[4], z3[4];
    const float y = mlp_forward(pk, P[kOB4], xb, z1, z2, z3);
    float dy = 0.f;
    if (lane < 16 && live) {
      const float r = softplus(y) - t[n];
      sloss += r * r;
      dy = 2.0f * r * sigmoid(y) * inv_n;
      DY[n] = (__bf16)dy;
    }
    const float dys = __shfl(dy, c);  // this lane's sample
    // 32-bit row offsets (N < 2^24 is checked by the host): every array's
    // rows share them
    const int Ni = (int)N, ni = (int)n;
    if (g == 0 && live) {
#pragma unroll
      for (int j = 0; j < 8; ++j) H0[j * Ni + ni] = xb[j];
    }
    // layer inputs (BF16, as the forward pass used them) and dZ3 = dy w4 GELU'(Z3)
    // (w4 rounded as the forward pass used it)
    const f32x4 *w4p = reinterpret_cast<const f32x4 *>(pk) + kPW4;
    f32x4 d3[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      const f32x4 w4 = w4p[mt * 64 + lane];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int u = 16 * mt + 4 * g + i;
        if (live) {
          H1[u * Ni + ni] = (__bf16)gelu(z1[mt][i]);
          H2[u * Ni + ni] = (__bf16)gelu(z2[mt][i]);
          H3[u * Ni + ni] = (__bf16)gelu(z3[mt][i]);
        }
        d3[mt][i] = dys * w4[i] * gelu_grad(z3[mt][i]);
      }
    }
    if (live) {
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int i = 0; i < 4; ++i) D3[(16 * mt + 4 * g + i) * Ni + ni] = (__bf16)d3[mt][i];
    }
    // dH2 = W3 dZ3^T, dZ2 = dH2 GELU'(Z2); dH1 = W2 dZ2^T, dZ1 = dH1 GELU'(Z1)
    f32x4 d2[4], d1[4];
    {
      const bf16x8 b[2] = {pack_b(d3, 0), pack_b(d3, 1)};
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) {
        f32x4 d = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < 2; ++kb)
          d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(pk[kPB3 + (mt * 2 + kb) * 64 + lane], b[kb], d, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          d2[mt][i] = d[i] * gelu_grad(z2[mt][i]);
          if (live) D2[(16 * mt + 4 * g + i) * Ni + ni] = (__bf16)d2[mt][i];
        }
      }
    }
    {
      const bf16x8 b[2] = {pack_b(d2, 0), pack_b(d2, 1)};
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) {
        f32x4 d = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < 2; ++kb)
          d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(pk[kPB2 + (mt * 2 + kb) * 64 + lane], b[kb], d, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          d1[mt][i] = d[i] * gelu_grad(z1[mt][i]);
          if (live) D1[(16 * mt + 4 * g + i) * Ni + ni] = (__bf16)d1[mt][i];
        }
      }
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) sloss += __shfl_xor(sloss, o);
  if (lane == 0) part[blockIdx.x] = sloss * inv_n;  // the loss is the mean over the batch
}

// Parameter gradients of layer L = blockIdx.y over the samples of slice
// blockIdx.x, as 16x16x32 BF16 MFMAs with the samples as K (lane l reads 8
// consecutive samples of one row for A and for B):
//   L = 0..2: dW[in][out] = sum_n Hin[in][n] dZ[out][n] (A rows in =
//             16 mt + l % 16, B rows out = 16 nt + l % 16; layer 0 has one
//             row tile, inputs 0..15, of which 0..7 are stored), and the
//             bias gradient sum_n dZ[out][n] from an A operand of ones;
//   L = 3:    dW4[u] = sum_n H3[u][n] dy[n] and db4 = sum_n dy[n] (B = the
//             dy row in column 0).
// Partial sums to wpart[slice][L][65][64]: rows 0..63 the weights (dW4 in
// column 0), row 64 the bias.
constexpr int kWRows = 65;

__global__ void __launch_bounds__(64) k_mlp_wgrad(const __bf16 *__restrict__ act, long long N, long long per_slice,
                                                float *__restrict__ wpart) {
  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  const int layer = blockIdx.y;
  const __bf16 *H0 = act, *H1 = H0 + 8 * N, *H2 = H1 + 64 * N, *H3 = H2 + 64 * N;
  const __bf16 *DY = H3 + 64 * N, *D1 = DY + 8 * N, *D2 = D1 + 64 * N, *D3 = D2 + 64 * N;
  const __bf16 *Hin = (layer == 0) ? H0 : (layer == 1) ? H1 : (layer == 2) ? H2 : H3;
  const __bf16 *Dz = (layer == 0) ? D1 : (layer == 1) ? D2 : (layer == 2) ? D3 : DY;
  const int mts = (layer == 0) ? 1 : 4, nts = (layer == 3) ? 1 : 4;
  f32x4 acc[5][4];  // row tiles 0..3 of the weights, 4: the bias (A of ones)
#pragma unroll
  for (int mt = 0; mt < 5; ++mt)
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
  bf16x8 ones;
#pragma unroll
  for (int j = 0; j < 8; ++j) ones[j] = (__bf16)1.0f;
  const long long n_lo = (long long)blockIdx.x * per_slice, n_hi = min(N, n_lo + per_slice);
  for (long long n0 = n_lo; n0 < n_hi; n0 += 32) {
    const long long n = n0 + 8 * g;  // N is a multiple of 32: whole 8-sample groups
    bf16x8 b[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      if (nt < nts) {
        if (layer == 3) {  // column 0: dy; the others zero
          if (c == 0) {
            b[nt] = *reinterpret_cast<const bf16x8 *>(Dz + n);
          } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) b[nt][j] = (__bf16)0.0f;
          }
        } else {
          b[nt] = *reinterpret_cast<const bf16x8 *>(Dz + (16 * nt + c) * N + n);
        }
      }
    }
#pragma unroll
    for (int mt = 0; mt < 5; ++mt) {
      if (mt < 4 && mt >= mts) continue;
      bf16x8 a;
      if (mt == 4) {
        a = ones;
      } else if (layer == 0) {
        if (c < 8) {
          a = *reinterpret_cast<const bf16x8 *>(Hin + c * N + n);
        } else {
#pragma unroll
          for (int j = 0; j < 8; ++j) a[j] = (__bf16)0.0f;
        }
      } else {
        a = *reinterpret_cast<const bf16x8 *>(Hin + (16 * mt + c) * N + n);
      }
#pragma unroll
      for (int nt = 0; nt < 4; ++nt)
        if (nt < nts) acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[nt], acc[mt][nt], 0, 0, 0);
    }
  }
  float *o = wpart + ((long long)blockIdx.x * 4 + layer) * (kWRows * kMlpH);
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    if (mt >= mts) continue;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
      if (nt < nts)
#pragma unroll
        for (int i = 0; i < 4; ++i) o[(16 * mt + 4 * g + i) * kMlpH + 16 * nt + c] = acc[mt][nt][i];
  }
  // the bias row: every row of the ones tile holds the column sums; take row 0
  if (g == 0) {
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
      if (nt < nts) o[64 * kMlpH + 16 * nt + c] = acc[4][nt][0];
  }
}

// grad[p] = the sum over slices of the partial sums, in the parameter
// vector's layout (W1 rows 8..31 are zero: their inputs are zero); loss =
// the sum over waves of their parts of the batch mean.
__global__ void k_mlp_reduce(const float *__restrict__ wpart, int nslice, const float *__restrict__ part, int nwave,
                             float *__restrict__ grad, float *__restrict__ loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > kMlpNP) return;
  if (p == kMlpNP) {
    float a = 0.f;
    for (int k = 0; k < nwave; ++k) a += part[k];
    *loss = a;
    return;
  }
  int layer, row, col;
  if (p < kOB1) { layer = 0; row = (p - kOW1) / kMlpH; col = (p - kOW1) % kMlpH; }
  else if (p < kOW2) { layer = 0; row = 64; col = p - kOB1; }
  else if (p < kOB2) { layer = 1; row = (p - kOW2) / kMlpH; col = (p - kOW2) % kMlpH; }
  else if (p < kOW3) { layer = 1; row = 64; col = p - kOB2; }
  else if (p < kOB3) { layer = 2; row = (p - kOW3) / kMlpH; col = (p - kOW3) % kMlpH; }
  else if (p < kOW4) { layer = 2; row = 64; col = p - kOB3; }
  else if (p < kOB4) { layer = 3; row = p - kOW4; col = 0; }
  else { layer = 3; row = 64; col = 0; }
  float a = 0.f;
  if (!(layer == 0 && row >= 16 && row < 64))
    for (int k = 0; k < nslice; ++k) a += wpart[((long long)k * 4 + layer) * (kWRows * kMlpH) + row * kMlpH + col];
  grad[p] = a;
}

// Adam (Kingma & Ba 2015) on the FP32 parameters; c1 = 1 - beta1^t,
// c2 = 1 - beta2^t come from the host
__global__ void k_mlp_adam(float *__restrict__ P, float *__restrict__ m, float *__restrict__ v,
                           const float *__restrict__ grad, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kMlpNP) return;
  const float gr = grad[p];
  const float mm = b1 * m[p] + (1.f - b1) * gr, vv = b2 * v[p] + (1.f - b2) * gr * gr;
  m[p] = mm;
  v[p] = vv;
  P[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

hipError_t launch_mlp_pack(const float *params, void *packed, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_pack, dim3((kPAll + 255) / 256), dim3(256), 0, st, params, static_cast<bf16x8 *>(packed));
  return hipGetLastError();
}

}  // namespace dh

using namespace dh;

namespace {
thread_local std::string g_mlp_err;
int mlp_fail(int code, const std::string &msg) {
  g_mlp_err = msg;
  return code;
}
constexpr int kWaves = 512;            // blocks of k_mlp_fwd_bwd (one wave each)
constexpr long long kSliceMin = 1024;  // samples per weight-gradient slice at least
long long n_slices(long long n) { return std::max<long long>(1, std::min<long long>(256, n / kSliceMin)); }
long long per_slice(long long n) {
  const long long s = n_slices(n);
  return ((n + s - 1) / s + 31) / 32 * 32;
}
}  // namespace

extern "C" {

const char *dh_mlp_last_error(void) { return g_mlp_err.c_str(); }

int32_t dh_mlp_nparams(void) { return kMlpNP; }

int64_t dh_mlp_packed_bytes(void) { return 16LL * kPAll; }

int64_t dh_mlp_work_bytes(int64_t n) {
  if (n <= 0) return 0;
  const long long act = 2LL * n * (16 + 6 * 64);
  const long long wp = 4LL * n_slices(n) * 4 * kWRows * kMlpH;
  const long long pp = 4LL * kWaves;
  return act + wp + pp;
}

int dh_mlp_pack_dev(const float *params, void *packed, void *stream) {
  if (!params || !packed) return mlp_fail(DH_EARG, "null pointer");
  hipLaunchKernelGGL(k_mlp_pack, dim3((kPAll + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream), params,
                     static_cast<bf16x8 *>(packed));
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? DH_OK : mlp_fail(DH_EHIP, hipGetErrorString(e));
}

int dh_mlp_infer_dev(const void *packed, const float *params, const float *x, int64_t ld, int64_t n, float scale,
                     float *out, void *stream) {
  if (!packed || !params || !x || !out) return mlp_fail(DH_EARG, "null pointer");
  if (n <= 0 || ld < n) return mlp_fail(DH_EARG, "need 0 < n <= ld");
  if (ld > (1LL << 27)) return mlp_fail(DH_EARG, "ld must be below 2^27");
  const long long waves = std::min<long long>((n + 15) / 16, 4096);
  hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, static_cast<hipStream_t>(stream),
                     static_cast<const bf16x8 *>(packed), params, x, (long long)ld, (long long)n, scale, out);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? DH_OK : mlp_fail(DH_EHIP, hipGetErrorString(e));
}

int dh_mlp_grad_dev(const float *params, const void *packed, const float *x, const float *target, int64_t n,
                    void *work, int64_t work_bytes, float *grad, float *loss, void *stream) {
  if (!params || !packed || !x || !target || !work || !grad || !loss) return mlp_fail(DH_EARG, "null pointer");
  if (n <= 0 || n % 32 != 0) return mlp_fail(DH_EARG, "the batch must be a positive multiple of 32 samples");
  if (n > (1LL << 24)) return mlp_fail(DH_EARG, "the batch must hold at most 2^24 samples");
  if (work_bytes < dh_mlp_work_bytes(n)) return mlp_fail(DH_EARG, "work buffer smaller than dh_mlp_work_bytes(n)");
  hipStream_t st = static_cast<hipStream_t>(stream);
  char *w = static_cast<char *>(work);
  __bf16 *act = reinterpret_cast<__bf16 *>(w);
  float *wp = reinterpret_cast<float *>(w + 2LL * n * (16 + 6 * 64));
  const long long ns = n_slices(n);
  float *part = wp + ns * 4 * kWRows * kMlpH;
  hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(kWaves), dim3(64), 0, st, static_cast<const bf16x8 *>(packed), params, x,
                     target, (long long)n, 1.0f / (float)n, act, part);
  hipLaunchKernelGGL(k_mlp_wgrad, dim3((unsigned)ns, 4), dim3(64), 0, st, act, (long long)n, per_slice(n), wp);
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kMlpNP + 1 + 255) / 256), dim3(256), 0, st, wp, (int)ns, part, kWaves, grad,
                     loss);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? DH_OK : mlp_fail(DH_EHIP, hipGetErrorString(e));
}

int dh_mlp_adam_dev(float *params, float *m, float *v, const float *grad, float lr, int32_t step, void *packed,
                    void *stream) {
  if (!params || !m || !v || !grad || !packed) return mlp_fail(DH_EARG, "null pointer");
  if (step < 1) return mlp_fail(DH_EARG, "Adam steps count from 1");
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpNP + 255) / 256), dim3(256), 0, st, params, m, v, grad, lr, b1, b2, 1e-8f,
                     c1, c2);
  hipLaunchKernelGGL(k_mlp_pack, dim3((kPAll + 255) / 256), dim3(256), 0, st, params, static_cast<bf16x8 *>(packed));
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? DH_OK : mlp_fail(DH_EHIP, hipGetErrorString(e));
}

}  // extern "C"
