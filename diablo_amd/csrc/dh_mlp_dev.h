// Device side of the learned eddy-viscosity closure's network (no
// counterpart in the reference; it stands in for the constant-Smagorinsky
// NU_T = C_s^2 Delta^2 |S| of les.f:196-227):
//   6 inputs -> 64 -> 64 -> 64 -> 1, GELU hidden layers, softplus output,
// BF16 operands and FP32 accumulation on gfx950's 16x16x32 BF16 MFMA.
//
// Every layer runs transposed, Z^T = W^T H^T, with the samples as the MFMA's
// N dimension: a wave takes 16 samples at a time.  The result tile of an
// MFMA (lane l holds rows 4(l/16) + i, i < 4, of column l % 16) is then the B
// operand of the next layer without any data movement, provided the K index
// of that layer's operands is permuted to match: K block kb, lane group g =
// l/16 and element j < 8 stand for the unit
//   u(kb, g, j) = 16 (2 kb + j/4) + 4 g + j % 4,
// i.e. elements 0..3 come from result tile 2 kb and 4..7 from tile 2 kb + 1.
// The weights are packed once per update into that order (k_mlp_pack), so a
// lane loads each A operand with one 16-byte read.  The first layer's K is
// the 6 inputs padded to 32 (features 8 g + j in lane group g; only g = 0
// holds inputs); the output layer is a 16 x 64 A operand whose row 0 is w4.
//
// The backward pass uses the same orientation: dH_prev^T = W dZ^T with the
// packed untransposed weights, so the gradient tiles chain the same way.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace dh {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kMlpIn = 6, kMlpInPad = 32, kMlpH = 64;
// FP32 parameter vector, natural layout W[in][out]
constexpr int kOW1 = 0;
constexpr int kOB1 = kOW1 + kMlpInPad * kMlpH;
constexpr int kOW2 = kOB1 + kMlpH;
constexpr int kOB2 = kOW2 + kMlpH * kMlpH;
constexpr int kOW3 = kOB2 + kMlpH;
constexpr int kOB3 = kOW3 + kMlpH * kMlpH;
constexpr int kOW4 = kOB3 + kMlpH;
constexpr int kOB4 = kOW4 + kMlpH;
constexpr int kMlpNP = kOB4 + 1;  // 10497
// packed operands, in 16-byte units (one per lane): forward A1[4][64],
// A2[4][2][64], A3[4][2][64], A4[2][64]; backward B2[4][2][64] (W2), B3 (W3)
// as bf16x8; then as f32x4 in the result-tile layout (units 16 mt + 4 g + i)
// the biases b1, b2, b3 [layer][mt][64] and w4 rounded to BF16 [mt][64]
constexpr int kPA1 = 0, kPA2 = kPA1 + 4 * 64, kPA3 = kPA2 + 8 * 64, kPA4 = kPA3 + 8 * 64;
constexpr int kPB2 = kPA4 + 2 * 64, kPB3 = kPB2 + 8 * 64, kPTot = kPB3 + 8 * 64;
constexpr int kPBias = kPTot, kPW4 = kPBias + 12 * 64, kPAll = kPW4 + 4 * 64;

__host__ __device__ constexpr int mlp_unit(int kb, int g, int j) { return 16 * (2 * kb + j / 4) + 4 * g + (j & 3); }

// GELU in its tanh form (Hendrycks & Gimpel 2016): 0.5 z (1 + tanh(a)),
// a = sqrt(2/pi) (z + 0.044715 z^3).  Since 0.5 (1 + tanh(a)) is the
// logistic function of 2a, GELU = z s with s = 1 / (1 + 2^(-2a log2 e)):
// one exp2, one add and one reciprocal (v_exp_f32 and v_rcp_f32, each within
// an FP32 ulp or so; the values go on to BF16), about 7 instructions where a
// tanh through an IEEE division takes about 22.  Far below zero the exp2
// overflows to +inf and s to 0; far above it s is 1.
__device__ __forceinline__ float gelu_s(float z) {
  // -2 log2(e) sqrt(2/pi) and the same times 0.044715
  const float c0 = -2.302208198144325f, c1 = -0.10294324f;
  const float e = __builtin_amdgcn_exp2f(z * fmaf(c1, z * z, c0));
  return __builtin_amdgcn_rcpf(e + 1.0f);
}
__device__ __forceinline__ float gelu(float z) { return z * gelu_s(z); }
// d GELU / dz = s + z s (1 - s) 2 a'(z), a'(z) = sqrt(2/pi) (1 + 3 0.044715 z^2)
__device__ __forceinline__ float gelu_grad(float z) {
  const float s = gelu_s(z);
  return fmaf(z * s * (1.0f - s), 1.5957691216057308f * fmaf(0.134145f, z * z, 1.0f), s);
}
__device__ __forceinline__ float softplus(float y) { return fmaxf(y, 0.0f) + log1pf(__expf(-fabsf(y))); }
__device__ __forceinline__ float sigmoid(float y) { return 1.0f / (1.0f + __expf(-y)); }

// B operand of K block kb from four result tiles (elements 0..3 of tile
// 2 kb, 4..7 of tile 2 kb + 1), rounded to BF16
__device__ __forceinline__ bf16x8 pack_b(const f32x4 (&h)[4], int kb) {
  bf16x8 b;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    b[i] = (__bf16)h[2 * kb][i];
    b[4 + i] = (__bf16)h[2 * kb + 1][i];
  }
  return b;
}

// One hidden layer: Z = W^T Hin + bias for the 4 output tiles (KB K
// blocks); bias: the layer's packed biases (f32x4 per lane and tile).
template <int KB>
__device__ __forceinline__ void mlp_layer(const bf16x8 *A, const bf16x8 (&B)[KB], const f32x4 *bias, f32x4 (&z)[4]) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    f32x4 d = bias[mt * 64 + lane];
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[(mt * KB + kb) * 64 + lane], B[kb], d, 0, 0, 0);
    z[mt] = d;
  }
}

// Forward pass of one 16-sample tile.  x: this lane's B operand of layer 1
// (features 8 g + j of sample l % 16).  z1..z3: the pre-activations
// (result-tile layout); returns the output y of sample l % 16 in lanes 0..15
// (other lanes hold the padding rows' results).
// b4: the output bias.
__device__ __forceinline__ float mlp_forward(const bf16x8 *pk, float b4, const bf16x8 &x, f32x4 (&z1)[4],
                                             f32x4 (&z2)[4], f32x4 (&z3)[4]) {
  const int lane = threadIdx.x & 63;
  const f32x4 *pb = reinterpret_cast<const f32x4 *>(pk) + kPBias;
  {
    const bf16x8 b[1] = {x};
    mlp_layer<1>(pk + kPA1, b, pb, z1);
  }
  f32x4 h[4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int i = 0; i < 4; ++i) h[mt][i] = gelu(z1[mt][i]);
  {
    const bf16x8 b[2] = {pack_b(h, 0), pack_b(h, 1)};
    mlp_layer<2>(pk + kPA2, b, pb + 4 * 64, z2);
  }
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int i = 0; i < 4; ++i) h[mt][i] = gelu(z2[mt][i]);
  {
    const bf16x8 b[2] = {pack_b(h, 0), pack_b(h, 1)};
    mlp_layer<2>(pk + kPA3, b, pb + 8 * 64, z3);
  }
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int i = 0; i < 4; ++i) h[mt][i] = gelu(z3[mt][i]);
  f32x4 d = {0.f, 0.f, 0.f, 0.f};
  d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(pk[kPA4 + lane], pack_b(h, 0), d, 0, 0, 0);
  d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(pk[kPA4 + 64 + lane], pack_b(h, 1), d, 0, 0, 0);
  return d[0] + b4;
}

// Inference holds this lane's forward operands in registers for a wave's
// whole loop (22 A operands and 12 bias tiles, 136 VGPRs): re-read in every
// 16-sample pass they are 35 KB per wave and pass, more than the vector L1
// holds, and the pass becomes bound by L2 reads.
struct MlpRegs {
  bf16x8 a[22];  // A1[4], A2[4][2], A3[4][2], A4[2]: packed units kPA1 + 64 i
  f32x4 b[12];   // b1, b2, b3 [mt]
};
__device__ __forceinline__ void mlp_load_regs(const bf16x8 *pk, MlpRegs &r) {
  const int lane = threadIdx.x & 63;
  const f32x4 *pb = reinterpret_cast<const f32x4 *>(pk) + kPBias;
#pragma unroll
  for (int i = 0; i < 22; ++i) r.a[i] = pk[kPA1 + i * 64 + lane];
#pragma unroll
  for (int i = 0; i < 12; ++i) r.b[i] = pb[i * 64 + lane];
}

// mlp_forward's output y (in lanes 0..15) from the operands in registers.
__device__ __forceinline__ float mlp_forward_regs(const MlpRegs &r, float b4, const bf16x8 &x) {
  static_assert(kPA2 - kPA1 == 4 * 64 && kPA3 - kPA1 == 12 * 64 && kPA4 - kPA1 == 20 * 64, "A operand order");
  f32x4 z[4], h[4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) z[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(r.a[mt], x, r.b[mt], 0, 0, 0);
#pragma unroll
  for (int layer = 1; layer < 3; ++layer) {
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int i = 0; i < 4; ++i) h[mt][i] = gelu(z[mt][i]);
    const bf16x8 b0 = pack_b(h, 0), b1 = pack_b(h, 1);
    const int a0 = 4 + 8 * (layer - 1);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      f32x4 d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(r.a[a0 + 2 * mt], b0, r.b[4 * layer + mt], 0, 0, 0);
      z[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(r.a[a0 + 2 * mt + 1], b1, d, 0, 0, 0);
    }
  }
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int i = 0; i < 4; ++i) h[mt][i] = gelu(z[mt][i]);
  f32x4 d = {0.f, 0.f, 0.f, 0.f};
  d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(r.a[20], pack_b(h, 0), d, 0, 0, 0);
  d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(r.a[21], pack_b(h, 1), d, 0, 0, 0);
  return d[0] + b4;
}

}  // namespace dh
