// The router's network on the matrix cores (inference).  Each layer is one
// v_mfma_f32_16x16x32_bf16 per 16 outputs and 32 inputs, for 16 samples at
// once (the 16 columns of the B operand); inside the routed posv every column
// holds the same sample, in the standalone inference kernel each column its
// own.  bfloat16 weights and activations, float accumulation and biases.
//
// Operand layouts of v_mfma_f32_16x16x32_bf16 (checked on the GPU):
//   A (16 x 32): lane l holds A[l % 16][8 (l / 16) + e], e = 0..7
//   B (32 x 16): lane l holds B[8 (l / 16) + e][l % 16]
//   D (16 x 16): lane l holds D[4 (l / 16) + v][l % 16], v = 0..3
// The weights are prepared once per call by k_router_prep (router.hip) as A
// fragments (layout in router_layout.hpp), the feature normalisation folded
// into layer 1: W1'[o][q] = W1[o][q] / scale_q, b1'[o] = b1[o] - sum_q
// W1'[o][q] mean_q.
#pragma once

#include "device_common.hpp"
#include "router_layout.hpp"

namespace tsdx_hip {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4fn __attribute__((ext_vector_type(4)));

// LDS row stride (bfloat16 entries) of the activations of one sample: 144
// bytes keeps the 16-byte reads aligned and spreads the 16 samples' 8-byte
// writes over the banks.
constexpr int kNetLd = 72;
constexpr int kNetLdsHalves = 16 * kNetLd;

__device__ __forceinline__ v4fn ld_bias(const float* __restrict__ b, int i) {
  const float4 q = *reinterpret_cast<const float4*>(b + i);
  return v4fn{q.x, q.y, q.z, q.w};
}

// ReLU of four outputs (rows 4g..4g+3 of tile t, sample j) in bfloat16, to LDS.
// The outputs are rounded to bfloat16 in pairs (v_cvt_pk_bf16_f32, nearest
// even), and the ReLU is taken on the packed halves as 16-bit signed integers
// (v_pk_max_i16 with 0): a bfloat16 with its sign bit set, -0 included, is a
// negative integer and becomes +0, every other value is kept.  Rounding keeps
// the sign, so this equals rounding the ReLU of the float output; a NaN output
// with the sign bit clear stays NaN (and reaches the logits, whose NaN test
// then picks the float64 tier).
__device__ __forceinline__ unsigned relu_pk_bf16(float a, float b) {
  typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
  typedef float v2f __attribute__((ext_vector_type(2)));
  typedef short v2s __attribute__((ext_vector_type(2)));
  const v2bf p = __builtin_convertvector((v2f){a, b}, v2bf);  // one v_cvt_pk_bf16_f32
  const v2s r = __builtin_elementwise_max(__builtin_bit_cast(v2s, p), (v2s){0, 0});
  return __builtin_bit_cast(unsigned, r);
}
__device__ __forceinline__ void net_store(v4fn h, __bf16* hs, int j, int g, int t) {
  const uint2 q = {relu_pk_bf16(h[0], h[1]), relu_pk_bf16(h[2], h[3])};
  *reinterpret_cast<uint2*>(hs + j * kNetLd + 16 * t + 4 * g) = q;
}

// Logits of the 16 samples: on return lane (j, 0) (l = j < 16) holds z[0..2]
// of sample j in its first three entries.  x: this lane's layer-1 operand
// (features 8 (l/16) .. +7 of sample l % 16, zero from feature 12 on).
// Wf: the prepared fragments, bias: b1', b2, b3 (prepared), hs: this wave's
// kNetLdsHalves of LDS.
__device__ __forceinline__ v4fn net_logits(const v8bf* __restrict__ Wf, const float* __restrict__ bias, v8bf x,
                                           __bf16* hs, int lane) {
  const int j = lane & 15, g = lane >> 4;
  v4fn h[4];
  // layer 1: 12 (padded to 32) -> 64
  static_for<0, 4>([&](auto t) {
    h[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Wf[t * 64 + lane], x, ld_bias(bias, 16 * t + 4 * g), 0, 0, 0);
  });
  static_for<0, 4>([&](auto t) { net_store(h[t], hs, j, g, t); });
  wave_sync();
  v8bf y0 = *reinterpret_cast<const v8bf*>(hs + j * kNetLd + 8 * g);
  v8bf y1 = *reinterpret_cast<const v8bf*>(hs + j * kNetLd + 32 + 8 * g);
  wave_sync();
  // layer 2: 64 -> 64
  static_for<0, 4>([&](auto t) {
    v4fn a = ld_bias(bias, kPrepB2 + 16 * t + 4 * g);
    a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Wf[(4 + 2 * t) * 64 + lane], y0, a, 0, 0, 0);
    h[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Wf[(5 + 2 * t) * 64 + lane], y1, a, 0, 0, 0);
  });
  static_for<0, 4>([&](auto t) { net_store(h[t], hs, j, g, t); });
  wave_sync();
  y0 = *reinterpret_cast<const v8bf*>(hs + j * kNetLd + 8 * g);
  y1 = *reinterpret_cast<const v8bf*>(hs + j * kNetLd + 32 + 8 * g);
  wave_sync();
  // layer 3: 64 -> 3 (padded to 16)
  v4fn z = ld_bias(bias, kPrepB3 + 4 * g);
  z = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Wf[12 * 64 + lane], y0, z, 0, 0, 0);
  z = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Wf[13 * 64 + lane], y1, z, 0, 0, 0);
  return z;
}

// Tier of logits z[0..2]: the first largest; a NaN logit gives tier 2.
__device__ __forceinline__ int net_tier(float z0, float z1, float z2) {
  int t = z1 > z0 ? 1 : 0;
  t = z2 > (t == 1 ? z1 : z0) ? 2 : t;
  if (!(z0 == z0 && z1 == z1 && z2 == z2)) t = 2;
  return t;
}

}  // namespace tsdx_hip
