// Sizes and parameter layout of the precision router's network (shared by the
// router kernels and the routed posv kernel).  Parameter block, floats:
//   [0, 12) feature means, [12, 24) feature scales (x_std = (x - mean) / scale),
//   W1 (64 x 12), b1 (64), W2 (64 x 64), b2 (64), W3 (3 x 64), b3 (3),
// weights row-major [out][in]; ReLU hidden layers; tiers 0 = bf16 + refinement,
// 1 = fp32 + refinement, 2 = float64 direct.
#pragma once

#include <stddef.h>

namespace tsdx_hip {

constexpr int kRF = 12, kRH = 64, kRT = 3;
constexpr int kOffW1 = 24, kOffB1 = kOffW1 + kRH * kRF, kOffW2 = kOffB1 + kRH, kOffB2 = kOffW2 + kRH * kRH,
              kOffW3 = kOffB2 + kRH, kOffB3 = kOffW3 + kRT * kRH, kRouterParams = kOffB3 + kRT;
static_assert(kRouterParams == 5211, "parameter layout");
// The network prepared for the matrix cores (k_router_prep, router_net.hpp):
// kPrepFrags bfloat16 A-operand fragments of v_mfma_f32_16x16x32_bf16, 16 bytes
// per lane, fragment f of lane l at byte (64 f + l) * 16:
//   f = t (0..3):            W1'[16t + l % 16][8 (l / 16) + e]   (inputs >= 12 zero)
//   f = 4 + 2t + c (c = 0,1): W2[16t + l % 16][32c + 8 (l / 16) + e]
//   f = 12 + c:              W3[l % 16][32c + 8 (l / 16) + e]     (outputs >= 3 zero)
// then float biases at kPrepBiasByte: b1' (64), b2 (64 from kPrepB2), b3
// (padded to 16 from kPrepB3), with W1' = W1 / scale and b1' = b1 - W1' mean
// (the feature normalisation folded into layer 1).
constexpr int kPrepFrags = 14;
constexpr size_t kPrepBiasByte = (size_t)kPrepFrags * 64 * 16;
constexpr int kPrepB2 = 64, kPrepB3 = 128, kPrepBiasFloats = 144;
constexpr size_t kRouterPrepBytes = kPrepBiasByte + kPrepBiasFloats * sizeof(float);

// log10 with non-positive (and tiny) arguments clamped to -30: log2 on the
// transcendental unit (arguments past the clamp are normal numbers, which it
// handles to about 1 ulp) times log10(2).
__device__ __forceinline__ float flog10(float x) {
  return x > 1e-30f ? __builtin_amdgcn_logf(x) * 0.30102999566398120f : -30.0f;
}

}  // namespace tsdx_hip
