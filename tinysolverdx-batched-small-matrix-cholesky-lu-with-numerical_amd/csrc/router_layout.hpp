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
// The regrouped W1 and W2 the in-wave network reads (k_router_tiles).
constexpr int kTileW1 = 0, kTileW2 = kRH * kRF, kTileFloats = kTileW2 + kRH * kRH;
constexpr size_t kRouterTileBytes = kTileFloats * sizeof(float);

// log10 with non-positive (and tiny) arguments clamped to -30: log2 on the
// transcendental unit (arguments past the clamp are normal numbers, which it
// handles to about 1 ulp) times log10(2).
__device__ __forceinline__ float flog10(float x) {
  return x > 1e-30f ? __builtin_amdgcn_logf(x) * 0.30102999566398120f : -30.0f;
}

}  // namespace tsdx_hip
