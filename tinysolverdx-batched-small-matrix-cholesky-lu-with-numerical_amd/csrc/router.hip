// Launchers of the router kernels (router_kernels.hpp) and of the routed posv.
#include "capi_internal.hpp"
#include "routed_kernels.hpp"
#include "router_api.hpp"
#include "router_kernels.hpp"
#include "router_net.hpp"

namespace tsdx_hip {

hipError_t launch_router_features(const double* A, int n, int64_t batch, float* feat, hipStream_t s) {
  const int64_t blocks = (batch + kFeatWPB - 1) / kFeatWPB;
  hipLaunchKernelGGL(k_router_features, dim3((unsigned)blocks), dim3(kWave * kFeatWPB), 0, s, A, n, batch, feat);
  return hipGetLastError();
}

hipError_t launch_router_forward(const float* P, const float* feat, int64_t count, float* logits, hipStream_t s) {
  const int64_t blocks = (count + kMlpThreads - 1) / kMlpThreads;
  hipLaunchKernelGGL(k_router_forward, dim3((unsigned)blocks), dim3(kMlpThreads), 0, s, P, feat, count, logits);
  return hipGetLastError();
}

size_t router_grad_workspace(int64_t count) {
  const int64_t nblk = (count + kGradChunk - 1) / kGradChunk;
  return (size_t)nblk * (kRouterParams + 1) * sizeof(float) + 256;
}

hipError_t launch_router_grad(const float* P, const float* feat, const int* label, int64_t count, float* grad,
                              float* loss, void* work, hipStream_t s) {
  const int64_t nblk = (count + kGradChunk - 1) / kGradChunk;
  float* partial = static_cast<float*>(work);
  float* lpart = partial + nblk * kRouterParams;
  hipLaunchKernelGGL(k_router_grad_partial, dim3((unsigned)nblk), dim3(kGradChunk), 0, s, P, feat, label, count,
                     partial, lpart);
  hipLaunchKernelGGL(k_router_reduce, dim3((kRouterParams + 255) / 256), dim3(256), 0, s, partial, lpart, (int)nblk,
                     1.0f / (float)count, grad, loss);
  return hipGetLastError();
}

hipError_t launch_router_adam(float* P, const float* grad, float* m, float* v, int step, float lr, hipStream_t s) {
  hipLaunchKernelGGL(k_router_adam, dim3((kRouterParams - kOffW1 + 255) / 256), dim3(256), 0, s, P, grad, m, v, step,
                     lr);
  return hipGetLastError();
}

// The network of the parameter block P prepared for the matrix cores
// (router_layout.hpp): bfloat16 A fragments (round to nearest even) with the
// feature normalisation folded into layer 1, and the biases; one thread per
// bfloat16 entry (kPrepBlocks workgroups).  When counts is given, the first
// thread also clears ncount counters (the routed solve's list counts).
constexpr int kPrepThreads = 256, kPrepBlocks = kPrepFrags * 64 * 8 / kPrepThreads;
__global__ __launch_bounds__(kPrepThreads) void k_router_prep(const float* __restrict__ P, void* __restrict__ prep,
                                                              int* __restrict__ counts, int ncount) {
  __bf16* W = static_cast<__bf16*>(prep);
  float* bias = reinterpret_cast<float*>(static_cast<char*>(prep) + kPrepBiasByte);
  const int e = blockIdx.x * kPrepThreads + threadIdx.x;  // < kPrepFrags * 64 * 8
  const int f = e / 512, l = (e / 8) % 64, x = e % 8;
  const int row = l % 16, col = 8 * (l / 16) + x;  // within the fragment's 16 x 32 tile
  float w = 0.0f;
  if (f < 4) {
    const int o = 16 * f + row;
    if (col < kRF) w = P[kOffW1 + o * kRF + col] / P[kRF + col];
  } else if (f < 12) {
    const int t = (f - 4) / 2, c = (f - 4) % 2;
    w = P[kOffW2 + (16 * t + row) * kRH + 32 * c + col];
  } else {
    const int c = f - 12;
    if (row < kRT) w = P[kOffW3 + row * kRH + 32 * c + col];
  }
  W[e] = (__bf16)w;
  if (blockIdx.x == 0 && threadIdx.x < kPrepBiasFloats) {
    const int i = threadIdx.x;
    float b = 0.0f;
    if (i < kRH) {  // b1' = b1 - sum_q (W1[o][q] / scale_q) mean_q, in the order of q
      b = P[kOffB1 + i];
      for (int q = 0; q < kRF; ++q) b -= (P[kOffW1 + i * kRF + q] / P[kRF + q]) * P[q];
    } else if (i < kPrepB3) {
      b = P[kOffB2 + i - kPrepB2];
    } else if (i - kPrepB3 < kRT) {
      b = P[kOffB3 + i - kPrepB3];
    }
    bias[i] = b;
  }
  if (counts && e < ncount) counts[e] = 0;
}
static_assert(kPrepFrags * 64 * 8 % kPrepThreads == 0 && kPrepBiasFloats <= kPrepThreads, "prep grid");

// Tiers and logits of stored feature vectors: one wave per 16 samples, sample
// 16 w + j in column j of the network (router_net.hpp).
constexpr int kInferWPB = 4;
__global__ __launch_bounds__(kWave* kInferWPB) void k_router_infer(const float* __restrict__ feat, int64_t count,
                                                                   const void* __restrict__ prep, int* __restrict__ tier,
                                                                   float* __restrict__ logits) {
  __shared__ __align__(16) __bf16 hs[kInferWPB][kNetLdsHalves];
  const int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave;
  const int64_t base = ((int64_t)blockIdx.x * kInferWPB + wv) * 16;
  if (base >= count) return;  // uniform per wave; waves never synchronise with each other
  const int j = lane & 15, kg = lane >> 4;
  const int64_t s = base + j;
  const float* fs = feat + (s < count ? s : count - 1) * kRF;  // a valid row for every lane
  v8bf x;
  static_for<0, 8>([&](auto ec) {
    constexpr int e = decltype(ec)::value;
    const int q = 8 * kg + e;
    const float v = fs[q < kRF ? q : 0];
    x[e] = (__bf16)((s < count && q < kRF) ? v : 0.0f);
  });
  const v4fn z = net_logits(static_cast<const v8bf*>(prep),
                            reinterpret_cast<const float*>(static_cast<const char*>(prep) + kPrepBiasByte), x, hs[wv],
                            lane);
  if (kg == 0 && s < count) {
    if (tier) tier[s] = net_tier(z[0], z[1], z[2]);
    if (logits) {
      logits[s * kRT + 0] = z[0];
      logits[s * kRT + 1] = z[1];
      logits[s * kRT + 2] = z[2];
    }
  }
}

hipError_t launch_router_infer(const float* P, const float* feat, int64_t count, int* tier, float* logits, void* prep,
                               hipStream_t s) {
  hipLaunchKernelGGL(k_router_prep, dim3(kPrepBlocks), dim3(kPrepThreads), 0, s, P, prep, nullptr, 0);
  const int64_t waves = (count + 15) / 16;
  hipLaunchKernelGGL(k_router_infer, dim3((unsigned)((waves + kInferWPB - 1) / kInferWPB)), dim3(kWave * kInferWPB),
                     0, s, feat, count, prep, tier, logits);
  return hipGetLastError();
}

hipError_t launch_posv_routed(const IrArgs& a0, const float* P, void* prep, int* tier, float* feat, hipStream_t s) {
  // prepared network, and the float64 and bfloat16 list counts (adjacent) cleared
  hipLaunchKernelGGL(k_router_prep, dim3(kPrepBlocks), dim3(kPrepThreads), 0, s, P, prep, a0.fcount, 2);
  const int nblk = (a0.n + 15) / 16;
  for (int64_t m0 = 0; m0 < a0.batch; m0 += kIrChunk) {
    IrArgs a = a0;
    a.m0 = m0;
    const dim3 grid((unsigned)(a0.batch - m0 < kIrChunk ? a0.batch - m0 : kIrChunk)), block(kWave);
    if (nblk == 1)
      hipLaunchKernelGGL((k_posv_mx_routed<1>), grid, block, 0, s, a, prep, tier, feat);
    else if (nblk == 2)
      hipLaunchKernelGGL((k_posv_mx_routed<2>), grid, block, 0, s, a, prep, tier, feat);
    else if (nblk == 3)
      hipLaunchKernelGGL((k_posv_mx_routed<3>), grid, block, 0, s, a, prep, tier, feat);
    else
      hipLaunchKernelGGL((k_posv_mx_routed<4>), grid, block, 0, s, a, prep, tier, feat);
  }
  // the bfloat16 tier's list (usually short or empty: a small grid striding
  // over it); its refinement failures join the float64 list
  const dim3 lg((unsigned)(a0.batch < 256 ? a0.batch : 256));
  if (nblk == 1)
    hipLaunchKernelGGL((k_posv_mx_list<1, Low::BF16>), lg, dim3(kWave), 0, s, a0);
  else if (nblk == 2)
    hipLaunchKernelGGL((k_posv_mx_list<2, Low::BF16>), lg, dim3(kWave), 0, s, a0);
  else if (nblk == 3)
    hipLaunchKernelGGL((k_posv_mx_list<3, Low::BF16>), lg, dim3(kWave), 0, s, a0);
  else
    hipLaunchKernelGGL((k_posv_mx_list<4, Low::BF16>), lg, dim3(kWave), 0, s, a0);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_posv_gather(a0, s);
}

}  // namespace tsdx_hip
