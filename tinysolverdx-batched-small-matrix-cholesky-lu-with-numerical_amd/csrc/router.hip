// Launchers of the router kernels (router_kernels.hpp) and of the routed posv.
#include "capi_internal.hpp"
#include "mfma_kernels.hpp"
#include "router_api.hpp"
#include "router_kernels.hpp"

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

// W1 and W2 of the parameter block P regrouped as the in-wave network reads
// them (mx_route): element (o, k) at ((k / 4) * 64 + o) * 4 + k % 4.
__global__ void k_router_tiles(const float* __restrict__ P, float* __restrict__ W) {
  for (int e = threadIdx.x; e < kTileFloats; e += blockDim.x) {
    const bool w1 = e < kTileW2;
    const int x = w1 ? e : e - kTileW2, K = w1 ? kRF : kRH;
    const int c = x / (kRH * 4), o = (x / 4) % kRH, k = 4 * c + x % 4;
    W[e] = P[(w1 ? kOffW1 : kOffW2) + o * K + k];
  }
}

__global__ void k_zero_counts(int* p, int k) {
  if ((int)threadIdx.x < k) p[threadIdx.x] = 0;
}

hipError_t launch_posv_routed(const IrArgs& a0, const float* P, float* Wt, int* tier, float* feat, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_counts, dim3(1), dim3(64), 0, s, a0.fcount, 1);
  hipLaunchKernelGGL(k_router_tiles, dim3(1), dim3(256), 0, s, P, Wt);
  const int nblk = (a0.n + 15) / 16;
  for (int64_t m0 = 0; m0 < a0.batch; m0 += kIrChunk) {
    IrArgs a = a0;
    a.m0 = m0;
    const dim3 grid((unsigned)(a0.batch - m0 < kIrChunk ? a0.batch - m0 : kIrChunk)), block(kWave);
    if (nblk == 1)
      hipLaunchKernelGGL((k_posv_mx_routed<1>), grid, block, 0, s, a, P, Wt, tier, feat);
    else if (nblk == 2)
      hipLaunchKernelGGL((k_posv_mx_routed<2>), grid, block, 0, s, a, P, Wt, tier, feat);
    else if (nblk == 3)
      hipLaunchKernelGGL((k_posv_mx_routed<3>), grid, block, 0, s, a, P, Wt, tier, feat);
    else
      hipLaunchKernelGGL((k_posv_mx_routed<4>), grid, block, 0, s, a, P, Wt, tier, feat);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_posv_gather(a0, s);
}

}  // namespace tsdx_hip
