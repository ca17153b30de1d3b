// Learned precision router for posv (no reference counterpart).
//
// Features: 12 numbers per matrix from one pass over A (only its lower
// triangle, the posv input convention), staged through LDS as float.
// Network: 12 -> 64 (ReLU) -> 64 (ReLU) -> 3 logits, float weights; the tiers
// are 0 = bf16 factor + refinement, 1 = fp32 factor + refinement, 2 = float64
// direct.  Parameter block (floats, kRouterParams in total):
//   [0, 12) feature means, [12, 24) feature scales (x_std = (x - mean) / scale),
//   W1 (64 x 12), b1 (64), W2 (64 x 64), b2 (64), W3 (3 x 64), b3 (3),
// weights row-major [out][in].  Gradients use the same layout (the first 24
// entries stay zero).
#pragma once

#include "device_common.hpp"
#include "router_layout.hpp"

namespace tsdx_hip {

// ------------------------------------------------------------------ features
// One wave per matrix; the symmetric matrix of A's lower triangle, in float.
//  0 n / 64                       6 log10 max_i r_i / a_ii  (r_i = sum_{j!=i} |a_ij|)
//  1 log10 min_i a_ii             7 log10 (max a_ii / min a_ii)
//  2 log10 max_i |a_ii|           8 log10 (||A||_F / mean a_ii)
//  3 log10 ||A||_F                9 mean_i log10 a_ii - log10 max a_ii
//  4 log10 ||A||_inf             10 1 if some a_ii <= 0 (or NaN) else 0
//  5 min_i (a_ii - r_i) / ||A||_F 11 fraction of rows with a_ii > r_i
// Logarithms of non-positive quantities are clamped to -30 (and ratios with a
// non-positive diagonal to +30).
constexpr int kFeatWPB = 2;  // 2 x 16.6 KB of LDS
constexpr int kFeatLd = 65;  // padded row stride of the LDS copy

__global__ __launch_bounds__(kWave* kFeatWPB) void k_router_features(const double* __restrict__ A, int n,
                                                                     int64_t batch, float* __restrict__ feat) {
  __shared__ float sm[kFeatWPB][64 * kFeatLd];
  const int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave;
  const int64_t m = (int64_t)blockIdx.x * kFeatWPB + wv;
  if (m >= batch) return;  // uniform per wave; waves never synchronise with each other
  float* L = sm[wv];
  const double* Am = A + m * (int64_t)n * n;
  for (int e = lane; e < n * n; e += kWave) {
    const int r = e / n, c = e - r * n;
    if (c <= r) L[r * kFeatLd + c] = (float)Am[e];
  }
  wave_sync();
  const int i = lane;
  const bool act = i < n;
  float d = 1.0f, rs = 0.0f, fro = 0.0f;
  if (act) {
    for (int j = 0; j < n; ++j) {
      const float v = j <= i ? L[i * kFeatLd + j] : L[j * kFeatLd + i];
      fro = __builtin_fmaf(v, v, fro);
      if (j == i)
        d = v;
      else
        rs += __builtin_fabsf(v);
    }
  }
  const float big = 3.0e38f;
  const bool pos = act && d > 0.0f;
  auto sum = [](float a, float b) { return a + b; };
  auto mn = [](float a, float b) { return a < b ? a : b; };
  auto mx = [](float a, float b) { return a > b ? a : b; };
  const float dmin = group_allreduce<64>(act ? d : big, mn);
  const float dmax = group_allreduce<64>(act ? __builtin_fabsf(d) : 0.0f, mx);
  const float fr = __builtin_sqrtf(group_allreduce<64>(fro, sum));
  const float ninf = group_allreduce<64>(act ? __builtin_fabsf(d) + rs : 0.0f, mx);
  const float gers = group_allreduce<64>(act ? d - rs : big, mn);
  const float dom = group_allreduce<64>(act ? (pos ? rs / d : big) : 0.0f, mx);
  const float dsum = group_allreduce<64>(act ? d : 0.0f, sum);
  const float lsum = group_allreduce<64>(act ? flog10(d) : 0.0f, sum);
  const float nonpos = group_allreduce<64>((act && !(d > 0.0f)) ? 1.0f : 0.0f, mx);
  const float ddom = group_allreduce<64>((act && d > rs) ? 1.0f : 0.0f, sum);
  float f = 0.0f;
  switch (lane) {
    case 0: f = (float)n / 64.0f; break;
    case 1: f = flog10(dmin); break;
    case 2: f = flog10(dmax); break;
    case 3: f = flog10(fr); break;
    case 4: f = flog10(ninf); break;
    case 5: f = fr > 0.0f ? gers / fr : 0.0f; f = f < -30.0f ? -30.0f : f; break;
    case 6: f = dom >= big ? 30.0f : flog10(dom); break;
    case 7: f = dmin > 0.0f ? flog10(dmax / dmin) : 30.0f; break;
    case 8: f = dsum > 0.0f ? flog10(fr / (dsum / (float)n)) : 30.0f; break;
    case 9: f = lsum / (float)n - flog10(dmax); break;
    case 10: f = nonpos; break;
    case 11: f = ddom / (float)n; break;
    default: break;
  }
  if (lane < kRF) feat[m * kRF + lane] = f;
}

// ------------------------------------------------------------------ network
// Forward pass of one sample with the parameters in LDS (every thread reads
// the same weight at the same time: LDS broadcasts).
struct MlpState {
  float x[kRF], h1[kRH], h2[kRH], z[kRT];
};
__device__ __forceinline__ void mlp_forward(const float* __restrict__ P, const float* __restrict__ fx, MlpState& s) {
  // (every loop that indexes the register arrays is fully unrolled, so the
  // arrays stay in registers)
#pragma unroll
  for (int f = 0; f < kRF; ++f) s.x[f] = (fx[f] - P[f]) / P[kRF + f];
#pragma unroll
  for (int o = 0; o < kRH; ++o) {
    float a = P[kOffB1 + o];
#pragma unroll
    for (int f = 0; f < kRF; ++f) a = __builtin_fmaf(P[kOffW1 + o * kRF + f], s.x[f], a);
    s.h1[o] = a > 0.0f ? a : 0.0f;
  }
#pragma unroll
  for (int o = 0; o < kRH; ++o) {
    float a = P[kOffB2 + o];
#pragma unroll
    for (int k = 0; k < kRH; ++k) a = __builtin_fmaf(P[kOffW2 + o * kRH + k], s.h1[k], a);
    s.h2[o] = a > 0.0f ? a : 0.0f;
  }
#pragma unroll
  for (int o = 0; o < kRT; ++o) {
    float a = P[kOffB3 + o];
#pragma unroll
    for (int k = 0; k < kRH; ++k) a = __builtin_fmaf(P[kOffW3 + o * kRH + k], s.h2[k], a);
    s.z[o] = a;
  }
}

__device__ __forceinline__ void load_params(float* sP, const float* __restrict__ P) {
  for (int e = threadIdx.x; e < kRouterParams; e += blockDim.x) sP[e] = P[e];
  __syncthreads();
}

constexpr int kMlpThreads = 256;

// logits (count x 3) of every sample
__global__ __launch_bounds__(kMlpThreads) void k_router_forward(const float* __restrict__ P, const float* __restrict__ feat,
                                                                int64_t count, float* __restrict__ logits) {
  __shared__ float sP[kRouterParams];
  load_params(sP, P);
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= count) return;
  MlpState st;
  mlp_forward(sP, feat + s * kRF, st);
  for (int o = 0; o < kRT; ++o) logits[s * kRT + o] = st.z[o];
}

// ------------------------------------------------------------------ training
// Gradient of the summed cross-entropy over a chunk of kGradChunk samples per
// workgroup (one thread per sample for the per-sample passes, the weight
// gradients as small GEMMs over the chunk from LDS); each workgroup writes its
// partial gradient and loss, k_router_reduce sums them in a fixed order.
constexpr int kGradChunk = 64;  // one wave; 58 KB of LDS

__global__ __launch_bounds__(kGradChunk) void k_router_grad_partial(const float* __restrict__ P,
                                                                    const float* __restrict__ feat,
                                                                    const int* __restrict__ label, int64_t count,
                                                                    float* __restrict__ partial,
                                                                    float* __restrict__ loss_partial) {
  __shared__ float sP[kRouterParams];
  __shared__ float sX[kGradChunk][kRF];
  __shared__ float sH1[kGradChunk][kRH + 1];
  __shared__ float sH2[kGradChunk][kRH + 1];
  __shared__ float sDZ[kGradChunk][kRT];
  __shared__ float sLoss[kGradChunk];
  load_params(sP, P);
  const int t = threadIdx.x;
  const int64_t s = (int64_t)blockIdx.x * kGradChunk + t;
  const bool live = s < count;
  {
    MlpState st;
    if (live) {
      mlp_forward(sP, feat + s * kRF, st);
      const int y = label[s];
      float zm = st.z[0] > st.z[1] ? st.z[0] : st.z[1];
      zm = zm > st.z[2] ? zm : st.z[2];
      float e[kRT], se = 0.0f;
#pragma unroll
      for (int o = 0; o < kRT; ++o) {
        e[o] = __expf(st.z[o] - zm);
        se += e[o];
      }
      const float zy = y == 0 ? st.z[0] : (y == 1 ? st.z[1] : st.z[2]);
      sLoss[t] = zm + __logf(se) - zy;
#pragma unroll
      for (int o = 0; o < kRT; ++o) sDZ[t][o] = e[o] / se - (o == y ? 1.0f : 0.0f);
    } else {
#pragma unroll
      for (int f = 0; f < kRF; ++f) st.x[f] = 0.0f;
#pragma unroll
      for (int k = 0; k < kRH; ++k) st.h1[k] = st.h2[k] = 0.0f;
      sLoss[t] = 0.0f;
      for (int o = 0; o < kRT; ++o) sDZ[t][o] = 0.0f;
    }
#pragma unroll
    for (int f = 0; f < kRF; ++f) sX[t][f] = st.x[f];
#pragma unroll
    for (int k = 0; k < kRH; ++k) {
      sH1[t][k] = st.h1[k];
      sH2[t][k] = st.h2[k];
    }
  }
  __syncthreads();
  float* G = partial + (int64_t)blockIdx.x * kRouterParams;
  // layer 3: dW3[o][k] = sum_s dz[s][o] h2[s][k], db3[o] = sum_s dz[s][o]
  for (int e = t; e < kRT * kRH + kRT; e += kGradChunk) {
    float a = 0.0f;
    if (e < kRT * kRH) {
      const int o = e / kRH, k = e % kRH;
      for (int q = 0; q < kGradChunk; ++q) a = __builtin_fmaf(sDZ[q][o], sH2[q][k], a);
      G[kOffW3 + e] = a;
    } else {
      const int o = e - kRT * kRH;
      for (int q = 0; q < kGradChunk; ++q) a += sDZ[q][o];
      G[kOffB3 + o] = a;
    }
  }
  __syncthreads();
  // dh2 = (W3^T dz) * [h2 > 0], in place of h2 (own sample)
  for (int k = 0; k < kRH; ++k) {
    float a = 0.0f;
    for (int o = 0; o < kRT; ++o) a = __builtin_fmaf(sP[kOffW3 + o * kRH + k], sDZ[t][o], a);
    sH2[t][k] = sH2[t][k] > 0.0f ? a : 0.0f;
  }
  __syncthreads();
  // layer 2: dW2[o][k] = sum_s dh2[s][o] h1[s][k], db2[o] = sum_s dh2[s][o]
  for (int e = t; e < kRH * kRH + kRH; e += kGradChunk) {
    float a = 0.0f;
    if (e < kRH * kRH) {
      const int o = e / kRH, k = e % kRH;
      for (int q = 0; q < kGradChunk; ++q) a = __builtin_fmaf(sH2[q][o], sH1[q][k], a);
      G[kOffW2 + e] = a;
    } else {
      const int o = e - kRH * kRH;
      for (int q = 0; q < kGradChunk; ++q) a += sH2[q][o];
      G[kOffB2 + o] = a;
    }
  }
  __syncthreads();
  // dh1 = (W2^T dh2) * [h1 > 0], in place of h1 (own sample)
  {
    float dh1[kRH];
#pragma unroll
    for (int k = 0; k < kRH; ++k) {
      float a = 0.0f;
      for (int o = 0; o < kRH; ++o) a = __builtin_fmaf(sP[kOffW2 + o * kRH + k], sH2[t][o], a);
      dh1[k] = sH1[t][k] > 0.0f ? a : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < kRH; ++k) sH1[t][k] = dh1[k];
  }
  __syncthreads();
  // layer 1: dW1[o][f] = sum_s dh1[s][o] x[s][f], db1[o] = sum_s dh1[s][o]
  for (int e = t; e < kRH * kRF + kRH; e += kGradChunk) {
    float a = 0.0f;
    if (e < kRH * kRF) {
      const int o = e / kRF, f = e % kRF;
      for (int q = 0; q < kGradChunk; ++q) a = __builtin_fmaf(sH1[q][o], sX[q][f], a);
      G[kOffW1 + e] = a;
    } else {
      const int o = e - kRH * kRF;
      for (int q = 0; q < kGradChunk; ++q) a += sH1[q][o];
      G[kOffB1 + o] = a;
    }
  }
  if (t < kOffW1) G[t] = 0.0f;
  if (t == 0) {
    float a = 0.0f;
    for (int q = 0; q < kGradChunk; ++q) a += sLoss[q];
    loss_partial[blockIdx.x] = a;
  }
}

// grad = scale * sum over workgroups (fixed order), loss likewise.
__global__ void k_router_reduce(const float* __restrict__ partial, const float* __restrict__ loss_partial, int nblk,
                                float scale, float* __restrict__ grad, float* __restrict__ loss) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < kRouterParams) {
    float a = 0.0f;
    for (int b = 0; b < nblk; ++b) a += partial[(int64_t)b * kRouterParams + e];
    grad[e] = a * scale;
  }
  if (e == 0 && loss) {
    float a = 0.0f;
    for (int b = 0; b < nblk; ++b) a += loss_partial[b];
    *loss = a * scale;
  }
}

// Adam (beta1 0.9, beta2 0.999, eps 1e-8) on the trainable entries.
__global__ void k_router_adam(float* __restrict__ P, const float* __restrict__ grad, float* __restrict__ mom,
                              float* __restrict__ vel, int step, float lr) {
  const int e = kOffW1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= kRouterParams) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float g = grad[e];
  const float m = b1 * mom[e] + (1.0f - b1) * g;
  const float v = b2 * vel[e] + (1.0f - b2) * g * g;
  mom[e] = m;
  vel[e] = v;
  const float mh = m / (1.0f - __powf(b1, (float)step)), vh = v / (1.0f - __powf(b2, (float)step));
  P[e] -= lr * mh / (__builtin_sqrtf(vh) + eps);
}

}  // namespace tsdx_hip
