// Machine probes of the GPU: HBM bandwidth (triad and copy kernels), VALU
// multiply-add peak (fp32, fp64) and matrix-core peak (MFMA fp32, fp64, bf16,
// fp16).  They replace the reference's CPU probes (src/runtime.cpp:262-310:
// an FMA loop and a triad over all threads) as the measured ceilings of the
// roofline: each probe keeps enough independent work in flight per SIMD that
// only the unit being measured limits it, and reports the best of several
// timed launches (HIP events).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "capi_internal.hpp"

namespace tsdx_hip {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

// c = a + 3 b over float4 chunks, grid-stride
__global__ void k_triad(const float4* __restrict__ a, const float4* __restrict__ b, float4* __restrict__ c,
                        int64_t n4) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (int64_t)gridDim.x * blockDim.x) {
    const float4 x = a[i], y = b[i];
    c[i] = make_float4(x.x + 3.0f * y.x, x.y + 3.0f * y.y, x.z + 3.0f * y.z, x.w + 3.0f * y.w);
  }
}
__global__ void k_copy(const float4* __restrict__ a, float4* __restrict__ c, int64_t n4) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (int64_t)gridDim.x * blockDim.x)
    c[i] = a[i];
}
__global__ void k_fill(float4* __restrict__ a, int64_t n4, float v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = make_float4(v, v, v, v);
}

// R independent multiply-add chains per lane; 2 R iters flops per lane.
template <class T, int R>
__global__ void k_fma_peak(T m, T c, int iters, T* __restrict__ out) {
  T acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = T(1) + T(threadIdx.x + r) * T(1e-3);
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = __builtin_fma(acc[r], m, c);
  }
  T s = 0;
#pragma unroll
  for (int r = 0; r < R; ++r) s += acc[r];
  out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// Four independent MFMA accumulator chains per wave.
constexpr int kChains = 4;
template <int KIND>
__global__ void k_mfma_peak(float seed, int iters, float* __restrict__ out) {
  const float x = seed + (float)threadIdx.x * 1e-3f;
  float s = 0.0f;
  if constexpr (KIND == 0) {  // v_mfma_f32_16x16x4f32: 2 * 16 * 16 * 4 flops
    v4f acc[kChains] = {};
    for (int i = 0; i < iters; ++i)
#pragma unroll
      for (int q = 0; q < kChains; ++q) acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(x, x, acc[q], 0, 0, 0);
#pragma unroll
    for (int q = 0; q < kChains; ++q) s += acc[q][0] + acc[q][3];
  } else if constexpr (KIND == 1) {  // v_mfma_f64_16x16x4f64
    v4d acc[kChains] = {};
    const double xd = x;
    for (int i = 0; i < iters; ++i)
#pragma unroll
      for (int q = 0; q < kChains; ++q) acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(xd, xd, acc[q], 0, 0, 0);
#pragma unroll
    for (int q = 0; q < kChains; ++q) s += (float)(acc[q][0] + acc[q][3]);
  } else if constexpr (KIND == 2) {  // v_mfma_f32_16x16x32_bf16: 2 * 16 * 16 * 32 flops
    v8bf a;
#pragma unroll
    for (int e = 0; e < 8; ++e) a[e] = (__bf16)(x * (float)(e + 1));
    v4f acc[kChains] = {};
    for (int i = 0; i < iters; ++i)
#pragma unroll
      for (int q = 0; q < kChains; ++q) acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, a, acc[q], 0, 0, 0);
#pragma unroll
    for (int q = 0; q < kChains; ++q) s += acc[q][0] + acc[q][3];
  } else {  // v_mfma_f32_16x16x32_f16
    v8h a;
#pragma unroll
    for (int e = 0; e < 8; ++e) a[e] = (_Float16)(x * (float)(e + 1));
    v4f acc[kChains] = {};
    for (int i = 0; i < iters; ++i)
#pragma unroll
      for (int q = 0; q < kChains; ++q) acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, a, acc[q], 0, 0, 0);
#pragma unroll
    for (int q = 0; q < kChains; ++q) s += acc[q][0] + acc[q][3];
  }
  out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

int cu_count() {
  int dev = 0, ncu = 0;
  hip_check(hipGetDevice(&dev), "device");
  hip_check(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev), "device attribute");
  return ncu;
}

// Best time (seconds) of `reps` launches of f on a private stream.
template <class F>
double best_seconds(int reps, F&& f) {
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream");
  double best = 1e30;
  try {
    hip_check(hipEventCreate(&e0), "event");
    hip_check(hipEventCreate(&e1), "event");
    f(s);  // warm-up
    hip_check(hipStreamSynchronize(s), "probe warm-up");
    for (int r = 0; r < reps; ++r) {
      hip_check(hipEventRecord(e0, s), "event");
      f(s);
      hip_check(hipEventRecord(e1, s), "event");
      hip_check(hipEventSynchronize(e1), "probe");
      float ms = 0.0f;
      hip_check(hipEventElapsedTime(&ms, e0, e1), "event time");
      best = std::min(best, (double)ms * 1e-3);
    }
  } catch (...) {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(s);
  return best;
}

// Device buffer freed on scope exit.
struct DevBuf {
  void* p = nullptr;
  explicit DevBuf(size_t bytes) { hip_check(hipMalloc(&p, bytes), "probe allocation"); }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

double bandwidth(int kind) {
  // 1 GiB per array: far beyond the 256 MiB of Infinity Cache
  const int64_t n4 = (int64_t(1) << 30) / 16;
  DevBuf a(n4 * 16), b(n4 * 16), c(n4 * 16);
  const unsigned grid = (unsigned)(cu_count() * 16);
  hipLaunchKernelGGL(k_fill, dim3(grid), dim3(256), 0, nullptr, (float4*)a.p, n4, 1.0f);
  hipLaunchKernelGGL(k_fill, dim3(grid), dim3(256), 0, nullptr, (float4*)b.p, n4, 2.0f);
  hip_check(hipDeviceSynchronize(), "probe fill");
  const double t = best_seconds(5, [&](hipStream_t s) {
    if (kind == TSDX_HIP_PROBE_COPY)
      hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, s, (const float4*)a.p, (float4*)c.p, n4);
    else
      hipLaunchKernelGGL(k_triad, dim3(grid), dim3(256), 0, s, (const float4*)a.p, (const float4*)b.p, (float4*)c.p,
                         n4);
    hip_check(hipGetLastError(), "probe launch");
  });
  const double bytes = (kind == TSDX_HIP_PROBE_COPY ? 2.0 : 3.0) * (double)n4 * 16.0;
  return bytes / t * 1e-9;
}

double valu_gflops(int prec) {
  constexpr int R = 16;
  const int iters = 20000;
  const unsigned grid = (unsigned)(cu_count() * 16), block = 256;
  DevBuf out((size_t)grid * block * 8);
  const double t = best_seconds(3, [&](hipStream_t s) {
    if (prec == TSDX_HIP_PROBE_FP64)
      hipLaunchKernelGGL((k_fma_peak<double, R>), dim3(grid), dim3(block), 0, s, 0.999999, 1e-7, iters,
                         (double*)out.p);
    else
      hipLaunchKernelGGL((k_fma_peak<float, R>), dim3(grid), dim3(block), 0, s, 0.999999f, 1e-7f, iters,
                         (float*)out.p);
    hip_check(hipGetLastError(), "probe launch");
  });
  return 2.0 * R * (double)iters * grid * block / t * 1e-9;
}

double mfma_tflops(int kind) {
  const int iters = kind == 1 ? 1000 : 4000;
  // four waves per SIMD
  const unsigned grid = (unsigned)(cu_count() * 4), block = 256;
  DevBuf out((size_t)grid * block * 4);
  const double t = best_seconds(3, [&](hipStream_t s) {
    switch (kind) {
      case TSDX_HIP_PROBE_MFMA_FP32:
        hipLaunchKernelGGL((k_mfma_peak<0>), dim3(grid), dim3(block), 0, s, 0.5f, iters, (float*)out.p);
        break;
      case TSDX_HIP_PROBE_MFMA_FP64:
        hipLaunchKernelGGL((k_mfma_peak<1>), dim3(grid), dim3(block), 0, s, 0.5f, iters, (float*)out.p);
        break;
      case TSDX_HIP_PROBE_MFMA_BF16:
        hipLaunchKernelGGL((k_mfma_peak<2>), dim3(grid), dim3(block), 0, s, 0.5f, iters, (float*)out.p);
        break;
      default:
        hipLaunchKernelGGL((k_mfma_peak<3>), dim3(grid), dim3(block), 0, s, 0.5f, iters, (float*)out.p);
        break;
    }
    hip_check(hipGetLastError(), "probe launch");
  });
  const double k = (kind == TSDX_HIP_PROBE_MFMA_BF16 || kind == TSDX_HIP_PROBE_MFMA_FP16) ? 32.0 : 4.0;
  const double flops_per_mfma = 2.0 * 16 * 16 * k;
  const double waves = (double)grid * block / 64.0;
  return flops_per_mfma * kChains * (double)iters * waves / t * 1e-12;
}

}  // namespace
}  // namespace tsdx_hip

using namespace tsdx_hip;

extern "C" {

int tsdx_hip_measure_bandwidth_gbs(int kind, double* gbs) {
  return guarded([&] {
    if (!gbs) throw InvalidArg("gbs must not be null");
    if (kind != TSDX_HIP_PROBE_TRIAD && kind != TSDX_HIP_PROBE_COPY) throw InvalidArg("unknown bandwidth probe");
    require_device();
    *gbs = bandwidth(kind);
  });
}

int tsdx_hip_measure_peak_gflops(int precision, double* gflops) {
  return guarded([&] {
    if (!gflops) throw InvalidArg("gflops must not be null");
    if (precision != TSDX_HIP_PROBE_FP32 && precision != TSDX_HIP_PROBE_FP64) throw InvalidArg("unknown VALU probe");
    require_device();
    *gflops = valu_gflops(precision);
  });
}

int tsdx_hip_measure_mfma_tflops(int kind, double* tflops) {
  return guarded([&] {
    if (!tflops) throw InvalidArg("tflops must not be null");
    if (kind < TSDX_HIP_PROBE_MFMA_FP32 || kind > TSDX_HIP_PROBE_MFMA_FP16) throw InvalidArg("unknown MFMA probe");
    require_device();
    *tflops = mfma_tflops(kind);
  });
}

}  // extern "C"
