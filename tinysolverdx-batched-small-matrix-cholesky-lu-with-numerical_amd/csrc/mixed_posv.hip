// Launchers of the mixed-precision refinement tier (mixed_kernels.hpp).
#include "capi_internal.hpp"
#include "launch_impl.hpp"
#include "mfma64_kernels.hpp"

namespace tsdx_hip {

__global__ void k_zero_int(int* p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}

// list[i] = i for i < batch, count = batch (every matrix takes the float64 solve)
__global__ void k_iota(int* list, int* count, int* stat, int64_t batch) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < batch) {
    list[i] = (int)i;
    if (stat) stat[i] = -1;
  }
  if (i == 0) *count = (int)batch;
}

// Matrix-core refinement kernel: one wave (one workgroup) per matrix.
template <int NBLK>
static void launch_mx(const IrArgs& a0, int low, hipStream_t s) {
  for (int64_t m0 = 0; m0 < a0.batch; m0 += kIrChunk) {
    IrArgs a = a0;
    a.m0 = m0;
    const dim3 grid((unsigned)(a0.batch - m0 < kIrChunk ? a0.batch - m0 : kIrChunk)), block(kWave);
    if (low == 2)
      hipLaunchKernelGGL((k_posv_mx<NBLK, Low::BF16>), grid, block, 0, s, a);
    else if (low == 3)
      hipLaunchKernelGGL((k_posv_mx<NBLK, Low::F16>), grid, block, 0, s, a);
    else
      hipLaunchKernelGGL((k_posv_mx<NBLK, Low::F32>), grid, block, 0, s, a);
  }
}

// float64 posv of the matrices on the fallback list (grid-stride over it):
// for n > 32 the float64 block kernel on the matrix cores (k_posv_mx64, which
// hands its own failures to the plain posv in the same wave) unless `plain`
// asks for the plain posv throughout; the plain posv for smaller sizes, whose
// buckets put several matrices in a wave.
// full: the list holds every matrix (a float64 factor was asked for); else it
// holds the refinement's failures and the router's float64 choices, usually
// few, so a smaller grid strides over it (workgroups of an empty list still
// cost their dispatch).
constexpr unsigned kGatherGridShort = 1024;
template <int NB>
static hipError_t gather(const IrArgs& a, hipStream_t s, bool plain, bool full) {
  if constexpr (NB == 64) {
    if (!plain) {
      // two waves per SIMD fit this kernel: 4096 workgroups fill the GPU twice.
      // A short list is latency-bound (one matrix takes a wave ~25 us), so it
      // gets one workgroup per entry up to 2048: up to about 1800 entries
      // (7 workgroups of 20.5 KiB LDS per CU) then finish in one round.
      const unsigned cap = full ? 4096u : 2048u;
      const dim3 gg((unsigned)(a.batch < (int64_t)cap ? a.batch : cap));
      if (a.nrhs <= 1)
        hipLaunchKernelGGL((k_posv_mx64<4, 1>), gg, dim3(kWave), 0, s, a);
      else
        hipLaunchKernelGGL((k_posv_mx64<4, 4>), gg, dim3(kWave), 0, s, a);
      return hipGetLastError();
    }
  }
  const dim3 g = grid_for<NB>(a.batch);
  const unsigned cap = full ? 16384u : kGatherGridShort;
  const dim3 gg(g.x < cap ? g.x : cap);
  if (a.nrhs <= 1)
    hipLaunchKernelGGL((k_posv_gather<NB, 1>), gg, block_for<NB>(), 0, s, a);
  else
    hipLaunchKernelGGL((k_posv_gather<NB, 4>), gg, block_for<NB>(), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_posv_gather(const IrArgs& a, hipStream_t s) {
  if (a.n <= 4) return gather<4>(a, s, false, false);
  if (a.n <= 8) return gather<8>(a, s, false, false);
  if (a.n <= 16) return gather<16>(a, s, false, false);
  if (a.n <= 32) return gather<32>(a, s, false, false);
  return gather<64>(a, s, false, false);
}

template <int NB>
static hipError_t launch_ir(const IrArgs& a, int low, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, s, a.fcount);
  if (low != 4 && low != 5) {
    launch_mx<(NB + 15) / 16>(a, low, s);
  } else {  // float64 for every matrix (4 plain, 5 the fallback tier): the list is 0..batch-1
    hipLaunchKernelGGL(k_iota, dim3((unsigned)((a.batch + 255) / 256)), dim3(256), 0, s, a.flist, a.fcount, a.stat, a.batch);
  }
  return gather<NB>(a, s, low == 4, low == 4 || low == 5);
}

hipError_t launch_posv_ir(const IrArgs& a, int low, hipStream_t s) {
  if (a.n <= 4) return launch_ir<4>(a, low, s);
  if (a.n <= 8) return launch_ir<8>(a, low, s);
  if (a.n <= 16) return launch_ir<16>(a, low, s);
  if (a.n <= 32) return launch_ir<32>(a, low, s);
  return launch_ir<64>(a, low, s);
}

}  // namespace tsdx_hip
