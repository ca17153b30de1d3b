// The allocation-free form of the refinement-tier posv
// (tsdx_hip_posv_ir_noalloc_f64, include/tsdx_hip.h).  Its scratch memory - the
// float64 fallback list and its counter, which tsdx_hip_posv_ir_f64 takes as a
// caller workspace - is a fixed file-scope device array of this library, so
// the call makes no allocation of any kind and can run where allocation is not
// allowed (HIP graph capture, serving loops).  Every GPU holds its own copy of
// the array; its address on a device is looked up once, with that device
// current, and kept.  All calls on one device share the array: calls that
// overlap on one GPU race.  The counter is cleared by a kernel at the start of
// every call (no memset node), so the contents left by an earlier call do not
// matter.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "capi_internal.hpp"

namespace tsdx_hip {
namespace {

// The workspace of tsdx_hip_posv_ir_f64 for the largest batch: 256 bytes of
// header (the counter) and one int per matrix, rounded up to 256 bytes.
constexpr int64_t kNoAllocMaxBatch = TSDX_HIP_NOALLOC_MAX_BATCH;
constexpr size_t kScratchBytes = 256 + ((size_t)kNoAllocMaxBatch * sizeof(int) + 255) / 256 * 256;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::once_flag g_once[kMaxDevices];
void* g_ptr[kMaxDevices] = {};
hipError_t g_rc[kMaxDevices] = {};

// Device address of this device's scratch array (looked up on first use).
void* scratch_on_current_device() {
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) throw RuntimeErr("device index beyond the library's table");
  std::call_once(g_once[dev], [dev] {
    void* p = nullptr;
    g_rc[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_ptr[dev] = p;
  });
  hip_check(g_rc[dev], "hipGetSymbolAddress(scratch)");
  if (!g_ptr[dev]) throw RuntimeErr("scratch array has no device address");
  return g_ptr[dev];
}

}  // namespace
}  // namespace tsdx_hip

using namespace tsdx_hip;

extern "C" {

int tsdx_hip_posv_ir_noalloc_f64(const double* A, int n, int64_t batch, const double* B, double* X, int nrhs,
                                 int factor, int max_iter, double tol, int* info, int* stat, void* stream) {
  return guarded([&] {
    if (batch > kNoAllocMaxBatch)
      throw InvalidArg("batch " + std::to_string(batch) + " above the allocation-free limit of " +
                       std::to_string(kNoAllocMaxBatch) + " matrices (use tsdx_hip_posv_ir_f64 with a workspace)");
    check_dims(n, batch);
    if (batch == 0) return;
    require_device();
    void* work = scratch_on_current_device();
    posv_ir(A, n, batch, B, X, nrhs, factor, max_iter, tol, info, stat, work, kScratchBytes,
            static_cast<hipStream_t>(stream));
  });
}

}  // extern "C"
