// C ABI device-memory helpers (declared in include/tsdx_hip.h): allocation,
// copies and synchronisation for callers that hold device arrays but no GPU
// framework of their own (the Python module's device-array path).
#include <hip/hip_runtime.h>

#include "../../include/tsdx_hip.h"
#include "capi_internal.hpp"

using namespace tsdx_hip;

extern "C" {

int tsdx_hip_malloc(void** ptr, size_t bytes) {
  return guarded([&] {
    if (!ptr) throw InvalidArg("ptr must not be null");
    *ptr = nullptr;
    require_device();
    hip_check(hipMalloc(ptr, bytes ? bytes : 16), "hipMalloc");
  });
}

int tsdx_hip_free(void* ptr) {
  return guarded([&] {
    if (ptr) hip_check(hipFree(ptr), "hipFree");
  });
}

int tsdx_hip_memcpy(void* dst, const void* src, size_t bytes) {
  return guarded([&] {
    if (bytes == 0) return;
    if (!dst || !src) throw InvalidArg("dst and src must not be null");
    require_device();
    hip_check(hipMemcpy(dst, src, bytes, hipMemcpyDefault), "hipMemcpy");
  });
}

int tsdx_hip_device_synchronize(void) {
  return guarded([&] {
    require_device();
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  });
}

int tsdx_hip_get_device(int* device) {
  return guarded([&] {
    if (!device) throw InvalidArg("device must not be null");
    require_device();
    hip_check(hipGetDevice(device), "hipGetDevice");
  });
}

int tsdx_hip_pointer_device(const void* ptr, int* device) {
  return guarded([&] {
    if (!ptr || !device) throw InvalidArg("ptr and device must not be null");
    require_device();
    hipPointerAttribute_t attr{};
    if (hipPointerGetAttributes(&attr, ptr) != hipSuccess) {
      (void)hipGetLastError();
      throw InvalidArg("pointer is not device memory known to HIP");
    }
    if (attr.type != hipMemoryTypeDevice && attr.type != hipMemoryTypeManaged)
      throw InvalidArg("pointer is not device memory");
    *device = attr.device;
  });
}

}  // extern "C"
