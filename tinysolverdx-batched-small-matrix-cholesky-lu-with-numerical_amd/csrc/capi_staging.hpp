// Host staging of the synchronous host-pointer calls: grow-only device
// buffers and one library stream per device.  Calls are serialised per
// process by the pool's mutex.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "capi_internal.hpp"

namespace tsdx_hip {

struct Staging {
  std::mutex mu;
  struct PerDevice {
    hipStream_t stream = nullptr;
    void* buf[4] = {nullptr, nullptr, nullptr, nullptr};
    size_t cap[4] = {0, 0, 0, 0};
  };
  std::vector<PerDevice> dev;
};
Staging& staging();  // the process-wide pool (capi.cpp)

struct HostCall {
  std::lock_guard<std::mutex> lock;
  Staging::PerDevice* pd;
  explicit HostCall(Staging& s) : lock(s.mu) {
    require_device();
    int d = 0;
    hip_check(hipGetDevice(&d), "hipGetDevice");
    if ((int)s.dev.size() <= d) s.dev.resize(d + 1);
    pd = &s.dev[d];
    if (!pd->stream) hip_check(hipStreamCreateWithFlags(&pd->stream, hipStreamNonBlocking), "hipStreamCreate");
  }
  void* buf(int slot, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (pd->cap[slot] < bytes) {
      if (pd->buf[slot]) hip_check(hipFree(pd->buf[slot]), "hipFree");
      pd->buf[slot] = nullptr;
      pd->cap[slot] = 0;
      hip_check(hipMalloc(&pd->buf[slot], bytes), "hipMalloc");
      pd->cap[slot] = bytes;
    }
    return pd->buf[slot];
  }
  void h2d(void* d, const void* h, size_t bytes) {
    if (bytes) hip_check(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, pd->stream), "H2D");
  }
  void d2h(void* h, const void* d, size_t bytes) {
    if (bytes) hip_check(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, pd->stream), "D2H");
  }
  void sync() { hip_check(hipStreamSynchronize(pd->stream), "hipStreamSynchronize"); }
};


}  // namespace tsdx_hip
