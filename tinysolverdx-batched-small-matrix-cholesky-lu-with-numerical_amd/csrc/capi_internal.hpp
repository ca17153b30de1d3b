// Shared host-side helpers of the C ABI units: error plumbing and constants.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <type_traits>

#include "../../include/tsdx_hip.h"

#define TSDX_HIP_VERSION 100  // 0.1.0

namespace tsdx_hip {

// Largest n for which the reference's Auto LU uses the SIMD tile kernel
// (kLuSimdMaxN, batched.hpp:33); that kernel divides by 1 on a zero pivot.
constexpr int kLuTileMaxN = 24;

template <int I>
using ic_t = std::integral_constant<int, I>;

struct InvalidArg : std::invalid_argument {
  using std::invalid_argument::invalid_argument;
};
struct RuntimeErr : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void set_error(const std::string& s);
const char* last_error();
void hip_check(hipError_t e, const char* what);
void check_dims(int n, int64_t batch);
void require_device();

// Refinement-tier posv over a caller's workspace (capi.cpp).
size_t ir_workspace_size(int64_t batch);
void posv_ir(const double* A, int n, int64_t batch, const double* B, double* X, int nrhs, int factor, int max_iter,
             double tol, int* info, int* stat, void* work, size_t work_bytes, hipStream_t s);

// Calls f with the size bucket of n (4, 8, 16, 32 or 64) as an ic_t.
template <class F>
hipError_t with_bucket(int n, F&& f) {
  if (n <= 4) return f(ic_t<4>{});
  if (n <= 8) return f(ic_t<8>{});
  if (n <= 16) return f(ic_t<16>{});
  if (n <= 32) return f(ic_t<32>{});
  return f(ic_t<64>{});
}

// Runs f and maps exceptions to the ABI's return codes.
template <class F>
int guarded(F&& f) {
  try {
    f();
    set_error("");
    return TSDX_HIP_OK;
  } catch (const std::invalid_argument& e) {
    set_error(e.what());
    return TSDX_HIP_EINVAL;
  } catch (const std::exception& e) {
    set_error(e.what());
    return TSDX_HIP_ERUNTIME;
  } catch (...) {
    set_error("unknown error");
    return TSDX_HIP_ERUNTIME;
  }
}

}  // namespace tsdx_hip
