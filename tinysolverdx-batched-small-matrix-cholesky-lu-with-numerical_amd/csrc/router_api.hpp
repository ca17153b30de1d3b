// Host-side declarations of the router launchers (router.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "mixed_api.hpp"

namespace tsdx_hip {

hipError_t launch_router_features(const double* A, int n, int64_t batch, float* feat, hipStream_t s);
hipError_t launch_router_forward(const float* P, const float* feat, int64_t count, float* logits, hipStream_t s);
size_t router_grad_workspace(int64_t count);
hipError_t launch_router_grad(const float* P, const float* feat, const int* label, int64_t count, float* grad,
                              float* loss, void* work, hipStream_t s);
hipError_t launch_router_adam(float* P, const float* grad, float* m, float* v, int step, float lr, hipStream_t s);
// Routed posv: features, network, per-tier refinement launches over the tier
// lists, then the float64 solve of the tier-2 list and of every fallback.
hipError_t launch_posv_routed(const IrArgs& a, const float* P, float* feat, int* tier, int* list0, int* list1,
                              int* counts, hipStream_t s);

}  // namespace tsdx_hip
