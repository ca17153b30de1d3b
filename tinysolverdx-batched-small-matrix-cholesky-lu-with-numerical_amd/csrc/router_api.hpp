// Host-side declarations of the router launchers (router.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "mixed_api.hpp"
#include "router_layout.hpp"

namespace tsdx_hip {

hipError_t launch_router_features(const double* A, int n, int64_t batch, float* feat, hipStream_t s);
hipError_t launch_router_forward(const float* P, const float* feat, int64_t count, float* logits, hipStream_t s);
size_t router_grad_workspace(int64_t count);
hipError_t launch_router_grad(const float* P, const float* feat, const int* label, int64_t count, float* grad,
                              float* loss, void* work, hipStream_t s);
hipError_t launch_router_adam(float* P, const float* grad, float* m, float* v, int step, float lr, hipStream_t s);
// Routed posv: one kernel reads each matrix, computes its features, runs the
// network and solves in the chosen tier; then the float64 solve of tier-2
// matrices and of every fallback.  tier / feat (optional) receive the choice
// and the features.
// prep: kRouterPrepBytes of device memory (256-byte aligned) for the network
// prepared for the matrix cores (router_layout.hpp).
hipError_t launch_posv_routed(const IrArgs& a, const float* P, void* prep, int* tier, float* feat, hipStream_t s);
// Tiers (and, if logits is not null, logits) of `count` feature vectors from
// the matrix-core network the routed posv runs.
hipError_t launch_router_infer(const float* P, const float* feat, int64_t count, int* tier, float* logits, void* prep,
                               hipStream_t s);

}  // namespace tsdx_hip
