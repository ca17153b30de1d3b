// C ABI of the learned precision router and the routed posv
// (include/tsdx_hip.h, tsdx_hip_router_* / tsdx_hip_posv_routed_f64).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/tsdx_hip.h"
#include "capi_internal.hpp"
#include "router_api.hpp"

using namespace tsdx_hip;

namespace {

hipStream_t S(void* s) { return static_cast<hipStream_t>(s); }

void check_count(int64_t count) {
  if (count < 0) throw InvalidArg("negative count");
  if (count > 0x7fffffff) throw InvalidArg("count above 2^31-1");
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// routed workspace: [float64 count, bfloat16 count (256 B)] [float64 list]
// [bfloat16 list] [prepared network]
size_t routed_list_bytes(int64_t batch) { return align256((batch < 0 ? 0 : (size_t)batch) * sizeof(int)); }
size_t routed_workspace(int64_t batch) { return 256 + 2 * routed_list_bytes(batch) + align256(kRouterPrepBytes); }

}  // namespace

extern "C" {

int tsdx_hip_router_features_f64(const double* A, int n, int64_t batch, float* feat, void* stream) {
  return guarded([&] {
    check_dims(n, batch);
    check_count(batch);
    if (batch == 0) return;
    if (!A || !feat) throw InvalidArg("A and feat must not be null");
    require_device();
    hip_check(launch_router_features(A, n, batch, feat, S(stream)), "router features launch");
  });
}

int tsdx_hip_router_forward(const float* params, const float* feat, int64_t count, float* logits, void* stream) {
  return guarded([&] {
    check_count(count);
    if (count == 0) return;
    if (!params || !feat || !logits) throw InvalidArg("params, feat and logits must not be null");
    require_device();
    hip_check(launch_router_forward(params, feat, count, logits, S(stream)), "router forward launch");
  });
}

size_t tsdx_hip_router_infer_workspace_size(void) { return align256(kRouterPrepBytes); }

int tsdx_hip_router_infer(const float* params, const float* feat, int64_t count, int* tier, float* logits, void* work,
                          size_t work_bytes, void* stream) {
  return guarded([&] {
    check_count(count);
    if (count == 0) return;
    if (!params || !feat || (!tier && !logits)) throw InvalidArg("params, feat and tier or logits must not be null");
    if (!work || work_bytes < kRouterPrepBytes) throw InvalidArg("workspace missing or too small");
    if (reinterpret_cast<uintptr_t>(work) % 16 != 0) throw InvalidArg("workspace must be 16-byte aligned");
    require_device();
    hip_check(launch_router_infer(params, feat, count, tier, logits, work, S(stream)), "router inference launch");
  });
}

size_t tsdx_hip_router_grad_workspace_size(int64_t count) { return count <= 0 ? 256 : router_grad_workspace(count); }

int tsdx_hip_router_grad(const float* params, const float* feat, const int* label, int64_t count, float* grad,
                         float* loss, void* work, size_t work_bytes, void* stream) {
  return guarded([&] {
    check_count(count);
    if (count == 0) throw InvalidArg("gradient over an empty batch");
    if (!params || !feat || !label || !grad) throw InvalidArg("params, feat, label and grad must not be null");
    if (!work || work_bytes < router_grad_workspace(count)) throw InvalidArg("workspace missing or too small");
    require_device();
    hip_check(launch_router_grad(params, feat, label, count, grad, loss, work, S(stream)), "router gradient launch");
  });
}

int tsdx_hip_router_adam(float* params, const float* grad, float* m, float* v, int step, float lr, void* stream) {
  return guarded([&] {
    if (!params || !grad || !m || !v) throw InvalidArg("params, grad, m and v must not be null");
    if (step < 1) throw InvalidArg("Adam step counts from 1");
    require_device();
    hip_check(launch_router_adam(params, grad, m, v, step, lr, S(stream)), "router Adam launch");
  });
}

size_t tsdx_hip_posv_routed_workspace_size(int64_t batch) { return routed_workspace(batch); }

int tsdx_hip_posv_routed_f64(const double* A, int n, int64_t batch, const double* B, double* X, int nrhs,
                             const float* params, int max_iter, double tol, int* info, int* stat, int* tier,
                             float* feat, void* work, size_t work_bytes, void* stream) {
  return guarded([&] {
    check_dims(n, batch);
    check_count(batch);
    if (nrhs < 0) throw InvalidArg("negative nrhs");
    if (max_iter < 0 || max_iter > 1000) throw InvalidArg("max_iter must be in [0, 1000]");
    if (!(tol > 0.0)) throw InvalidArg("tol must be positive");
    if (batch == 0) return;
    if (!A || !params || (nrhs > 0 && (!B || !X))) throw InvalidArg("A, params, B and X must not be null");
    if (!work || work_bytes < routed_workspace(batch)) throw InvalidArg("workspace missing or too small");
    require_device();
    char* w = static_cast<char*>(work);
    int* fcount = reinterpret_cast<int*>(w);
    int* flist = reinterpret_cast<int*>(w + 256);
    IrArgs a;
    a.A = A;
    a.B = B;
    a.X = X;
    a.n = n;
    a.batch = batch;
    a.nrhs = nrhs;
    a.max_iter = max_iter;
    a.tol = tol;
    a.info = info;
    a.stat = stat;
    a.flist = flist;
    a.fcount = fcount;
    a.lcount = fcount + 1;
    a.llist = reinterpret_cast<int*>(w + 256 + routed_list_bytes(batch));
    if (reinterpret_cast<uintptr_t>(work) % 16 != 0) throw InvalidArg("workspace must be 16-byte aligned");
    void* prep = w + 256 + 2 * routed_list_bytes(batch);
    hip_check(launch_posv_routed(a, params, prep, tier, feat, S(stream)), "routed posv launch");
  });
}

}  // extern "C"
