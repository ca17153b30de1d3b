// Mixed-precision gesv with float64 refinement: the C ABI entry points
// tsdx_hip_gesv_ir_f64 / tsdx_hip_gesv_ir_workspace_size.  The kernels are in
// gesv_ir.hpp, their per-bucket launchers in the inst_gesvir_nb<NB>.hip units.
#include "capi_internal.hpp"
#include "launch.hpp"
#include "mixed_api.hpp"

namespace tsdx_hip {

namespace {

__global__ void k_gesv_ir_begin(int* count) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *count = 0;
}

__global__ void k_fill_int(int* p, int v, int64_t count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = v;
}

hipError_t launch_gesv_ir(const IrArgs& a, int low, hipStream_t s) {
  hipLaunchKernelGGL(k_gesv_ir_begin, dim3(1), dim3(64), 0, s, a.fcount);
  const bool zt = a.n <= kLuTileMaxN;
  if (a.n <= 4) return launch_gesv_ir_nb<4>(a, low, true, s);
  if (a.n <= 8) return launch_gesv_ir_nb<8>(a, low, true, s);
  if (a.n <= 16) return launch_gesv_ir_nb<16>(a, low, true, s);
  if (a.n <= 32) return launch_gesv_ir_nb<32>(a, low, zt, s);
  return launch_gesv_ir_nb<64>(a, low, false, s);
}

// Every matrix by the plain float64 gesv (factor TSDX_HIP_FACTOR_FP64).
hipError_t launch_all_f64(const IrArgs& a, hipStream_t s) {
  const int vec = (reinterpret_cast<uintptr_t>(a.A) % 16 == 0) && (a.n % 2 == 0);
  const bool zt = a.n <= kLuTileMaxN;
  hipError_t e;
  if (a.n <= 4)
    e = launch_gesv<double, 4>(a.A, nullptr, a.n, a.batch, a.B, a.X, a.nrhs, nullptr, a.info, vec, zt, s);
  else if (a.n <= 8)
    e = launch_gesv<double, 8>(a.A, nullptr, a.n, a.batch, a.B, a.X, a.nrhs, nullptr, a.info, vec, zt, s);
  else if (a.n <= 16)
    e = launch_gesv<double, 16>(a.A, nullptr, a.n, a.batch, a.B, a.X, a.nrhs, nullptr, a.info, vec, zt, s);
  else if (a.n <= 32)
    e = launch_gesv<double, 32>(a.A, nullptr, a.n, a.batch, a.B, a.X, a.nrhs, nullptr, a.info, vec, zt, s);
  else
    e = launch_gesv<double, 64>(a.A, nullptr, a.n, a.batch, a.B, a.X, a.nrhs, nullptr, a.info, vec, zt, s);
  if (e != hipSuccess || !a.stat) return e;
  hipLaunchKernelGGL(k_fill_int, dim3((unsigned)((a.batch + 255) / 256)), dim3(256), 0, s, a.stat, -1, a.batch);
  return hipGetLastError();
}

void gesv_ir(const double* A, int n, int64_t batch, const double* B, double* X, int nrhs, int factor, int max_iter,
             double tol, int* info, int* stat, void* work, size_t work_bytes, hipStream_t s) {
  check_dims(n, batch);
  if (nrhs < 0) throw InvalidArg("negative nrhs");
  if (factor < TSDX_HIP_FACTOR_AUTO || factor > TSDX_HIP_FACTOR_FP64)
    throw InvalidArg("unknown factor precision (gesv_ir takes AUTO, FP32, BF16, FP16 or FP64)");
  if (max_iter < 0 || max_iter > 1000) throw InvalidArg("max_iter must be in [0, 1000]");
  if (!(tol > 0.0)) throw InvalidArg("tol must be positive");
  if (batch > 0x7fffffff) throw InvalidArg("batch above 2^31-1 is not supported by the refinement tier");
  if (batch == 0) return;
  if (!A || (nrhs > 0 && (!B || !X))) throw InvalidArg("A, B and X must not be null");
  if (!work || work_bytes < ir_workspace_size(batch)) throw InvalidArg("workspace missing or too small");
  require_device();
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
  a.fcount = static_cast<int*>(work);
  a.flist = reinterpret_cast<int*>(static_cast<char*>(work) + 256);
  if (factor == TSDX_HIP_FACTOR_FP64)
    hip_check(launch_all_f64(a, s), "gesv_ir launch");
  else
    hip_check(launch_gesv_ir(a, factor == TSDX_HIP_FACTOR_AUTO ? TSDX_HIP_FACTOR_FP32 : factor, s), "gesv_ir launch");
}

}  // namespace

}  // namespace tsdx_hip

using namespace tsdx_hip;

extern "C" {

size_t tsdx_hip_gesv_ir_workspace_size(int64_t batch) { return ir_workspace_size(batch); }

int tsdx_hip_gesv_ir_f64(const double* A, int n, int64_t batch, const double* B, double* X, int nrhs, int factor,
                         int max_iter, double tol, int* info, int* stat, void* work, size_t work_bytes, void* s) {
  return guarded([&] {
    gesv_ir(A, n, batch, B, X, nrhs, factor, max_iter, tol, info, stat, work, work_bytes,
            static_cast<hipStream_t>(s));
  });
}

}  // extern "C"
