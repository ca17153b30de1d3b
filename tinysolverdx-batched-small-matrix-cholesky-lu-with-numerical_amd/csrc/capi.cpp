// C ABI of libtsdx_hip.so (declared in include/tsdx_hip.h): argument checks
// mirroring the reference (batched.hpp:39-43, :184), dispatch of n to a size
// bucket, stream-ordered device-pointer calls, and the synchronous host-pointer
// calls that stage through a per-device buffer pool.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/tsdx_hip.h"
#include "capi_internal.hpp"
#include "capi_staging.hpp"
#include "launch.hpp"
#include "mixed_api.hpp"

namespace tsdx_hip {

namespace {
thread_local std::string g_err;
}

void set_error(const std::string& s) { g_err = s; }
const char* last_error() { return g_err.c_str(); }

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw RuntimeErr(std::string(what) + ": " + hipGetErrorString(e));
}

void check_dims(int n, int64_t batch) {
  if (n < 1 || n > 64)
    throw InvalidArg("matrix size n=" + std::to_string(n) + " outside supported range [1, 64]");
  if (batch < 0) throw InvalidArg("negative batch");
}

void require_device() {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess || c <= 0) {
    (void)hipGetLastError();
    throw RuntimeErr("HIP backend unavailable: no HIP device visible");
  }
}

template <class T>
int vec_ok(const void* p, int n) {
  return (reinterpret_cast<uintptr_t>(p) % 16 == 0) && ((size_t)n * sizeof(T)) % 16 == 0;
}

// ------------------------------------------------------------ device calls
template <class T>
void potrf(T* A, int n, int64_t batch, int* info, hipStream_t s) {
  check_dims(n, batch);
  if (batch == 0) return;
  if (!A) throw InvalidArg("A must not be null");
  require_device();
  const int vec = vec_ok<T>(A, n);
  hip_check(with_bucket(n, [&](auto nb) { return launch_potrf<T, decltype(nb)::value>(A, n, batch, info, vec, s); }),
            "potrf launch");
}

template <class T>
void getrf(T* A, int n, int64_t batch, int* piv, int* info, hipStream_t s) {
  check_dims(n, batch);
  if (!piv) throw InvalidArg("piv must not be null");
  if (batch == 0) return;
  if (!A) throw InvalidArg("A must not be null");
  require_device();
  const int vec = vec_ok<T>(A, n);
  hip_check(with_bucket(n,
                        [&](auto nb) {
                          return launch_getrf<T, decltype(nb)::value>(A, n, batch, piv, info, vec, n <= kLuTileMaxN, s);
                        }),
            "getrf launch");
}

template <class T>
void potrs(const T* L, int n, int64_t batch, T* B, int nrhs, hipStream_t s) {
  check_dims(n, batch);
  if (nrhs < 0) throw InvalidArg("negative nrhs");
  if (batch == 0 || nrhs == 0) return;
  if (!L || !B) throw InvalidArg("L and B must not be null");
  require_device();
  const int vec = vec_ok<T>(L, n);
  hip_check(with_bucket(n, [&](auto nb) { return launch_potrs<T, decltype(nb)::value>(L, n, batch, B, nrhs, vec, s); }),
            "potrs launch");
}

template <class T>
void getrs(const T* LU, const int* piv, int n, int64_t batch, T* B, int nrhs, hipStream_t s) {
  check_dims(n, batch);
  if (nrhs < 0) throw InvalidArg("negative nrhs");
  if (batch == 0 || nrhs == 0) return;
  if (!LU || !piv || !B) throw InvalidArg("LU, piv and B must not be null");
  require_device();
  const int vec = vec_ok<T>(LU, n);
  hip_check(with_bucket(n,
                        [&](auto nb) {
                          return launch_getrs<T, decltype(nb)::value>(LU, piv, n, batch, B, nrhs, vec, s);
                        }),
            "getrs launch");
}

template <class T>
void posv(const T* A, T* Lout, int n, int64_t batch, const T* B, T* X, int nrhs, int* info, hipStream_t s) {
  check_dims(n, batch);
  if (nrhs < 0) throw InvalidArg("negative nrhs");
  if (batch == 0) return;
  if (!A || (nrhs > 0 && (!B || !X))) throw InvalidArg("A, B and X must not be null");
  require_device();
  const int vec = vec_ok<T>(A, n) && (!Lout || vec_ok<T>(Lout, n));
  hip_check(with_bucket(n,
                        [&](auto nb) {
                          return launch_posv<T, decltype(nb)::value>(A, Lout, n, batch, B, X, nrhs, info, vec, s);
                        }),
            "posv launch");
}

template <class T>
void gesv(const T* A, T* LUout, int n, int64_t batch, const T* B, T* X, int nrhs, int* piv, int* info, bool need_piv,
          hipStream_t s) {
  check_dims(n, batch);
  if (need_piv && !piv) throw InvalidArg("piv must not be null");
  if (nrhs < 0) throw InvalidArg("negative nrhs");
  if (batch == 0) return;
  if (!A || (nrhs > 0 && (!B || !X))) throw InvalidArg("A, B and X must not be null");
  require_device();
  const int vec = vec_ok<T>(A, n) && (!LUout || vec_ok<T>(LUout, n));
  hip_check(with_bucket(n,
                        [&](auto nb) {
                          return launch_gesv<T, decltype(nb)::value>(A, LUout, n, batch, B, X, nrhs, piv, info, vec,
                                                                     n <= kLuTileMaxN, s);
                        }),
            "gesv launch");
}

// ------------------------------------------------------ refinement tier
size_t ir_workspace_size(int64_t batch) {
  const size_t b = batch < 0 ? 0 : (size_t)batch;
  return 256 + ((b * sizeof(int) + 255) / 256) * 256;
}

void posv_ir(const double* A, int n, int64_t batch, const double* B, double* X, int nrhs, int factor, int max_iter,
             double tol, int* info, int* stat, void* work, size_t work_bytes, hipStream_t s) {
  check_dims(n, batch);
  if (nrhs < 0) throw InvalidArg("negative nrhs");
  if (factor < TSDX_HIP_FACTOR_AUTO || factor > TSDX_HIP_FACTOR_FP64_BLOCK) throw InvalidArg("unknown factor precision");
  if (max_iter < 0 || max_iter > 1000) throw InvalidArg("max_iter must be in [0, 1000]");
  if (!(tol > 0.0)) throw InvalidArg("tol must be positive");
  // one workgroup per matrix and int entries in the fallback list
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
  const int low = factor == TSDX_HIP_FACTOR_AUTO ? TSDX_HIP_FACTOR_FP32 : factor;
  hip_check(launch_posv_ir(a, low, s), "posv_ir launch");
}

Staging& staging() {
  static Staging s;
  return s;
}

template <class T>
void cholesky_host(T* A, int n, int64_t batch, int* info) {
  check_dims(n, batch);
  if (batch == 0) return;
  HostCall c(staging());
  const size_t ab = sizeof(T) * (size_t)n * n * batch, ib = sizeof(int) * (size_t)batch;
  T* dA = (T*)c.buf(0, ab);
  int* dI = (int*)c.buf(1, ib);
  c.h2d(dA, A, ab);
  potrf<T>(dA, n, batch, dI, c.pd->stream);
  c.d2h(A, dA, ab);
  if (info) c.d2h(info, dI, ib);
  c.sync();
}

template <class T>
void lu_host(T* A, int n, int64_t batch, int* piv, int* info) {
  check_dims(n, batch);
  if (!piv) throw InvalidArg("piv must not be null");
  if (batch == 0) return;
  HostCall c(staging());
  const size_t ab = sizeof(T) * (size_t)n * n * batch, pb = sizeof(int) * (size_t)n * batch,
               ib = sizeof(int) * (size_t)batch;
  T* dA = (T*)c.buf(0, ab);
  int* dP = (int*)c.buf(1, pb);
  int* dI = (int*)c.buf(2, ib);
  c.h2d(dA, A, ab);
  getrf<T>(dA, n, batch, dP, dI, c.pd->stream);
  c.d2h(A, dA, ab);
  c.d2h(piv, dP, pb);
  if (info) c.d2h(info, dI, ib);
  c.sync();
}

template <class T>
void chol_solve_host(const T* L, int n, int64_t batch, T* B, int nrhs) {
  check_dims(n, batch);
  if (nrhs < 0) throw InvalidArg("negative nrhs");
  if (batch == 0 || nrhs == 0) return;
  HostCall c(staging());
  const size_t lb = sizeof(T) * (size_t)n * n * batch, bb = sizeof(T) * (size_t)n * nrhs * batch;
  T* dL = (T*)c.buf(0, lb);
  T* dB = (T*)c.buf(1, bb);
  c.h2d(dL, L, lb);
  c.h2d(dB, B, bb);
  potrs<T>(dL, n, batch, dB, nrhs, c.pd->stream);
  c.d2h(B, dB, bb);
  c.sync();
}

template <class T>
void lu_solve_host(const T* LU, const int* piv, int n, int64_t batch, T* B, int nrhs) {
  check_dims(n, batch);
  if (nrhs < 0) throw InvalidArg("negative nrhs");
  if (batch == 0 || nrhs == 0) return;
  if (!piv) throw InvalidArg("piv must not be null");
  HostCall c(staging());
  const size_t lb = sizeof(T) * (size_t)n * n * batch, bb = sizeof(T) * (size_t)n * nrhs * batch,
               pb = sizeof(int) * (size_t)n * batch;
  T* dL = (T*)c.buf(0, lb);
  T* dB = (T*)c.buf(1, bb);
  int* dP = (int*)c.buf(2, pb);
  c.h2d(dL, LU, lb);
  c.h2d(dB, B, bb);
  c.h2d(dP, piv, pb);
  getrs<T>(dL, dP, n, batch, dB, nrhs, c.pd->stream);
  c.d2h(B, dB, bb);
  c.sync();
}

template <class T>
void posv_host(T* A, int n, int64_t batch, T* B, int nrhs, int* info) {
  check_dims(n, batch);
  if (nrhs < 0) throw InvalidArg("negative nrhs");
  if (batch == 0) return;
  HostCall c(staging());
  const size_t ab = sizeof(T) * (size_t)n * n * batch, bb = sizeof(T) * (size_t)n * nrhs * batch,
               ib = sizeof(int) * (size_t)batch;
  T* dA = (T*)c.buf(0, ab);
  T* dB = (T*)c.buf(1, bb);
  int* dI = (int*)c.buf(2, ib);
  c.h2d(dA, A, ab);
  c.h2d(dB, B, bb);
  posv<T>(dA, dA, n, batch, dB, dB, nrhs, dI, c.pd->stream);
  c.d2h(A, dA, ab);
  c.d2h(B, dB, bb);
  if (info) c.d2h(info, dI, ib);
  c.sync();
}

template <class T>
void gesv_host(T* A, int n, int64_t batch, T* B, int nrhs, int* piv, int* info) {
  check_dims(n, batch);
  if (!piv) throw InvalidArg("piv must not be null");
  if (nrhs < 0) throw InvalidArg("negative nrhs");
  if (batch == 0) return;
  HostCall c(staging());
  const size_t ab = sizeof(T) * (size_t)n * n * batch, bb = sizeof(T) * (size_t)n * nrhs * batch,
               pb = sizeof(int) * (size_t)n * batch, ib = sizeof(int) * (size_t)batch;
  T* dA = (T*)c.buf(0, ab);
  T* dB = (T*)c.buf(1, bb);
  int* dP = (int*)c.buf(2, pb);
  int* dI = (int*)c.buf(3, ib);
  c.h2d(dA, A, ab);
  c.h2d(dB, B, bb);
  gesv<T>(dA, dA, n, batch, dB, dB, nrhs, dP, dI, true, c.pd->stream);
  c.d2h(A, dA, ab);
  c.d2h(B, dB, bb);
  c.d2h(piv, dP, pb);
  if (info) c.d2h(info, dI, ib);
  c.sync();
}

}  // namespace tsdx_hip

using namespace tsdx_hip;

#define S(x) static_cast<hipStream_t>(x)

extern "C" {

int tsdx_hip_available(void) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return c > 0 ? 1 : 0;
}
const char* tsdx_hip_last_error(void) { return last_error(); }
int tsdx_hip_version(void) { return TSDX_HIP_VERSION; }

int tsdx_hip_potrf_f32(float* A, int n, int64_t batch, int* info, void* s) {
  return guarded([&] { potrf<float>(A, n, batch, info, S(s)); });
}
int tsdx_hip_potrf_f64(double* A, int n, int64_t batch, int* info, void* s) {
  return guarded([&] { potrf<double>(A, n, batch, info, S(s)); });
}
int tsdx_hip_getrf_f32(float* A, int n, int64_t batch, int* piv, int* info, void* s) {
  return guarded([&] { getrf<float>(A, n, batch, piv, info, S(s)); });
}
int tsdx_hip_getrf_f64(double* A, int n, int64_t batch, int* piv, int* info, void* s) {
  return guarded([&] { getrf<double>(A, n, batch, piv, info, S(s)); });
}
int tsdx_hip_potrs_f32(const float* L, int n, int64_t batch, float* B, int nrhs, void* s) {
  return guarded([&] { potrs<float>(L, n, batch, B, nrhs, S(s)); });
}
int tsdx_hip_potrs_f64(const double* L, int n, int64_t batch, double* B, int nrhs, void* s) {
  return guarded([&] { potrs<double>(L, n, batch, B, nrhs, S(s)); });
}
int tsdx_hip_getrs_f32(const float* LU, const int* piv, int n, int64_t batch, float* B, int nrhs, void* s) {
  return guarded([&] { getrs<float>(LU, piv, n, batch, B, nrhs, S(s)); });
}
int tsdx_hip_getrs_f64(const double* LU, const int* piv, int n, int64_t batch, double* B, int nrhs, void* s) {
  return guarded([&] { getrs<double>(LU, piv, n, batch, B, nrhs, S(s)); });
}
int tsdx_hip_posv_f32(float* A, int n, int64_t batch, float* B, int nrhs, int* info, void* s) {
  return guarded([&] { posv<float>(A, A, n, batch, B, B, nrhs, info, S(s)); });
}
int tsdx_hip_posv_f64(double* A, int n, int64_t batch, double* B, int nrhs, int* info, void* s) {
  return guarded([&] { posv<double>(A, A, n, batch, B, B, nrhs, info, S(s)); });
}
int tsdx_hip_gesv_f32(float* A, int n, int64_t batch, float* B, int nrhs, int* piv, int* info, void* s) {
  return guarded([&] { gesv<float>(A, A, n, batch, B, B, nrhs, piv, info, true, S(s)); });
}
int tsdx_hip_gesv_f64(double* A, int n, int64_t batch, double* B, int nrhs, int* piv, int* info, void* s) {
  return guarded([&] { gesv<double>(A, A, n, batch, B, B, nrhs, piv, info, true, S(s)); });
}
int tsdx_hip_posv_solve_f32(const float* A, int n, int64_t batch, const float* B, float* X, int nrhs, int* info,
                            void* s) {
  return guarded([&] { posv<float>(A, nullptr, n, batch, B, X, nrhs, info, S(s)); });
}
int tsdx_hip_posv_solve_f64(const double* A, int n, int64_t batch, const double* B, double* X, int nrhs, int* info,
                            void* s) {
  return guarded([&] { posv<double>(A, nullptr, n, batch, B, X, nrhs, info, S(s)); });
}
int tsdx_hip_gesv_solve_f32(const float* A, int n, int64_t batch, const float* B, float* X, int nrhs, int* piv,
                            int* info, void* s) {
  return guarded([&] { gesv<float>(A, nullptr, n, batch, B, X, nrhs, piv, info, false, S(s)); });
}
int tsdx_hip_gesv_solve_f64(const double* A, int n, int64_t batch, const double* B, double* X, int nrhs, int* piv,
                            int* info, void* s) {
  return guarded([&] { gesv<double>(A, nullptr, n, batch, B, X, nrhs, piv, info, false, S(s)); });
}

size_t tsdx_hip_posv_ir_workspace_size(int64_t batch) { return ir_workspace_size(batch); }
int tsdx_hip_posv_ir_f64(const double* A, int n, int64_t batch, const double* B, double* X, int nrhs, int factor,
                         int max_iter, double tol, int* info, int* stat, void* work, size_t work_bytes, void* s) {
  return guarded(
      [&] { posv_ir(A, n, batch, B, X, nrhs, factor, max_iter, tol, info, stat, work, work_bytes, S(s)); });
}

int tsdx_hip_cholesky_batched_f32(float* A, int n, int64_t batch, int* info) {
  return guarded([&] { cholesky_host<float>(A, n, batch, info); });
}
int tsdx_hip_cholesky_batched_f64(double* A, int n, int64_t batch, int* info) {
  return guarded([&] { cholesky_host<double>(A, n, batch, info); });
}
int tsdx_hip_lu_batched_f32(float* A, int n, int64_t batch, int* piv, int* info) {
  return guarded([&] { lu_host<float>(A, n, batch, piv, info); });
}
int tsdx_hip_lu_batched_f64(double* A, int n, int64_t batch, int* piv, int* info) {
  return guarded([&] { lu_host<double>(A, n, batch, piv, info); });
}
int tsdx_hip_cholesky_solve_batched_f32(const float* L, int n, int64_t batch, float* B, int nrhs) {
  return guarded([&] { chol_solve_host<float>(L, n, batch, B, nrhs); });
}
int tsdx_hip_cholesky_solve_batched_f64(const double* L, int n, int64_t batch, double* B, int nrhs) {
  return guarded([&] { chol_solve_host<double>(L, n, batch, B, nrhs); });
}
int tsdx_hip_lu_solve_batched_f32(const float* LU, const int* piv, int n, int64_t batch, float* B, int nrhs) {
  return guarded([&] { lu_solve_host<float>(LU, piv, n, batch, B, nrhs); });
}
int tsdx_hip_lu_solve_batched_f64(const double* LU, const int* piv, int n, int64_t batch, double* B, int nrhs) {
  return guarded([&] { lu_solve_host<double>(LU, piv, n, batch, B, nrhs); });
}
int tsdx_hip_posv_batched_f32(float* A, int n, int64_t batch, float* B, int nrhs, int* info) {
  return guarded([&] { posv_host<float>(A, n, batch, B, nrhs, info); });
}
int tsdx_hip_posv_batched_f64(double* A, int n, int64_t batch, double* B, int nrhs, int* info) {
  return guarded([&] { posv_host<double>(A, n, batch, B, nrhs, info); });
}
int tsdx_hip_gesv_batched_f32(float* A, int n, int64_t batch, float* B, int nrhs, int* piv, int* info) {
  return guarded([&] { gesv_host<float>(A, n, batch, B, nrhs, piv, info); });
}
int tsdx_hip_gesv_batched_f64(double* A, int n, int64_t batch, double* B, int nrhs, int* piv, int* info) {
  return guarded([&] { gesv_host<double>(A, n, batch, B, nrhs, piv, info); });
}

}  // extern "C"
