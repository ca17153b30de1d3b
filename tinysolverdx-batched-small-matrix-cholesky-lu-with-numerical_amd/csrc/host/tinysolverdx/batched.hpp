// Batched small-matrix solvers, host API (counterpart of the reference's
// include/tinysolverdx/batched.hpp:121-280).
//
// Same layout and semantics: A is (batch, n, n) row-major, matrices contiguous;
// Cholesky returns the lower factor with the upper triangle zeroed; LU returns
// packed L\U and sequential 0-based pivots piv (batch, n); info (batch) holds
// LAPACK codes and may be nullptr; B is (batch, n, nrhs) and is overwritten by
// the solution.  Errors: std::invalid_argument for n outside [1, 64], a
// negative batch or a null piv; std::runtime_error when the backend cannot run.
//
// Backend::Auto and Backend::Hip run on the GPU through the C ABI of
// libtsdx_hip.so (include/tsdx_hip.h).  The reference's CPU backends (Naive,
// Scalar, Simd) and Cuda are not part of this library and raise
// std::runtime_error: no call silently falls back to the CPU.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>

#include "precision.hpp"
#include "tsdx_hip.h"

namespace tsdx {

constexpr int kMinN = 1;
constexpr int kMaxN = 64;

namespace detail {

inline void check_dims(int n, std::int64_t batch) {
  if (n < kMinN || n > kMaxN)
    throw std::invalid_argument("matrix size n=" + std::to_string(n) + " outside supported range [1, 64]");
  if (batch < 0) throw std::invalid_argument("negative batch");
}

inline const char* backend_name(Backend b) {
  switch (b) {
    case Backend::Auto: return "auto";
    case Backend::Naive: return "naive";
    case Backend::Scalar: return "scalar";
    case Backend::Simd: return "simd";
    case Backend::Cuda: return "cuda";
    case Backend::Hip: return "hip";
  }
  return "?";
}

// Auto resolves to Hip; every other backend is rejected before any work.
inline void require_hip(Backend b) {
  if (b == Backend::Auto || b == Backend::Hip) return;
  if (b == Backend::Cuda) throw std::runtime_error("CUDA backend unavailable for this build/precision");
  throw std::runtime_error(std::string("backend '") + backend_name(b) +
                           "' is a CPU backend of the reference library and is not part of tinysolverdx-hip");
}

// Maps a C ABI status to the reference's exception types.
inline void check(int rc) {
  if (rc == TSDX_HIP_OK) return;
  const std::string msg = tsdx_hip_last_error();
  if (rc == TSDX_HIP_EINVAL) throw std::invalid_argument(msg);
  throw std::runtime_error(msg.empty() ? std::string("HIP backend failure") : msg);
}

inline int hip_chol(float* A, int n, std::int64_t b, int* i) { return tsdx_hip_cholesky_batched_f32(A, n, b, i); }
inline int hip_chol(double* A, int n, std::int64_t b, int* i) { return tsdx_hip_cholesky_batched_f64(A, n, b, i); }
inline int hip_lu(float* A, int n, std::int64_t b, int* p, int* i) { return tsdx_hip_lu_batched_f32(A, n, b, p, i); }
inline int hip_lu(double* A, int n, std::int64_t b, int* p, int* i) { return tsdx_hip_lu_batched_f64(A, n, b, p, i); }
inline int hip_chol_solve(const float* L, int n, std::int64_t b, float* B, int r) {
  return tsdx_hip_cholesky_solve_batched_f32(L, n, b, B, r);
}
inline int hip_chol_solve(const double* L, int n, std::int64_t b, double* B, int r) {
  return tsdx_hip_cholesky_solve_batched_f64(L, n, b, B, r);
}
inline int hip_lu_solve(const float* LU, const int* p, int n, std::int64_t b, float* B, int r) {
  return tsdx_hip_lu_solve_batched_f32(LU, p, n, b, B, r);
}
inline int hip_lu_solve(const double* LU, const int* p, int n, std::int64_t b, double* B, int r) {
  return tsdx_hip_lu_solve_batched_f64(LU, p, n, b, B, r);
}
inline int hip_posv(float* A, int n, std::int64_t b, float* B, int r, int* i) {
  return tsdx_hip_posv_batched_f32(A, n, b, B, r, i);
}
inline int hip_posv(double* A, int n, std::int64_t b, double* B, int r, int* i) {
  return tsdx_hip_posv_batched_f64(A, n, b, B, r, i);
}
inline int hip_gesv(float* A, int n, std::int64_t b, float* B, int r, int* p, int* i) {
  return tsdx_hip_gesv_batched_f32(A, n, b, B, r, p, i);
}
inline int hip_gesv(double* A, int n, std::int64_t b, double* B, int r, int* p, int* i) {
  return tsdx_hip_gesv_batched_f64(A, n, b, B, r, p, i);
}

// Emulated fp16 / bf16 (fmt: TSDX_HIP_EMU_FP16 / _BF16) on float32 or float64 storage.
inline int hip_chol_emu(float* A, int n, std::int64_t b, int* i, int f) {
  return tsdx_hip_cholesky_emu_batched_f32(A, n, b, i, f);
}
inline int hip_chol_emu(double* A, int n, std::int64_t b, int* i, int f) {
  return tsdx_hip_cholesky_emu_batched_f64(A, n, b, i, f);
}
inline int hip_lu_emu(float* A, int n, std::int64_t b, int* p, int* i, int f) {
  return tsdx_hip_lu_emu_batched_f32(A, n, b, p, i, f);
}
inline int hip_lu_emu(double* A, int n, std::int64_t b, int* p, int* i, int f) {
  return tsdx_hip_lu_emu_batched_f64(A, n, b, p, i, f);
}
inline int hip_chol_solve_emu(const float* L, int n, std::int64_t b, float* B, int r, int f) {
  return tsdx_hip_cholesky_solve_emu_batched_f32(L, n, b, B, r, f);
}
inline int hip_chol_solve_emu(const double* L, int n, std::int64_t b, double* B, int r, int f) {
  return tsdx_hip_cholesky_solve_emu_batched_f64(L, n, b, B, r, f);
}
inline int hip_lu_solve_emu(const float* LU, const int* p, int n, std::int64_t b, float* B, int r, int f) {
  return tsdx_hip_lu_solve_emu_batched_f32(LU, p, n, b, B, r, f);
}
inline int hip_lu_solve_emu(const double* LU, const int* p, int n, std::int64_t b, double* B, int r, int f) {
  return tsdx_hip_lu_solve_emu_batched_f64(LU, p, n, b, B, r, f);
}
inline int hip_posv_emu(float* A, int n, std::int64_t b, float* B, int r, int* i, int f) {
  return tsdx_hip_posv_emu_batched_f32(A, n, b, B, r, i, f);
}
inline int hip_posv_emu(double* A, int n, std::int64_t b, double* B, int r, int* i, int f) {
  return tsdx_hip_posv_emu_batched_f64(A, n, b, B, r, i, f);
}
inline int hip_gesv_emu(float* A, int n, std::int64_t b, float* B, int r, int* p, int* i, int f) {
  return tsdx_hip_gesv_emu_batched_f32(A, n, b, B, r, p, i, f);
}
inline int hip_gesv_emu(double* A, int n, std::int64_t b, double* B, int r, int* p, int* i, int f) {
  return tsdx_hip_gesv_emu_batched_f64(A, n, b, B, r, p, i, f);
}

template <class T>
void require_native() {
  static_assert(is_native<T>::value, "the HIP backend computes in float or double");
}

}  // namespace detail

// ------------------------------------------------------------------ factor
template <class T>
void cholesky_batched(T* A, int n, std::int64_t batch, int* info, Backend backend = Backend::Auto) {
  detail::require_native<T>();
  detail::check_dims(n, batch);
  detail::require_hip(backend);
  if (batch == 0) return;
  detail::check(detail::hip_chol(A, n, batch, info));
}

template <class T>
void lu_batched(T* A, int n, std::int64_t batch, int* piv, int* info, Backend backend = Backend::Auto) {
  detail::require_native<T>();
  detail::check_dims(n, batch);
  if (!piv) throw std::invalid_argument("piv must not be null");
  detail::require_hip(backend);
  if (batch == 0) return;
  detail::check(detail::hip_lu(A, n, batch, piv, info));
}

// ------------------------------------------------------------------ solves
template <class T>
void cholesky_solve_batched(const T* L, int n, std::int64_t batch, T* B, int nrhs, Backend backend = Backend::Auto) {
  detail::require_native<T>();
  detail::check_dims(n, batch);
  detail::require_hip(backend);
  if (batch == 0 || nrhs == 0) return;
  detail::check(detail::hip_chol_solve(L, n, batch, B, nrhs));
}

template <class T>
void lu_solve_batched(const T* LU, const int* piv, int n, std::int64_t batch, T* B, int nrhs,
                      Backend backend = Backend::Auto) {
  detail::require_native<T>();
  detail::check_dims(n, batch);
  detail::require_hip(backend);
  if (batch == 0 || nrhs == 0) return;
  detail::check(detail::hip_lu_solve(LU, piv, n, batch, B, nrhs));
}

// One-shot drivers: A is overwritten by its factor, B by the solution.  Unlike
// the reference (two passes, the solve on the CPU) both run in one launch.
template <class T>
void posv_batched(T* A, int n, std::int64_t batch, T* B, int nrhs, int* info, Backend be = Backend::Auto) {
  detail::require_native<T>();
  detail::check_dims(n, batch);
  detail::require_hip(be);
  if (batch == 0) return;
  detail::check(detail::hip_posv(A, n, batch, B, nrhs, info));
}
template <class T>
void gesv_batched(T* A, int n, std::int64_t batch, T* B, int nrhs, int* piv, int* info,
                  Backend be = Backend::Auto) {
  detail::require_native<T>();
  detail::check_dims(n, batch);
  if (!piv) throw std::invalid_argument("piv must not be null");
  detail::require_hip(be);
  if (batch == 0) return;
  detail::check(detail::hip_gesv(A, n, batch, B, nrhs, piv, info));
}

}  // namespace tsdx
