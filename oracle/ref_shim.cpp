// ref_shim.cpp -- C entry points over the reference's own headers and
// src/runtime.cpp, compiled from the sources where they lie (see Makefile,
// target ref).  TEST INFRASTRUCTURE ONLY: used to pin the oracle restatement
// and as the "reference" CPU baseline; never linked into the product.
//
// Each function forwards to the reference API named beside it and maps its
// exceptions to the oracle's codes: invalid_argument -> -1, runtime_error -> -2.
#include <stdexcept>

#include "tinysolverdx/tinysolverdx.hpp"

namespace {
template <class F>
int guard(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::invalid_argument&) {
    return -1;
  } catch (const std::runtime_error&) {
    return -2;
  }
}
tsdx::Backend be(int b) { return static_cast<tsdx::Backend>(b); }
tsdx::Precision prec(int p) {
  switch (p) {
    case 0: return tsdx::Precision::FP32;
    case 1: return tsdx::Precision::FP64;
    case 2: return tsdx::Precision::FP16;
    default: return tsdx::Precision::BF16;
  }
}
}  // namespace

#define REF_TYPED(T, SUF)                                                                                   \
  int ref_cholesky_batched_##SUF(T* A, int n, long long batch, int* info, int b) {                        \
    return guard([&] { tsdx::cholesky_batched<T>(A, n, batch, info, be(b)); });                             \
  }                                                                                                         \
  int ref_lu_batched_##SUF(T* A, int n, long long batch, int* piv, int* info, int b) {                    \
    return guard([&] { tsdx::lu_batched<T>(A, n, batch, piv, info, be(b)); });                              \
  }                                                                                                         \
  int ref_chol_solve_batched_##SUF(const T* L, int n, long long batch, T* B, int nrhs) {                  \
    return guard([&] { tsdx::cholesky_solve_batched<T>(L, n, batch, B, nrhs); });                           \
  }                                                                                                         \
  int ref_lu_solve_batched_##SUF(const T* LU, const int* piv, int n, long long batch, T* B, int nrhs) {   \
    return guard([&] { tsdx::lu_solve_batched<T>(LU, piv, n, batch, B, nrhs); });                           \
  }                                                                                                         \
  int ref_posv_batched_##SUF(T* A, int n, long long batch, T* B, int nrhs, int* info, int b) {            \
    return guard([&] { tsdx::posv_batched<T>(A, n, batch, B, nrhs, info, be(b)); });                       \
  }                                                                                                         \
  int ref_gesv_batched_##SUF(T* A, int n, long long batch, T* B, int nrhs, int* piv, int* info, int b) {  \
    return guard([&] { tsdx::gesv_batched<T>(A, n, batch, B, nrhs, piv, info, be(b)); });                   \
  }                                                                                                         \
  int ref_cholesky_##SUF(int p, T* A, int n, long long batch, int* info, int b) {                         \
    return guard([&] { tsdx::cholesky<T>(prec(p), A, n, batch, info, be(b)); });                            \
  }                                                                                                         \
  int ref_lu_##SUF(int p, T* A, int n, long long batch, int* piv, int* info, int b) {                     \
    return guard([&] { tsdx::lu<T>(prec(p), A, n, batch, piv, info, be(b)); });                            \
  }                                                                                                         \
  int ref_cholesky_solve_##SUF(int p, const T* L, int n, long long batch, T* B, int nrhs) {               \
    return guard([&] { tsdx::cholesky_solve<T>(prec(p), L, n, batch, B, nrhs); });                         \
  }                                                                                                         \
  int ref_lu_solve_##SUF(int p, const T* LU, const int* piv, int n, long long batch, T* B, int nrhs) {    \
    return guard([&] { tsdx::lu_solve<T>(prec(p), LU, piv, n, batch, B, nrhs); });                          \
  }                                                                                                         \
  int ref_round_to_##SUF(int p, T* x, long long count) {                                                  \
    return guard([&] { tsdx::round_to<T>(prec(p), x, count); });                                            \
  }

extern "C" {
REF_TYPED(float, f32)
REF_TYPED(double, f64)

float ref_round_fp16(float x) { return tsdx::detail::round_fp16(x); }
float ref_round_bf16(float x) { return tsdx::detail::round_bf16(x); }
double ref_unit_roundoff(int p) { return tsdx::unit_roundoff(prec(p)); }
void ref_set_num_threads(int n) { tsdx::set_num_threads(n); }
int ref_num_threads(void) { return tsdx::num_threads(); }
int ref_has_openmp(void) { return tsdx::has_openmp() ? 1 : 0; }

void ref_chol_backward_error(const double* A, const double* L, int n, long long batch, double* out) {
  tsdx::chol_backward_error(A, L, n, batch, out);
}
void ref_lu_backward_error(const double* A, const double* LU, const int* piv, int n, long long batch, double* out) {
  tsdx::lu_backward_error(A, LU, piv, n, batch, out);
}
void ref_solve_backward_error(const double* A, const double* x, const double* b, int n, long long batch, double* out) {
  tsdx::solve_backward_error(A, x, b, n, batch, out);
}
void ref_solve_forward_error(const double* A, const double* x, const double* b, int n, long long batch, double* out) {
  tsdx::solve_forward_error(A, x, b, n, batch, out);
}
void ref_reference_solve(const double* A, const double* b, int n, long long batch, double* x) {
  tsdx::reference_solve(A, b, n, batch, x);
}
}
