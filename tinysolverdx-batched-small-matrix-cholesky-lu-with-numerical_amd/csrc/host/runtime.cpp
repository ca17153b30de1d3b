// Host runtime of the C++ API: runtime-precision dispatch over the HIP
// backend, fp16/bf16 rounding, long-double error metrics, thread control for
// the metric loops.  Counterpart of the reference's src/runtime.cpp.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <thread>
#include <type_traits>
#include <vector>

#include "tinysolverdx/tinysolverdx.hpp"

namespace tsdx {

// ------------------------------------------------------------- devices
bool has_hip() { return tsdx_hip_available() != 0; }
bool has_cuda() { return false; }

// ------------------------------------------------------- machine probes
namespace {
// Runs one probe of the C ABI (rc = f(&value)) and returns its value.
template <class F>
double probe(F&& f) {
  double v = 0.0;
  if (f(&v) != TSDX_HIP_OK) throw std::runtime_error(tsdx_hip_last_error());
  return v;
}
}  // namespace

double measure_peak_gflops(Precision p) {
  const int kind = p == Precision::FP64 ? TSDX_HIP_PROBE_FP64 : TSDX_HIP_PROBE_FP32;
  return probe([&](double* v) { return tsdx_hip_measure_peak_gflops(kind, v); });
}
double measure_bandwidth_gbs() {
  return probe([](double* v) { return tsdx_hip_measure_bandwidth_gbs(TSDX_HIP_PROBE_TRIAD, v); });
}
double measure_mfma_tflops(Precision p) {
  const int kind = p == Precision::FP64   ? TSDX_HIP_PROBE_MFMA_FP64
                   : p == Precision::BF16 ? TSDX_HIP_PROBE_MFMA_BF16
                   : p == Precision::FP16 ? TSDX_HIP_PROBE_MFMA_FP16
                                          : TSDX_HIP_PROBE_MFMA_FP32;
  return probe([&](double* v) { return tsdx_hip_measure_mfma_tflops(kind, v); });
}

// ------------------------------------------------------------- threads
namespace {
int env_threads() {
  if (const char* e = std::getenv("TSDX_NUM_THREADS")) {
    const int v = std::atoi(e);
    if (v > 0) return v;
  }
  const unsigned h = std::thread::hardware_concurrency();
  return h ? static_cast<int>(h) : 1;
}
std::atomic<int> g_threads{0};

// Runs f(0..count-1) on up to num_threads() std::threads (work split by an
// atomic counter); used only by the long-double metric loops.
void for_each_index(std::int64_t count, const std::function<void(std::int64_t)>& f) {
  const int nt = static_cast<int>(std::min<std::int64_t>(num_threads(), count));
  if (nt <= 1) {
    for (std::int64_t i = 0; i < count; ++i) f(i);
    return;
  }
  std::atomic<std::int64_t> next{0};
  std::vector<std::thread> pool;
  pool.reserve(nt);
  for (int t = 0; t < nt; ++t)
    pool.emplace_back([&] {
      for (std::int64_t i = next++; i < count; i = next++) f(i);
    });
  for (auto& th : pool) th.join();
}
}  // namespace

int num_threads() {
  int t = g_threads.load();
  if (t <= 0) {
    t = env_threads();
    g_threads = t;
  }
  return t;
}
void set_num_threads(int n) { g_threads = n > 0 ? n : env_threads(); }

// ----------------------------------------------------------- precision
namespace detail {
namespace {
std::uint32_t bits_of(float x) {
  std::uint32_t u;
  std::memcpy(&u, &x, 4);
  return u;
}
float float_of(std::uint32_t u) {
  float x;
  std::memcpy(&x, &u, 4);
  return x;
}
}  // namespace

float round_fp16(float x) {
  const float a = std::fabs(x);
  if (!(a < INFINITY)) return x;  // infinities and NaN pass through
  const std::uint32_t sign = bits_of(x) & 0x80000000u;
  if (a >= 65520.0f) return float_of(sign | 0x7f800000u);  // half-way above 65504 and beyond: overflow
  if (a < 0x1p-14f) {                                      // binary16 subnormals: integer multiples of 2^-24
    const float q = std::nearbyint(a * 0x1p24f) * 0x1p-24f;
    return float_of(sign | bits_of(q));
  }
  // normal range: keep 10 fraction bits, round to nearest, ties to even
  std::uint32_t m = bits_of(a);
  const std::uint32_t keep_lsb = (m >> 13) & 1u;
  m = (m + 0x0fffu + keep_lsb) & ~0x1fffu;
  return float_of(sign | m);
}

float round_bf16(float x) {
  const std::uint32_t u = bits_of(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return x;  // NaN passes through
  // keep the upper 16 bits, round to nearest, ties to even (the carry may
  // move a finite value onto infinity, which is the overflow rule)
  return float_of((u + 0x7fffu + ((u >> 16) & 1u)) & 0xffff0000u);
}
}  // namespace detail

double unit_roundoff(Precision p) {
  switch (p) {
    case Precision::FP32: return std::ldexp(1.0, -24);
    case Precision::FP64: return std::ldexp(1.0, -53);
    case Precision::FP16: return std::ldexp(1.0, -11);
    case Precision::BF16: return std::ldexp(1.0, -8);
  }
  return 0.0;
}

// ------------------------------------------------ runtime precision
// FP32 / FP64 run the native kernels (a converting copy when the storage type
// differs, runtime.cpp:81-140 of the reference); FP16 / BF16 run the emulated
// kernels of the library on the caller's storage directly (each entry rounded
// on load, results written back as the rounded float values), which is the
// reference's Emu copy-in / copy-out.
namespace {
bool emulated(Precision p) { return p == Precision::FP16 || p == Precision::BF16; }
int emu_format(Precision p) { return p == Precision::FP16 ? TSDX_HIP_EMU_FP16 : TSDX_HIP_EMU_BF16; }

template <class F>
void with_native(Precision p, F&& f) {
  if (p == Precision::FP32) return f(float{});
  return f(double{});
}

template <class T, class S>
std::vector<T> copy_as(const S* x, std::int64_t count) {
  return std::vector<T>(x, x + count);
}
template <class T, class S>
void copy_back(const std::vector<T>& v, S* x) {
  std::transform(v.begin(), v.end(), x, [](T a) { return static_cast<S>(a); });
}
}  // namespace

template <class S>
void cholesky(Precision p, S* A, int n, std::int64_t batch, int* info, Backend be) {
  detail::check_dims(n, batch);
  if (emulated(p)) {
    detail::require_hip(be);
    return detail::check(detail::hip_chol_emu(A, n, batch, info, emu_format(p)));
  }
  with_native(p, [&](auto tag) {
    using T = decltype(tag);
    if constexpr (std::is_same<T, S>::value) {
      cholesky_batched<T>(A, n, batch, info, be);
    } else {
      auto v = copy_as<T>(A, batch * n * n);
      cholesky_batched<T>(v.data(), n, batch, info, be);
      copy_back(v, A);
    }
  });
}

template <class S>
void lu(Precision p, S* A, int n, std::int64_t batch, int* piv, int* info, Backend be) {
  detail::check_dims(n, batch);
  if (emulated(p)) {
    detail::require_hip(be);
    return detail::check(detail::hip_lu_emu(A, n, batch, piv, info, emu_format(p)));
  }
  with_native(p, [&](auto tag) {
    using T = decltype(tag);
    if constexpr (std::is_same<T, S>::value) {
      lu_batched<T>(A, n, batch, piv, info, be);
    } else {
      auto v = copy_as<T>(A, batch * n * n);
      lu_batched<T>(v.data(), n, batch, piv, info, be);
      copy_back(v, A);
    }
  });
}

template <class S>
void cholesky_solve(Precision p, const S* L, int n, std::int64_t batch, S* B, int nrhs) {
  detail::check_dims(n, batch);
  if (emulated(p)) return detail::check(detail::hip_chol_solve_emu(L, n, batch, B, nrhs, emu_format(p)));
  with_native(p, [&](auto tag) {
    using T = decltype(tag);
    if constexpr (std::is_same<T, S>::value) {
      cholesky_solve_batched<T>(L, n, batch, B, nrhs);
    } else {
      auto l = copy_as<T>(L, batch * n * n);
      auto b = copy_as<T>(B, batch * n * nrhs);
      cholesky_solve_batched<T>(l.data(), n, batch, b.data(), nrhs);
      copy_back(b, B);
    }
  });
}

template <class S>
void lu_solve(Precision p, const S* LU, const int* piv, int n, std::int64_t batch, S* B, int nrhs) {
  detail::check_dims(n, batch);
  if (emulated(p)) return detail::check(detail::hip_lu_solve_emu(LU, piv, n, batch, B, nrhs, emu_format(p)));
  with_native(p, [&](auto tag) {
    using T = decltype(tag);
    if constexpr (std::is_same<T, S>::value) {
      lu_solve_batched<T>(LU, piv, n, batch, B, nrhs);
    } else {
      auto l = copy_as<T>(LU, batch * n * n);
      auto b = copy_as<T>(B, batch * n * nrhs);
      lu_solve_batched<T>(l.data(), piv, n, batch, b.data(), nrhs);
      copy_back(b, B);
    }
  });
}

template <class S>
void posv(Precision p, S* A, int n, std::int64_t batch, S* B, int nrhs, int* info, Backend be) {
  detail::check_dims(n, batch);
  if (emulated(p)) {
    detail::require_hip(be);
    return detail::check(detail::hip_posv_emu(A, n, batch, B, nrhs, info, emu_format(p)));
  }
  with_native(p, [&](auto tag) {
    using T = decltype(tag);
    if constexpr (std::is_same<T, S>::value) {
      posv_batched<T>(A, n, batch, B, nrhs, info, be);
    } else {
      auto a = copy_as<T>(A, batch * n * n);
      auto b = copy_as<T>(B, batch * n * nrhs);
      posv_batched<T>(a.data(), n, batch, b.data(), nrhs, info, be);
      copy_back(a, A);
      copy_back(b, B);
    }
  });
}

template <class S>
void gesv(Precision p, S* A, int n, std::int64_t batch, S* B, int nrhs, int* piv, int* info, Backend be) {
  detail::check_dims(n, batch);
  if (emulated(p)) {
    detail::require_hip(be);
    return detail::check(detail::hip_gesv_emu(A, n, batch, B, nrhs, piv, info, emu_format(p)));
  }
  with_native(p, [&](auto tag) {
    using T = decltype(tag);
    if constexpr (std::is_same<T, S>::value) {
      gesv_batched<T>(A, n, batch, B, nrhs, piv, info, be);
    } else {
      auto a = copy_as<T>(A, batch * n * n);
      auto b = copy_as<T>(B, batch * n * nrhs);
      gesv_batched<T>(a.data(), n, batch, b.data(), nrhs, piv, info, be);
      copy_back(a, A);
      copy_back(b, B);
    }
  });
}

template <class S>
void round_to(Precision p, S* x, std::int64_t count) {
  if (p == Precision::FP64) return;
  for (std::int64_t i = 0; i < count; ++i) {
    const float f = static_cast<float>(x[i]);
    const float r = p == Precision::FP32 ? f : (p == Precision::FP16 ? detail::round_fp16(f) : detail::round_bf16(f));
    x[i] = static_cast<S>(r);
  }
}

#define TSDX_HOST_INST(S)                                                                          \
  template void cholesky<S>(Precision, S*, int, std::int64_t, int*, Backend);                     \
  template void lu<S>(Precision, S*, int, std::int64_t, int*, int*, Backend);                     \
  template void cholesky_solve<S>(Precision, const S*, int, std::int64_t, S*, int);               \
  template void lu_solve<S>(Precision, const S*, const int*, int, std::int64_t, S*, int);         \
  template void posv<S>(Precision, S*, int, std::int64_t, S*, int, int*, Backend);                \
  template void gesv<S>(Precision, S*, int, std::int64_t, S*, int, int*, int*, Backend);          \
  template void round_to<S>(Precision, S*, std::int64_t);
TSDX_HOST_INST(float)
TSDX_HOST_INST(double)
#undef TSDX_HOST_INST

// ---------------------------------------------------- long-double metrics
namespace {
using LD = long double;

LD norm2(const double* a, int count) {
  LD s = 0;
  for (int i = 0; i < count; ++i) s += static_cast<LD>(a[i]) * a[i];
  return std::sqrt(s);
}

// Gaussian elimination with partial pivoting in long double on copies of A
// and b, then back substitution (the ground-truth solve of runtime.cpp:167-189).
void solve_ld(const double* A, const double* b, int n, LD* x) {
  std::vector<LD> a(A, A + static_cast<std::size_t>(n) * n);
  for (int i = 0; i < n; ++i) x[i] = b[i];
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(a[i * n + k]) > std::fabs(a[p * n + k])) p = i;
    if (p != k) {
      std::swap_ranges(a.begin() + k * n, a.begin() + (k + 1) * n, a.begin() + p * n);
      std::swap(x[k], x[p]);
    }
    for (int i = k + 1; i < n; ++i) {
      const LD f = a[i * n + k] / a[k * n + k];
      for (int j = k + 1; j < n; ++j) a[i * n + j] -= f * a[k * n + j];
      x[i] -= f * x[k];
    }
  }
  for (int i = n - 1; i >= 0; --i) {
    LD s = x[i];
    for (int j = i + 1; j < n; ++j) s -= a[i * n + j] * x[j];
    x[i] = s / a[i * n + i];
  }
}
}  // namespace

void chol_backward_error(const double* A, const double* L, int n, std::int64_t batch, double* out) {
  for_each_index(batch, [&](std::int64_t m) {
    const double* a = A + m * n * n;
    const double* l = L + m * n * n;
    LD num = 0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        LD s = 0;
        for (int k = 0; k <= std::min(i, j); ++k) s += static_cast<LD>(l[i * n + k]) * l[j * n + k];
        const LD d = static_cast<LD>(a[i * n + j]) - s;
        num += d * d;
      }
    out[m] = static_cast<double>(std::sqrt(num) / norm2(a, n * n));
  });
}

void lu_backward_error(const double* A, const double* LU, const int* piv, int n, std::int64_t batch, double* out) {
  for_each_index(batch, [&](std::int64_t m) {
    const double* a = A + m * n * n;
    const double* f = LU + m * n * n;
    const int* pv = piv + m * n;
    std::vector<int> row(n);  // row i of P*A is row row[i] of A
    for (int i = 0; i < n; ++i) row[i] = i;
    for (int k = 0; k < n; ++k) std::swap(row[k], row[pv[k]]);
    LD num = 0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        LD s = 0;
        for (int k = 0; k <= std::min(i, j); ++k) {
          const LD lik = (k == i) ? LD(1) : static_cast<LD>(f[i * n + k]);
          s += lik * f[k * n + j];
        }
        const LD d = static_cast<LD>(a[row[i] * n + j]) - s;
        num += d * d;
      }
    out[m] = static_cast<double>(std::sqrt(num) / norm2(a, n * n));
  });
}

void solve_backward_error(const double* A, const double* x, const double* b, int n, std::int64_t batch, double* out) {
  for_each_index(batch, [&](std::int64_t m) {
    const double* a = A + m * n * n;
    const double* xv = x + m * n;
    const double* bv = b + m * n;
    LD num = 0;
    for (int i = 0; i < n; ++i) {
      LD s = bv[i];
      for (int j = 0; j < n; ++j) s -= static_cast<LD>(a[i * n + j]) * xv[j];
      num += s * s;
    }
    const LD den = norm2(a, n * n) * norm2(xv, n) + norm2(bv, n);
    out[m] = static_cast<double>(std::sqrt(num) / den);
  });
}

void reference_solve(const double* A, const double* b, int n, std::int64_t batch, double* x) {
  for_each_index(batch, [&](std::int64_t m) {
    std::vector<LD> xr(n);
    solve_ld(A + m * n * n, b + m * n, n, xr.data());
    for (int i = 0; i < n; ++i) x[m * n + i] = static_cast<double>(xr[i]);
  });
}

void solve_forward_error(const double* A, const double* x, const double* b, int n, std::int64_t batch, double* out) {
  for_each_index(batch, [&](std::int64_t m) {
    std::vector<LD> xr(n);
    solve_ld(A + m * n * n, b + m * n, n, xr.data());
    LD num = 0, den = 0;
    for (int i = 0; i < n; ++i) {
      const LD d = static_cast<LD>(x[m * n + i]) - xr[i];
      num += d * d;
      den += xr[i] * xr[i];
    }
    out[m] = static_cast<double>(std::sqrt(num / den));
  });
}

}  // namespace tsdx
