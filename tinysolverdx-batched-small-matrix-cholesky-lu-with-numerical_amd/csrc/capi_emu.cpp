// C ABI of the emulated fp16 / bfloat16 path (declared in include/tsdx_hip.h):
// argument checks as for the plain calls, dispatch of n to a size bucket,
// stream-ordered device-pointer calls and synchronous host-pointer calls over
// the staging pool.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/tsdx_hip.h"
#include "capi_internal.hpp"
#include "capi_staging.hpp"
#include "emu_api.hpp"

namespace tsdx_hip {
namespace {

void check_format(int fmt) {
  if (fmt != TSDX_HIP_EMU_FP16 && fmt != TSDX_HIP_EMU_BF16)
    throw InvalidArg("format must be TSDX_HIP_EMU_FP16 or TSDX_HIP_EMU_BF16");
}

// ------------------------------------------------------------ device calls
template <class S>
void potrf_emu(S* A, int n, int64_t batch, int* info, int fmt, hipStream_t s) {
  check_dims(n, batch);
  check_format(fmt);
  if (batch == 0) return;
  if (!A) throw InvalidArg("A must not be null");
  require_device();
  hip_check(with_bucket(n,
                        [&](auto nb) {
                          return launch_potrf_emu<S, decltype(nb)::value>(A, n, batch, info, fmt, s);
                        }),
            "potrf_emu launch");
}

template <class S>
void getrf_emu(S* A, int n, int64_t batch, int* piv, int* info, int fmt, hipStream_t s) {
  check_dims(n, batch);
  check_format(fmt);
  if (!piv) throw InvalidArg("piv must not be null");
  if (batch == 0) return;
  if (!A) throw InvalidArg("A must not be null");
  require_device();
  hip_check(with_bucket(n,
                        [&](auto nb) {
                          return launch_getrf_emu<S, decltype(nb)::value>(A, n, batch, piv, info, fmt, s);
                        }),
            "getrf_emu launch");
}

template <class S>
void potrs_emu(const S* L, int n, int64_t batch, S* B, int nrhs, int fmt, hipStream_t s) {
  check_dims(n, batch);
  check_format(fmt);
  if (nrhs < 0) throw InvalidArg("negative nrhs");
  if (batch == 0 || nrhs == 0) return;
  if (!L || !B) throw InvalidArg("L and B must not be null");
  require_device();
  hip_check(with_bucket(n,
                        [&](auto nb) {
                          return launch_potrs_emu<S, decltype(nb)::value>(L, n, batch, B, nrhs, fmt, s);
                        }),
            "potrs_emu launch");
}

template <class S>
void getrs_emu(const S* LU, const int* piv, int n, int64_t batch, S* B, int nrhs, int fmt, hipStream_t s) {
  check_dims(n, batch);
  check_format(fmt);
  if (nrhs < 0) throw InvalidArg("negative nrhs");
  if (batch == 0 || nrhs == 0) return;
  if (!LU || !piv || !B) throw InvalidArg("LU, piv and B must not be null");
  require_device();
  hip_check(with_bucket(n,
                        [&](auto nb) {
                          return launch_getrs_emu<S, decltype(nb)::value>(LU, piv, n, batch, B, nrhs, fmt, s);
                        }),
            "getrs_emu launch");
}

// ------------------------------------------------------------- host calls
// One staging round trip per call: the factor of posv / gesv stays on the
// device between its two launches.
template <class S>
void factor_solve_host(S* A, int n, int64_t batch, S* B, int nrhs, int* piv, int* info, int fmt, bool lu, bool factor,
                       bool solve) {
  check_dims(n, batch);
  check_format(fmt);
  if (lu && !piv) throw InvalidArg("piv must not be null");
  if (solve && nrhs < 0) throw InvalidArg("negative nrhs");
  if (batch == 0 || (!factor && nrhs == 0)) return;
  HostCall c(staging());
  const size_t ab = sizeof(S) * (size_t)n * n * batch, bb = sizeof(S) * (size_t)n * (solve ? nrhs : 0) * batch,
               pb = sizeof(int) * (size_t)n * batch, ib = sizeof(int) * (size_t)batch;
  S* dA = (S*)c.buf(0, ab);
  S* dB = (S*)c.buf(1, bb);
  int* dP = lu ? (int*)c.buf(2, pb) : nullptr;
  int* dI = (int*)c.buf(3, ib);
  c.h2d(dA, A, ab);
  if (solve) c.h2d(dB, B, bb);
  if (lu && !factor) c.h2d(dP, piv, pb);
  if (factor) {
    if (lu)
      getrf_emu<S>(dA, n, batch, dP, dI, fmt, c.pd->stream);
    else
      potrf_emu<S>(dA, n, batch, dI, fmt, c.pd->stream);
  }
  if (solve && nrhs > 0) {
    if (lu)
      getrs_emu<S>(dA, dP, n, batch, dB, nrhs, fmt, c.pd->stream);
    else
      potrs_emu<S>(dA, n, batch, dB, nrhs, fmt, c.pd->stream);
  }
  if (factor) {
    c.d2h(A, dA, ab);
    if (lu) c.d2h(piv, dP, pb);
    if (info) c.d2h(info, dI, ib);
  }
  if (solve) c.d2h(B, dB, bb);
  c.sync();
}

}  // namespace
}  // namespace tsdx_hip

using namespace tsdx_hip;

#define S(x) static_cast<hipStream_t>(x)

extern "C" {

#define TSDX_EMU_ABI(T, SUF)                                                                                         \
  int tsdx_hip_potrf_emu_##SUF(T* A, int n, int64_t batch, int* info, int fmt, void* s) {                            \
    return guarded([&] { potrf_emu<T>(A, n, batch, info, fmt, S(s)); });                                             \
  }                                                                                                                  \
  int tsdx_hip_getrf_emu_##SUF(T* A, int n, int64_t batch, int* piv, int* info, int fmt, void* s) {                  \
    return guarded([&] { getrf_emu<T>(A, n, batch, piv, info, fmt, S(s)); });                                        \
  }                                                                                                                  \
  int tsdx_hip_potrs_emu_##SUF(const T* L, int n, int64_t batch, T* B, int nrhs, int fmt, void* s) {                 \
    return guarded([&] { potrs_emu<T>(L, n, batch, B, nrhs, fmt, S(s)); });                                          \
  }                                                                                                                  \
  int tsdx_hip_getrs_emu_##SUF(const T* LU, const int* piv, int n, int64_t batch, T* B, int nrhs, int fmt, void* s) { \
    return guarded([&] { getrs_emu<T>(LU, piv, n, batch, B, nrhs, fmt, S(s)); });                                    \
  }                                                                                                                  \
  int tsdx_hip_cholesky_emu_batched_##SUF(T* A, int n, int64_t batch, int* info, int fmt) {                          \
    return guarded([&] { factor_solve_host<T>(A, n, batch, nullptr, 0, nullptr, info, fmt, false, true, false); });  \
  }                                                                                                                  \
  int tsdx_hip_lu_emu_batched_##SUF(T* A, int n, int64_t batch, int* piv, int* info, int fmt) {                      \
    return guarded([&] { factor_solve_host<T>(A, n, batch, nullptr, 0, piv, info, fmt, true, true, false); });       \
  }                                                                                                                  \
  int tsdx_hip_cholesky_solve_emu_batched_##SUF(const T* L, int n, int64_t batch, T* B, int nrhs, int fmt) {         \
    return guarded([&] {                                                                                             \
      factor_solve_host<T>(const_cast<T*>(L), n, batch, B, nrhs, nullptr, nullptr, fmt, false, false, true);         \
    });                                                                                                              \
  }                                                                                                                  \
  int tsdx_hip_lu_solve_emu_batched_##SUF(const T* LU, const int* piv, int n, int64_t batch, T* B, int nrhs,         \
                                          int fmt) {                                                                 \
    return guarded([&] {                                                                                             \
      factor_solve_host<T>(const_cast<T*>(LU), n, batch, B, nrhs, const_cast<int*>(piv), nullptr, fmt, true, false,  \
                           true);                                                                                    \
    });                                                                                                              \
  }                                                                                                                  \
  int tsdx_hip_posv_emu_batched_##SUF(T* A, int n, int64_t batch, T* B, int nrhs, int* info, int fmt) {              \
    return guarded([&] { factor_solve_host<T>(A, n, batch, B, nrhs, nullptr, info, fmt, false, true, true); });      \
  }                                                                                                                  \
  int tsdx_hip_gesv_emu_batched_##SUF(T* A, int n, int64_t batch, T* B, int nrhs, int* piv, int* info, int fmt) {    \
    return guarded([&] { factor_solve_host<T>(A, n, batch, B, nrhs, piv, info, fmt, true, true, true); });           \
  }

TSDX_EMU_ABI(float, f32)
TSDX_EMU_ABI(double, f64)
#undef TSDX_EMU_ABI

}  // extern "C"
