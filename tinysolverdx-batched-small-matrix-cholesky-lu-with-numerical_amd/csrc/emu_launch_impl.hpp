// Definitions of the launchers declared in emu_api.hpp.  Included only by the
// inst_emu_*.hip units.
#pragma once

#include "../../include/tsdx_hip.h"
#include "emu_api.hpp"
#include "emu_kernels.hpp"
#include "launch_impl.hpp"

namespace tsdx_hip {

// Runs f with the kernel's format: Low::F16 for TSDX_HIP_EMU_FP16, else Low::BF16.
template <class F>
hipError_t with_format(int fmt, F&& f) {
  if (fmt == TSDX_HIP_EMU_FP16) f(std::integral_constant<Low, Low::F16>{});
  else f(std::integral_constant<Low, Low::BF16>{});
  return hipGetLastError();
}

template <class S, int NB>
hipError_t launch_potrf_emu(S* A, int n, int64_t batch, int* info, int fmt, hipStream_t s) {
  return with_format(fmt, [&](auto lf) {
    hipLaunchKernelGGL((k_potrf_emu<S, NB, decltype(lf)::value>), grid_for<NB>(batch), block_for<NB>(), 0, s, A, n,
                       batch, info);
  });
}
template <class S, int NB>
hipError_t launch_getrf_emu(S* A, int n, int64_t batch, int* piv, int* info, int fmt, hipStream_t s) {
  return with_format(fmt, [&](auto lf) {
    hipLaunchKernelGGL((k_getrf_emu<S, NB, decltype(lf)::value>), grid_for<NB>(batch), block_for<NB>(), 0, s, A, n,
                       batch, piv, info);
  });
}
template <class S, int NB>
hipError_t launch_potrs_emu(const S* L, int n, int64_t batch, S* B, int nrhs, int fmt, hipStream_t s) {
  return with_format(fmt, [&](auto lf) {
    hipLaunchKernelGGL((k_potrs_emu<S, NB, decltype(lf)::value>), grid_for<NB>(batch), block_for<NB>(), 0, s, L, n,
                       batch, B, nrhs);
  });
}
template <class S, int NB>
hipError_t launch_getrs_emu(const S* LU, const int* piv, int n, int64_t batch, S* B, int nrhs, int fmt,
                            hipStream_t s) {
  return with_format(fmt, [&](auto lf) {
    hipLaunchKernelGGL((k_getrs_emu<S, NB, decltype(lf)::value>), grid_for<NB>(batch), block_for<NB>(), 0, s, LU, piv,
                       n, batch, B, nrhs);
  });
}

}  // namespace tsdx_hip

#define TSDX_INSTANTIATE_EMU(S, NB)                                                                            \
  namespace tsdx_hip {                                                                                         \
  template hipError_t launch_potrf_emu<S, NB>(S*, int, int64_t, int*, int, hipStream_t);                       \
  template hipError_t launch_getrf_emu<S, NB>(S*, int, int64_t, int*, int*, int, hipStream_t);                 \
  template hipError_t launch_potrs_emu<S, NB>(const S*, int, int64_t, S*, int, int, hipStream_t);              \
  template hipError_t launch_getrs_emu<S, NB>(const S*, const int*, int, int64_t, S*, int, int, hipStream_t);  \
  }
