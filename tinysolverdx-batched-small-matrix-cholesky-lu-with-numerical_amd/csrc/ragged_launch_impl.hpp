// Definitions of the ragged launchers; included only by the inst_ragged_*.hip
// units, each of which instantiates one (type, class, variant).
#pragma once

#include "ragged_api.hpp"
#include "ragged_kernels.hpp"

namespace tsdx_hip {

template <class T, int CLS, int VAR>
hipError_t launch_ragged(const RaggedArgs<T>& a, const RTask* tasks, int ntasks, int* counter, int grid, hipStream_t s) {
  constexpr int WPB = RaggedVariant<T, CLS, VAR>::WPB;
  hipLaunchKernelGGL((k_ragged<T, CLS, VAR>), dim3((unsigned)grid), dim3(kWave * WPB), 0, s, a, tasks, ntasks, counter);
  return hipGetLastError();
}
template <class T, int CLS, int VAR>
int ragged_blocks_per_cu() {
  int b = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_ragged<T, CLS, VAR>, kWave * RaggedVariant<T, CLS, VAR>::WPB,
                                                   0) != hipSuccess)
    return 1;
  return b > 0 ? b : 1;
}
template <class T, int CLS, int VAR>
int ragged_capacity(int code) {
  return rtask_capacity<RaggedVariant<T, CLS, VAR>>(code);
}
template <class T, int CLS, int VAR>
int ragged_wpb() {
  return RaggedVariant<T, CLS, VAR>::WPB;
}

}  // namespace tsdx_hip

#define TSDX_INSTANTIATE_RAGGED(T, CLS, VAR)                                                                   \
  namespace tsdx_hip {                                                                                         \
  template hipError_t launch_ragged<T, CLS, VAR>(const RaggedArgs<T>&, const RTask*, int, int*, int, hipStream_t); \
  template int ragged_blocks_per_cu<T, CLS, VAR>();                                                            \
  template int ragged_capacity<T, CLS, VAR>(int);                                                              \
  template int ragged_wpb<T, CLS, VAR>();                                                                      \
  }
