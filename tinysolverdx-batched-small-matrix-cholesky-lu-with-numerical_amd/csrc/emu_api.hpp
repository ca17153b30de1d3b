// Host-side launchers of the emulated fp16 / bfloat16 kernels (emu_kernels.hpp),
// one set per (storage type S, bucket NB), instantiated by the
// inst_emu_<type>_nb<NB>.hip units.  fmt: TSDX_HIP_EMU_FP16 or TSDX_HIP_EMU_BF16.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tsdx_hip {

template <class S, int NB>
hipError_t launch_potrf_emu(S* A, int n, int64_t batch, int* info, int fmt, hipStream_t s);
template <class S, int NB>
hipError_t launch_getrf_emu(S* A, int n, int64_t batch, int* piv, int* info, int fmt, hipStream_t s);
template <class S, int NB>
hipError_t launch_potrs_emu(const S* L, int n, int64_t batch, S* B, int nrhs, int fmt, hipStream_t s);
template <class S, int NB>
hipError_t launch_getrs_emu(const S* LU, const int* piv, int n, int64_t batch, S* B, int nrhs, int fmt,
                            hipStream_t s);

}  // namespace tsdx_hip
