// Host-side launchers for the plain kernels, one set per (type, bucket NB).
// Declared here, defined in launch_impl.hpp and explicitly instantiated by the
// inst_<type>_nb<NB>.hip units (one unit per bucket keeps compiles parallel).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tsdx_hip {

// vec = rows are 16-byte aligned (128-bit loads/stores allowed).
// ztile = zero LU pivot divides by 1 and continues (else the column is skipped).
template <class T, int NB>
hipError_t launch_potrf(T* A, int n, int64_t batch, int* info, int vec, hipStream_t s);
template <class T, int NB>
hipError_t launch_getrf(T* A, int n, int64_t batch, int* piv, int* info, int vec, bool ztile, hipStream_t s);
template <class T, int NB>
hipError_t launch_potrs(const T* L, int n, int64_t batch, T* B, int nrhs, int vec, hipStream_t s);
template <class T, int NB>
hipError_t launch_getrs(const T* LU, const int* piv, int n, int64_t batch, T* B, int nrhs, int vec, hipStream_t s);
template <class T, int NB>
hipError_t launch_posv(const T* A, T* Lout, int n, int64_t batch, const T* B, T* X, int nrhs, int* info, int vec,
                       hipStream_t s);
template <class T, int NB>
hipError_t launch_gesv(const T* A, T* LUout, int n, int64_t batch, const T* B, T* X, int nrhs, int* piv, int* info,
                       int vec, bool ztile, hipStream_t s);

}  // namespace tsdx_hip
