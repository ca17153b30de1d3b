// Definitions of the launchers declared in launch.hpp.  Included only by the
// inst_*.hip units, each of which instantiates one (family, type, NB) triple.
#pragma once

#include "launch.hpp"
#include "plain_kernels.hpp"
#include "rows_kernels.hpp"

namespace tsdx_hip {

template <int NB>
inline dim3 grid_for(int64_t batch) {
  constexpr int64_t per_block = (int64_t)waves_per_block<NB>() * (kWave / NB);
  return dim3((unsigned)((batch + per_block - 1) / per_block));
}
template <int NB>
inline dim3 block_for() {
  return dim3(kWave * waves_per_block<NB>());
}

template <class T, int NB, bool COLS>
static void potrf_go(T* A, int n, int64_t batch, int* info, int vec, hipStream_t s) {
  using S = PotrfShape<T, NB>;
  if constexpr (S::R == 1) {
    hipLaunchKernelGGL((k_potrf<T, NB, COLS>), grid_for<NB>(batch), block_for<NB>(), 0, s, A, n, batch, info, vec);
  } else {
    constexpr int64_t per_block = (int64_t)S::WPB * (kWave / S::L);
    const dim3 grid((unsigned)((batch + per_block - 1) / per_block));
    hipLaunchKernelGGL((k_potrf_rows<T, S::L, S::R, S::WPB, S::OCC, S::LOWER, S::STAGE, S::CHUNK, COLS>), grid,
                       dim3(kWave * S::WPB), 0, s, A, n, batch, info, vec);
  }
}
// In the n <= 64 bucket, the one-row kernel skips the columns past n, 16 at a
// time, for sizes up to 48 (COLS, as the LU launchers below; fp64 n=40 0.465
// -> 0.435 ms at 512 MiB of A).  The multi-row kernel gained 2-8 % at n=40..48
// but ran slower at n=33 with it, so it keeps every column.
template <class T, int NB>
hipError_t launch_potrf(T* A, int n, int64_t batch, int* info, int vec, hipStream_t s) {
  if constexpr (NB == 64 && PotrfShape<T, NB>::R == 1) {
    if (n <= 48) {
      potrf_go<T, NB, true>(A, n, batch, info, vec, s);
      return hipGetLastError();
    }
  }
  potrf_go<T, NB, false>(A, n, batch, info, vec, s);
  return hipGetLastError();
}

// Below this batch the one-row LU kernel is used even where a multi-row shape
// exists: with fewer matrices than a few rounds of waves the longer wave of
// the multi-row kernel costs more than its throughput gains (fp32 n=64 on
// MI355X: 89 vs 67 us at 1024 matrices, 216 vs 199 us at 8192, 350 vs 356 us
// at 16384, 1175 vs 1291 us at 65536; both give identical results).
constexpr int64_t kGetrfRowsMinBatch = 16384;

// In the n <= 64 bucket, sizes up to 48 take the kernels that skip the columns
// past n, 16 at a time (COLS); at 49..64 every run of 16 columns holds data.
template <class T, int NB, bool Z, bool COLS>
static void getrf_go_cols(T* A, int n, int64_t batch, int* piv, int* info, int vec, hipStream_t s) {
  using S = GetrfShape<T, NB>;
  auto one_row = [&] {
    hipLaunchKernelGGL((k_getrf<T, NB, Z, COLS>), grid_for<NB>(batch), block_for<NB>(), 0, s, A, n, batch, piv, info,
                       vec);
  };
  if constexpr (S::R == 1) {
    one_row();
  } else {
    if (batch < kGetrfRowsMinBatch) {
      one_row();
      return;
    }
    constexpr int64_t per_block = (int64_t)S::WPB * (kWave / S::L);
    const dim3 grid((unsigned)((batch + per_block - 1) / per_block));
    hipLaunchKernelGGL((k_getrf_rows<T, S::L, S::R, S::WPB, Z, COLS>), grid, dim3(kWave * S::WPB), 0, s, A, n, batch,
                       piv, info, vec);
  }
}
template <class T, int NB, bool Z>
static void getrf_go(T* A, int n, int64_t batch, int* piv, int* info, int vec, hipStream_t s) {
  if constexpr (NB == 64) {
    if (n <= 48) {
      getrf_go_cols<T, NB, Z, true>(A, n, batch, piv, info, vec, s);
      return;
    }
  }
  getrf_go_cols<T, NB, Z, false>(A, n, batch, piv, info, vec, s);
}

template <class T, int NB>
hipError_t launch_getrf(T* A, int n, int64_t batch, int* piv, int* info, int vec, bool ztile, hipStream_t s) {
  if constexpr (NB <= 16) {  // n <= 16: always the zero-pivot tile rule
    getrf_go<T, NB, true>(A, n, batch, piv, info, vec, s);
  } else if constexpr (NB == 64) {  // n >= 33: always the skip rule
    getrf_go<T, NB, false>(A, n, batch, piv, info, vec, s);
  } else {
    if (ztile)
      getrf_go<T, NB, true>(A, n, batch, piv, info, vec, s);
    else
      getrf_go<T, NB, false>(A, n, batch, piv, info, vec, s);
  }
  return hipGetLastError();
}

template <class T, int NB>
hipError_t launch_potrs(const T* L, int n, int64_t batch, T* B, int nrhs, int vec, hipStream_t s) {
  if (nrhs == 1)
    hipLaunchKernelGGL((k_potrs<T, NB, 1>), grid_for<NB>(batch), block_for<NB>(), 0, s, L, n, batch, B, nrhs, vec);
  else
    hipLaunchKernelGGL((k_potrs<T, NB, 4>), grid_for<NB>(batch), block_for<NB>(), 0, s, L, n, batch, B, nrhs, vec);
  return hipGetLastError();
}

template <class T, int NB>
hipError_t launch_getrs(const T* LU, const int* piv, int n, int64_t batch, T* B, int nrhs, int vec, hipStream_t s) {
  if (nrhs == 1)
    hipLaunchKernelGGL((k_getrs<T, NB, 1>), grid_for<NB>(batch), block_for<NB>(), 0, s, LU, piv, n, batch, B, nrhs, vec);
  else
    hipLaunchKernelGGL((k_getrs<T, NB, 4>), grid_for<NB>(batch), block_for<NB>(), 0, s, LU, piv, n, batch, B, nrhs, vec);
  return hipGetLastError();
}

template <class T, int NB>
hipError_t launch_posv(const T* A, T* Lout, int n, int64_t batch, const T* B, T* X, int nrhs, int* info, int vec,
                       hipStream_t s) {
  if (nrhs <= 1)
    hipLaunchKernelGGL((k_posv<T, NB, 1>), grid_for<NB>(batch), block_for<NB>(), 0, s, A, Lout, n, batch, B, X, nrhs, info,
                       vec);
  else
    hipLaunchKernelGGL((k_posv<T, NB, 4>), grid_for<NB>(batch), block_for<NB>(), 0, s, A, Lout, n, batch, B, X, nrhs, info,
                       vec);
  return hipGetLastError();
}

template <class T, int NB, int NR>
static void gesv_go(const T* A, T* LUout, int n, int64_t batch, const T* B, T* X, int nrhs, int* piv, int* info,
                    int vec, bool ztile, hipStream_t s) {
  if constexpr (NB <= 16) {
    hipLaunchKernelGGL((k_gesv<T, NB, NR, true>), grid_for<NB>(batch), block_for<NB>(), 0, s, A, LUout, n, batch, B, X,
                       nrhs, piv, info, vec);
  } else if constexpr (NB == 64) {  // n <= 48: the LU skips the columns past n (as getrf_go)
    if (n <= 48)
      hipLaunchKernelGGL((k_gesv<T, NB, NR, false, true>), grid_for<NB>(batch), block_for<NB>(), 0, s, A, LUout, n,
                         batch, B, X, nrhs, piv, info, vec);
    else
      hipLaunchKernelGGL((k_gesv<T, NB, NR, false>), grid_for<NB>(batch), block_for<NB>(), 0, s, A, LUout, n, batch, B,
                         X, nrhs, piv, info, vec);
  } else {
    if (ztile)
      hipLaunchKernelGGL((k_gesv<T, NB, NR, true>), grid_for<NB>(batch), block_for<NB>(), 0, s, A, LUout, n, batch, B, X,
                         nrhs, piv, info, vec);
    else
      hipLaunchKernelGGL((k_gesv<T, NB, NR, false>), grid_for<NB>(batch), block_for<NB>(), 0, s, A, LUout, n, batch, B, X,
                         nrhs, piv, info, vec);
  }
}

template <class T, int NB>
hipError_t launch_gesv(const T* A, T* LUout, int n, int64_t batch, const T* B, T* X, int nrhs, int* piv, int* info,
                       int vec, bool ztile, hipStream_t s) {
  if (nrhs <= 1)
    gesv_go<T, NB, 1>(A, LUout, n, batch, B, X, nrhs, piv, info, vec, ztile, s);
  else
    gesv_go<T, NB, 4>(A, LUout, n, batch, B, X, nrhs, piv, info, vec, ztile, s);
  return hipGetLastError();
}

}  // namespace tsdx_hip

// One macro per kernel family so that each (type, NB, family) triple compiles
// in its own translation unit.
#define TSDX_INSTANTIATE_FACTOR(T, NB)                                                                   \
  namespace tsdx_hip {                                                                                   \
  template hipError_t launch_potrf<T, NB>(T*, int, int64_t, int*, int, hipStream_t);                     \
  template hipError_t launch_getrf<T, NB>(T*, int, int64_t, int*, int*, int, bool, hipStream_t);         \
  }
#define TSDX_INSTANTIATE_SOLVE(T, NB)                                                                    \
  namespace tsdx_hip {                                                                                   \
  template hipError_t launch_potrs<T, NB>(const T*, int, int64_t, T*, int, int, hipStream_t);            \
  template hipError_t launch_getrs<T, NB>(const T*, const int*, int, int64_t, T*, int, int, hipStream_t);\
  }
#define TSDX_INSTANTIATE_POSV(T, NB)                                                                     \
  namespace tsdx_hip {                                                                                   \
  template hipError_t launch_posv<T, NB>(const T*, T*, int, int64_t, const T*, T*, int, int*, int,       \
                                         hipStream_t);                                                   \
  }
#define TSDX_INSTANTIATE_GESV(T, NB)                                                                     \
  namespace tsdx_hip {                                                                                   \
  template hipError_t launch_gesv<T, NB>(const T*, T*, int, int64_t, const T*, T*, int, int*, int*, int, \
                                         bool, hipStream_t);                                             \
  }
