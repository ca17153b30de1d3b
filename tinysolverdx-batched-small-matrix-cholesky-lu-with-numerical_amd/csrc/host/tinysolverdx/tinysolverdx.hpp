// Umbrella header of the C++ host API (counterpart of the reference's
// include/tinysolverdx/tinysolverdx.hpp): typed batched templates plus the
// runtime-precision entry points, the long-double error metrics and thread
// control for the host-side metric loops.
#pragma once

#include <cstdint>

#include "batched.hpp"
#include "precision.hpp"

namespace tsdx {

bool has_hip();   // a HIP device is visible to libtsdx_hip.so
bool has_cuda();  // always false: this library has no CUDA backend

int num_threads();
void set_num_threads(int n);  // n <= 0 restores the default (TSDX_NUM_THREADS or hardware concurrency)

// Runtime-precision API (reference runtime.cpp:95-150).  S is the storage type
// (float or double).  FP32/FP64 run on the HIP backend, in place when the
// storage already has that type, through a converting copy otherwise.  FP16 /
// BF16 run the emulated kernels of the HIP backend (the reference's Emu<F>
// arithmetic under its Auto backend, precision.hpp:71-97): every entry is
// rounded to the format on the way in, every operation rounds its result, and
// the rounded values come back in S, bit-identical to the reference.  posv /
// gesv are the factor followed by the solve, as posv_batched / gesv_batched.
template <class S>
void cholesky(Precision p, S* A, int n, std::int64_t batch, int* info, Backend be = Backend::Auto);
template <class S>
void lu(Precision p, S* A, int n, std::int64_t batch, int* piv, int* info, Backend be = Backend::Auto);
template <class S>
void cholesky_solve(Precision p, const S* L, int n, std::int64_t batch, S* B, int nrhs);
template <class S>
void lu_solve(Precision p, const S* LU, const int* piv, int n, std::int64_t batch, S* B, int nrhs);
template <class S>
void posv(Precision p, S* A, int n, std::int64_t batch, S* B, int nrhs, int* info, Backend be = Backend::Auto);
template <class S>
void gesv(Precision p, S* A, int n, std::int64_t batch, S* B, int nrhs, int* piv, int* info,
          Backend be = Backend::Auto);
template <class S>
void round_to(Precision p, S* x, std::int64_t count);

// Long-double ground truth (inputs are doubles, one value per matrix),
// definitions of the reference's runtime.cpp:162-260.
void chol_backward_error(const double* A, const double* L, int n, std::int64_t batch, double* out);
void lu_backward_error(const double* A, const double* LU, const int* piv, int n, std::int64_t batch, double* out);
void solve_backward_error(const double* A, const double* x, const double* b, int n, std::int64_t batch, double* out);
void solve_forward_error(const double* A, const double* x, const double* b, int n, std::int64_t batch, double* out);
void reference_solve(const double* A, const double* b, int n, std::int64_t batch, double* x);

// Machine probes of the GPU (the reference's measure_peak_gflops /
// measure_bandwidth_gbs measure its CPU, tinysolverdx.hpp:31-32).
// measure_peak_gflops: multiply-add loop on the vector ALUs, float64 for FP64
// and float otherwise (the reference's rule); measure_bandwidth_gbs: the triad
// c = a + 3b over 1 GiB arrays in HBM; measure_mfma_tflops: one matrix-core
// instruction of the format (FP32 16x16x4, FP64 16x16x4, BF16 / FP16 16x16x32).
// std::runtime_error when no device is visible.
double measure_peak_gflops(Precision p);
double measure_bandwidth_gbs();
double measure_mfma_tflops(Precision p);

}  // namespace tsdx
