// Pieces of the mixed-precision posv shared with the matrix-core kernel
// (mfma_kernels.hpp): the low-precision formats, and the float64 fallback
// that solves, with the plain float64 posv, every matrix the refinement did
// not bring to the tolerance (gathered from a device list, so the fallback
// costs nothing when the list is empty).
#pragma once

#include "mixed_api.hpp"
#include "plain_kernels.hpp"

namespace tsdx_hip {

// Low-precision formats of the factorization.
enum class Low : int { F32 = 1, BF16 = 2, F16 = 3 };

__device__ __forceinline__ float round_low(float x, Low f) {
  if (f == Low::BF16) {  // nearest bfloat16, ties to even (NaN kept)
    const unsigned u = __builtin_bit_cast(unsigned, x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return x;
    return __builtin_bit_cast(float, (u + 0x7fffu + ((u >> 16) & 1u)) & 0xffff0000u);
  }
  if (f == Low::F16) return (float)(_Float16)x;  // hardware round-to-nearest-even
  return x;
}

// Plain float64 posv of matrix m by one group of NB lanes (lane i holds row
// i): factor, info, and every right-hand side.  live = false leaves the
// group idle (a partial wave).  col_s: 2 * kWave, tb_s: kWave * min(NB, 8)
// doubles of the wave's LDS.
template <int NB, int NR>
__device__ __forceinline__ void posv_plain_group(const IrArgs& a, int64_t m, bool live, int i, int g, double* col_s,
                                                 double* tb_s) {
  const int n = a.n;
  const int64_t nn = (int64_t)n * n;
  const bool io = live && i < n;
  double r[NB];
  load_row<double, NB, false>(r, a.A + m * nn + (int64_t)i * n, n, io, false);
  const int inf = chol_regs<double, NB>(r, n, n, i, g, col_s);
  if (live && i == 0 && a.info) a.info[m] = inf;
  const double dinv = diag_inv<double, NB>(r, i);
  for (int c0 = 0; c0 < a.nrhs; c0 += NR) {
    double b[NR];
    const int li = opaque(i);
    const int64_t off = (m * n + li) * (int64_t)a.nrhs + c0;
    load_rhs<double, NR>(b, a.B + off, a.nrhs - c0, io);
    forward_rows<double, NB, NR, false>(r, b, n, li, dinv);
    backward_rows_lt<double, NB, NR>(r, b, n, li, g, tb_s, dinv);
    store_rhs<double, NR>(b, a.X + off, a.nrhs - c0, io);
  }
}

// Plain float64 posv over a gathered list of matrices (the fallback).
template <int NB, int NR>
__global__ __launch_bounds__(kWave * waves_per_block<NB>()) void k_posv_gather(IrArgs a) {
  constexpr int WPB = waves_per_block<NB>();
  constexpr int CH = NB < 8 ? NB : 8;
  __shared__ double col_s[WPB][2 * kWave];
  __shared__ double tb_s[WPB][kWave * CH];
  const int count = *a.fcount;
  const int64_t slots = (int64_t)gridDim.x * WPB * (kWave / NB);
  const int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave, g = lane / NB, i = lane % NB;
  // grid-stride over the list; the loop bound is uniform per wave
  for (int64_t base = ((int64_t)blockIdx.x * WPB + wv) * (kWave / NB); base < count; base += slots) {
    const int64_t pos = base + g;
    const bool live = pos < count;
    const int64_t m = live ? a.flist[pos] : 0;
    posv_plain_group<NB, NR>(a, m, live, i, g, col_s[wv], tb_s[wv]);
  }
}

}  // namespace tsdx_hip
