// Mixed-precision gesv with float64 iterative refinement (no reference
// counterpart: the reference has only the plain driver gesv_batched,
// batched.hpp:276-280, whose solve is lu_solve_rm, kernels.hpp:204-227).
//
// One row per lane as in the plain LU kernels (NB lanes per matrix, 64 / NB
// matrices per wave): lane i keeps row i of the matrix and tracks its position
// pos.  The matrix is factored with partial pivoting in float (bfloat16 and
// half round every factor entry to their format: the arithmetic of hardware
// with no wider accumulator), then each right-hand side is refined:
//     x0 = (LU)^-1 P b                     (float)
//     r = b - A x   (float64; lane i forms its own row's dot product from
//                    its row of A, kept in registers, and x through LDS)
//     stop when ||r|| <= tol/2 (||A||_F ||x|| + ||b||)  (as the posv tiers)
//     x += (LU)^-1 P r                     (correction in float, sum in float64)
// A matrix whose low-precision factorization meets a zero pivot, or whose
// refinement stalls or runs out of steps, goes to the float64 list, which the
// plain float64 gesv then solves (with the reference's info).  Only x and info
// are written: the refinement tier keeps no float64 factor.
#pragma once

#include "mixed_kernels.hpp"

namespace tsdx_hip {

// LU with partial pivoting in float, every new factor entry rounded to LF.
// Same pivot rule as lu_regs (first largest |a|, a NaN never wins); a zero
// pivot leaves its column (info > 0: the matrix goes to the float64 tier).
template <int NB, Low LF>
__device__ __forceinline__ void lu_low(float (&r)[NB], int n, int i, int g, float* rowb, int& pos, int& info) {
  pos = i;
  info = 0;
  const bool act = i < n;
  static_for<0, NB>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    if (k < n) {
      const bool cand = act && pos >= k;
      const int cp = holds_max_key<NB>(__builtin_fabsf(r[k]), cand, false) ? pos : 0x7fffffff;
      const int p = group_allreduce<NB>(cp, [](int a, int b) { return a < b ? a : b; });
      const bool isp = act && pos == p;
      pos = (pos == k) ? p : (isp ? k : pos);
      float* rb = rowb + (k & 1) * kWave + g * NB;
      if (isp) static_for<k, NB>([&](auto jc) { rb[jc] = r[jc]; });
      wave_sync();
      const float pk = rb[k];
      const bool zero = !(pk != 0.0f) || pk != pk;
      info = (zero && info == 0) ? k + 1 : info;
      if (act && pos > k && !zero) {
        const float f = round_low(r[k] / pk, LF);
        r[k] = f;
        static_for<k + 1, NB>([&](auto jc) {
          constexpr int j = decltype(jc)::value;
          r[j] = round_low(__builtin_fmaf(-f, rb[j], r[j]), LF);
        });
      }
    }
  });
}

// Refinement kernel: factor in LF, refine every right-hand side, record the
// outcome (X for converged matrices, the float64 list for the rest; info 0
// and stat for all).
template <int NB, Low LF>
__global__ __launch_bounds__(kWave * waves_per_block<NB>()) void k_gesv_mx(IrArgs a) {
  constexpr int WPB = waves_per_block<NB>();
  __shared__ float rowb_s[WPB][2 * kWave];
  __shared__ __align__(16) double xs_s[WPB][kWave];
  const Tile t = tile_of<NB>(a.batch - a.m0);
  const int64_t m = a.m0 + t.m;
  const int n = a.n, i = t.i, g = t.g;
  const bool io = t.live && i < n;
  double* xs = xs_s[t.w] + g * NB;
  const double* Arow = a.A + (t.live ? m : 0) * (int64_t)n * n + (int64_t)(io ? i : 0) * n;
  // row i kept in float64 for the residuals, in float (rounded to LF) for the
  // factor, and its contribution to ||A||_F^2
  double ar[NB];
  float r[NB];
  double aa = 0.0;
  static_for<0, NB>([&](auto j) {
    ar[j] = (io && j < n) ? Arow[j] : 0.0;
    aa = __builtin_fma(ar[j], ar[j], aa);
    r[j] = round_low((float)ar[j], LF);
  });
  aa = group_allreduce<NB>(aa, [](double u, double v) { return u + v; });
  int pos, inf;
  lu_low<NB, LF>(r, n, i, g, rowb_s[t.w], pos, inf);
  bool ok = inf == 0;
  float upiv = 1.0f;
  static_for<0, NB>([&](auto j) { upiv = (pos == j) ? r[j] : upiv; });
  const float dinv = 1.0f / upiv;
  int used = 0;
  const double ctol = 0.25 * a.tol * a.tol;
  for (int c = 0; c < a.nrhs && ok; ++c) {
    const double bi = io ? a.B[(m * n + i) * (int64_t)a.nrhs + c] : 0.0;
    const double bb3 = 3.0 * group_allreduce<NB>(bi * bi, [](double u, double v) { return u + v; });
    // step 0 solves for x0 (from x = 0, r = b), step it > 0 applies correction it
    float z[1] = {(float)bi};
    double xp = 0.0;  // x[pos]
    bool conv = false;
    double prev = __builtin_huge_val();
    for (int it = 0;; ++it) {
      wave_sync();
      lu_solve_pos<float, NB, 1>(r, z, n, opaque(pos), g, rowb_s[t.w], dinv);
      xp += io ? (double)z[0] : 0.0;
      // r_i = b_i - sum_j a_ij x_j in float64, x through LDS
      wave_sync();
      if (io) xs[pos] = xp;
      wave_sync();
      double ri = bi;
      static_for<0, NB>([&](auto j) {
        if (j < n) ri = __builtin_fma(-ar[j], xs[j], ri);
      });
      const double rr = group_allreduce<NB>(ri * ri, [](double u, double v) { return u + v; });
      const double xx = group_allreduce<NB>(xp * xp, [](double u, double v) { return u + v; });
      const double thr = ctol * __builtin_fma(aa, xx, bb3);  // an overflowed bar never passes
      if (rr <= thr && thr < __builtin_huge_val()) {
        conv = true;
        used = used > it ? used : it;
        break;
      }
      if (it == a.max_iter || !(rr < 0.25 * prev)) break;
      prev = rr;
      z[0] = (float)ri;
    }
    ok = conv;
    if (conv && io) a.X[(m * n + pos) * (int64_t)a.nrhs + c] = xp;
  }
  if (t.live && i == 0) {
    if (!ok) a.flist[atomicAdd(a.fcount, 1)] = (int)m;
    if (a.info) a.info[m] = 0;
    if (a.stat) a.stat[m] = ok ? used : -1;
  }
}

// Plain float64 gesv (the plain kernel's chain; ZTILE: a zero pivot divides by
// one, as for n <= 24) of the matrices on the float64 list: x and info,
// grid-stride over the list.
template <int NB, bool ZTILE>
__global__ __launch_bounds__(kWave * waves_per_block<NB>()) void k_gesv_gather(IrArgs a) {
  constexpr int WPB = waves_per_block<NB>();
  __shared__ double rowb_s[WPB][2 * kWave];
  const int count = *a.fcount;
  const int64_t slots = (int64_t)gridDim.x * WPB * (kWave / NB);
  const int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave, g = lane / NB, i = lane % NB;
  const int n = a.n;
  for (int64_t base = ((int64_t)blockIdx.x * WPB + wv) * (kWave / NB); base < count; base += slots) {  // uniform
    const int64_t p = base + g;
    const bool live = p < count;
    const int64_t m = live ? a.flist[p] : 0;
    const bool io = live && i < n;
    double r[NB];
    load_row<double, NB, false>(r, a.A + m * (int64_t)n * n + (int64_t)i * n, n, io, false);
    int pos, mypiv, inf;
    lu_regs<double, NB, ZTILE>(r, n, n, i, g, rowb_s[wv], pos, mypiv, inf);
    if (live && i == 0 && a.info) a.info[m] = inf;
    double upiv = 1.0;
    static_for<0, NB>([&](auto j) { upiv = (pos == j) ? r[j] : upiv; });
    const double dinv = 1.0 / upiv;
    for (int c = 0; c < a.nrhs; ++c) {
      double b[1] = {io ? a.B[(m * n + i) * (int64_t)a.nrhs + c] : 0.0};
      wave_sync();
      lu_solve_pos<double, NB, 1>(r, b, n, opaque(pos), g, rowb_s[wv], dinv);
      if (io) a.X[(m * n + pos) * (int64_t)a.nrhs + c] = b[0];
    }
    wave_sync();
  }
}

// Launches of at most kIrChunk matrices keep every grid below 2^32 work items.
template <int NB>
hipError_t launch_gesv_ir_nb(const IrArgs& a0, int low, bool ztile, hipStream_t s) {
  constexpr int64_t per_block = (int64_t)waves_per_block<NB>() * (kWave / NB);
  const dim3 block(kWave * waves_per_block<NB>());
  for (int64_t m0 = 0; m0 < a0.batch; m0 += kIrChunk) {
    IrArgs a = a0;
    a.m0 = m0;
    const int64_t cnt = a0.batch - m0 < kIrChunk ? a0.batch - m0 : kIrChunk;
    const dim3 grid((unsigned)((cnt + per_block - 1) / per_block));
    if (low == 2)
      hipLaunchKernelGGL((k_gesv_mx<NB, Low::BF16>), grid, block, 0, s, a);
    else if (low == 3)
      hipLaunchKernelGGL((k_gesv_mx<NB, Low::F16>), grid, block, 0, s, a);
    else
      hipLaunchKernelGGL((k_gesv_mx<NB, Low::F32>), grid, block, 0, s, a);
  }
  // the list holds the failures, usually few: a short grid strides over it
  constexpr int64_t kShort = 1024;
  const int64_t blocks = (a0.batch + per_block - 1) / per_block;
  const dim3 gg((unsigned)(blocks < kShort ? blocks : kShort));
  if constexpr (NB <= 16) {  // n <= 16: always the divide-by-one rule
    hipLaunchKernelGGL((k_gesv_gather<NB, true>), gg, block, 0, s, a0);
  } else if constexpr (NB == 64) {  // n > 32: always the skip rule
    hipLaunchKernelGGL((k_gesv_gather<NB, false>), gg, block, 0, s, a0);
  } else {
    if (ztile)
      hipLaunchKernelGGL((k_gesv_gather<NB, true>), gg, block, 0, s, a0);
    else
      hipLaunchKernelGGL((k_gesv_gather<NB, false>), gg, block, 0, s, a0);
  }
  return hipGetLastError();
}

}  // namespace tsdx_hip
