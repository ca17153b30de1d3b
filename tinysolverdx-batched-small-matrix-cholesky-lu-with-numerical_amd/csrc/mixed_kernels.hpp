// Mixed-precision posv with iterative refinement (no reference counterpart:
// the reference has only the plain drivers, batched.hpp:271-275).
//
// Per matrix (one group of NB lanes, lane i = row i, as the plain kernels):
//   1. load row i of A (float64), keep ||A||_F, factor a rounded copy in the
//      low precision LOW (float, or float holding bf16 / fp16 values) with
//      float arithmetic; pivots by v_rsq (no IEEE division: this factor only
//      seeds the refinement, the float64 residual judges the answer);
//   2. per right-hand side: x0 = (L L^T)^-1 b in float, then up to max_iter
//      steps of  r = b - A x (float64, A re-read from memory),
//      stop when ||r|| / (||A||_F ||x|| + ||b||) <= tol,  d = (L L^T)^-1 r
//      (float), x += d (float64);
//   3. a matrix whose low-precision factorization fails (a non-positive
//      pivot) or whose refinement does not reach tol is appended to a device
//      list; the plain float64 posv kernel then solves exactly those
//      (gathered launch), so every returned x comes from a converged
//      refinement or from the float64 direct solve.
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

// Float Cholesky of the row-per-lane rounded copy; every stored entry of L is
// rounded to the low format (the precision an MFMA factorization would keep).
// Returns 0, or k+1 for the first non-positive pivot.
template <int NB>
__device__ __forceinline__ int chol_low(float (&r)[NB], int n, int i, int g, float* col, Low f) {
  int info = 0;
  auto pivot = [&](auto kc) {
    constexpr int k = decltype(kc)::value;
    const float akk = bcast<NB, k>(r[k]);
    const bool bad = !(akk > 0.0f);
    info = (bad && info == 0 && k < n) ? k + 1 : info;
    const float inv = __builtin_amdgcn_rsqf(bad ? 1.0f : akk);
    const float d = round_low((bad ? 1.0f : akk) * inv, f);
    const float l = round_low(r[k] * inv, f);
    r[k] = (i == k) ? d : ((i > k) ? l : r[k]);
  };
  pivot(ic<0>{});
  static_for<0, NB - 1>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    if (k + 1 < n) {
      float* cb = col + (k & 1) * kWave + g * NB;
      cb[i] = r[k];
      wave_sync();
      r[k + 1] = __builtin_fmaf(-r[k], cb[k + 1], r[k + 1]);
      pivot(ic<k + 1>{});
      static_for<k + 2, NB>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        r[j] = __builtin_fmaf(-r[k], cb[j], r[j]);
      });
    }
  });
  return info;
}

template <int NB>
__device__ __forceinline__ double group_sum(double x) {
  return group_allreduce<NB>(x, [](double a, double b) { return a + b; });
}

template <int NB>
__global__ __launch_bounds__(kWave * waves_per_block<NB>()) void k_posv_ir(IrArgs a, int low) {
  constexpr int WPB = waves_per_block<NB>();
  constexpr int CH = NB < 8 ? NB : 8;
  __shared__ float col_s[WPB][2 * kWave];
  __shared__ float tb_s[WPB][kWave * CH];
  __shared__ double xb_s[WPB][kWave];
  const Tile t = tile_of<NB>(a.batch);
  const int wv = t.w;
  const Low f = static_cast<Low>(low);
  const int n = a.n;
  const int64_t nn = (int64_t)n * n;
  const bool io = t.live && t.i < n;
  const double* arow = a.A + t.m * nn + (int64_t)t.i * n;

  // load, norm, round
  double ad[NB];
  load_row<double, NB>(ad, arow, n, io, ((reinterpret_cast<uintptr_t>(a.A) & 15) == 0) && (n % 2 == 0));
  double ss = 0.0;
  float r[NB];
  static_for<0, NB>([&](auto j) {
    ss = fma(ad[j], ad[j], ss);
    r[j] = round_low((float)ad[j], f);
  });
  const double anorm = sqrt(group_sum<NB>(ss));
  int info = chol_low<NB>(r, n, t.i, t.g, col_s[wv], f);
  float dsel = 1.0f;
  static_for<0, NB>([&](auto j) { dsel = (t.i == j) ? r[j] : dsel; });
  const float dinv = 1.0f / dsel;

  bool ok = info == 0;
  int used = 0;
  for (int c = 0; c < a.nrhs && ok; ++c) {
    const int li = opaque(t.i);
    const int64_t off = (t.m * n + li) * (int64_t)a.nrhs + c;
    const double bi = io ? a.B[off] : 0.0;
    const double bnorm = sqrt(group_sum<NB>(bi * bi));
    float b32[1] = {(float)bi};
    forward_rows<float, NB, 1, false>(r, b32, n, li, dinv);
    backward_rows_lt<float, NB, 1>(r, b32, n, li, t.g, tb_s[wv], dinv);
    double x = io ? (double)b32[0] : 0.0;
    bool conv = false;
    double prev = __builtin_huge_val();
    for (int it = 0; it <= a.max_iter; ++it) {
      // r = b - A x in float64: x is broadcast through LDS, A re-read
      double* xg = xb_s[wv] + t.g * NB;
      xg[li] = x;
      wave_sync();
      double res = bi;
      if (io) {
        static_for<0, NB>([&](auto j) {
          if (j < n) res = fma(-arow[j], xg[j], res);
        });
      } else {
        res = 0.0;
      }
      wave_sync();
      const double rn = sqrt(group_sum<NB>(res * res));
      const double xn = sqrt(group_sum<NB>(x * x));
      // half the tolerance: the residual computed here and a recomputation by
      // the caller round differently, the answer must pass either
      if (rn <= 0.5 * a.tol * (anorm * xn + bnorm)) {
        conv = true;
        used = used > it ? used : it;
        break;
      }
      // stop when refinement no longer contracts (kappa * u_low too large)
      if (it == a.max_iter || !(rn < 0.5 * prev)) break;
      prev = rn;
      float d32[1] = {(float)res};
      forward_rows<float, NB, 1, false>(r, d32, n, li, dinv);
      backward_rows_lt<float, NB, 1>(r, d32, n, li, t.g, tb_s[wv], dinv);
      x += io ? (double)d32[0] : 0.0;
    }
    ok = conv;
    if (conv && io) a.X[off] = x;
  }
  if (t.live && t.i == 0) {
    if (!ok) {
      const int slot = atomicAdd(a.fcount, 1);
      a.flist[slot] = (int)t.m;
    }
    if (a.info) a.info[t.m] = 0;
    if (a.stat) a.stat[t.m] = ok ? used : -1;
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
  const int n = a.n;
  const int64_t nn = (int64_t)n * n;
  // grid-stride over the list; the loop bound is uniform per wave
  for (int64_t base = ((int64_t)blockIdx.x * WPB + wv) * (kWave / NB); base < count; base += slots) {
    const int64_t pos = base + g;
    const bool live = pos < count;
    const int64_t m = live ? a.flist[pos] : 0;
    const bool io = live && i < n;
    double r[NB];
    load_row<double, NB>(r, a.A + m * nn + (int64_t)i * n, n, io, false);
    const int inf = chol_regs<double, NB>(r, n, i, g, col_s[wv]);
    if (live && i == 0 && a.info) a.info[m] = inf;
    const double dinv = diag_inv<double, NB>(r, i);
    for (int c0 = 0; c0 < a.nrhs; c0 += NR) {
      double b[NR];
      const int li = opaque(i);
      const int64_t off = (m * n + li) * (int64_t)a.nrhs + c0;
      load_rhs<double, NR>(b, a.B + off, a.nrhs - c0, io);
      forward_rows<double, NB, NR, false>(r, b, n, li, dinv);
      backward_rows_lt<double, NB, NR>(r, b, n, li, g, tb_s[wv], dinv);
      store_rhs<double, NR>(b, a.X + off, a.nrhs - c0, io);
    }
  }
}

__global__ void k_zero_int(int* p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}

// list[i] = i for i < batch, count = batch (every matrix takes the float64 solve)
__global__ void k_iota(int* list, int* count, int* stat, int64_t batch) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < batch) {
    list[i] = (int)i;
    if (stat) stat[i] = -1;
  }
  if (i == 0) *count = (int)batch;
}

}  // namespace tsdx_hip
