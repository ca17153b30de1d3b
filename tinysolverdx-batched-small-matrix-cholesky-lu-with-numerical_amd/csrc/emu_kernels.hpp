// Emulated fp16 / bfloat16 kernels: the reference's Emu<F> arithmetic
// (precision.hpp:71-97) under its Auto backend, which for these types is the
// scalar kernels (batched.hpp:46-50): cholesky_rm, lu_rm, cholesky_solve_rm and
// lu_solve_rm (kernels.hpp:18-37, 91-119, 184-227), reached through the
// runtime-precision calls (runtime.cpp:72-140).
//
// Arithmetic contract (bit-identical results):
//   * a value is a float representable in the format; every +, -, *, / and
//     sqrt is computed in float (IEEE, correctly rounded) and rounded to the
//     format by round_low; `x -= y*z` is two rounded operations, never fused;
//   * input entries are rounded on load (float64 storage through float first)
//     and the float values are stored back in the caller's storage type;
//   * Cholesky runs right-looking with one row per lane: every entry receives
//     its updates in ascending step order, the sequence of the reference's
//     left-looking dot products.  At the first failing pivot f (!(s > 0)) the
//     reference returns at once, so rows < f keep their factor rows (upper
//     entries untouched), row f its entries left of the diagonal, rows > f
//     nothing: those untouched entries are written back as the rounded input,
//     and the upper triangle is zeroed only on success;
//   * LU pivots as the plain kernels (first largest |a|, a NaN diagonal keeps
//     the step's row, a NaN elsewhere never wins) and skips a zero-pivot
//     column (lu_rm);
//   * forward substitution is a column sweep (each row's updates in ascending
//     order, then its division); backward substitution keeps the reference's
//     row order: row t, once every x_k (k > t) is known, subtracts u_tk x_k
//     for k = t+1 .. n-1 ascending, then divides.  That chain is serial in t
//     (n^2/2 rounded steps per right-hand side): one lane per matrix works at a
//     time, the cost of keeping the reference's summation order.
#pragma once

#include "mixed_kernels.hpp"

namespace tsdx_hip {

template <Low LF>
__device__ __forceinline__ float emu(float x) {
  return round_low(x, LF);
}

// Row i of a matrix, rounded into the format (zero outside the n x n block).
template <class S, int NB, Low LF>
__device__ __forceinline__ void load_emu(float (&r)[NB], const S* __restrict__ p, int n, bool ok) {
  static_for<0, NB>([&](auto j) { r[j] = (ok && j < n) ? emu<LF>((float)p[j]) : 0.0f; });
}

// Right-looking Cholesky of the group's matrix (lane i holds row i).  Returns
// info; r holds the factor row for rows up to the failing pivot.
template <int NB, Low LF>
__device__ __forceinline__ int chol_emu(float (&r)[NB], int n, int i, int g, float* col) {
  int info = 0;
  static_for<0, NB>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    if (k < n) {
      const float akk = bcast<NB, k>(r[k]);
      info = (!(akk > 0.0f) && info == 0) ? k + 1 : info;
      const float d = emu<LF>(__builtin_sqrtf(akk));
      r[k] = (i == k) ? d : (i > k ? emu<LF>(r[k] / d) : r[k]);
      float* cb = col + (k & 1) * kWave + g * NB;
      cb[i] = r[k];
      wave_sync();
      static_for<k + 1, NB>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        r[j] = emu<LF>(r[j] - emu<LF>(r[k] * cb[j]));
      });
    }
  });
  return info;
}

// LU with partial pivoting of the group's matrix, zero pivots skipped (lu_rm).
// Lane i keeps original row i at position pos; lane k ends with piv[k].
template <int NB, Low LF>
__device__ __forceinline__ void lu_emu(float (&r)[NB], int n, int i, int g, float* rowb, int& pos, int& mypiv,
                                       int& info) {
  pos = i;
  mypiv = 0;
  info = 0;
  const bool act = i < n;
  static_for<0, NB>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    if (k < n) {
      const bool cand = act && pos >= k;
      const int cp = holds_max_key<NB>(__builtin_fabsf(r[k]), cand, pos == k) ? pos : 0x7fffffff;
      const int p = group_allreduce<NB>(cp, [](int a, int b) { return a < b ? a : b; });
      if (i == k) mypiv = p;
      const bool isp = act && pos == p;
      pos = (pos == k) ? p : (isp ? k : pos);
      float* rb = rowb + (k & 1) * kWave + g * NB;
      if (isp) static_for<k, NB>([&](auto jc) { rb[jc] = r[jc]; });
      wave_sync();
      const float pk = rb[k];
      info = (pk == 0.0f && info == 0) ? k + 1 : info;
      if (act && pos > k && pk != 0.0f) {
        const float f = emu<LF>(r[k] / pk);
        r[k] = f;
        static_for<k + 1, NB>([&](auto jc) {
          constexpr int j = decltype(jc)::value;
          r[j] = emu<LF>(r[j] - emu<LF>(f * rb[j]));
        });
      }
    }
  });
}

// Backward substitution in the reference's row order.  b: this lane's entry
// after the forward pass (row i); fs: the group's n x NB factor rows in LDS
// (row t at fs + t * NB); TRANS: the coefficient of row t, column k is
// fs[k][t] (L^T) instead of fs[t][k] (U).  xs: the group's NB floats of LDS.
template <int NB, Low LF, bool TRANS>
__device__ __forceinline__ float backward_emu(float b, int n, int i, const float* fs, float* xs) {
  for (int t = n - 1; t >= 0; --t) {  // uniform
    if (i == t) {
      float s = b;
      for (int k = t + 1; k < n; ++k) s = emu<LF>(s - emu<LF>((TRANS ? fs[k * NB + t] : fs[t * NB + k]) * xs[k]));
      s = emu<LF>(s / fs[t * NB + t]);
      xs[t] = s;
      b = s;
    }
    wave_sync();
  }
  return b;
}

template <class S, int NB, Low LF>
__global__ __launch_bounds__(kWave * waves_per_block<NB>()) void k_potrf_emu(S* __restrict__ A, int n, int64_t batch,
                                                                             int* __restrict__ info) {
  __shared__ float col_s[waves_per_block<NB>()][2 * kWave];
  const Tile t = tile_of<NB>(batch);
  const bool io = t.live && t.i < n;
  S* row = A + (t.live ? t.m : 0) * (int64_t)n * n + (int64_t)(io ? t.i : 0) * n;
  float r[NB];
  load_emu<S, NB, LF>(r, row, n, io);
  const int inf = chol_emu<NB, LF>(r, n, t.i, t.g, col_s[t.w]);
  if (io) {
    const int f = inf > 0 ? inf - 1 : n;  // rows before f are complete
    const int i = t.i;
    static_for<0, NB>([&](auto j) {
      if (j < n) {
        const bool done = (i < f) ? (j <= i) : (i == f && j < i);
        const float v = done ? r[j] : (inf == 0 ? 0.0f : emu<LF>((float)row[j]));
        row[j] = (S)v;
      }
    });
  }
  if (t.live && t.i == 0 && info) info[t.m] = inf;
}

template <class S, int NB, Low LF>
__global__ __launch_bounds__(kWave * waves_per_block<NB>()) void k_getrf_emu(S* __restrict__ A, int n, int64_t batch,
                                                                             int* __restrict__ piv,
                                                                             int* __restrict__ info) {
  __shared__ float rowb_s[waves_per_block<NB>()][2 * kWave];
  const Tile t = tile_of<NB>(batch);
  const bool io = t.live && t.i < n;
  S* base = A + (t.live ? t.m : 0) * (int64_t)n * n;
  float r[NB];
  load_emu<S, NB, LF>(r, base + (int64_t)(io ? t.i : 0) * n, n, io);
  int pos, mypiv, inf;
  lu_emu<NB, LF>(r, n, t.i, t.g, rowb_s[t.w], pos, mypiv, inf);
  if (io) {
    S* dst = base + (int64_t)pos * n;
    static_for<0, NB>([&](auto j) {
      if (j < n) dst[j] = (S)r[j];
    });
    piv[t.m * n + t.i] = mypiv;
  }
  if (t.live && t.i == 0 && info) info[t.m] = inf;
}

// potrs: lane i loads row i of L (lower triangle read) into registers and LDS.
template <class S, int NB, Low LF>
__global__ __launch_bounds__(kWave * waves_per_block<NB>()) void k_potrs_emu(const S* __restrict__ L, int n,
                                                                             int64_t batch, S* __restrict__ B,
                                                                             int nrhs) {
  constexpr int WPB = waves_per_block<NB>();
  __shared__ float fs_s[WPB][kWave * NB];
  __shared__ float xs_s[WPB][kWave];
  const Tile t = tile_of<NB>(batch);
  const bool io = t.live && t.i < n;
  const int i = t.i;
  float r[NB];
  load_emu<S, NB, LF>(r, L + (t.live ? t.m : 0) * (int64_t)n * n + (int64_t)(io ? i : 0) * n, n, io);
  float* fs = fs_s[t.w] + t.g * NB * NB;
  float* xs = xs_s[t.w] + t.g * NB;
  static_for<0, NB>([&](auto j) { fs[i * NB + j] = r[j]; });
  float dii = 1.0f;
  static_for<0, NB>([&](auto j) { dii = (i == j) ? r[j] : dii; });
  wave_sync();
  for (int c = 0; c < nrhs; ++c) {
    S* bp = B + (t.m * n + i) * (int64_t)nrhs + c;
    float b = io ? emu<LF>((float)*bp) : 0.0f;
    static_for<0, NB>([&](auto kc) {
      constexpr int k = decltype(kc)::value;
      if (k < n) {
        const float yk = bcast<NB, k>(emu<LF>(b / dii));
        b = (i == k) ? yk : (i > k ? emu<LF>(b - emu<LF>(r[k] * yk)) : b);
      }
    });
    b = backward_emu<NB, LF, true>(b, n, i, fs, xs);
    if (io) *bp = (S)b;
  }
}

// getrs: row t of LU at lane t; right-hand-side row t is row src of B, where
// src undoes the sequential swaps piv[0..n-1].
template <class S, int NB, Low LF>
__global__ __launch_bounds__(kWave * waves_per_block<NB>()) void k_getrs_emu(const S* __restrict__ LU,
                                                                             const int* __restrict__ piv, int n,
                                                                             int64_t batch, S* __restrict__ B,
                                                                             int nrhs) {
  constexpr int WPB = waves_per_block<NB>();
  __shared__ float fs_s[WPB][kWave * NB];
  __shared__ float xs_s[WPB][kWave];
  __shared__ int pv_s[WPB][kWave];
  const Tile t = tile_of<NB>(batch);
  const bool io = t.live && t.i < n;
  const int i = t.i;
  int* pv = pv_s[t.w];
  pv[t.lane] = io ? piv[t.m * n + i] : 0;
  wave_sync();
  int src = i;
  for (int k = n - 1; k >= 0; --k) {
    int p = pv[t.g * NB + k];
    p = (p < 0 || p >= n) ? k : p;  // an out-of-range pivot (invalid input) is read as "no swap"
    src = (src == k) ? p : (src == p ? k : src);
  }
  float r[NB];
  load_emu<S, NB, LF>(r, LU + (t.live ? t.m : 0) * (int64_t)n * n + (int64_t)(io ? i : 0) * n, n, io);
  float* fs = fs_s[t.w] + t.g * NB * NB;
  float* xs = xs_s[t.w] + t.g * NB;
  static_for<0, NB>([&](auto j) { fs[i * NB + j] = r[j]; });
  wave_sync();
  for (int c = 0; c < nrhs; ++c) {
    float b = io ? emu<LF>((float)B[(t.m * n + src) * (int64_t)nrhs + c]) : 0.0f;
    wave_sync();  // every lane has read its source row before any row is rewritten
    static_for<0, NB>([&](auto kc) {
      constexpr int k = decltype(kc)::value;
      if (k < n) {
        const float yk = bcast<NB, k>(b);
        b = (i > k) ? emu<LF>(b - emu<LF>(r[k] * yk)) : b;
      }
    });
    b = backward_emu<NB, LF, false>(b, n, i, fs, xs);
    if (io) B[(t.m * n + i) * (int64_t)nrhs + c] = (S)b;
  }
}

}  // namespace tsdx_hip
