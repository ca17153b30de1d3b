// Plain fp32 / fp64 kernels: potrf, getrf, potrs, getrs and the fused
// posv / gesv drivers, for one bucket NB of matrix sizes (n <= NB).
//
// Arithmetic contract (what makes factors, pivots and info bit-identical to the
// reference CPU kernels, include/tinysolverdx/kernels.hpp:18-180):
//   * every trailing update is one fma(-x, y, acc), applied to each element in
//     ascending step order k, exactly the chain of the reference's contracted
//     `acc -= x*y` (right-looking order visits k in the same order as the
//     reference's left-looking dot products);
//   * multipliers and off-diagonal Cholesky entries come from true IEEE
//     division, diagonals from IEEE sqrt;
//   * Cholesky failure (!(s > 0)) records the first failing index, substitutes
//     1 and continues (the reference's SIMD tile semantics, kernels.hpp:70-76,
//     which its Auto backend uses for fp32/fp64);
//   * LU pivot = first position holding the largest |a_ik| among non-NaN values,
//     a NaN diagonal keeps p = k (kernels.hpp:95-100); a zero pivot records info
//     and either divides by 1 (ZTILE, the tile kernel used by the reference's
//     Auto up to n = 24, kernels.hpp:165-170) or skips the column (the scalar
//     kernel Auto uses from n = 25, kernels.hpp:109).
// Columns j >= n are updated like the others (no per-column guard): they only
// ever feed columns >= n, which are never stored, and lanes i >= n hold zero
// rows, so the padded work cannot reach a stored value.
// The triangular solves keep the forward passes in the reference's order and
// run the backward passes right-looking (column sweeps), which is a different
// summation order: solutions agree to rounding, not bit for bit.
#pragma once

#include "device_common.hpp"

namespace tsdx_hip {

// ------------------------------------------------------------ Cholesky
// r: row i of A on entry, row i of L on exit (entries j > i are garbage).
// col: per-wave LDS, 2*64 elements.  Returns the group's LAPACK info.
//
// Step k: column k (already scaled) goes through LDS, every lane applies
// r[j] -= l_ik * l_jk.  The update of column k+1 is issued first and the
// pivot work of column k+1 (broadcast of the diagonal, sqrt, the division of
// the column) right after it, so that dependent chain overlaps the remaining
// independent updates of step k instead of following them.
// nloop: wave-uniform step bound (the largest size among the wave's groups);
// n: this group's size (info only counts steps k < n; steps past a group's
// size touch only its zero padding).  COLS: columns at or past nloop (rounded
// up to 16) are not updated; they lie right of every row below nloop, which is
// all the factor a caller stores.
template <class T, int NB, bool COLS = false>
__device__ __forceinline__ int chol_regs(T (&r)[NB], int nloop, int n, int i, int g, T* col) {
  const int bound = COLS ? __builtin_amdgcn_readfirstlane(nloop) : NB;
  int info = 0;
  auto pivot = [&](auto kc) {
    constexpr int k = decltype(kc)::value;
    const T akk = bcast<NB, k>(r[k]);
    const bool bad = !(akk > T(0));
    info = (bad && info == 0 && k < n) ? k + 1 : info;
    const T d = sqrtT(bad ? T(1) : akk);
    const T l = r[k] / d;
    r[k] = (i == k) ? d : ((i > k) ? l : r[k]);
  };
  pivot(ic<0>{});
  static_for<0, NB - 1>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    if (k + 1 < nloop) {
      T* cb = col + (k & 1) * kWave + g * NB;
      cb[i] = r[k];
      wave_sync();
      r[k + 1] = fmaT(-r[k], cb[k + 1], r[k + 1]);
      pivot(ic<k + 1>{});
      static_for_cols<k + 2, NB, COLS>(bound, [&](auto jc) {
        constexpr int j = decltype(jc)::value;
        r[j] = fmaT(-r[k], cb[j], r[j]);
      });
    }
  });
  return info;
}

// Reciprocal of the diagonal entry this lane holds (row i of a factor whose
// diagonal sits in register i): the substitution passes multiply by it.
template <class T, int NB>
__device__ __forceinline__ T diag_inv(const T (&r)[NB], int i) {
  T d = T(1);
  static_for<0, NB>([&](auto j) { d = (i == j) ? r[j] : d; });
  return T(1) / d;
}

// ------------------------------------------------------------------- LU
// Rows never move between lanes: lane i keeps original row i and tracks its
// current position `pos`.  On exit row `pos` of the packed L\U factor is in r,
// lane k holds piv[k] in mypiv.  rowb: per-wave LDS, 2*64 elements.
// nloop / n as for chol_regs.  COLS: columns at or past nloop (rounded up to
// 16) are neither copied nor updated.  Those columns are zero padding of every
// matrix of the group (nloop >= its n), which no later step reads for its own
// rows and which is never stored, so the results do not change; ragged tasks,
// whose nloop is the largest n of the task, skip that work.
template <class T, int NB, bool ZTILE, bool COLS = false>
__device__ __forceinline__ void lu_regs(T (&r)[NB], int nloop, int n, int i, int g, T* rowb, int& pos, int& mypiv,
                                        int& info) {
  const int bound = COLS ? __builtin_amdgcn_readfirstlane(nloop) : NB;
  pos = i;
  mypiv = 0;
  info = 0;
  const bool act = i < n;
  // The update without a branch (BF) measured faster for float up to n = 32
  // and float64 up to n = 16 (1 GiB of A on MI355X: fp32 n=16 0.537 -> 0.502
  // ms, n=32 0.807 -> 0.750 with getrf_occupancy; fp64 n=16 0.581 -> 0.514),
  // slower for the larger float64 rows (n=32 0.684 -> 0.711, n=64 1.167 ->
  // 1.235), which keep the branch.
  constexpr bool BF = sizeof(T) == 4 ? NB <= 32 : NB <= 16;
  __shared__ __align__(16) T zraw[BF ? NB + 128 / sizeof(T) : 1];
  const T* zb = BF ? zero_row_beside<T, NB>(zraw, rowb) : nullptr;
  static_for<0, NB>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    if (k < nloop) {
      const bool cand = act && pos >= k;
      const int cp = holds_max_key<NB>(absT(r[k]), cand, pos == k) ? pos : 0x7fffffff;
      const int p = group_allreduce<NB>(cp, [](int a, int b) { return a < b ? a : b; });
      if (i == k) mypiv = p;
      const bool isp = act && pos == p;
      pos = (pos == k) ? p : (isp ? k : pos);
      T* rb = rowb + (k & 1) * kWave + g * NB;
      if (isp) {
        // from column k; for float64 from the 16-byte chunk that holds it, which
        // lets the registers go out as the pairs they are held in (the column
        // left of k is never read).  Either start gave the fewer register moves
        // for its type.
        static_for_cols<(sizeof(T) == 8 ? k / 2 * 2 : k), NB, COLS>(bound, [&](auto jc) { rb[jc] = r[jc]; });
      }
      wave_sync();
      const T pk = rb[k];
      const bool zero = (pk == T(0));
      info = (zero && info == 0 && k < n) ? k + 1 : info;
      const bool upd = act && pos > k && (ZTILE || !zero);
      if constexpr (BF) {
        // Every lane runs the update; a row that must stay as it is reads the
        // pivot row from zb (zeros) with the multiplier +0, and
        // fma(-0, +0, x) = x + (-0) = x exactly for every x, signed zeros
        // included (a divergent branch here makes the compiler copy the
        // whole row array twice per step).
        const T* y = upd ? rb : zb;
        const T q = r[k] / (zero ? T(1) : pk);
        r[k] = upd ? q : r[k];
        const T f = upd ? q : T(0);
        static_for_cols<k + 1, NB, COLS>(bound, [&](auto jc) {
          constexpr int j = decltype(jc)::value;
          r[j] = fmaT(-f, y[j], r[j]);
        });
      } else if (upd) {
        const T f = r[k] / (zero ? T(1) : pk);
        r[k] = f;
        static_for_cols<k + 1, NB, COLS>(bound, [&](auto jc) {
          constexpr int j = decltype(jc)::value;
          r[j] = fmaT(-f, rb[j], r[j]);
        });
      }
    }
  });
}

// ---------------------------------------------------------------- solves
// Substitution passes.  dinv is 1 / (this lane's diagonal): y_k = b_k * dinv
// instead of b_k / d_k keeps the per-step dependence chain short (one extra
// rounding per step, the same backward stability).
//
// Forward pass with L held row-per-lane (lane == row): y_i = (b_i - sum_{k<i}
// L_ik y_k) / L_ii, updates applied in ascending k as the reference does.
template <class T, int NB, int NR, bool UNIT>
__device__ __forceinline__ void forward_rows(const T (&r)[NB], T (&b)[NR], int n, int i, T dinv) {
  static_for<0, NB>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    if (k < n) {
      static_for<0, NR>([&](auto c) {
        const T yk = UNIT ? bcast<NB, k>(b[c]) : bcast<NB, k>(b[c] * dinv);
        b[c] = (i == k) ? yk : (i > k ? fmaT(-r[k], yk, b[c]) : b[c]);
      });
    }
  });
}

// Backward pass with an upper factor held row-per-lane (U_ik = r[k] in lane i):
// column sweep k = n-1..0, x_k = b_k / U_kk, then b_i -= U_ik x_k for i < k.
template <class T, int NB, int NR>
__device__ __forceinline__ void backward_rows_upper(const T (&r)[NB], T (&b)[NR], int n, int i, T dinv) {
  static_for_rev<0, NB>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    if (k < n) {
      static_for<0, NR>([&](auto c) {
        const T xk = bcast<NB, k>(b[c] * dinv);
        b[c] = (i == k) ? xk : (i < k ? fmaT(-r[k], xk, b[c]) : b[c]);
      });
    }
  });
}

// Backward pass L^T x = y with L held row-per-lane: row k of L (lane k) feeds
// rows i < k, so blocks of CH rows of L are transposed through LDS.
// tb: per-wave LDS, 64*CH elements.
template <class T, int NB, int NR>
__device__ __forceinline__ void backward_rows_lt(const T (&r)[NB], T (&b)[NR], int n, int i, int g, T* tb, T dinv) {
  constexpr int CH = NB < 8 ? NB : 8;
  T* tg = tb + g * (CH * NB);
  static_for_rev<0, NB / CH>([&](auto qc) {
    constexpr int c0 = decltype(qc)::value * CH;
    if (c0 < n) {
      if (i >= c0 && i < c0 + CH) {
        T* dst = tg + (i - c0) * NB;
        static_for<0, c0 + CH>([&](auto jc) { dst[jc] = r[jc]; });
      }
      wave_sync();
      static_for_rev<c0, c0 + CH>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        if (k < n) {
          const T lki = tg[(k - c0) * NB + i];
          static_for<0, NR>([&](auto c) {
            const T xk = bcast<NB, k>(b[c] * dinv);
            b[c] = (i == k) ? xk : (i < k ? fmaT(-lki, xk, b[c]) : b[c]);
          });
        }
      });
      wave_sync();
    }
  });
}

// Forward / backward passes for LU factors whose rows sit at dynamic positions
// (lane i holds factor row pos, and right-hand-side row i, which is row pos of
// P*B).  yb: per-wave LDS, 2*64 elements.
template <class T, int NB, int NR>
__device__ __forceinline__ void lu_solve_pos(const T (&r)[NB], T (&b)[NR], int n, int pos, int g, T* yb, T dinv) {
  static_for<0, NB>([&](auto tc) {
    constexpr int t = decltype(tc)::value;
    if (t < n) {
      T* yg = yb + (t & 1) * kWave + g * NB;
      if (pos == t) static_for<0, NR>([&](auto c) { yg[c] = b[c]; });
      wave_sync();
      static_for<0, NR>([&](auto c) {
        const T y = yg[c];
        if (pos > t) b[c] = fmaT(-r[t], y, b[c]);
      });
    }
  });
  static_for_rev<0, NB>([&](auto tc) {
    constexpr int t = decltype(tc)::value;
    if (t < n) {
      T* yg = yb + (t & 1) * kWave + g * NB;
      if (pos == t) static_for<0, NR>([&](auto c) {
        b[c] = b[c] * dinv;
        yg[c] = b[c];
      });
      wave_sync();
      static_for<0, NR>([&](auto c) {
        const T x = yg[c];
        if (pos < t) b[c] = fmaT(-r[t], x, b[c]);
      });
    }
  });
}

template <class T, int NR>
__device__ __forceinline__ void load_rhs(T (&b)[NR], const T* __restrict__ p, int cnt, bool ok) {
  static_for<0, NR>([&](auto c) { b[c] = (ok && c < cnt) ? p[c] : T(0); });
}
template <class T, int NR>
__device__ __forceinline__ void store_rhs(const T (&b)[NR], T* __restrict__ p, int cnt, bool ok) {
  if (ok) static_for<0, NR>([&](auto c) {
      if (c < cnt) p[c] = b[c];
    });
}

// ---------------------------------------------------------------- kernels
// Thread -> (wave w of the block, group g of the wave, row i of the group);
// matrix m = (block * waves + w) * groups + g.
struct Tile {
  int lane, w, g, i;
  int64_t m;
  bool live;
};
template <int NB>
__device__ __forceinline__ Tile tile_of(int64_t batch) {
  Tile t;
  t.lane = threadIdx.x % kWave;
  t.w = threadIdx.x / kWave;
  t.g = t.lane / NB;
  t.i = t.lane % NB;
  t.m = ((int64_t)blockIdx.x * waves_per_block<NB>() + t.w) * (kWave / NB) + t.g;
  t.live = t.m < batch;
  return t;
}

// Line-ordered stores of the LU factor (n == NB, 16-byte aligned rows).  Lane
// i ends with row pos of its matrix; written row by row, one store instruction
// touches 64 rows 16 bytes at a time, and those partial-line writes bound the
// kernel (the same store pattern alone, in a plain copy, takes as long as the
// whole factorization at small n).  The rows go through LDS in ROUNDS rounds of NB/ROUNDS
// output rows per matrix (at most 4 KiB of LDS per wave): the lanes whose row
// falls in the round write it (chunk c of staged row sr at position
// c ^ (sr mod C)), then the wave reads the round back in memory order and
// stores it with consecutive lanes on consecutive 16-byte chunks.
template <class T, int NB>
constexpr int lu_stage_rounds() {
  constexpr int bytes = kWave * NB * (int)sizeof(T);
  // Measured on MI355X (1 GiB of A, bit-identical): fp32 n=8 0.494 -> 0.475 ms,
  // n=16 0.627 -> 0.558; fp64 n=8 0.618 -> 0.431, n=16 (two rounds) 0.693 ->
  // 0.585.  At n >= 32 the pivot chain, not the stores, sets the pace and the
  // rounds cost more than they save (fp32 n=32 0.848 -> 0.905 ms, fp64 n=32
  // 0.682 -> 0.749), so those sizes keep the row stores.
  return (NB < 8 || NB > 16) ? 0 : (bytes <= 4096 ? 1 : bytes / 4096);
}
template <class T, int NB>
constexpr int lu_stage_elems() {
  return lu_stage_rounds<T, NB>() ? kWave * NB / lu_stage_rounds<T, NB>() : 4;
}
template <class T, int NB>
__device__ __forceinline__ void store_rows_lines(const T (&r)[NB], int g, int pos, bool live, int lane, T* st, T* A,
                                                 int64_t m0, int64_t batch) {
  constexpr int ROUNDS = lu_stage_rounds<T, NB>(), RPR = NB / ROUNDS;
  constexpr int VE = Vec16<T>::N, C = NB / VE, MC = RPR * C;  // chunks per row, per matrix and round
  static_assert(ROUNDS > 0 && (C & (C - 1)) == 0 && (kWave * C / ROUNDS) % kWave == 0, "stage shape");
  typedef T vt __attribute__((ext_vector_type(VE)));
  vt* sv = reinterpret_cast<vt*>(st);
  static_for<0, ROUNDS>([&](auto rc) {
    constexpr int rd = decltype(rc)::value;
    if (rd > 0) wave_sync();  // the previous round's reads are done
    if (live && pos / RPR == rd) {
      const int sr = g * RPR + (pos - rd * RPR);
      static_for<0, C>([&](auto c) {
        vt v;
        static_for<0, VE>([&](auto e) { v[e()] = r[c * VE + e]; });
        sv[sr * C + (c ^ (sr & (C - 1)))] = v;
      });
    }
    wave_sync();
    static_for<0, C / ROUNDS>([&](auto i) {
      const int t = i * kWave + lane;
      const int gg = t / MC, u = t % MC, k = u / C, j = u % C, sr = gg * RPR + k;
      const vt v = sv[sr * C + (j ^ (sr & (C - 1)))];
      if (m0 + gg < batch)
        *reinterpret_cast<vt*>(A + (m0 + gg) * (int64_t)(NB * NB) + (int64_t)(rd * RPR + k) * NB + j * VE) = v;
    });
  });
}

template <class T, int NB, bool COLS = false>
__global__ __launch_bounds__(kWave * waves_per_block<NB>()) void k_potrf(T* __restrict__ A, int n, int64_t batch, int* __restrict__ info, int vec) {
  __shared__ T col_s[waves_per_block<NB>()][2 * kWave];
  __shared__ __align__(16) T stage_s[waves_per_block<NB>()][lu_stage_elems<T, NB>()];
  const Tile t = tile_of<NB>(batch);
  [[maybe_unused]] const int wv = t.w;
  const int64_t nn = (int64_t)n * n;
  T* row = A + t.m * nn + (int64_t)t.i * n;
  const bool io = t.live && t.i < n;
  T r[NB];
  load_row<T, NB>(r, row, n, io, vec != 0);
  const int inf = chol_regs<T, NB, COLS>(r, n, n, t.i, t.g, col_s[wv]);
  if constexpr (lu_stage_rounds<T, NB>() > 0) {
    if (vec && n == NB) {  // uniform: line-ordered stores as for the LU factor
      static_for<0, NB>([&](auto j) { r[j] = j > t.i ? T(0) : r[j]; });
      store_rows_lines<T, NB>(r, t.g, t.i, io, t.lane, stage_s[wv], A, t.m - t.g, batch);
      if (t.live && t.i == 0 && info) info[t.m] = inf;
      return;
    }
  }
  if (io) store_row<T, NB>(r, row, n, vec != 0, true, t.i);
  if (t.live && t.i == 0 && info) info[t.m] = inf;
}

// Waves per SIMD the register allocation of k_getrf is asked to allow: left
// to itself the compiler spends registers it does not need (fp32 n = 32: 107
// VGPRs, 4 waves per SIMD; asked for 5 it fits 53, 8 waves, with no spills),
// and the kernel is bound by the latency of its per-column chain, which more
// waves hide.  1 = no request.
template <class T, int NB>
constexpr int getrf_occupancy() {
  return NB == 32 ? 5 : 1;
}

// COLS: the columns at or past n (rounded up to 16) are skipped (lu_regs),
// for n well below the bucket's width.
template <class T, int NB, bool ZTILE, bool COLS = false>
__global__ __launch_bounds__(kWave * waves_per_block<NB>()) __attribute__((amdgpu_waves_per_eu(getrf_occupancy<T, NB>()))) void k_getrf(T* __restrict__ A, int n, int64_t batch, int* __restrict__ piv,
                                                 int* __restrict__ info, int vec) {
  __shared__ __align__(16) T rowb_s[waves_per_block<NB>()][2 * kWave];
  __shared__ __align__(16) T stage_s[waves_per_block<NB>()][lu_stage_elems<T, NB>()];
  const Tile t = tile_of<NB>(batch);
  [[maybe_unused]] const int wv = t.w;
  const int64_t nn = (int64_t)n * n;
  T* base = A + t.m * nn;
  const bool io = t.live && t.i < n;
  T r[NB];
  load_row<T, NB>(r, base + (int64_t)t.i * n, n, io, vec != 0);
  int pos, mypiv, inf;
  lu_regs<T, NB, ZTILE, COLS>(r, n, n, t.i, t.g, rowb_s[wv], pos, mypiv, inf);
  if constexpr (lu_stage_rounds<T, NB>() > 0) {
    if (vec && n == NB) {  // uniform
      store_rows_lines<T, NB>(r, t.g, pos, io && (unsigned)pos < (unsigned)NB, t.lane, stage_s[wv], A, t.m - t.g,
                              batch);
      if (io) piv[t.m * n + t.i] = mypiv;
      if (t.live && t.i == 0 && info) info[t.m] = inf;
      return;
    }
  }
  if (io) {
    store_row<T, NB>(r, base + (int64_t)pos * n, n, vec != 0, false, 0);
    piv[t.m * n + t.i] = mypiv;
  }
  if (t.live && t.i == 0 && info) info[t.m] = inf;
}

// potrs: L from memory.  Forward pass on rows of L, backward pass on columns of
// L (lane i loads column i, which is row i of L^T; loads are coalesced across
// lanes because lane i reads element i of each row).
template <class T, int NB, int NR>
__global__ __launch_bounds__(kWave * waves_per_block<NB>()) void k_potrs(const T* __restrict__ L, int n, int64_t batch, T* __restrict__ B,
                                                 int nrhs, int vec) {
  const Tile t = tile_of<NB>(batch);
  [[maybe_unused]] const int wv = t.w;
  const int64_t nn = (int64_t)n * n;
  const T* base = L + t.m * nn;
  const bool io = t.live && t.i < n;
  T r[NB];
  for (int c0 = 0; c0 < nrhs; c0 += NR) {
    T b[NR];
    T* brow = B + (t.m * n + t.i) * (int64_t)nrhs + c0;
    load_rhs<T, NR>(b, brow, nrhs - c0, io);
    const int li = opaque(t.i);
    load_row<T, NB>(r, base + (int64_t)li * n, n, io, vec != 0);
    const T dinv = diag_inv<T, NB>(r, li);
    forward_rows<T, NB, NR, false>(r, b, n, li, dinv);
    static_for<0, NB>([&](auto k) { r[k] = (io && k < n) ? base[(int64_t)k * n + li] : T(0); });
    backward_rows_upper<T, NB, NR>(r, b, n, li, dinv);
    store_rhs<T, NR>(b, brow, nrhs - c0, io);
  }
}

// getrs: LU from memory (row t at lane t); right-hand-side row t is row
// perm[t] of B where perm applies the sequential swaps piv[0..n-1].
template <class T, int NB, int NR>
__global__ __launch_bounds__(kWave * waves_per_block<NB>()) void k_getrs(const T* __restrict__ LU, const int* __restrict__ piv, int n,
                                                 int64_t batch, T* __restrict__ B, int nrhs, int vec) {
  __shared__ int pv_s[waves_per_block<NB>()][kWave];
  const Tile t = tile_of<NB>(batch);
  [[maybe_unused]] const int wv = t.w;
  const int64_t nn = (int64_t)n * n;
  const bool io = t.live && t.i < n;
  int* pv = pv_s[wv];
  pv[t.lane] = io ? piv[t.m * n + t.i] : 0;
  wave_sync();
  // source row of P*B for position i: undo the swaps from the last one back
  int src = t.i;
  for (int k = n - 1; k >= 0; --k) {
    int p = pv[t.g * NB + k];
    p = (p < 0 || p >= n) ? k : p;  // an out-of-range pivot (invalid input) is read as "no swap"
    src = (src == k) ? p : (src == p ? k : src);
  }
  T r[NB];
  load_row<T, NB>(r, LU + t.m * nn + (int64_t)t.i * n, n, io, vec != 0);
  const T dinv = diag_inv<T, NB>(r, t.i);
  for (int c0 = 0; c0 < nrhs; c0 += NR) {
    T b[NR];
    load_rhs<T, NR>(b, B + (t.m * n + src) * (int64_t)nrhs + c0, nrhs - c0, io);
    wave_sync();  // every lane has read its source row before any row is rewritten
    const int li = opaque(t.i);
    forward_rows<T, NB, NR, true>(r, b, n, li, T(1));
    backward_rows_upper<T, NB, NR>(r, b, n, li, dinv);
    store_rhs<T, NR>(b, B + (t.m * n + t.i) * (int64_t)nrhs + c0, nrhs - c0, io);
  }
}

// posv: potrf + potrs with the factor kept in registers.  Lout == nullptr
// leaves A untouched (solve-only use); otherwise the factor is written there.
template <class T, int NB, int NR>
__global__ __launch_bounds__(kWave * waves_per_block<NB>()) void k_posv(const T* __restrict__ A, T* __restrict__ Lout, int n, int64_t batch,
                                                const T* __restrict__ Bin, T* __restrict__ X, int nrhs,
                                                int* __restrict__ info, int vec) {
  constexpr int CH = NB < 8 ? NB : 8;
  __shared__ T col_s[waves_per_block<NB>()][2 * kWave];
  __shared__ T tb_s[waves_per_block<NB>()][kWave * CH];
  const Tile t = tile_of<NB>(batch);
  [[maybe_unused]] const int wv = t.w;
  const int64_t nn = (int64_t)n * n;
  const bool io = t.live && t.i < n;
  T r[NB];
  load_row<T, NB, false>(r, A + t.m * nn + (int64_t)t.i * n, n, io, vec != 0);
  const int inf = chol_regs<T, NB>(r, n, n, t.i, t.g, col_s[wv]);
  if (Lout && io) store_row<T, NB, false>(r, Lout + t.m * nn + (int64_t)t.i * n, n, vec != 0, true, t.i);
  if (t.live && t.i == 0 && info) info[t.m] = inf;
  const T dinv = diag_inv<T, NB>(r, t.i);
  for (int c0 = 0; c0 < nrhs; c0 += NR) {
    T b[NR];
    const int64_t off = (t.m * n + t.i) * (int64_t)nrhs + c0;
    load_rhs<T, NR>(b, Bin + off, nrhs - c0, io);
    const int li = opaque(t.i);
    forward_rows<T, NB, NR, false>(r, b, n, li, dinv);
    backward_rows_lt<T, NB, NR>(r, b, n, li, t.g, tb_s[wv], dinv);
    store_rhs<T, NR>(b, X + off, nrhs - c0, io);
  }
}

// gesv: getrf + getrs with the factor kept in registers.
template <class T, int NB, int NR, bool ZTILE, bool COLS = false>
__global__ __launch_bounds__(kWave * waves_per_block<NB>()) void k_gesv(const T* __restrict__ A, T* __restrict__ LUout, int n, int64_t batch,
                                                const T* __restrict__ Bin, T* __restrict__ X, int nrhs,
                                                int* __restrict__ piv, int* __restrict__ info, int vec) {
  __shared__ T rowb_s[waves_per_block<NB>()][2 * kWave];
  const Tile t = tile_of<NB>(batch);
  [[maybe_unused]] const int wv = t.w;
  const int64_t nn = (int64_t)n * n;
  const bool io = t.live && t.i < n;
  T r[NB];
  load_row<T, NB, false>(r, A + t.m * nn + (int64_t)t.i * n, n, io, vec != 0);
  int pos, mypiv, inf;
  lu_regs<T, NB, ZTILE, COLS>(r, n, n, t.i, t.g, rowb_s[wv], pos, mypiv, inf);
  if (io) {
    if (LUout) store_row<T, NB, false>(r, LUout + t.m * nn + (int64_t)pos * n, n, vec != 0, false, 0);
    if (piv) piv[t.m * n + t.i] = mypiv;
  }
  if (t.live && t.i == 0 && info) info[t.m] = inf;
  // this lane's row of U sits at position pos: its pivot is r[pos]
  T upiv = T(1);
  static_for<0, NB>([&](auto j) { upiv = (pos == j) ? r[j] : upiv; });
  const T dinv = T(1) / upiv;
  for (int c0 = 0; c0 < nrhs; c0 += NR) {
    T b[NR];
    load_rhs<T, NR>(b, Bin + (t.m * n + t.i) * (int64_t)nrhs + c0, nrhs - c0, io);
    wave_sync();
    lu_solve_pos<T, NB, NR>(r, b, n, opaque(pos), t.g, rowb_s[wv], dinv);
    store_rhs<T, NR>(b, X + (t.m * n + pos) * (int64_t)nrhs + c0, nrhs - c0, io);
  }
}

}  // namespace tsdx_hip
