// Ragged batches: matrices of different sizes and kinds (Cholesky or LU) in
// one graph launch.  No reference counterpart: the reference drivers take one
// size per call (batched.hpp:121-244).
//
// Storage: matrix m (row-major, n_m x n_m) starts at element off[m] of A; LU
// pivots of matrix m at element poff[m] of piv; info[m] as for the plain
// drivers.  Matrices are grouped on the host into buckets (kind, n <= NB); a
// bucket's matrices, sorted by decreasing n, are cut into wave tasks of at
// most 64 / L matrices (L lanes per matrix).  Three persistent kernels - one
// per register class (n <= 16, n <= 32, n <= 64), so small matrices do not pay
// for the registers of large ones - pull tasks from their own atomic counter.
//
// A task runs the same per-matrix code as the per-size kernels (chol_rows,
// lu_regs, lu_rows): the steps run to the largest size in the task, and steps past a
// matrix's own size only touch its zero padding (info is counted per matrix),
// so every factor, pivot and info is bit-identical to a per-size launch.  LU
// keeps the reference's zero-pivot rule per matrix (divide by 1 for n <= 24,
// skip from n = 25): the n <= 32 bucket is split at 24.
#pragma once

#include "plain_kernels.hpp"
#include "rows_kernels.hpp"

namespace tsdx_hip {

template <class T>
struct RaggedArgs {
  T* A;
  const int64_t* off;
  const int* size;
  int* piv;             // may be null
  const int64_t* poff;  // pivot offsets (LU matrices), may be null when piv is
  int* info;            // may be null
  const int* order;     // matrix ids, bucket by bucket
};

struct RTask {
  int code;   // bucket
  int nmax;   // largest n in the task
  int first;  // index into order
  int count;  // matrices in the task
};

// Buckets.  RC_LU32T: 17 <= n <= 24 (zero pivot divides by 1), RC_LU32S: 25..32.
enum : int {
  RC_CHOL4 = 0, RC_CHOL8, RC_CHOL16, RC_CHOL32, RC_CHOL64,
  RC_LU4, RC_LU8, RC_LU16, RC_LU32T, RC_LU32S, RC_LU64,
  RC_NCODES
};
// Register class of a bucket: 0 (n <= 16), 1 (n <= 32), 2 (n <= 64).
constexpr int rclass(int code) {
  return (code == RC_CHOL32 || code == RC_LU32T || code == RC_LU32S) ? 1
         : (code == RC_CHOL64 || code == RC_LU64)                     ? 2
                                                                      : 0;
}
constexpr int rbucket_nb(int code) {
  return (code == RC_CHOL4 || code == RC_LU4)     ? 4
         : (code == RC_CHOL8 || code == RC_LU8)   ? 8
         : (code == RC_CHOL16 || code == RC_LU16) ? 16
         : (code == RC_CHOL64 || code == RC_LU64) ? 64
                                                  : 32;
}

// Compiled variants per (type, class): waves per workgroup and rows per lane of
// the Cholesky buckets (matrices per wave = 64 / (NB / R)), and of the LU
// bucket n <= 64 (RL64); OCC, waves per SIMD the register allocation must
// allow (1: no request).  The autotuner times the variants of each class on
// the plan's size mix and keeps the best.
template <class T, int CLS, int VAR>
struct RaggedVariant;
// class 0 (n <= 16)
template <class T> struct RaggedVariant<T, 0, 0> { static constexpr int WPB = 1, R4 = 1, R8 = 1, R16 = 1, R32 = 1, R64 = 1, RL64 = 1, OCC = 1; };
template <class T> struct RaggedVariant<T, 0, 1> { static constexpr int WPB = 4, R4 = 1, R8 = 1, R16 = 1, R32 = 1, R64 = 1, RL64 = 1, OCC = 1; };
template <class T> struct RaggedVariant<T, 0, 2> { static constexpr int WPB = 1, R4 = 1, R8 = 2, R16 = 4, R32 = 1, R64 = 1, RL64 = 1, OCC = 1; };
// class 1 (n <= 32)
template <class T> struct RaggedVariant<T, 1, 0> { static constexpr int WPB = 1, R4 = 1, R8 = 1, R16 = 1, R32 = 1, R64 = 1, RL64 = 1, OCC = 1; };
template <class T> struct RaggedVariant<T, 1, 1> { static constexpr int WPB = 1, R4 = 1, R8 = 1, R16 = 1, R32 = 4, R64 = 1, RL64 = 1, OCC = 1; };
// class 2 (n <= 64); two rows per lane only for float (registers)
template <class T> struct RaggedVariant<T, 2, 0> { static constexpr int WPB = 1, R4 = 1, R8 = 1, R16 = 1, R32 = 1, R64 = 1, RL64 = 1, OCC = 1; };
template <class T> struct RaggedVariant<T, 2, 1> { static constexpr int WPB = 1, R4 = 1, R8 = 1, R16 = 1, R32 = 1, R64 = 2, RL64 = 1, OCC = 1; };
template <class T> struct RaggedVariant<T, 2, 2> { static constexpr int WPB = 1, R4 = 1, R8 = 1, R16 = 1, R32 = 1, R64 = 2, RL64 = 2, OCC = 1; };
// float64: variant 0 asked to fit 3 waves per SIMD (168 VGPRs; the Cholesky
// path alone needs 158, the LU path more)
template <class T> struct RaggedVariant<T, 2, 3> { static constexpr int WPB = 1, R4 = 1, R8 = 1, R16 = 1, R32 = 1, R64 = 1, RL64 = 1, OCC = 3; };

// Matrices per wave task for a bucket under a variant.
template <class V>
constexpr int rtask_capacity(int code) {
  const int nb = rbucket_nb(code);
  const bool chol = code <= RC_CHOL64;
  const int r = !chol ? (nb == 64 ? V::RL64 : 1)
                      : nb == 4 ? V::R4 : nb == 8 ? V::R8 : nb == 16 ? V::R16 : nb == 32 ? V::R32 : V::R64;
  return kWave / (nb / r);
}

template <class T>
struct RMat {
  bool live;
  int m, n;
  T* A;
  bool vec;
};
template <class T>
__device__ __forceinline__ RMat<T> rmat(const RaggedArgs<T>& a, const RTask& t, int g) {
  RMat<T> r;
  r.live = g < t.count;
  r.m = r.live ? a.order[t.first + g] : 0;
  r.n = r.live ? a.size[r.m] : 0;
  r.A = a.A + (r.live ? a.off[r.m] : 0);
  r.vec = ((reinterpret_cast<uintptr_t>(r.A) & 15) == 0) && ((r.n * (int)sizeof(T)) % 16 == 0);
  return r;
}

// Cholesky of one task, group shape L lanes x R rows (as k_potrf_rows).
template <class T, int L, int R>
__device__ __forceinline__ void ragged_chol(const RaggedArgs<T>& a, const RTask& t, int lane, T* col) {
  constexpr int N = L * R;
  const int g = lane / L, q = lane % L;
  const RMat<T> x = rmat(a, t, g);
  T r[R][N];
  static_for<0, R>([&](auto s) {
    const int row = q + L * s;
    load_row<T, N>(r[s], x.A + (int64_t)row * x.n, x.n, x.live && row < x.n, x.vec);
  });
  const int inf = chol_rows<T, L, R, NoSlotHook, (N > 32)>(r, t.nmax, x.n, q, g, col);
  static_for<0, R>([&](auto s) {
    const int row = q + L * s;
    if (x.live && row < x.n) store_row<T, N>(r[s], x.A + (int64_t)row * x.n, x.n, x.vec, true, row);
  });
  if (x.live && q == 0 && a.info) a.info[x.m] = inf;
}

// LU of one task (as k_getrf).
template <class T, int NB, bool ZTILE>
__device__ __forceinline__ void ragged_lu(const RaggedArgs<T>& a, const RTask& t, int lane, T* rowb) {
  const int g = lane / NB, i = lane % NB;
  const RMat<T> x = rmat(a, t, g);
  const bool io = x.live && i < x.n;
  T r[NB];
  load_row<T, NB>(r, x.A + (int64_t)i * x.n, x.n, io, x.vec);
  int pos, mypiv, inf;
  lu_regs<T, NB, ZTILE, (NB > 32)>(r, t.nmax, x.n, i, g, rowb, pos, mypiv, inf);
  if (io) {
    store_row<T, NB>(r, x.A + (int64_t)pos * x.n, x.n, x.vec, false, 0);
    if (a.piv) a.piv[a.poff[x.m] + i] = mypiv;
  }
  if (x.live && i == 0 && a.info) a.info[x.m] = inf;
}

// LU of one task, L lanes x R rows per matrix (as k_getrf_rows).
template <class T, int L, int R, bool ZTILE>
__device__ __forceinline__ void ragged_lu_rows(const RaggedArgs<T>& a, const RTask& t, int lane, T* rowb) {
  constexpr int N = L * R;
  const int g = lane / L, q = lane % L;
  const RMat<T> x = rmat(a, t, g);
  T r[R][N];
  static_for<0, R>([&](auto s) {
    const int row = q + L * s;
    load_row<T, N>(r[s], x.A + (int64_t)row * x.n, x.n, x.live && row < x.n, x.vec);
  });
  int pos[R], mypiv[R], inf;
  lu_rows<T, L, R, ZTILE, (L * R > 32)>(r, t.nmax, x.n, q, g, rowb, pos, mypiv, inf);
  static_for<0, R>([&](auto s) {
    const int row = q + L * s;
    if (x.live && row < x.n) {
      store_row<T, N>(r[s], x.A + (int64_t)pos[s] * x.n, x.n, x.vec, false, 0);
      if (a.piv) a.piv[a.poff[x.m] + row] = mypiv[s];
    }
  });
  if (x.live && q == 0 && a.info) a.info[x.m] = inf;
}

template <class V>
constexpr int rvariant_lds_rows() {  // per-wave LDS in elements / (2 * 64)
  int r = 1;
  r = V::R4 > r ? V::R4 : r;
  r = V::R8 > r ? V::R8 : r;
  r = V::R16 > r ? V::R16 : r;
  r = V::R32 > r ? V::R32 : r;
  r = V::R64 > r ? V::R64 : r;
  r = V::RL64 > r ? V::RL64 : r;
  return r;
}

// Persistent kernel of one register class: each wave takes tasks from the
// class counter until the list is exhausted (every wave reaches the exit).
template <class T, int CLS, int VAR>
__global__ __launch_bounds__((kWave * RaggedVariant<T, CLS, VAR>::WPB))
__attribute__((amdgpu_waves_per_eu(RaggedVariant<T, CLS, VAR>::OCC))) void k_ragged(RaggedArgs<T> a,
                                                                                 const RTask* __restrict__ tasks,
                                                                                 int ntasks, int* __restrict__ counter) {
  using V = RaggedVariant<T, CLS, VAR>;
  __shared__ T lds[V::WPB][2 * kWave * rvariant_lds_rows<V>()];
  const int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave;
  T* scratch = lds[wv];
  for (;;) {
    int t = 0;
    if (lane == 0) t = atomicAdd(counter, 1);
    t = __builtin_amdgcn_readfirstlane(t);
    if (t >= ntasks) break;
    const RTask tk = tasks[t];
    const int code = __builtin_amdgcn_readfirstlane(tk.code);
    if constexpr (CLS == 0) {
      switch (code) {
        case RC_CHOL4: ragged_chol<T, 4 / V::R4, V::R4>(a, tk, lane, scratch); break;
        case RC_CHOL8: ragged_chol<T, 8 / V::R8, V::R8>(a, tk, lane, scratch); break;
        case RC_CHOL16: ragged_chol<T, 16 / V::R16, V::R16>(a, tk, lane, scratch); break;
        case RC_LU4: ragged_lu<T, 4, true>(a, tk, lane, scratch); break;
        case RC_LU8: ragged_lu<T, 8, true>(a, tk, lane, scratch); break;
        default: ragged_lu<T, 16, true>(a, tk, lane, scratch); break;
      }
    } else if constexpr (CLS == 1) {
      switch (code) {
        case RC_CHOL32: ragged_chol<T, 32 / V::R32, V::R32>(a, tk, lane, scratch); break;
        case RC_LU32T: ragged_lu<T, 32, true>(a, tk, lane, scratch); break;
        default: ragged_lu<T, 32, false>(a, tk, lane, scratch); break;
      }
    } else {
      switch (code) {
        case RC_CHOL64: ragged_chol<T, 64 / V::R64, V::R64>(a, tk, lane, scratch); break;
        default:
          if constexpr (V::RL64 == 1)
            ragged_lu<T, 64, false>(a, tk, lane, scratch);
          else
            ragged_lu_rows<T, 64 / V::RL64, V::RL64, false>(a, tk, lane, scratch);
          break;
      }
    }
  }
}

// Synthetic inputs for the autotuner: matrix m gets n*I plus a symmetric
// pattern in [-0.5, 0.5] (SPD and well pivoted), so timing runs never touch
// the caller's data.
template <class T>
__global__ void k_ragged_synth(T* A, const int64_t* off, const int* size, int count) {
  const int m = blockIdx.x;
  if (m >= count) return;
  const int n = size[m];
  T* p = A + off[m];
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int i = e / n, j = e % n;
    const unsigned h = (unsigned)((i < j ? i : j) * 131 + (i < j ? j : i) * 7919 + m * 2654435761u);
    const T x = (T)((h >> 8) & 0xffff) / (T)65536 - (T)0.5;
    p[e] = (i == j) ? (T)n + x : x;
  }
}

}  // namespace tsdx_hip
