// Cholesky with several rows per lane ("L x R" groups).
//
// A matrix of size n <= N = L*R is owned by a group of L lanes; lane q holds
// rows q, q+L, ..., q+(R-1)L (cyclic, so every lane keeps a share of the
// long rows), slot s = row / L, column index static.  A wave therefore works
// on 64/L matrices at once, and the per-column fixed cost (pivot broadcast,
// square root, the LDS hand-off) is paid once for R rows per lane instead of
// one.  Slots whose rows are all finished drop out of the update statically.
//
// The per-element operation chain is the reference's (kernels.hpp:18-37, the
// same contract as plain_kernels.hpp): ascending-k fused updates, IEEE sqrt,
// correctly rounded quotients.  For float the quotient a/d is formed as
// RN_float(a * y) with y a Newton-refined float64 reciprocal of d: the float64
// product is within 2^-52 of a/d, and a/d never lies closer than 2^-49
// (relative) to a float rounding boundary unless it is one exactly, so the
// rounding to float is the correctly rounded quotient.  The exceptions
// (quotients in the float subnormal range, non-finite d) take IEEE division
// for the whole wave.
#pragma once

#include "device_common.hpp"


namespace tsdx_hip {

// Newton-refined reciprocal: v_rcp_f64 then two corrections.
__device__ __forceinline__ double rcp_refined(double d) {
  double y = __builtin_amdgcn_rcp(d);
  double e = __builtin_fma(-d, y, 1.0);
  y = __builtin_fma(y, e, y);
  e = __builtin_fma(-d, y, 1.0);
  return __builtin_fma(y, e, y);
}

// Correctly rounded a / d for every row slot of column k (see the header).
template <class T, int R>
struct ColumnDivider;

template <int R>
struct ColumnDivider<float, R> {
  template <class Rows, class F>
  __device__ __forceinline__ static void apply(Rows& r, float d, F&& store) {
    const double y = rcp_refined((double)d);
    float q[R];
    bool slow = !(__builtin_fabsf(d) < __builtin_huge_valf());
    static_for<0, R>([&](auto s) {
      const double p = (double)r[s] * y;
      const double ap = __builtin_fabs(p);
      slow = slow || (ap != 0.0 && ap < 0x1p-125);
      q[s] = (float)p;
    });
    if (__builtin_amdgcn_ballot_w64(slow) != 0) {
      static_for<0, R>([&](auto s) { q[s] = r[s] / d; });
    }
    static_for<0, R>([&](auto s) { store(s, q[s]); });
  }
};

template <int R>
struct ColumnDivider<double, R> {
  template <class Rows, class F>
  __device__ __forceinline__ static void apply(Rows& r, double d, F&& store) {
    static_for<0, R>([&](auto s) { store(s, r[s] / d); });
  }
};

// Factor in registers.  r[s][j] = A(q + L*s, j) on entry, L on exit (entries
// above the diagonal are garbage).  Columns j >= L*(s+1), right of every row
// of slot s, are neither read nor written, so a caller that does not touch
// them either leaves them to the compiler to drop: slot s then needs L*(s+1)
// registers, not L*R.  col: per-wave LDS, 2 * 64 * R elements.
// nloop / n as for chol_regs (plain_kernels.hpp).
// done(ic<s>) is called as soon as slot s < R - 1 holds final rows (after the
// pivot of row L*s + L - 1), so that a caller can store them while the
// factorization goes on; later steps leave those rows' lower parts alone.
// COLS: columns at or past nloop (rounded up to 16) are not updated: they are
// zero padding of every matrix of the group (nloop >= its n), never read for
// a row below nloop and never stored, so the results do not change (ragged
// tasks, whose nloop is the largest n of the task).
struct NoSlotHook {
  template <class S>
  __device__ __forceinline__ void operator()(S) const {}
};
template <class T, int L, int R, class Done = NoSlotHook, bool COLS = false>
__device__ __forceinline__ int chol_rows(T (&r)[R][L * R], int nloop, int n, int q, int g, T* col,
                                         Done done = Done{}) {
  constexpr int N = L * R;
  const int bound = COLS ? __builtin_amdgcn_readfirstlane(nloop) : N;
  int info = 0;
  auto pivot = [&](auto kc) {
    constexpr int k = decltype(kc)::value;
    constexpr int kq = k % L, ks = k / L;
    const T akk = bcast<L, kq>(r[ks][k]);
    const bool bad = !(akk > T(0));
    info = (bad && info == 0 && k < n) ? k + 1 : info;
    const T d = sqrtT(bad ? T(1) : akk);
    constexpr int S0 = ks;  // slots below ks hold rows < k only
    T a[R - S0];
    static_for<S0, R>([&](auto s) { a[s - S0] = r[s][k]; });
    ColumnDivider<T, R - S0>::apply(a, d, [&](auto sl, T l) {
      constexpr int s = decltype(sl)::value + S0;
      const int row = q + L * s;
      r[s][k] = (row == k) ? d : ((row > k) ? l : r[s][k]);
    });
  };
  pivot(ic<0>{});
  static_for<0, N - 1>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    if (k + 1 < nloop) {
      T* cb = col + (k & 1) * (kWave * R) + g * N;
      static_for<k / L, R>([&](auto s) { cb[q + L * s] = r[s][k]; });
      wave_sync();
      constexpr int SA = (k + 1) / L;  // first slot with a row > k
      static_for<SA, R>([&](auto s) { r[s][k + 1] = fmaT(-r[s][k], cb[k + 1], r[s][k + 1]); });
      pivot(ic<k + 1>{});
      static_for_cols<k + 2, N, COLS>(bound, [&](auto jc) {
        constexpr int j = decltype(jc)::value;
        const T c = cb[j];
        static_for<SA, R>([&](auto s) {
          // columns right of slot s's last row hold nothing the factor needs
          if constexpr (j < L * (decltype(s)::value + 1)) r[s][j] = fmaT(-r[s][k], c, r[s][j]);
        });
      });
    }
    // rows L*s .. L*s + L - 1 are final once pivot(L*s + L - 1) has run (in
    // step k = L*s + L - 2, or never when nloop stops earlier, in which case
    // those rows are at or beyond nloop and only padding)
    if constexpr ((k + 2) % L == 0 && (k + 2) / L - 1 < R - 1) done(ic<(k + 2) / L - 1>{});
  });
  return info;
}

// LU with partial pivoting, L lanes x R rows per matrix (rows cyclic over
// lanes as in chol_rows).  Rows never move: each row keeps its registers and
// tracks its current position pos[s]; on exit row pos[s] of the packed L\U
// factor is in r[s], and lane q holds piv[k] for k = q + L*s in mypiv[s].  The
// per-element chain, the pivot choice (key = |a_ik|, NaN -> -1, a NaN at the
// current position -> +inf, then the smallest position among the maxima) and
// the zero-pivot rule (ZTILE: divide by 1, else skip the column) are those of
// lu_regs (plain_kernels.hpp), so factors, pivots and info are identical.
// What R rows per lane buy: the pivot search reduces over the lane's R rows
// first and over L lanes after, and one refined reciprocal of the pivot
// gives the lane's R quotients (ColumnDivider: correctly rounded).
// rowb: per-wave LDS, 2 * 64 * R elements.  COLS as for lu_regs.
template <class T, int L, int R, bool ZTILE, bool COLS = false>
__device__ __forceinline__ void lu_rows(T (&r)[R][L * R], int nloop, int n, int q, int g, T* rowb, int (&pos)[R],
                                        int (&mypiv)[R], int& info) {
  constexpr int N = L * R;
  const int bound = COLS ? __builtin_amdgcn_readfirstlane(nloop) : N;
  info = 0;
  bool act[R];
  static_for<0, R>([&](auto s) {
    pos[s] = q + L * s;
    mypiv[s] = 0;
    act[s] = q + L * s < n;
  });
  static_for<0, N>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    if (k < nloop) {
      // keys of the pivot rule (holds_max_key); float keys as integers
      using K = typename std::conditional<sizeof(T) == 4, int, T>::type;
      K key[R];
      K best = K(-2);
      static_for<0, R>([&](auto s) {
        const bool cand = act[s] && pos[s] >= k;
        if constexpr (sizeof(T) == 4) {
          key[s] = pivot_key_i32<L>(absT(r[s][k]), cand, pos[s] == k);
        } else {
          const T v = absT(r[s][k]);
          const T kk = (v != v) ? (pos[s] == k ? infT<T>() : T(-1)) : v;
          key[s] = cand ? kk : T(-2);
        }
        best = key[s] > best ? key[s] : best;
      });
      const K mx = group_allreduce<L>(best, [](K a, K b) { return a > b ? a : b; });
      int cpl = 0x7fffffff;
      static_for<0, R>([&](auto s) {
        const bool c = act[s] && pos[s] >= k && key[s] == mx;
        cpl = (c && pos[s] < cpl) ? pos[s] : cpl;
      });
      const int p = group_allreduce<L>(cpl, [](int a, int b) { return a < b ? a : b; });
      if (q == k % L) mypiv[k / L] = p;
      static_for<0, R>([&](auto s) {
        const bool isp = act[s] && pos[s] == p;
        pos[s] = (pos[s] == k) ? p : (isp ? k : pos[s]);
      });
      T* rb = rowb + (k & 1) * (kWave * R) + g * N;
      static_for<0, R>([&](auto s) {
        if (act[s] && pos[s] == k) static_for_cols<k, N, COLS>(bound, [&](auto jc) { rb[jc] = r[s][jc]; });
      });
      wave_sync();
      const T pk = rb[k];
      const bool zero = (pk == T(0));
      info = (zero && info == 0 && k < n) ? k + 1 : info;
      if (ZTILE || !zero) {
        T a[R], f[R];
        static_for<0, R>([&](auto s) { a[s] = r[s][k]; });
        ColumnDivider<T, R>::apply(a, zero ? T(1) : pk, [&](auto sl, T x) { f[decltype(sl)::value] = x; });
        static_for<0, R>([&](auto s) {
          if (act[s] && pos[s] > k) {
            r[s][k] = f[s];
            static_for_cols<k + 1, N, COLS>(bound, [&](auto jc) {
              constexpr int j = decltype(jc)::value;
              r[s][j] = fmaT(-f[s], rb[j], r[s][j]);
            });
          }
        });
      }
    }
  });
}

// Group shape per (type, bucket): L lanes x R rows, WPB waves per workgroup;
// R = 1 selects the one-row kernel of plain_kernels.hpp.  OCC: waves per SIMD
// the register allocation must allow.  LOWER: load only up to the diagonal
// (load_row_lower).  STAGE: store through LDS with whole-line writes
// (store_slot_lines).  Measured on MI355X, bit-identical throughout: fp32
// n=32, 65536 matrices: 301 us one-row, 172 us 8 x 4; a copy kernel with the
// row-per-lane store pattern alone takes the same 172 us, with line-ordered
// stores 142 us; the kernel with STAGE 136 us.  OCC 4 (128 VGPRs, possible
// once slots keep only the columns their rows reach) measured 1 % slower than
// the default 3 waves per SIMD there, so OCC stays 1.  LOWER pays at fp32 n=64
// (rows of two lines; 0.580 -> 0.572 ms at 65536) and not where a row is one
// line.  fp32 n=64 keeps the row-per-lane stores: 16 KiB of staging would
// cost it half its waves.  CHUNK: rows that are not whole 16-byte chunks
// load and store as chunks (load_row, store_row); fp32 n=32 keeps the
// aligned-or-element form, whose 136 VGPRs hold it at the 3 waves per SIMD
// measured fastest (chunked, 122 VGPRs and 4 waves: 0.1385 against 0.1339 ms).
template <class T, int NB>
struct PotrfShape {
  static constexpr int L = NB, R = 1, WPB = waves_per_block<NB>(), OCC = 1;
  static constexpr bool LOWER = false, STAGE = false, CHUNK = true;
};
template <>
struct PotrfShape<float, 16> {
  static constexpr int L = 4, R = 4, WPB = 1, OCC = 1;
  static constexpr bool LOWER = false, STAGE = true, CHUNK = true;
};
template <>
struct PotrfShape<float, 32> {
  static constexpr int L = 8, R = 4, WPB = 1, OCC = 1;
  static constexpr bool LOWER = false, STAGE = true, CHUNK = false;
};
template <>
struct PotrfShape<float, 64> {
  static constexpr int L = 32, R = 2, WPB = 1, OCC = 1;
  static constexpr bool LOWER = true, STAGE = false, CHUNK = true;
};
template <>
struct PotrfShape<double, 32> {
  static constexpr int L = 8, R = 4, WPB = 1, OCC = 1;
  static constexpr bool LOWER = false, STAGE = true, CHUNK = true;
};

// Buffer resource word 3 for gfx9 (raw 32-bit data format, no swizzle), and
// an offset past the end of every resource built here (at most 2 MiB): a
// buffer load at such an offset returns zeros without a memory access.
constexpr int kBufferWord3 = 0x00020000;
constexpr uint32_t kBufferOutOfRange = 0x7ffffff0u;

// Loads row `row` of a lower Cholesky input, only as far as the diagonal:
// the 16-byte chunks that hold an entry j <= row (the upper triangle is never
// read by the factorization, and its output is zeros).  rs covers the wave's
// matrices; off is the row's byte offset in it.  The chunks above the
// diagonal, and every chunk of a row that is not loaded (ok false), take the
// out-of-range offset, so the loads are issued without a branch and all of
// them are in flight at once.  Entries not loaded are zero.  Needs rows of
// whole 16-byte chunks (n * sizeof(T) a multiple of 16, 16-byte aligned A).
template <class T, int N>
__device__ __forceinline__ void load_row_lower(T (&r)[N], __amdgpu_buffer_rsrc_t rs, uint32_t off, int row, bool ok) {
  constexpr int VE = Vec16<T>::N;
  static_assert(N % VE == 0, "whole chunks");
  typedef T vt __attribute__((ext_vector_type(VE)));
  static_for<0, N / VE>([&](auto c) {
    const bool in = ok && c * VE <= row;
    const uint32_t o = in ? off + 16u * (uint32_t)c : kBufferOutOfRange;
    const vt v = __builtin_bit_cast(vt, __builtin_amdgcn_raw_buffer_load_b128(rs, o, 0, 0));
    static_for<0, VE>([&](auto e) { r[c * VE + e] = v[e()]; });
  });
}

// Stores slot s (rows L*s .. L*s + L - 1) of every matrix of the wave with
// whole-line writes.  A lane's own row (entries above the diagonal written as
// zeros) spans L*(s+1) <= N columns of a 128-byte-multiple row; written by
// one lane per row, every store instruction would touch 64 lines 16 bytes
// at a time, and those partial-line writes, not the factorization, bound the
// kernel.  So the rows go to LDS first (chunk c of row k at position
// c ^ (k mod C), which keeps the 8 lanes of a 128-byte LDS window on distinct
// banks both ways) and are read back in memory order: consecutive lanes then
// write consecutive 16-byte chunks of the same rows.  Needs n == N (the
// wave's matrices are N x N and contiguous) and 16-byte aligned rows.
//   st: 64 * N elements of LDS (one slot of each of the wave's 64/L matrices).
template <class T, int L, int N, int S>
__device__ __forceinline__ void store_slot_lines(const T (&w)[N], int row, int g, int q, int lane, T* st, T* A,
                                                 int64_t m0, int64_t batch) {
  constexpr int VE = Vec16<T>::N, C = N / VE, MC = L * C;  // chunks per row, per matrix slot
  static_assert((C & (C - 1)) == 0 && (kWave * C) % MC == 0, "power-of-two chunks per row");
  typedef T vt __attribute__((ext_vector_type(VE)));
  vt* sv = reinterpret_cast<vt*>(st);
  static_for<0, C>([&](auto c) {
    vt v;
    static_for<0, VE>([&](auto e) {
      constexpr int j = decltype(c)::value * VE + decltype(e)::value;
      v[e()] = j > row ? T(0) : w[j];
    });
    sv[(g * L + q) * C + (c ^ (q & (C - 1)))] = v;
  });
  wave_sync();
  static_for<0, C>([&](auto i) {
    const int t = i * kWave + lane;
    const int gg = t / MC, u = t % MC, k = u / C, j = u % C;
    const vt v = sv[(gg * L + k) * C + (j ^ (k & (C - 1)))];
    if (m0 + gg < batch)
      *reinterpret_cast<vt*>(A + (m0 + gg) * (int64_t)(N * N) + (int64_t)(L * S + k) * N + j * VE) = v;
  });
}

template <class T, int L, int R, int WPB, int OCC, bool LOWER, bool STAGE, bool CHUNK, bool COLS = false>
__global__ __launch_bounds__(kWave * WPB) __attribute__((amdgpu_waves_per_eu(OCC))) void k_potrf_rows(
    T* __restrict__ A, int n, int64_t batch, int* __restrict__ info, int vec) {
  constexpr int N = L * R;
  __shared__ T col_s[WPB][2 * kWave * R];
  __shared__ __align__(16) T stage_s[WPB][STAGE ? kWave * N : 4];
  const int lane = threadIdx.x % kWave, wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int g = lane / L, q = lane % L;
  const int64_t m0 = ((int64_t)blockIdx.x * WPB + wv) * (kWave / L);  // the wave's first matrix
  const int64_t m = m0 + g;
  const bool live = m < batch;
  const int64_t nn = (int64_t)n * n;
  T r[R][N];
  if (LOWER && vec) {
    // one resource over the wave's matrices (at most 64 / L of them)
    const int64_t left = batch - m0;
    const int cnt = left <= 0 ? 0 : (left < kWave / L ? (int)left : kWave / L);
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        A + (left <= 0 ? 0 : m0 * nn), (short)0, (int)(cnt * nn * (int64_t)sizeof(T)), kBufferWord3);
    static_for<0, R>([&](auto s) {
      const int row = q + L * s;
      const uint32_t off = (uint32_t)(((int64_t)g * nn + (int64_t)row * n) * (int64_t)sizeof(T));
      load_row_lower<T, N>(r[s], rs, off, row, live && row < n);
    });
  } else {
    static_for<0, R>([&](auto s) {
      const int row = q + L * s;
      load_row<T, N, CHUNK>(r[s], A + m * nn + (int64_t)row * n, n, live && row < n, vec != 0);
    });
  }
  // each slot is stored as soon as its rows are final: the writes of the
  // first R - 1 slots overlap the rest of the factorization
  auto store_slot = [&](auto sc) {
    constexpr int s = decltype(sc)::value;
    const int row = q + L * s;
    T w[N];  // the columns right of the slot's rows are zeros by position
    static_for<0, N>([&](auto j) { w[j] = j < L * (s + 1) ? r[s][j] : T(0); });
    if (STAGE && vec && n == N) {
      store_slot_lines<T, L, N, s>(w, row, g, q, lane, stage_s[wv], A, m0, batch);
      return;
    }
    if (live && row < n) store_row<T, N, CHUNK>(w, A + m * nn + (int64_t)row * n, n, vec != 0, true, row);
  };
  const int inf = chol_rows<T, L, R, decltype(store_slot), COLS>(r, n, n, q, g, col_s[wv], store_slot);
  store_slot(ic<R - 1>{});
  if (live && q == 0 && info) info[m] = inf;
}

// Group shape of the LU kernel per (type, bucket), as PotrfShape; R = 1
// selects the one-row kernel of plain_kernels.hpp.  Measured on MI355X (fp32,
// bit-identical results): n=64, 65536 matrices 1.321 ms one-row, 1.203 ms
// 32 x 2; but n=32 (262144) 0.850 ms one-row vs 0.985 ms 8 x 4 and n=16
// (2^20) 0.630 vs 0.773 ms 4 x 4: with few columns per step the per-row
// quotients and the per-row masked updates cost more than the shared pivot
// search saves, so the small buckets keep one row per lane.
template <class T, int NB>
struct GetrfShape {
  static constexpr int L = NB, R = 1, WPB = waves_per_block<NB>();
};
template <>
struct GetrfShape<float, 64> {
  static constexpr int L = 32, R = 2, WPB = 1;
};

template <class T, int L, int R, int WPB, bool ZTILE, bool COLS = false>
__global__ __launch_bounds__(kWave * WPB) void k_getrf_rows(T* __restrict__ A, int n, int64_t batch,
                                                           int* __restrict__ piv, int* __restrict__ info, int vec) {
  constexpr int N = L * R;
  __shared__ T rowb_s[WPB][2 * kWave * R];
  const int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave;
  const int g = lane / L, q = lane % L;
  const int64_t m = ((int64_t)blockIdx.x * WPB + wv) * (kWave / L) + g;
  const bool live = m < batch;
  const int64_t nn = (int64_t)n * n;
  T r[R][N];
  static_for<0, R>([&](auto s) {
    const int row = q + L * s;
    load_row<T, N>(r[s], A + m * nn + (int64_t)row * n, n, live && row < n, vec != 0);
  });
  int pos[R], mypiv[R], inf;
  lu_rows<T, L, R, ZTILE, COLS>(r, n, n, q, g, rowb_s[wv], pos, mypiv, inf);
  static_for<0, R>([&](auto s) {
    const int row = q + L * s;
    if (live && row < n) {
      store_row<T, N>(r[s], A + m * nn + (int64_t)pos[s] * n, n, vec != 0, false, 0);
      piv[m * n + row] = mypiv[s];
    }
  });
  if (live && q == 0 && info) info[m] = inf;
}

}  // namespace tsdx_hip
