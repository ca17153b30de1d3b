// Mixed-precision posv with iterative refinement on the matrix cores (no
// reference counterpart: the reference has only the plain drivers,
// batched.hpp:271-275).
//
// One wave per matrix, n <= 16*NBLK.  The symmetric matrix is handled as
// NBLK x NBLK blocks of 16 x 16, upper blocks only (I <= J), each block in the
// accumulator layout of v_mfma_f32_16x16x4f32: lane l = 16g + j holds column j
// and rows 4g + v (v = 0..3) of the block, four registers.  In that layout the
// product X^T Y of two stored blocks is four MFMAs whose A / B operands are the
// registers of X and Y as they stand, so the whole trailing update of a right-
// looking block Cholesky (A = U^T U)
//     U_KJ = U_KK^-T W_KJ,   W_IJ -= U_KI^T U_KJ      (K < I <= J)
// runs on MFMA with no data movement.  The diagonal blocks are eliminated on
// the VALU (16 steps each) and give the inverse V_KK = U_KK^-1 directly
// (column operations on an identity), which is what the panel product and the
// triangular solves use.
//
// Data kept per wave for the whole solve:
//   ad   float64 upper blocks of A (80 VGPRs at NBLK = 4): the residual;
//   wl   float factor in LDS: V_KK in the diagonal slots, U_KJ elsewhere
//        (10 KiB per wave at NBLK = 4, each lane owning its 16-byte slot of
//        every block).
// A is therefore read from HBM once.  With the factor in LDS the kernels fit
// 168 registers, i.e. 3 waves per SIMD (amdgpu_waves_per_eu(3)); the few
// values the compiler then spills are output addresses, reloaded only when X
// is stored.  Only the lower triangle of A is read
// (the reference's convention for the Cholesky input): element (r, c) of the
// symmetric matrix comes from A[max(r,c)][min(r,c)].  Rows and columns past n
// are identity padding.
//
// Per right-hand side: x0 = U^-1 U^-T b (float), then up to max_iter steps of
//   r = b - A x (float64), stop when ||r|| <= tol/2 (||A||_F ||x|| + ||b||),
//   d = U^-1 U^-T r (float), x += d (float64).
// Matrices whose factorization meets a non-positive pivot, or whose
// refinement stalls or runs out of steps, go to the float64 fallback list.
//
// Vector forms used by the solves (values of a length-16*NBLK vector):
//   column form  c[J]    lane (g, j) holds entry 16J + j (same for every g);
//   row form     R[I][v()] lane (g, j) holds entry 16I + 4g + v (same for every j).
#pragma once

#include "mixed_kernels.hpp"

namespace tsdx_hip {

typedef float v4f __attribute__((ext_vector_type(4)));

template <int NBLK>
constexpr int nupper() {
  return NBLK * (NBLK + 1) / 2;
}
// Slot of upper block (I, J), I <= J, blocks stored row by row.
template <int NBLK>
constexpr int bslot(int I, int J) {
  return I * NBLK - I * (I - 1) / 2 + (J - I);
}

// ------------------------------------------------------------ lane moves
// Every control used here reads a valid lane for every lane, so the move
// needs no "old" value (bound_ctrl set), which lets the compiler fold it into
// the consuming VALU operation.
template <int CTRL>
__device__ __forceinline__ int dpp_i32(int x) {
  return __builtin_amdgcn_mov_dpp(x, CTRL, 0xF, 0xF, true);
}
template <int CTRL>
__device__ __forceinline__ float dppf(float x) {
  return __builtin_bit_cast(float, dpp_i32<CTRL>(__builtin_bit_cast(int, x)));
}
template <int CTRL>
__device__ __forceinline__ double dppd(double x) {
  return mk64(dpp_i32<CTRL>(lo32(x)), dpp_i32<CTRL>(hi32(x)));
}
// DPP controls (rows of 16 lanes): row_newbcast:K = 0x150 + K (every lane of
// the row reads lane K of the row), row_ror:8 = 0x128 (partner j ^ 8),
// row_half_mirror = 0x141 (partner j ^ 7), row_mirror = 0x140 (partner
// 15 - j), quad_perm xor 2 = 0x4E, xor 1 = 0xB1.
template <int K>
__device__ __forceinline__ float row_bcast(float x) {
  return dppf<0x150 + K>(x);
}
__device__ __forceinline__ float bperm(float x, int src) {
  return __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(src << 2, __builtin_bit_cast(int, x)));
}
__device__ __forceinline__ double bperm(double x, int src) {
  return mk64(__builtin_amdgcn_ds_bpermute(src << 2, lo32(x)), __builtin_amdgcn_ds_bpermute(src << 2, hi32(x)));
}
// Sum over the four lanes that share j (the row groups), on the VALU:
// v_permlane16_swap(x, x) returns (x with rows 1 and 3 replaced by rows 0 and
// 2, x with rows 0 and 2 replaced by rows 1 and 3), so the sum of the pair is
// the sum over lanes l and l ^ 16; v_permlane32_swap does the same for halves.
__device__ __forceinline__ float sum_g(float x) {
  const unsigned u = __builtin_bit_cast(unsigned, x);
  const auto a = __builtin_amdgcn_permlane16_swap(u, u, false, false);
  x = __builtin_bit_cast(float, (unsigned)a[0]) + __builtin_bit_cast(float, (unsigned)a[1]);
  const unsigned w = __builtin_bit_cast(unsigned, x);
  const auto b = __builtin_amdgcn_permlane32_swap(w, w, false, false);
  return __builtin_bit_cast(float, (unsigned)b[0]) + __builtin_bit_cast(float, (unsigned)b[1]);
}
__device__ __forceinline__ double sum_g(double x) {
  auto swap16 = [](double y) {
    const unsigned lo = (unsigned)lo32(y), hi = (unsigned)hi32(y);
    const auto l = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    const auto h = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    return mk64((int)l[0], (int)h[0]) + mk64((int)l[1], (int)h[1]);
  };
  auto swap32 = [](double y) {
    const unsigned lo = (unsigned)lo32(y), hi = (unsigned)hi32(y);
    const auto l = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto h = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    return mk64((int)l[0], (int)h[0]) + mk64((int)l[1], (int)h[1]);
  };
  return swap32(swap16(x));
}
// Sum over the 16 lanes of a row group: partners j^1, j^2, j^7 (row half
// mirror) and 15-j (row mirror), which together reach every lane of the row.
__device__ __forceinline__ float sum_j(float x) {
  x += dppf<0xB1>(x);
  x += dppf<0x4E>(x);
  x += dppf<0x141>(x);
  return x + dppf<0x140>(x);
}
__device__ __forceinline__ double sum_j(double x) {
  x += dppd<0xB1>(x);
  x += dppd<0x4E>(x);
  x += dppd<0x141>(x);
  return x + dppd<0x140>(x);
}

// Transposed sum over the 16 lanes of a row group: on return lane j holds the
// row-group sum of p[j >> 2].  Butterfly steps with partners j^8 and j^7 (a
// lane keeps the half of its values selected by bit 3, then bit 2, of j and
// receives the same half from its partner), then j^2 and j^1 on the single
// value left; the masks 8, 7, 2, 1 are independent, so every lane of the row
// group is summed once.
// (Operands go through an empty asm first: a select between two elements of a
// private array would otherwise become a load from a computed address, which
// keeps the array out of registers.)
template <class V>
__device__ __forceinline__ void in_vgpr(V& x) {
  asm volatile("" : "+v"(x));
}
template <class V>
__device__ __forceinline__ V tsum4(const V (&p)[4], int j) {
  const bool b3 = j & 8, b2 = j & 4;
  V q[4] = {p[0], p[1], p[2], p[3]};
  static_for<0, 4>([&](auto t) { in_vgpr(q[t]); });
  V a[2];
  static_for<0, 2>([&](auto t) {
    const V send = b3 ? q[t] : q[t + 2], keep = b3 ? q[t + 2] : q[t];
    if constexpr (sizeof(V) == 8) a[t] = keep + dppd<0x128>(send); else a[t] = keep + dppf<0x128>(send);
  });
  in_vgpr(a[0]);
  in_vgpr(a[1]);
  const V send = b2 ? a[0] : a[1], keep = b2 ? a[1] : a[0];
  V c;
  if constexpr (sizeof(V) == 8) {
    c = keep + dppd<0x141>(send);
    c += dppd<0x4E>(c);
    c += dppd<0xB1>(c);
  } else {
    c = keep + dppf<0x141>(send);
    c += dppf<0x4E>(c);
    c += dppf<0xB1>(c);
  }
  return c;
}

// x[v] += bcast_K(x[v]) * s, v = 0..3, where bcast_K reads lane K of the
// lane's row of 16 (DPP row_newbcast:K): four v_fmac_f32 with the lane move as
// their DPP source operand.  The compiler folds DPP moves into adds and
// multiplies but not into FMAs, where the move would cost a second VALU
// instruction; the s_nop 1 covers the two wait states a DPP read needs after a
// VALU write of its source.  Same arithmetic as fmaf(bcast_K(x), s, x).
template <int K>
__device__ __forceinline__ void bcast_fmac4(v4f& x, float s);
#define TSDX_BCAST_FMAC4(K)                                                                        \
  template <>                                                                                      \
  __device__ __forceinline__ void bcast_fmac4<K>(v4f & x, float s) {                              \
    float x0 = x[0], x1 = x[1], x2 = x[2], x3 = x[3];                                              \
    asm volatile(                                                                                  \
        "s_nop 1\n\t"                                                                              \
        "v_fmac_f32_dpp %0, %0, %4 row_newbcast:" #K " row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t" \
        "v_fmac_f32_dpp %1, %1, %4 row_newbcast:" #K " row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t" \
        "v_fmac_f32_dpp %2, %2, %4 row_newbcast:" #K " row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t" \
        "v_fmac_f32_dpp %3, %3, %4 row_newbcast:" #K " row_mask:0xf bank_mask:0xf bound_ctrl:1"     \
        : "+v"(x0), "+v"(x1), "+v"(x2), "+v"(x3)                                                   \
        : "v"(s));                                                                                 \
    x[0] = x0;                                                                                     \
    x[1] = x1;                                                                                     \
    x[2] = x2;                                                                                     \
    x[3] = x3;                                                                                     \
  }
TSDX_BCAST_FMAC4(0)
TSDX_BCAST_FMAC4(1)
TSDX_BCAST_FMAC4(2)
TSDX_BCAST_FMAC4(3)
TSDX_BCAST_FMAC4(4)
TSDX_BCAST_FMAC4(5)
TSDX_BCAST_FMAC4(6)
TSDX_BCAST_FMAC4(7)
TSDX_BCAST_FMAC4(8)
TSDX_BCAST_FMAC4(9)
TSDX_BCAST_FMAC4(10)
TSDX_BCAST_FMAC4(11)
TSDX_BCAST_FMAC4(12)
TSDX_BCAST_FMAC4(13)
TSDX_BCAST_FMAC4(14)
TSDX_BCAST_FMAC4(15)
#undef TSDX_BCAST_FMAC4

// The same for the registers LO..3 only (LO = 1, 2, 3) or 0..HI only (HI =
// 0, 1, 2): x[v] += bcast_K(x[v]) * s for those v.
template <int K>
__device__ __forceinline__ void bcast_fmac1(float& x, float s) {
  asm volatile("s_nop 1\n\tv_fmac_f32_dpp %0, %0, %1 row_newbcast:%2 row_mask:0xf bank_mask:0xf bound_ctrl:1"
               : "+v"(x)
               : "v"(s), "n"(K));
}
template <int K, int LO, int HI>
__device__ __forceinline__ void bcast_fmac_range(v4f& x, float s) {
  static_assert(0 <= LO && LO <= HI && HI <= 3, "register range");
  if constexpr (LO == 0 && HI == 3) {
    bcast_fmac4<K>(x, s);
  } else {
    float y[4] = {x[0], x[1], x[2], x[3]};
    static_for<LO, HI + 1>([&](auto v) { bcast_fmac1<K>(y[v], s); });
    static_for<LO, HI + 1>([&](auto v) { x[v()] = y[v]; });
  }
}

// ------------------------------------------------- diagonal block elimination
// W: symmetric 16 x 16 block (accumulator layout).  Returns V = U^-1 where
// W = U^T U (U upper, entries rounded to the low format), and sets bad when a
// pivot is not positive.  The pivots are not tested step by step: a pivot
// that is not positive (or NaN) makes its 1/sqrt NaN or infinite, and every
// lane holds 1/U[j][j] of its column at the end, so one test of those 16
// values after the last step finds it (a NaN may then spread through the
// block, which is dropped with the matrix).
//
// With U = D T (D = diag(U), T unit upper, T[k][j] = W_k[k][j] / W_k[k][k]
// for the partly reduced W_k), step k does
//     W[i][j] -= W[i][k] T[k][j]              (the Schur update)
//     F[:,j]  -= T[k][j] F[:,k]   (j > k)     (F = identity at the start)
// which leaves F = T^-1; V = T^-1 D^-1 scales column j by 1/U[j][j] at the end.
//
// float (F32): both updates use T[k][j] masked to j > k, so column j of W is
// last changed at step j - 1 and keeps its pivot W_j[j][j] on the diagonal to
// the end; 1/U[j][j] = 1/sqrt(W[j][j]) is formed once after the last step
// instead of step by step, and each step needs one reciprocal of its pivot.
// Rows <= k of the columns > k still take the update and hold garbage, which
// is never read again (row k' is read at step k' only, right of its pivot).
// The low formats (BF16, F16) round every entry of U, which needs the square
// root at each step; they keep the unmasked Schur update of W.
template <Low LF>
__device__ __forceinline__ v4f diag_inverse(v4f W, int g, int j, bool& bad) {
  // fresh lane indices per call: the lane masks of the steps are formed where
  // they are used instead of being kept live across calls
  g = opaque(g);
  j = opaque(j);
  v4f F;
  static_for<0, 4>([&](auto v) { F[v()] = (4 * g + v == j) ? 1.0f : 0.0f; });
  float rdl = 1.0f;  // 1 / U[j][j] of this lane's column
  // Register v holds rows v, 4 + v, 8 + v, 12 + v.  Step k changes W only
  // where a later step reads it (rows > k: registers v > k - 12) and F only
  // where column k of F = T^-1 (upper triangular) can be nonzero (rows <= k:
  // registers v <= k); the skipped products are of rows no step reads again,
  // or exact zeros.  The float path's last step has no column right of its
  // pivot (its multipliers are all zero) and is left out.
  static_for<0, (LF == Low::F32 ? 15 : 16)>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    constexpr int kg = k / 4, kv = k % 4;
    constexpr int wlo = k > 12 ? k - 12 : 0, fhi = k < 3 ? k : 3;
    // (the element is copied to a scalar first: __builtin_bit_cast applied to
    // an element of a vector type reads element 0)
    const float wkv = W[kv];
    float nt;  // -T[k][j] (masked to j > k)
    const float akk = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, wkv), 16 * kg + k));
    const float wk = bperm(wkv, 16 * kg + j);  // W[k][j], from row k (lane 16 kg + j)
    if constexpr (LF == Low::F32) {
      nt = (j > k) ? wk * -__builtin_amdgcn_rcpf(akk) : 0.0f;
      bcast_fmac_range<k, wlo, 3>(W, nt);
    } else {  // U entries rounded to the low format
      const float rd = __builtin_amdgcn_rsqf(akk);
      const float uj = round_low(wk * rd, LF);
      nt = (j > k) ? uj * -rd : 0.0f;
      static_for<wlo, 4>([&](auto v) {
        const float ui = round_low(row_bcast<k>(W[v()]) * rd, LF);
        W[v()] = __builtin_fmaf(-ui, uj, W[v()]);
      });
      rdl = (j == k) ? rd : rdl;
    }
    bcast_fmac_range<k, 0, fhi>(F, nt);
  });
  if constexpr (LF == Low::F32) {
    // pivot j sits in register j % 4 of lane (j / 4, j)
    const float dv = (j & 3) == 0 ? W[0] : (j & 3) == 1 ? W[1] : (j & 3) == 2 ? W[2] : W[3];
    rdl = __builtin_amdgcn_rsqf(bperm(dv, 16 * (j >> 2) + j));
  }
  // every pivot positive and finite <=> every 1/sqrt positive and finite
  bad = bad || __builtin_amdgcn_ballot_w64(!(rdl > 0.0f && rdl < __builtin_huge_valf())) != 0;
  static_for<0, 4>([&](auto v) { F[v()] *= rdl; });
  return F;
}

// ------------------------------------------------------------------ kernel
// Loads the upper blocks of one matrix (see the file comment) into ad and
// returns the sum of squares of this lane's entries of the symmetric matrix.
// Every load is issued unconditionally (entries past n read a valid address,
// Am[0], and are replaced afterwards), so all of them are in flight at once:
// a load under a divergent branch is waited for before the branch closes,
// which would make the wave pay one memory latency per entry.  FULL
// (n = 16 NBLK) drops the padding selects.
// Off-diagonal block (I < J): lane (g, j) holds rows 16I + 4g + v, column
// 16J + j, i.e. A[16J + j][16I + 4g .. 4g + 3] (four consecutive entries of a
// lower row: two 16-byte loads when VEC, i.e. rows are 16-byte aligned).
// Diagonal blocks read A[max(r,c)][min(r,c)] entry by entry.
template <int NBLK, bool VEC, bool FULL>
__device__ __forceinline__ double mx_load_blocks(const double* __restrict__ Am, int n, int g, int j,
                                                 double (&ad)[nupper<NBLK>()][4]) {
  static_for<0, NBLK>([&](auto Ic) {
    static_for<Ic, NBLK>([&](auto Jc) {
      constexpr int I = decltype(Ic)::value, J = decltype(Jc)::value, b = bslot<NBLK>(I, J);
      const int c = 16 * J + j;
      if constexpr (I < J && VEC) {
        const int r0 = 16 * I + 4 * g;
        const bool ok = FULL || (r0 < n && c < n);  // r0 < n implies r0 + 3 < n (n % 4 == 0)
        const double2* p = reinterpret_cast<const double2*>(Am + (ok ? c * n + r0 : 0));  // < 4096
        const double2 q0 = p[0], q1 = p[1];
        ad[b][0] = ok ? q0.x : 0.0;
        ad[b][1] = ok ? q0.y : 0.0;
        ad[b][2] = ok ? q1.x : 0.0;
        ad[b][3] = ok ? q1.y : 0.0;
      } else {
        static_for<0, 4>([&](auto v) {
          const int r = 16 * I + 4 * g + v;
          const bool ok = FULL || (r < n && c < n);
          const int hi = r > c ? r : c, lo = r > c ? c : r;
          const double x = Am[ok ? hi * n + lo : 0];
          ad[b][v()] = ok ? x : ((r == c) ? 1.0 : 0.0);  // past n: identity padding
        });
      }
    });
  });
  // ||A||_F^2 of this lane's entries: diagonal blocks once, the others twice
  double sd = 0.0, so = 0.0;
  static_for<0, NBLK>([&](auto Ic) {
    static_for<Ic, NBLK>([&](auto Jc) {
      constexpr int I = decltype(Ic)::value, J = decltype(Jc)::value, b = bslot<NBLK>(I, J);
      static_for<0, 4>([&](auto v) {
        if constexpr (I == J) {  // the padding's unit diagonal is not part of A
          const double x = (FULL || 16 * I + 4 * g + v < n) ? ad[b][v()] : 0.0;
          sd = __builtin_fma(x, x, sd);
        } else {
          so = __builtin_fma(ad[b][v()], ad[b][v()], so);
        }
      });
    });
  });
  return __builtin_fma(2.0, so, sd);
}

// Loads the upper blocks of matrix m of a.A and returns ||A||_F of the
// symmetric matrix.  Only the lower triangle of A is read.
template <int NBLK>
__device__ __forceinline__ double mx_load(const IrArgs& a, int64_t m, int g, int j, double (&ad)[nupper<NBLK>()][4]) {
  const int n = a.n;
  const double* Am = a.A + m * (int64_t)n * n;
  const bool vec = (n % 4 == 0) && ((reinterpret_cast<uintptr_t>(a.A) & 15) == 0);
  double ss;
  if (vec && n == 16 * NBLK)
    ss = mx_load_blocks<NBLK, true, true>(Am, n, g, j, ad);
  else if (vec)
    ss = mx_load_blocks<NBLK, true, false>(Am, n, g, j, ad);
  else
    ss = mx_load_blocks<NBLK, false, false>(Am, n, g, j, ad);
  return __builtin_sqrt(sum_j(sum_g(ss)));
}

// Factor in the low format LF, refine every right-hand side, and record the
// outcome: X written for converged matrices, the others appended to the
// float64 fallback list; info 0 and stat (corrections, or -1) for all.
template <int NBLK, Low LF>
__device__ __forceinline__ void mx_solve(const IrArgs& a, int64_t m, int g, int j,
                                         const double (&ad)[nupper<NBLK>()][4], double anorm, double* xs, float* fs,
                                         v4f* wl) {
  constexpr int NU = nupper<NBLK>();
  const int lane = threadIdx.x;
  const int n = a.n;
  // ---- factor: float copy, block right-looking, MFMA trailing updates.
  // The blocks live in this wave's LDS (NU KiB), each lane owning its 16-byte
  // slot of every block (only that lane ever reads or writes it, so no
  // synchronisation); registers hold the float64 blocks, the pivot block's
  // inverse and one block row at a time, which keeps the kernel at 3 waves
  // per SIMD at n = 64.
  auto slot = [&](int b) -> v4f& { return wl[b * kWave + lane]; };
  static_for<0, NU>([&](auto b) {
    v4f x;
    static_for<0, 4>([&](auto v) { x[v()] = round_low((float)ad[b][v()], LF); });
    slot(b) = x;
  });
  bool bad = false;
  static_for<0, NBLK>([&](auto Kc) {
    constexpr int K = decltype(Kc)::value;
    const v4f V = diag_inverse<LF>(slot(bslot<NBLK>(K, K)), g, j, bad);
    slot(bslot<NBLK>(K, K)) = V;
    v4f row[NBLK];  // U_KJ, J > K
    static_for<K + 1, NBLK>([&](auto Jc) {
      constexpr int J = decltype(Jc)::value;
      const v4f wkj = slot(bslot<NBLK>(K, J));
      v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
      static_for<0, 4>([&](auto v) { acc = __builtin_amdgcn_mfma_f32_16x16x4f32(V[v()], wkj[v()], acc, 0, 0, 0); });
      static_for<0, 4>([&](auto v) { acc[v()] = round_low(acc[v()], LF); });
      row[J] = acc;
      slot(bslot<NBLK>(K, J)) = acc;
    });
    static_for<K + 1, NBLK>([&](auto Ic) {
      static_for<Ic, NBLK>([&](auto Jc) {
        constexpr int I = decltype(Ic)::value, J = decltype(Jc)::value;
        v4f acc = slot(bslot<NBLK>(I, J));
        static_for<0, 4>([&](auto v) {
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(-row[I][v()], row[J][v()], acc, 0, 0, 0);
        });
        slot(bslot<NBLK>(I, J)) = acc;
      });
    });
  });
  wave_sync();

  // ---- solves with the float factor: t (column form) -> z (column form)
  // one block, column form -> row form, through LDS
  auto c2r = [&](float c, float (&R)[4]) __attribute__((always_inline)) {
    if (g == 0) fs[j] = c;
    wave_sync();
    const float4 q = *reinterpret_cast<const float4*>(fs + 4 * g);
    R[0] = q.x; R[1] = q.y; R[2] = q.z; R[3] = q.w;
    wave_sync();
  };
  // source lane of entry j (column form) in the "lane j holds row 4g + (j >> 2)" form
  const int src_mixed = 16 * (j >> 2) + 4 * (j & 3);
  auto solve = [&](const float (&t)[NBLK], float (&z)[NBLK]) __attribute__((always_inline)) {
    float yc[NBLK], yR[NBLK][4];
    static_for<0, NBLK>([&](auto Kc) {
      constexpr int K = decltype(Kc)::value;
      float acc = 0.0f;
      static_for<0, K>([&](auto I) {
        const v4f blk = wl[bslot<NBLK>(I, K) * kWave + lane];
        static_for<0, 4>([&](auto v) { acc = __builtin_fmaf(blk[v()], yR[I][v()], acc); });
      });
      float sR[4];
      c2r(t[K] - sum_g(acc), sR);
      float p = 0.0f;
      const v4f dk = wl[bslot<NBLK>(K, K) * kWave + lane];
      static_for<0, 4>([&](auto v) { p = __builtin_fmaf(dk[v()], sR[v()], p); });
      yc[K] = sum_g(p);
      c2r(yc[K], yR[K]);
    });
    static_for_rev<0, NBLK>([&](auto Kc) {
      constexpr int K = decltype(Kc)::value;
      float wc = yc[K];  // column form
      if constexpr (K + 1 < NBLK) {
        float p[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        static_for<K + 1, NBLK>([&](auto J) {
          const v4f blk = wl[bslot<NBLK>(K, J) * kWave + lane];
          static_for<0, 4>([&](auto v) { p[v()] = __builtin_fmaf(blk[v()], z[J], p[v()]); });
        });
        wc -= bperm(tsum4(p, j), src_mixed);
      }
      float q[4];
      const v4f dk = wl[bslot<NBLK>(K, K) * kWave + lane];
      static_for<0, 4>([&](auto v) { q[v()] = dk[v()] * wc; });
      z[K] = bperm(tsum4(q, j), src_mixed);
    });
  };

  bool ok = !bad;
  int used = 0;
  const double* Bm = a.B + m * (int64_t)n * a.nrhs;
  double* Xm = a.X + m * (int64_t)n * a.nrhs;
  for (int c = 0; c < a.nrhs && ok; ++c) {
    double bc[NBLK], xc[NBLK];
    float t[NBLK], z[NBLK];
    double bb = 0.0;
    static_for<0, NBLK>([&](auto J) {
      const int row = 16 * J + j;
      const double bv = Bm[row < n ? (int64_t)row * a.nrhs + c : 0];  // unconditional: loads overlap
      bc[J] = row < n ? bv : 0.0;
      bb = __builtin_fma(bc[J], bc[J], bb);
      t[J] = (float)bc[J];
    });
    const double bb2 = sum_j(bb);  // ||b||^2
    solve(t, z);
    static_for<0, NBLK>([&](auto J) { xc[J] = (16 * J + j < n) ? (double)z[J] : 0.0; });
    bool conv = false;
    double prev = __builtin_huge_val();
    // The test ||r|| <= tol/2 (||A||_F ||x|| + ||b||) on squares, with no
    // square root: where it can pass, r is tiny and ||A||_F ||x|| >= ||A x||
    // >= ||b|| - ||r||, so (||A|| ||x|| + ||b||)^2 >= ||A||^2 ||x||^2 +
    // 3 ||b||^2 (to rounding), and ||r||^2 <= tol^2/4 (||A||^2 ||x||^2 +
    // 3 ||b||^2) implies the test.
    const double ctol = 0.25 * a.tol * a.tol, aa2 = anorm * anorm, bb3 = 3.0 * bb2;
    for (int it = 0; it <= a.max_iter; ++it) {
      // r = b - A x in float64 with the stored upper blocks.  A diagonal block
      // is symmetric, so its product is taken as column sums (a sum over the
      // lane's four rows and then over the four row groups); an off-diagonal
      // block serves both as A_IJ (row sums over the 16 lanes of a row group)
      // and as A_JI = A_IJ^T (column sums).
      if (g == 0) static_for<0, NBLK>([&](auto J) { xs[16 * J + j] = xc[J]; });
      wave_sync();
      double pc[NBLK], rc[NBLK];
      static_for<0, NBLK>([&](auto J) { pc[J] = 0.0; });
      static_for<0, NBLK>([&](auto Ic) {
        constexpr int I = decltype(Ic)::value;
        const double2 q0 = *reinterpret_cast<const double2*>(xs + 16 * I + 4 * g);
        const double2 q1 = *reinterpret_cast<const double2*>(xs + 16 * I + 4 * g + 2);
        const double xR[4] = {q0.x, q0.y, q1.x, q1.y};  // row form of block I
        static_for<0, 4>([&](auto v) { pc[I] = __builtin_fma(ad[bslot<NBLK>(I, I)][v()], xR[v], pc[I]); });
        if constexpr (I + 1 < NBLK) {
          double p[4] = {0.0, 0.0, 0.0, 0.0};
          static_for<I + 1, NBLK>([&](auto Jc) {
            constexpr int J = decltype(Jc)::value, b = bslot<NBLK>(I, J);
            static_for<0, 4>([&](auto v) {
              p[v] = __builtin_fma(ad[b][v()], xc[J], p[v]);
              pc[J] = __builtin_fma(ad[b][v()], xR[v], pc[J]);
            });
          });
          rc[I] = bperm(tsum4(p, j), src_mixed);  // rows of block I right of the diagonal, column form
        } else {
          rc[I] = 0.0;
        }
      });
      wave_sync();
      double rr = 0.0, xx = 0.0;
      static_for<0, NBLK>([&](auto J) {
        rc[J] = (16 * J + j < n) ? bc[J] - rc[J] - sum_g(pc[J]) : 0.0;
        rr = __builtin_fma(rc[J], rc[J], rr);
        xx = __builtin_fma(xc[J], xc[J], xx);
      });
      const double rn2 = sum_j(rr), xn2 = sum_j(xx);
      if (rn2 <= ctol * __builtin_fma(aa2, xn2, bb3)) {
        conv = true;
        used = used > it ? used : it;
        break;
      }
      if (it == a.max_iter || !(rn2 < 0.25 * prev)) break;
      prev = rn2;
      static_for<0, NBLK>([&](auto J) { t[J] = (float)rc[J]; });
      solve(t, z);
      static_for<0, NBLK>([&](auto J) { xc[J] += (16 * J + j < n) ? (double)z[J] : 0.0; });
    }
    ok = conv;
    if (conv && g == 0) {
      // the store offsets are formed here (opaque lane index): shared with
      // the loads of B they would stay live through the refinement and spill
      const int jj = opaque(j);
      static_for<0, NBLK>([&](auto J) {
        const int row = 16 * J + jj;
        if (row < n) Xm[(int64_t)row * a.nrhs + c] = xc[J];
      });
    }
  }
  if (lane == 0) {
    if (!ok) {
      const int slot = atomicAdd(a.fcount, 1);
      a.flist[slot] = (int)m;
    }
    if (a.info) a.info[m] = 0;
    if (a.stat) a.stat[m] = ok ? used : -1;
  }
}

// The same over the matrices of a list (the routed solve's bfloat16 tier):
// workgroups stride over llist[0 .. *lcount).
template <int NBLK, Low LF>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(3))) void k_posv_mx_list(IrArgs a) {
  __shared__ __align__(16) double xs[16 * NBLK];
  __shared__ __align__(16) float fs[16];
  __shared__ v4f wl[nupper<NBLK>() * kWave];
  const int lane = threadIdx.x, g = lane >> 4, j = lane & 15;
  const int count = *a.lcount;
  for (int w = blockIdx.x; w < count; w += gridDim.x) {  // uniform
    const int64_t m = a.llist[w];
    double ad[nupper<NBLK>()][4];
    const double anorm = mx_load<NBLK>(a, m, g, j, ad);
    mx_solve<NBLK, LF>(a, m, g, j, ad, anorm, xs, fs, wl);
    wave_sync();  // the next matrix reuses the LDS buffers
  }
}

template <int NBLK, Low LF>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(3))) void k_posv_mx(IrArgs a) {
  __shared__ __align__(16) double xs[16 * NBLK];
  __shared__ __align__(16) float fs[16];
  __shared__ v4f wl[nupper<NBLK>() * kWave];
  const int lane = threadIdx.x, g = lane >> 4, j = lane & 15;
  const int64_t m = a.m0 + blockIdx.x;
  double ad[nupper<NBLK>()][4];
  const double anorm = mx_load<NBLK>(a, m, g, j, ad);
  mx_solve<NBLK, LF>(a, m, g, j, ad, anorm, xs, fs, wl);
}

}  // namespace tsdx_hip
