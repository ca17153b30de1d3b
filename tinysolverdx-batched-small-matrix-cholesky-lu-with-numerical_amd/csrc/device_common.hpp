// Device-side building blocks shared by every kernel of the library.
//
// Work mapping used throughout ("row-per-lane groups"): a wave of 64 lanes is
// split into G = 64/NB groups of NB lanes; group g owns one matrix and lane i of
// the group holds row i of that matrix in registers (T r[NB], column index
// static after unrolling).  NB is a compile-time bucket (4, 8, 16, 32, 64); the
// runtime size n <= NB.  Lanes i >= n and groups past the end of the batch
// still execute every step (so wave-wide shuffles and barriers stay uniform)
// but never load or store.
//
// Cross-lane traffic inside a group:
//   * broadcast of lane k (compile-time k): v_readlane for NB = 64,
//     ds_swizzle (bit mode, and/or masks) for NB <= 32;
//   * butterfly reductions: DPP quad_perm / row_half_mirror / row_mirror for
//     the xor masks 1, 2, 7, 15, ds_swizzle for xor 16, ds_bpermute for xor 32;
//   * row / column broadcasts of a whole vector: LDS, one buffer per wave.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>
#include <utility>

namespace tsdx_hip {

constexpr int kWave = 64;

template <int I>
using ic = std::integral_constant<int, I>;

// Calls f(ic<B>), f(ic<B+1>), ..., f(ic<E-1>): indices stay constant
// expressions, so register arrays indexed by them never touch scratch.
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(ic<B>{});
    static_for<B + 1, E>(f);
  }
}
// Calls f(ic<E-1>), ..., f(ic<B>).
template <int B, int E, class F>
__device__ __forceinline__ void static_for_rev(F&& f) {
  if constexpr (B < E) {
    f(ic<E - 1>{});
    static_for_rev<B, E - 1>(f);
  }
}

// f(ic<j>) for the columns j in [B, E), 16 at a time; with CH, a run of 16
// columns that starts at or past `bound` (wave-uniform) is skipped with one
// scalar branch.
template <int B, int E, bool CH, class F>
__device__ __forceinline__ void static_for_cols(int bound, F&& f) {
  if constexpr (!CH) {
    static_for<B, E>(f);
  } else {
    static_for<B / 16, (E + 15) / 16>([&](auto cc) {
      constexpr int c = decltype(cc)::value;
      constexpr int lo = 16 * c > B ? 16 * c : B, hi = 16 * c + 16 < E ? 16 * c + 16 : E;
      if (16 * c < bound) static_for<lo, hi>(f);
    });
  }
}

// ------------------------------------------------------------- bit helpers
__device__ __forceinline__ int lo32(double x) { return (int)(__builtin_bit_cast(unsigned long long, x) & 0xffffffffull); }
__device__ __forceinline__ int hi32(double x) { return (int)(__builtin_bit_cast(unsigned long long, x) >> 32); }
__device__ __forceinline__ double mk64(int lo, int hi) {
  return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// ---------------------------------------------------- broadcast of lane K
template <int NB, int K>
__device__ __forceinline__ int bcast_i32(int x) {
  static_assert(K < NB, "lane outside group");
  if constexpr (NB == 64) {
    return __builtin_amdgcn_readlane(x, K);
  } else {
    // ds_swizzle bit mode: src = (lane & and_mask) | or_mask within 32 lanes
    return __builtin_amdgcn_ds_swizzle(x, (0x1F & ~(NB - 1)) | (K << 5));
  }
}
template <int NB, int K>
__device__ __forceinline__ float bcast(float x) {
  return __builtin_bit_cast(float, bcast_i32<NB, K>(__builtin_bit_cast(int, x)));
}
template <int NB, int K>
__device__ __forceinline__ double bcast(double x) {
  return mk64(bcast_i32<NB, K>(lo32(x)), bcast_i32<NB, K>(hi32(x)));
}

// Broadcast from a lane chosen at run time (any lane of the wave).
__device__ __forceinline__ int bcast_dyn_i32(int x, int src_lane) {
  return __builtin_amdgcn_ds_bpermute(src_lane << 2, x);
}
__device__ __forceinline__ float bcast_dyn(float x, int src_lane) {
  return __builtin_bit_cast(float, bcast_dyn_i32(__builtin_bit_cast(int, x), src_lane));
}
__device__ __forceinline__ double bcast_dyn(double x, int src_lane) {
  return mk64(bcast_dyn_i32(lo32(x), src_lane), bcast_dyn_i32(hi32(x), src_lane));
}

// ------------------------------------------------------ xor shuffles
// M in {1, 2, 7, 15, 16, 32}: lane l receives lane l ^ M.  The first four are
// DPP moves inside rows of 16 (every lane reads a lane of its own row, so
// bound_ctrl changes nothing and lets the compiler fold the move into the
// instruction that consumes it).
template <int M>
__device__ __forceinline__ int xshfl_i32(int x) {
  if constexpr (M == 1) {
    return __builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, true);  // quad_perm [1,0,3,2]
  } else if constexpr (M == 2) {
    return __builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, true);  // quad_perm [2,3,0,1]
  } else if constexpr (M == 7) {
    return __builtin_amdgcn_mov_dpp(x, 0x141, 0xF, 0xF, true);  // row_half_mirror
  } else if constexpr (M == 15) {
    return __builtin_amdgcn_mov_dpp(x, 0x140, 0xF, 0xF, true);  // row_mirror
  } else if constexpr (M == 16) {
    return __builtin_amdgcn_ds_swizzle(x, 0x401F);  // bit mode, xor_mask 16
  } else {
    static_assert(M == 32, "unsupported xor mask");
    return __builtin_amdgcn_ds_bpermute((int)((__lane_id() ^ 32u) << 2), x);
  }
}
template <int M>
__device__ __forceinline__ float xshfl(float x) {
  return __builtin_bit_cast(float, xshfl_i32<M>(__builtin_bit_cast(int, x)));
}
template <int M>
__device__ __forceinline__ double xshfl(double x) {
  return mk64(xshfl_i32<M>(lo32(x)), xshfl_i32<M>(hi32(x)));
}
template <int M>
__device__ __forceinline__ int xshfl(int x) {
  return xshfl_i32<M>(x);
}

// op(x of lane l, x of lane l ^ M) for M = 16, 32 on the VALU:
// v_permlane16_swap / v_permlane32_swap of x with itself gives every lane its
// own value and its partner's (which of the two comes first depends on the
// lane), so for a commutative op every lane gets op(own, partner).
template <int M>
__device__ __forceinline__ void xpair_u32(unsigned x, unsigned& a, unsigned& b) {
  static_assert(M == 16 || M == 32, "row pairs");
  if constexpr (M == 16) {
    const auto p = __builtin_amdgcn_permlane16_swap(x, x, false, false);
    a = p[0];
    b = p[1];
  } else {
    const auto p = __builtin_amdgcn_permlane32_swap(x, x, false, false);
    a = p[0];
    b = p[1];
  }
}
template <int M, class V, class Op>
__device__ __forceinline__ V xop(V x, Op op) {
  if constexpr (sizeof(V) == 4) {
    unsigned a, b;
    xpair_u32<M>(__builtin_bit_cast(unsigned, x), a, b);
    return op(__builtin_bit_cast(V, a), __builtin_bit_cast(V, b));
  } else {
    static_assert(sizeof(V) == 8, "32- or 64-bit values");
    unsigned la, lb, ha, hb;
    xpair_u32<M>((unsigned)lo32(__builtin_bit_cast(double, x)), la, lb);
    xpair_u32<M>((unsigned)hi32(__builtin_bit_cast(double, x)), ha, hb);
    return op(__builtin_bit_cast(V, mk64((int)la, (int)ha)), __builtin_bit_cast(V, mk64((int)lb, (int)hb)));
  }
}

// Butterfly all-reduce over the NB lanes of a group.  The xor masks
// {1, 2, 7, 15, 16, 32} restricted to the group span every lane index of the
// group, so after the last step every lane holds the reduction of the group.
// op must be commutative (every caller's is: sums, and minima / maxima of
// keys that hold no NaN), which the steps across rows of 16 rely on.
template <int NB, class V, class Op>
__device__ __forceinline__ V group_allreduce(V x, Op op) {
  static_assert(NB >= 4 && NB <= 64 && (NB & (NB - 1)) == 0, "group size");
  x = op(x, xshfl<1>(x));
  x = op(x, xshfl<2>(x));
  if constexpr (NB >= 8) x = op(x, xshfl<7>(x));
  if constexpr (NB >= 16) x = op(x, xshfl<15>(x));
  if constexpr (NB >= 32) x = xop<16>(x, op);
  if constexpr (NB >= 64) x = xop<32>(x, op);
  return x;
}

// ------------------------------------------------------------ arithmetic
// Fused multiply-add, one rounding (the reference's contracted `x -= y*z`).
__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }
// IEEE square root (correctly rounded; the fp32 form relies on hipcc's default
// -fhip-fp32-correctly-rounded-divide-sqrt).
__device__ __forceinline__ float sqrtT(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double sqrtT(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float absT(float x) { return __builtin_fabsf(x); }
__device__ __forceinline__ double absT(double x) { return __builtin_fabs(x); }
template <class T>
__device__ __forceinline__ T infT() {
  return (T)__builtin_huge_val();
}

// Pivot search of partial pivoting: whether this lane holds the largest key of
// its group of NB lanes.  The key of a candidate row is |a| = v (a NaN counts
// as +inf when nan_is_inf - the NaN at the current position - else as -1),
// a row that is not a candidate -2.  For float the keys are compared as
// 32-bit integers: the bits of a float that is not negative order like the
// float, and -1, -2 are integers below every such key, so each step of the
// reduction is a single DPP integer maximum.
template <int NB, class T>
__device__ __forceinline__ int pivot_key_i32(T v, bool cand, bool nan_is_inf) {
  static_assert(sizeof(T) == 4, "float keys");
  int key = (v != v) ? (nan_is_inf ? 0x7f800000 : -1) : __builtin_bit_cast(int, v);
  return cand ? key : -2;
}
template <int NB, class T>
__device__ __forceinline__ bool holds_max_key(T v, bool cand, bool nan_is_inf) {
  if constexpr (sizeof(T) == 4) {
    const int key = pivot_key_i32<NB>(v, cand, nan_is_inf);
    const int mx = group_allreduce<NB>(key, [](int a, int b) { return a > b ? a : b; });
    return cand && key == mx;
  } else {
    T key = (v != v) ? (nan_is_inf ? infT<T>() : T(-1)) : v;
    key = cand ? key : T(-2);
    const T mx = group_allreduce<NB>(key, [](T a, T b) { return a > b ? a : b; });
    return cand && key == mx;
  }
}


// Returns x unchanged but hides its value from the optimiser.  Used on the
// lane index at the top of the right-hand-side chunk loops: otherwise LLVM
// hoists the ~3*NB lane masks of the substitution passes (i == k, i < k, ...)
// out of the loop and spills them, which costs more than recomputing them.
__device__ __forceinline__ int opaque(int x) {
  asm volatile("" : "+v"(x));
  return x;
}

// Waves per workgroup for a size bucket.  Every wave works on its own
// matrices with its own LDS region, so this only changes how waves are packed
// into workgroups.  Measured on MI355X: four-wave workgroups are faster for
// the small buckets (many short waves), one-wave workgroups for n > 16.
template <int NB>
constexpr int waves_per_block() {
  return NB <= 16 ? 4 : 1;
}

// Wave-local LDS hand-off: orders this wave's LDS writes before its later LDS
// reads (other lanes of the same wave read what a lane wrote).  The fences stop
// the compiler from moving LDS accesses across; the wait makes the writes land.
// No s_barrier: the other waves of the workgroup are never involved.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0), vmcnt/expcnt untouched
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A row of NB zeros in LDS that starts 32 bytes (8 of the 32 four-byte banks)
// past `row`, modulo 128 bytes: lanes that read it in the same instruction as
// other lanes read `row` (or, for 16-element float rows, the next group's row,
// 64 bytes on) at the same element then never meet in a bank.  zraw holds
// NB + 128 / sizeof(T) elements; every wave writes the zeros and waits for its
// own writes.
template <class T, int NB>
__device__ __forceinline__ const T* zero_row_beside(T* zraw, const T* row) {
  constexpr int total = NB + 128 / (int)sizeof(T);
  for (int e = (int)(threadIdx.x % kWave); e < total; e += kWave) zraw[e] = T(0);
  wave_sync();
  const int off = ((int)((uintptr_t)row - (uintptr_t)zraw) + 32) & 127;
  return reinterpret_cast<const T*>(reinterpret_cast<const char*>(zraw) + off);
}


// ------------------------------------------------------- row load / store
template <class T>
struct Vec16;
template <>
struct Vec16<float> {
  using type = float4;
  static constexpr int N = 4;
};
template <>
struct Vec16<double> {
  using type = double2;
  static constexpr int N = 2;
};

// A 16-byte vector of T that needs only T's alignment: global loads and
// stores of whole 16-byte chunks at any element offset (the gfx9 memory path
// takes them; the compiler itself emits such accesses for element-aligned
// vectors).
template <class T>
using VecU = T __attribute__((ext_vector_type(Vec16<T>::N), aligned(sizeof(T))));

// Loads r[0..n-1] from p (contiguous); other entries are zero.  CHUNKED: as
// 16-byte chunks at the row's own element offset, the partial last chunk
// element by element, whatever the alignment (a row read one element per
// instruction makes each load touch as many lines as the wave has rows,
// which cost 2-2.5x at n = 49 or 57).  Otherwise aligned chunks when vec (p
// 16-byte aligned, the row whole chunks), else element by element: the posv
// and gesv kernels keep that form, which the compiler fits in fewer registers
// there (gesv fp32 n=32: 61 against 89 VGPRs, posv 81 against 87).
template <class T, int NB, bool CHUNKED = true>
__device__ __forceinline__ void load_row(T (&r)[NB], const T* __restrict__ p, int n, bool ok, bool vec) {
  static_for<0, NB>([&](auto j) { r[j] = T(0); });
  if (!ok) return;
  constexpr int VE = Vec16<T>::N;
  if constexpr (!CHUNKED && NB % VE == 0) {
    if (vec) {
      using V = typename Vec16<T>::type;
      static_for<0, NB / VE>([&](auto q) {
        if (q * VE < n) {
          const V v = *reinterpret_cast<const V*>(p + q * VE);
          if constexpr (VE == 4) {
            r[q * 4 + 0] = v.x; r[q * 4 + 1] = v.y; r[q * 4 + 2] = v.z; r[q * 4 + 3] = v.w;
          } else {
            r[q * 2 + 0] = v.x; r[q * 2 + 1] = v.y;
          }
        }
      });
      return;
    }
  }
  if constexpr (CHUNKED && NB % VE == 0) {
    static_for<0, NB / VE>([&](auto q) {
      if (q * VE + VE <= n) {  // n is uniform: no divergence
        const VecU<T> v = *reinterpret_cast<const VecU<T>*>(p + q * VE);
        static_for<0, VE>([&](auto e) { r[q * VE + e] = v[e()]; });
      }
    });
    if (n % VE != 0) {  // the partial last chunk, element by element
      const int q0 = n - n % VE;
      static_for<0, NB>([&](auto j) {
        if (j >= q0 && j < n) r[j] = p[j];
      });
    }
    return;
  }
  static_for<0, NB>([&](auto j) {
    if (j < n) r[j] = p[j];
  });
}

// Stores r[0..n-1] to p, chunked like load_row; with lower_only, entries
// j > i are written as zero.
template <class T, int NB, bool CHUNKED = true>
__device__ __forceinline__ void store_row(const T (&r)[NB], T* __restrict__ p, int n, bool vec, bool lower_only, int i) {
  constexpr int VE = Vec16<T>::N;
  auto val = [&](int j, T x) { return (lower_only && j > i) ? T(0) : x; };
  if constexpr (!CHUNKED && NB % VE == 0) {
    if (vec) {
      using V = typename Vec16<T>::type;
      static_for<0, NB / VE>([&](auto q) {
        if (q * VE < n) {
          V v;
          if constexpr (VE == 4) {
            v.x = val(q * 4 + 0, r[q * 4 + 0]); v.y = val(q * 4 + 1, r[q * 4 + 1]);
            v.z = val(q * 4 + 2, r[q * 4 + 2]); v.w = val(q * 4 + 3, r[q * 4 + 3]);
          } else {
            v.x = val(q * 2 + 0, r[q * 2 + 0]); v.y = val(q * 2 + 1, r[q * 2 + 1]);
          }
          *reinterpret_cast<V*>(p + q * VE) = v;
        }
      });
      return;
    }
  }
  if constexpr (CHUNKED && NB % VE == 0) {
    static_for<0, NB / VE>([&](auto q) {
      if (q * VE + VE <= n) {
        VecU<T> v;
        static_for<0, VE>([&](auto e) { v[e()] = val(q * VE + e, r[q * VE + e]); });
        *reinterpret_cast<VecU<T>*>(p + q * VE) = v;
      }
    });
    if (n % VE != 0) {
      const int q0 = n - n % VE;
      static_for<0, NB>([&](auto j) {
        if (j >= q0 && j < n) p[j] = val(j, r[j]);
      });
    }
    return;
  }
  static_for<0, NB>([&](auto j) {
    if (j < n) p[j] = val(j, r[j]);
  });
}

}  // namespace tsdx_hip
