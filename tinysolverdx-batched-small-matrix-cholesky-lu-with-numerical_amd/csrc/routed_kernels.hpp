// The routed posv kernel (no reference counterpart): the learned router's
// features and network in the wave that then solves the matrix in the tier
// the network picks.  The network on the matrix cores is in router_net.hpp,
// the fp32 tier in mfma_kernels.hpp.
#pragma once

#include "mfma_kernels.hpp"
#include "router_layout.hpp"
#include "router_net.hpp"

namespace tsdx_hip {

// ------------------------------------------------------------ in-wave router
// Reductions over the 64 lanes of the wave, several in lockstep (independent
// instructions between each lane move and its use), each result read into a
// scalar register: rows of 16 by DPP butterflies (partners j^1, j^2, j^7,
// 15-j: every lane of a row ends with its row's value), then row_bcast:15
// (rows 1 and 3 take lane 15 of the row before) and row_bcast:31 (rows 2 and
// 3 take lane 31), after which lane 63 holds the whole wave's.  Rows that a
// broadcast does not write receive the identity of the operation.
// Minima and maxima run on 32-bit integer keys: a float's bits compare like
// the float as signed integers when the float is not negative, and
// skey(x) = bits ^ ((bits >> 31) & 0x7fffffff) orders every non-NaN float
// (skey is its own inverse).  An integer minimum or maximum folds the lane
// move into one instruction; a float one would add a NaN quieting of its
// operands.  Kinds: 0 integer minimum, 1 integer maximum, 2 float sum.
__device__ __forceinline__ int skey(float x) {
  const int b = __builtin_bit_cast(int, x);
  return b ^ ((b >> 31) & 0x7fffffff);
}
__device__ __forceinline__ float skey_inv(int k) { return __builtin_bit_cast(float, k ^ ((k >> 31) & 0x7fffffff)); }

template <int NV>
__device__ __forceinline__ void wave_reduce_lockstep(int (&v)[NV], const int (&kind)[NV]) {
  auto op = [&](int q, int x, int y) {
    if (kind[q] == 0) return x < y ? x : y;
    if (kind[q] == 1) return x > y ? x : y;
    return __builtin_bit_cast(int, __builtin_bit_cast(float, x) + __builtin_bit_cast(float, y));
  };
  auto ident = [&](int q) { return kind[q] == 0 ? 0x7fffffff : kind[q] == 1 ? (int)0x80000000 : 0; };
  static_for<0, NV>([&](auto q) { v[q] = op(q, v[q], dpp_i32<0xB1>(v[q])); });
  static_for<0, NV>([&](auto q) { v[q] = op(q, v[q], dpp_i32<0x4E>(v[q])); });
  static_for<0, NV>([&](auto q) { v[q] = op(q, v[q], dpp_i32<0x141>(v[q])); });
  static_for<0, NV>([&](auto q) { v[q] = op(q, v[q], dpp_i32<0x140>(v[q])); });
  static_for<0, NV>([&](auto q) { v[q] = op(q, v[q], __builtin_amdgcn_update_dpp(ident(q), v[q], 0x142, 0xA, 0xF, false)); });
  static_for<0, NV>([&](auto q) { v[q] = op(q, v[q], __builtin_amdgcn_update_dpp(ident(q), v[q], 0x143, 0xC, 0xF, false)); });
  static_for<0, NV>([&](auto q) { v[q] = __builtin_amdgcn_readlane(v[q], 63); });
}

// Lane Q of f set to the wave-uniform x.
template <int Q>
__device__ __forceinline__ void put_lane(float& f, float x) {
  static_assert(Q >= 0 && Q < 64, "lane");
  f = (int)threadIdx.x == Q ? x : f;
}

// The 12 router features (definitions in router_kernels.hpp, k_router_features)
// of the symmetric matrix held in ad; on return lane q < 12 holds feature q and
// the other lanes hold 0.
// |a| sums per row r_i (diagonal left out) without moving the blocks: a stored
// block (I, J), I <= J, adds its column sums (over the lane's four rows, then
// the four row groups) to rows 16 J + j, and, for I < J, its row sums (over
// the 16 lanes of a row group: transposed butterfly) to rows 16 I + 4 g + v;
// by symmetry the two cover every entry of every row once.  The per-row work
// then runs with one row per lane (lane (g, j): row 16 g + j).
//   dl: 64 floats of LDS (the diagonal).
template <int NBLK>
__device__ __forceinline__ float mx_features(const double (&ad)[nupper<NBLK>()][4], int n, int g, int j, double anorm,
                                             float* dl) {
  // block row by block row: the row sums of block row I are complete after
  // its blocks and are reduced at once (few registers live next to ad)
  const int src_mixed = 16 * (j >> 2) + 4 * (j & 3);
  double pc[NBLK];
  float rr[NBLK];  // row 16 I + j (column form): |a| right of the diagonal block
  static_for<0, NBLK>([&](auto J) { pc[J] = 0.0; });
  static_for<0, NBLK>([&](auto Ic) {
    constexpr int I = decltype(Ic)::value;
    double p[4] = {0.0, 0.0, 0.0, 0.0};
    static_for<I, NBLK>([&](auto Jc) {
      constexpr int J = decltype(Jc)::value, b = bslot<NBLK>(I, J);
      static_for<0, 4>([&](auto v) {
        const double x = __builtin_fabs(ad[b][v()]);
        pc[J] += x;
        if constexpr (I < J) p[v] += x;
      });
    });
    if constexpr (I + 1 < NBLK) {
      const float q[4] = {(float)p[0], (float)p[1], (float)p[2], (float)p[3]};
      rr[I] = bperm(tsum4(q, j), src_mixed);
    } else {
      rr[I] = 0.0f;
    }
  });
  // the diagonal: a_{16J+j} is register j % 4 of lane (j / 4, j) in block
  // (J, J); every lane selects its register j % 4 (no branch), the lanes with
  // g = j / 4 write it
  {
    const int vj = j & 3;
    float dv[NBLK];
    static_for<0, NBLK>([&](auto J) {
      constexpr int b = bslot<NBLK>(J, J);
      const double lo = (vj & 1) ? ad[b][1] : ad[b][0], hi = (vj & 1) ? ad[b][3] : ad[b][2];
      dv[J] = (float)((vj & 2) ? hi : lo);
    });
    if (g == (j >> 2)) static_for<0, NBLK>([&](auto J) { dl[16 * J + j] = dv[J]; });
  }
  float rowsum = 0.0f;  // sum_c |a_rc| of this lane's row r = 16 g + j, diagonal included
  static_for<0, NBLK>([&](auto J) {
    const float s = sum_g((float)pc[J]) + rr[J];
    rowsum = (g == J) ? s : rowsum;
  });
  wave_sync();
  const int row = 16 * g + j;
  const bool act = row < n;
  const float d = dl[act ? row : 0];
  wave_sync();
  const float r = rowsum - __builtin_fabsf(d);
  const float big = 3.0e38f;
  const float ra = r + __builtin_fabsf(d);
  // counts on the scalar unit
  const float nonpos = __builtin_amdgcn_ballot_w64(act && !(d > 0.0f)) != 0 ? 1.0f : 0.0f;
  const float ddom = (float)__builtin_popcountll(__builtin_amdgcn_ballot_w64(act && d > r));
  // min d, max |d|, max (|d| + r), min (d - r), max r / d, sum d, sum log10 d
  int v[7] = {act ? skey(d) : skey(big),
              act ? __builtin_bit_cast(int, __builtin_fabsf(d)) : 0,
              act ? __builtin_bit_cast(int, ra) : 0,
              act ? skey(d - r) : skey(big),
              __builtin_bit_cast(int, act ? (d > 0.0f ? r * __builtin_amdgcn_rcpf(d) : big) : 0.0f),
              __builtin_bit_cast(int, act ? d : 0.0f),
              __builtin_bit_cast(int, act ? flog10(d) : 0.0f)};
  constexpr int kind[7] = {0, 1, 1, 0, 1, 2, 2};
  wave_reduce_lockstep(v, kind);
  const float dmin = skey_inv(v[0]), dmax = __builtin_bit_cast(float, v[1]), ninf = __builtin_bit_cast(float, v[2]);
  const float gers = skey_inv(v[3]), dom = __builtin_bit_cast(float, v[4]);
  const float dsum = __builtin_bit_cast(float, v[5]), lsum = __builtin_bit_cast(float, v[6]);
  const float fr = (float)anorm, rn = __builtin_amdgcn_rcpf((float)n);
  // one logarithm instruction for the seven logarithmic features: argument q
  // in lane q, 1 (log 0) in every other lane
  float arg = 1.0f;
  put_lane<1>(arg, dmin);
  put_lane<2>(arg, dmax);
  put_lane<3>(arg, fr);
  put_lane<4>(arg, ninf);
  put_lane<6>(arg, dom);
  put_lane<7>(arg, dmax * __builtin_amdgcn_rcpf(dmin));
  put_lane<8>(arg, fr * (float)n * __builtin_amdgcn_rcpf(dsum));
  const float lg = flog10(arg);
  auto lane_of = [](float f, int q) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, f), q));
  };
  float f = lg;  // features 1 2 3 4, and 0 past feature 11
  put_lane<0>(f, (float)n * (1.0f / 64.0f));
  float f5 = fr > 0.0f ? gers * __builtin_amdgcn_rcpf(fr) : 0.0f;
  f5 = f5 < -30.0f ? -30.0f : f5;
  put_lane<5>(f, f5);
  if (dom >= big) put_lane<6>(f, 30.0f);
  if (!(dmin > 0.0f)) put_lane<7>(f, 30.0f);
  if (!(dsum > 0.0f)) put_lane<8>(f, 30.0f);
  put_lane<9>(f, lsum * rn - lane_of(lg, 2));
  put_lane<10>(f, nonpos);
  put_lane<11>(f, ddom * rn);
  return f;
}

// Routed posv: features and network in the wave that then solves the matrix
// (A is read from HBM once): tier 1 factors in fp32 and refines here, tier 0
// goes to the bfloat16 list (a.llist, solved by k_posv_mx_list: with both
// refinement tiers inline the kernel would not fit its registers), tier 2 to
// the float64 list (stat -1).  prep: the network
// prepared for the matrix cores (router_net.hpp); tier_out / feat_out
// (optional) receive the choice and the features.
template <int NBLK>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(3))) void k_posv_mx_routed(
    IrArgs a, const void* __restrict__ prep, int* __restrict__ tier_out, float* __restrict__ feat_out) {
  __shared__ __align__(16) double xs[16 * NBLK];
  __shared__ __align__(16) float fs[16];
  // The diagonal (features) and the network's activations are dead before
  // mx_solve writes the factor, so they live in the factor's LDS: at n > 16
  // the kernel then needs no more LDS than k_posv_mx, and 12 workgroups (3 per
  // SIMD, the register limit) fit the 160 KiB of a CU.
  constexpr int kScratchVecs = (sizeof(float) * 64 + sizeof(__bf16) * kNetLdsHalves + sizeof(v4f) - 1) / sizeof(v4f);
  constexpr int kWlVecs = nupper<NBLK>() * kWave > kScratchVecs ? nupper<NBLK>() * kWave : kScratchVecs;
  __shared__ v4f wl[kWlVecs];
  float* dl = reinterpret_cast<float*>(wl);
  __bf16* hs = reinterpret_cast<__bf16*>(reinterpret_cast<char*>(wl) + sizeof(float) * 64);
  const int lane = threadIdx.x, g = lane >> 4, j = lane & 15;
  const int64_t m = a.m0 + blockIdx.x;
  double ad[nupper<NBLK>()][4];
  const double anorm = mx_load<NBLK>(a, m, g, j, ad);
  const float f = mx_features<NBLK>(ad, a.n, g, j, anorm, dl);
  if (feat_out && lane < kRF) feat_out[m * kRF + lane] = f;
  // the network's layer-1 operand: features 8 (l / 16) .. +7 in lane l
  // (lanes past feature 11 hold 0)
  v8bf x;
  static_for<0, 8>([&](auto e) { x[e()] = (__bf16)bperm(f, 8 * g + e); });
  const v4fn z = net_logits(static_cast<const v8bf*>(prep),
                            reinterpret_cast<const float*>(static_cast<const char*>(prep) + kPrepBiasByte), x, hs,
                            lane);
  // every column holds this matrix: lane 0 has its logits
  const float z0 = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, (float)z[0])));
  const float z1 = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, (float)z[1])));
  const float z2 = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, (float)z[2])));
  const int t = net_tier(z0, z1, z2);
  if (tier_out && lane == 0) tier_out[m] = t;
  if (t == 2) {
    if (lane == 0) {
      a.flist[atomicAdd(a.fcount, 1)] = (int)m;
      if (a.stat) a.stat[m] = -1;
    }
    return;
  }
  if (t == 0) {
    if (lane == 0) a.llist[atomicAdd(a.lcount, 1)] = (int)m;
    return;
  }
  mx_solve<NBLK, Low::F32>(a, m, g, j, ad, anorm, xs, fs, wl);
}

}  // namespace tsdx_hip
