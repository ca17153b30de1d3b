// Float64 posv by blocks on the matrix cores: the float64 tier of the
// mixed-precision driver (no reference counterpart; the reference's posv is
// batched.hpp:271-275).  It solves the matrices of a device list (those the
// low-precision refinement could not finish, or that the router sent to
// float64), one wave per matrix, 32 < n <= 16 NBLK.
//
// The block scheme of the low-precision kernel (mfma_kernels.hpp): upper
// 16 x 16 blocks of the symmetric matrix, right-looking block Cholesky
// A = U^T U with U_KJ = V_KK^T W_KJ and W_IJ -= U_KI^T U_KJ on
// v_mfma_f64_16x16x4f64, the diagonal blocks eliminated on the VALU into
// V_KK = U_KK^-1.  Here everything is float64, the factor stays in registers
// in LDS next to the float64 A in registers, and a block is held in the layout of the float64
// MFMA accumulator: lane l = 16g + j holds column j, rows g + 4v (v = 0..3)
// (the float MFMA uses rows 4g + v; this one was checked on the GPU).  The
// product X^T Y of two stored blocks is again four MFMAs on the registers as
// they stand, since the contraction runs over the rows in whatever order.
//
// The solution is refined in float64 (residual with A, correction with the
// factor) until the backward-error test of the low-precision tiers passes.
// A matrix whose block factorization meets a pivot that is not positive, or
// whose refinement stalls, is solved by the plain float64 posv in the same
// wave (posv_plain_group): that path is backward stable and reports the
// reference's info for a matrix that is not positive definite.
//
// Vector forms (length 16 NBLK):
//   column form  c[J]      lane (g, j) holds entry 16J + j;
//   row form     R[I][v]   lane (g, j) holds entry 16I + g + 4v.
#pragma once

#include "mfma_kernels.hpp"

namespace tsdx_hip {

typedef double v4d __attribute__((ext_vector_type(4)));

// Loads the upper blocks (float64 accumulator layout) of the symmetric matrix
// whose lower triangle is Am (element (r, c) from Am[max(r,c)][min(r,c)],
// identity padding past n); returns ||A||_F.
template <int NBLK>
__device__ __forceinline__ double m64_load(const double* __restrict__ Am, int n, int g, int j,
                                           double (&ad)[nupper<NBLK>()][4]) {
  double sd = 0.0, so = 0.0;
  static_for<0, NBLK>([&](auto Ic) {
    static_for<Ic, NBLK>([&](auto Jc) {
      constexpr int I = decltype(Ic)::value, J = decltype(Jc)::value, b = bslot<NBLK>(I, J);
      const int c = 16 * J + j;
      static_for<0, 4>([&](auto v) {
        const int r = 16 * I + g + 4 * v;
        const bool ok = r < n && c < n;
        const int hi = r > c ? r : c, lo = r > c ? c : r;
        const double x = Am[ok ? hi * n + lo : 0];  // unconditional: the loads overlap
        const double y = ok ? x : 0.0;
        ad[b][v()] = ok ? x : ((r == c) ? 1.0 : 0.0);
        if constexpr (I == J)
          sd = __builtin_fma(y, y, sd);
        else
          so = __builtin_fma(y, y, so);
      });
    });
  });
  return __builtin_sqrt(sum_j(sum_g(__builtin_fma(2.0, so, sd))));
}

// V = U^-1 of a symmetric 16 x 16 block W = U^T U (float64 layout), the
// elimination of diag_inverse (mfma_kernels.hpp) in float64: step k reads the
// pivot W[k][k] from lane 16 (k % 4) + k, register k / 4.  1/sqrt of the pivot
// is v_rsq_f64 refined by one Newton step.  bad is set when a pivot is not
// positive (tested once, on the 16 reciprocals, after the last step).
__device__ __forceinline__ v4d m64_diag_inverse(v4d W, int g, int j, bool& bad) {
  g = opaque(g);
  j = opaque(j);
  v4d F;
  static_for<0, 4>([&](auto v) { F[v()] = (g + 4 * v == j) ? 1.0 : 0.0; });
  double rdl = 1.0;  // 1 / U[j][j] of this lane's column
  static_for<0, 16>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    constexpr int kg = k % 4, kv = k / 4;
    const double wkv = W[kv];
    const double akk = mk64(__builtin_amdgcn_readlane(lo32(wkv), 16 * kg + k),
                            __builtin_amdgcn_readlane(hi32(wkv), 16 * kg + k));
    double rd = __builtin_amdgcn_rsq(akk);
    rd = rd * __builtin_fma(-0.5 * akk, rd * rd, 1.5);
    const double wk = bperm(wkv, 16 * kg + j);  // W[k][j], from row k (lane 16 kg + j, register kv)
    const double nt = wk * (-rd * rd);          // -T[k][j]
    static_for<0, 4>([&](auto v) { W[v()] = __builtin_fma(dppd<0x150 + k>(W[v()]), nt, W[v()]); });
    const double ntm = (j > k) ? nt : 0.0;
    static_for<0, 4>([&](auto v) { F[v()] = __builtin_fma(dppd<0x150 + k>(F[v()]), ntm, F[v()]); });
    rdl = (j == k) ? rd : rdl;
  });
  bad = bad || __builtin_amdgcn_ballot_w64(!(rdl > 0.0 && rdl < __builtin_huge_val())) != 0;
  static_for<0, 4>([&](auto v) { F[v()] *= rdl; });
  return F;
}

// Block Cholesky in this wave's LDS: slot(b) (each lane its own 32-byte slot
// of block b, touched by no other lane) holds the upper blocks of A on entry,
// V_KK in the diagonal slots and U_KJ elsewhere on exit.  Returns true when a
// pivot was not positive.
template <int NBLK, class Slot>
__device__ __forceinline__ bool m64_factor(Slot slot, int g, int j) {
  bool bad = false;
  static_for<0, NBLK>([&](auto Kc) {
    constexpr int K = decltype(Kc)::value;
    const v4d V = m64_diag_inverse(slot(bslot<NBLK>(K, K)), g, j, bad);
    slot(bslot<NBLK>(K, K)) = V;
    v4d row[NBLK];  // U_KJ, J > K
    static_for<K + 1, NBLK>([&](auto Jc) {
      constexpr int J = decltype(Jc)::value;
      const v4d w = slot(bslot<NBLK>(K, J));
      v4d acc = {0.0, 0.0, 0.0, 0.0};
      static_for<0, 4>([&](auto v) { acc = __builtin_amdgcn_mfma_f64_16x16x4f64(V[v()], w[v()], acc, 0, 0, 0); });
      row[J] = acc;
      slot(bslot<NBLK>(K, J)) = acc;
    });
    static_for<K + 1, NBLK>([&](auto Ic) {
      static_for<Ic, NBLK>([&](auto Jc) {
        constexpr int I = decltype(Ic)::value, J = decltype(Jc)::value;
        v4d acc = slot(bslot<NBLK>(I, J));
        static_for<0, 4>([&](auto v) {
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(-row[I][v()], row[J][v()], acc, 0, 0, 0);
        });
        slot(bslot<NBLK>(I, J)) = acc;
      });
    });
  });
  return bad;
}

// Float64 tier over the list a.flist[0 .. *a.fcount): one workgroup (one
// wave) per entry, grid-stride.  On success X is written and info[m] = 0;
// otherwise the plain float64 posv of the same wave writes X and info.  stat
// is left as the caller set it (-1 for the float64 tier).
template <int NBLK, int NR>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(2))) void k_posv_mx64(IrArgs a) {
  constexpr int NU = nupper<NBLK>();
  static_assert(NU * 4 >= 2 + 8, "the plain fallback's buffers live in the factor's LDS");
  __shared__ __align__(16) double xs[16 * NBLK];
  __shared__ v4d wl[NU * kWave];  // the factor (20 KiB at NBLK = 4)
  const int lane = threadIdx.x, g = lane >> 4, j = lane & 15;
  const int n = a.n;
  const int count = *a.fcount;
  // source lane of column-form entry j among the outputs of tsum4 (lane j'
  // holds row g + 4 (j' >> 2) of its row group g)
  const int src64 = 16 * (j & 3) + 4 * (j >> 2);
  auto slot = [&](int b) -> v4d& { return wl[b * kWave + lane]; };
  for (int64_t w = blockIdx.x; w < count; w += gridDim.x) {  // uniform
    const int64_t m = a.flist[w];
    double ad[NU][4];
    // (lane indices and n pass through opaque() so that the 40 element
    // offsets are formed here, not lifted out of the loop into registers)
    const double anorm = m64_load<NBLK>(a.A + m * (int64_t)n * n, opaque(n), opaque(g), opaque(j), ad);
    static_for<0, NU>([&](auto b) {
      v4d x;
      static_for<0, 4>([&](auto v) { x[v()] = ad[b][v()]; });
      slot(b) = x;
    });
    bool ok = !m64_factor<NBLK>(slot, g, j);
    wave_sync();

    // one block, column form -> row form, through LDS
    auto c2r = [&](double c, double (&R)[4]) __attribute__((always_inline)) {
      if (g == 0) xs[j] = c;
      wave_sync();
      static_for<0, 4>([&](auto v) { R[v()] = xs[g + 4 * v]; });
      wave_sync();
    };
    // z = U^-1 U^-T t, column form in and out
    auto solve = [&](const double (&t)[NBLK], double (&z)[NBLK]) __attribute__((always_inline)) {
      // the factor is read from LDS where it is used: without this barrier the
      // compiler lifts the reads out of the refinement loop into 80 registers
      asm volatile("" ::: "memory");
      double yc[NBLK], yR[NBLK][4];
      static_for<0, NBLK>([&](auto Kc) {
        constexpr int K = decltype(Kc)::value;
        double acc = 0.0;
        static_for<0, K>([&](auto I) {
          const v4d blk = slot(bslot<NBLK>(I, K));
          static_for<0, 4>([&](auto v) { acc = __builtin_fma(blk[v()], yR[I][v()], acc); });
        });
        double sR[4];
        c2r(t[K] - sum_g(acc), sR);
        const v4d dk = slot(bslot<NBLK>(K, K));
        double p = 0.0;
        static_for<0, 4>([&](auto v) { p = __builtin_fma(dk[v()], sR[v()], p); });
        yc[K] = sum_g(p);
        c2r(yc[K], yR[K]);
      });
      static_for_rev<0, NBLK>([&](auto Kc) {
        constexpr int K = decltype(Kc)::value;
        double p[4] = {0.0, 0.0, 0.0, 0.0};
        static_for<K + 1, NBLK>([&](auto J) {
          const v4d blk = slot(bslot<NBLK>(K, J));
          static_for<0, 4>([&](auto v) { p[v()] = __builtin_fma(blk[v()], z[J], p[v()]); });
        });
        const double wc = yc[K] - bperm(tsum4(p, j), src64);
        const v4d dk = slot(bslot<NBLK>(K, K));
        double q[4];
        static_for<0, 4>([&](auto v) { q[v()] = dk[v()] * wc; });
        z[K] = bperm(tsum4(q, j), src64);
      });
    };

    const double* Bm = a.B + m * (int64_t)n * a.nrhs;
    double* Xm = a.X + m * (int64_t)n * a.nrhs;
    const double ctol = 0.25 * a.tol * a.tol, aa2 = anorm * anorm;
    for (int c = 0; c < a.nrhs && ok; ++c) {
      double bc[NBLK], xc[NBLK], t[NBLK], z[NBLK];
      double bb = 0.0;
      static_for<0, NBLK>([&](auto J) {
        const int row = 16 * J + j;
        const double bv = Bm[row < n ? (int64_t)row * a.nrhs + c : 0];
        bc[J] = row < n ? bv : 0.0;
        bb = __builtin_fma(bc[J], bc[J], bb);
      });
      const double bb3 = 3.0 * sum_j(bb);
      solve(bc, z);
      static_for<0, NBLK>([&](auto J) { xc[J] = (16 * J + j < n) ? z[J] : 0.0; });
      bool conv = false;
      double prev = __builtin_huge_val();
      for (int it = 0; it <= a.max_iter; ++it) {
        // r = b - A x: diagonal blocks as column sums, off-diagonal blocks as
        // row sums (A_IJ) and column sums (A_JI = A_IJ^T), as in mx_solve
        if (g == 0) static_for<0, NBLK>([&](auto J) { xs[16 * J + j] = xc[J]; });
        wave_sync();
        double pc[NBLK], rc[NBLK];
        static_for<0, NBLK>([&](auto J) { pc[J] = 0.0; });
        static_for<0, NBLK>([&](auto Ic) {
          constexpr int I = decltype(Ic)::value;
          double xR[4];
          static_for<0, 4>([&](auto v) { xR[v()] = xs[16 * I + g + 4 * v]; });
          static_for<0, 4>([&](auto v) { pc[I] = __builtin_fma(ad[bslot<NBLK>(I, I)][v()], xR[v()], pc[I]); });
          if constexpr (I + 1 < NBLK) {
            double p[4] = {0.0, 0.0, 0.0, 0.0};
            static_for<I + 1, NBLK>([&](auto Jc) {
              constexpr int J = decltype(Jc)::value, b = bslot<NBLK>(I, J);
              static_for<0, 4>([&](auto v) {
                p[v()] = __builtin_fma(ad[b][v()], xc[J], p[v()]);
                pc[J] = __builtin_fma(ad[b][v()], xR[v()], pc[J]);
              });
            });
            rc[I] = bperm(tsum4(p, j), src64);
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
        if (rn2 <= ctol * __builtin_fma(aa2, xn2, bb3)) {  // the test of mx_solve
          conv = true;
          break;
        }
        if (it == a.max_iter || !(rn2 < 0.25 * prev)) break;
        prev = rn2;
        static_for<0, NBLK>([&](auto J) { t[J] = rc[J]; });
        solve(t, z);
        static_for<0, NBLK>([&](auto J) { xc[J] += (16 * J + j < n) ? z[J] : 0.0; });
      }
      ok = conv;
      if (conv && g == 0) {
        static_for<0, NBLK>([&](auto J) {
          const int row = 16 * J + j;
          if (row < n) Xm[(int64_t)row * a.nrhs + c] = xc[J];
        });
      }
    }
    wave_sync();
    if (ok) {
      if (lane == 0 && a.info) a.info[m] = 0;
    } else {  // the factor's LDS is free now: the plain posv's buffers go there
      double* col_s = reinterpret_cast<double*>(wl);
      posv_plain_group<64, NR>(a, m, true, lane, 0, col_s, col_s + 2 * kWave);
    }
    wave_sync();  // the LDS is reused by the next entry
  }
}

}  // namespace tsdx_hip
