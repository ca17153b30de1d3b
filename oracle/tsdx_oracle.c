/* tsdx_oracle.c -- CPU restatement of the reference's batched small-matrix
 * Cholesky / LU / solve path (TinySolverDx), used as the parity oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the C++
 * host API, the Python package) links, loads or calls this file.  Only the
 * tests, the smoke check and the CPU-baseline leg of bench.py load the shared
 * library built from it, and only to check or time against it.
 *
 * What it restates (file:line into the reference):
 *   include/tinysolverdx/kernels.hpp:18-227   per-matrix and tile kernels
 *   include/tinysolverdx/batched.hpp:27-280   batch drivers, Auto rules, padding
 *   include/tinysolverdx/precision.hpp:41-65, 118-126  fp16/bf16 rounding, u
 *   include/tinysolverdx/precision.hpp:71-97  Emu<F> (emulated fp16 / bf16 arithmetic)
 *   src/runtime.cpp:72-140                     runtime-precision calls (Emu copies)
 *   src/runtime.cpp:143-150                    round_to
 *   src/runtime.cpp:162-260                    long-double error metrics
 *   src/runtime.cpp:15-58                      thread count (TSDX_NUM_THREADS)
 *
 * Arithmetic: built with -ffp-contract=off; every `x -= y*z` of the reference
 * (which the reference's default g++/clang build contracts to one fused
 * multiply-add) is written as an explicit fma(-y, z, x); divisions and square
 * roots are IEEE (correctly rounded).  Pinned against the compiled reference
 * (oracle/_ref) and the golden fixtures under tests/golden.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_EINVAL (-1)   /* reference: std::invalid_argument */
#define ORC_EUNAVAIL (-2) /* reference: std::runtime_error (GPU backend absent) */

/* Backend enum order of the reference (precision.hpp:33) */
#define ORC_BE_AUTO 0
#define ORC_BE_NAIVE 1
#define ORC_BE_SCALAR 2
#define ORC_BE_SIMD 3
#define ORC_BE_CUDA 4

#define ORC_MIN_N 1
#define ORC_MAX_N 64
#define ORC_LU_SIMD_MAX_N 24 /* batched.hpp:33 */

/* ------------------------------------------------------------- threading */
static int g_threads = 0;

static int orc_default_threads(void) {
  const char *e = getenv("TSDX_NUM_THREADS");
  if (e) {
    int v = atoi(e);
    if (v > 0) return v;
  }
#ifdef _OPENMP
  return omp_get_num_procs();
#else
  return 1;
#endif
}

int orc_num_threads(void) {
  if (g_threads <= 0) g_threads = orc_default_threads();
  return g_threads;
}
void orc_set_num_threads(int n) { g_threads = n > 0 ? n : orc_default_threads(); }

static int orc_threads_for(int64_t chunks) {
  int t = orc_num_threads();
  if (chunks < t) t = (int)(chunks > 0 ? chunks : 1);
  return t;
}

/* ------------------------------------------------------- driver helpers */
static int orc_check_dims(int n, int64_t batch) { /* batched.hpp:39-43 */
  return (n < ORC_MIN_N || n > ORC_MAX_N || batch < 0) ? 1 : 0;
}
static int orc_resolve(int b) { /* batched.hpp:46-50, native types only */
  return b == ORC_BE_AUTO ? ORC_BE_SIMD : b;
}
static int64_t orc_chunk_size(int n, int multiple) { /* batched.hpp:53-57 */
  int64_t c = 32768 / ((int64_t)n * n);
  if (c < multiple) c = multiple;
  return (c + multiple - 1) / multiple * multiple;
}

/* ------------------------------------------------------ typed instances */
#define T float
#define SUF f32
#define FMA fmaf
#define SQRT sqrtf
#define LANES 8
#include "oracle_t.inc"
#undef T
#undef SUF
#undef FMA
#undef SQRT
#undef LANES

#define T double
#define SUF f64
#define FMA fma
#define SQRT sqrt
#define LANES 4
#include "oracle_t.inc"
#undef T
#undef SUF
#undef FMA
#undef SQRT
#undef LANES

/* ------------------------------------------------------------ precision */
static uint32_t f2u(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
static float u2f(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }

/* precision.hpp:41-55: nearest binary16 value, ties to even, subnormals kept,
 * magnitudes from 65520 up overflow to infinity; inf/NaN pass through. */
float orc_round_fp16(float x) {
  const uint32_t u = f2u(x), sign = u & 0x80000000u, mag = u & 0x7fffffffu;
  if (mag >= 0x7f800000u) return x;
  const float ax = u2f(mag);
  if (ax >= 65520.0f) return u2f(sign | 0x7f800000u);
  if (ax < 6.103515625e-05f) { /* below the smallest normal: multiples of 2^-24 */
    const float q = 5.9604644775390625e-08f;
    return u2f(sign | f2u(nearbyintf(ax / q) * q));
  }
  uint32_t r = mag + 0xfffu + ((mag >> 13) & 1u); /* drop 13 bits, RNE */
  r &= ~0x1fffu;
  return u2f(sign | r);
}

/* precision.hpp:58-65: nearest bfloat16 value, ties to even; NaN passes. */
float orc_round_bf16(float x) {
  const uint32_t u = f2u(x), mag = u & 0x7fffffffu;
  if (mag > 0x7f800000u) return x;
  uint32_t r = mag + 0x7fffu + ((mag >> 16) & 1u);
  r &= ~0xffffu;
  return u2f((u & 0x80000000u) | r);
}

/* precision.hpp:118-126; p: 0 fp32, 1 fp64, 2 fp16, 3 bf16 */
double orc_unit_roundoff(int p) {
  switch (p) {
    case 0: return ldexp(1.0, -24);
    case 1: return ldexp(1.0, -53);
    case 2: return ldexp(1.0, -11);
    case 3: return ldexp(1.0, -8);
  }
  return 0.0;
}

/* runtime.cpp:143-150 for double storage: fp64 leaves values alone, the other
 * formats go through float first. */
void orc_round_to_f64(int p, double *x, int64_t count) {
  if (p == 1) return;
  for (int64_t i = 0; i < count; ++i) {
    const float f = (float)x[i];
    x[i] = (double)(p == 0 ? f : p == 2 ? orc_round_fp16(f) : orc_round_bf16(f));
  }
}
void orc_round_to_f32(int p, float *x, int64_t count) {
  if (p == 1) return;
  for (int64_t i = 0; i < count; ++i) x[i] = p == 0 ? x[i] : p == 2 ? orc_round_fp16(x[i]) : orc_round_bf16(x[i]);
}

/* ------------------------------------------- emulated fp16 / bfloat16 */
/* precision.hpp:71-97 (Emu<F>): a value is a float representable in the
 * format; each of + - * / and sqrt is computed in float and its result rounded
 * to the format (no fused multiply-add: `x -= y*z` is two rounded operations).
 * runtime.cpp:72-140 (runtime-precision calls): the caller's array is copied
 * into Emu values (double storage through float first), the batched kernel
 * runs (Auto resolves to the scalar kernels for these types, batched.hpp:46-50),
 * and the float values are copied back, every entry (untouched ones come back
 * rounded).  p: 2 fp16, 3 bf16. */
typedef float (*orc_rnd)(float);

/* cholesky_rm with Emu operations (kernels.hpp:18-37): stops at the first
 * failing pivot, leaving the rest of the matrix as it is. */
static int emu_chol_rm(float *a, int n, orc_rnd R) {
  for (int i = 0; i < n; ++i) {
    float *ri = a + (size_t)i * n;
    for (int j = 0; j <= i; ++j) {
      const float *rj = a + (size_t)j * n;
      float s = ri[j];
      for (int k = 0; k < j; ++k) s = R(s - R(ri[k] * rj[k]));
      if (j == i) {
        if (!(s > 0.0f)) return i + 1;
        ri[i] = R(sqrtf(s));
      } else {
        ri[j] = R(s / rj[j]);
      }
    }
  }
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) a[(size_t)i * n + j] = 0.0f;
  return 0;
}

/* lu_rm with Emu operations (kernels.hpp:91-119): first largest |a_ik| by
 * strict >, whole-row swaps, a zero pivot records info and skips the column. */
static int emu_lu_rm(float *a, int n, int *piv, orc_rnd R) {
  int info = 0;
  for (int k = 0; k < n; ++k) {
    int p = k;
    float best = a[(size_t)k * n + k] < 0.0f ? -a[(size_t)k * n + k] : a[(size_t)k * n + k];
    for (int i = k + 1; i < n; ++i) {
      const float x = a[(size_t)i * n + k], v = x < 0.0f ? -x : x;
      if (v > best) {
        best = v;
        p = i;
      }
    }
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < n; ++j) {
        const float t = a[(size_t)k * n + j];
        a[(size_t)k * n + j] = a[(size_t)p * n + j];
        a[(size_t)p * n + j] = t;
      }
    const float pk = a[(size_t)k * n + k];
    if (pk == 0.0f) {
      if (info == 0) info = k + 1;
      continue;
    }
    const float *rk = a + (size_t)k * n;
    for (int i = k + 1; i < n; ++i) {
      float *ri = a + (size_t)i * n;
      const float f = R(ri[k] / pk);
      ri[k] = f;
      for (int j = k + 1; j < n; ++j) ri[j] = R(ri[j] - R(f * rk[j]));
    }
  }
  return info;
}

/* cholesky_solve_rm with Emu operations (kernels.hpp:184-202): forward rows
 * i ascending (updates k ascending, then the division), backward rows i
 * descending with the updates k = i+1 .. n-1 ascending. */
static void emu_chol_solve_rm(const float *L, int n, float *B, int nrhs, orc_rnd R) {
  for (int i = 0; i < n; ++i) {
    for (int k = 0; k < i; ++k)
      for (int r = 0; r < nrhs; ++r)
        B[(size_t)i * nrhs + r] = R(B[(size_t)i * nrhs + r] - R(L[(size_t)i * n + k] * B[(size_t)k * nrhs + r]));
    for (int r = 0; r < nrhs; ++r) B[(size_t)i * nrhs + r] = R(B[(size_t)i * nrhs + r] / L[(size_t)i * n + i]);
  }
  for (int i = n - 1; i >= 0; --i) {
    for (int k = i + 1; k < n; ++k)
      for (int r = 0; r < nrhs; ++r)
        B[(size_t)i * nrhs + r] = R(B[(size_t)i * nrhs + r] - R(L[(size_t)k * n + i] * B[(size_t)k * nrhs + r]));
    for (int r = 0; r < nrhs; ++r) B[(size_t)i * nrhs + r] = R(B[(size_t)i * nrhs + r] / L[(size_t)i * n + i]);
  }
}

/* lu_solve_rm with Emu operations (kernels.hpp:204-227): the swaps piv[0..n-1]
 * in order, unit-lower forward pass, upper backward pass as above. */
static void emu_lu_solve_rm(const float *LU, const int *piv, int n, float *B, int nrhs, orc_rnd R) {
  for (int k = 0; k < n; ++k)
    if (piv[k] != k)
      for (int r = 0; r < nrhs; ++r) {
        const float t = B[(size_t)k * nrhs + r];
        B[(size_t)k * nrhs + r] = B[(size_t)piv[k] * nrhs + r];
        B[(size_t)piv[k] * nrhs + r] = t;
      }
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < i; ++k)
      for (int r = 0; r < nrhs; ++r)
        B[(size_t)i * nrhs + r] = R(B[(size_t)i * nrhs + r] - R(LU[(size_t)i * n + k] * B[(size_t)k * nrhs + r]));
  for (int i = n - 1; i >= 0; --i) {
    for (int k = i + 1; k < n; ++k)
      for (int r = 0; r < nrhs; ++r)
        B[(size_t)i * nrhs + r] = R(B[(size_t)i * nrhs + r] - R(LU[(size_t)i * n + k] * B[(size_t)k * nrhs + r]));
    for (int r = 0; r < nrhs; ++r) B[(size_t)i * nrhs + r] = R(B[(size_t)i * nrhs + r] / LU[(size_t)i * n + i]);
  }
}

static orc_rnd emu_rounding(int p) { return p == 2 ? orc_round_fp16 : p == 3 ? orc_round_bf16 : NULL; }

/* Batched runtime-precision entry points for float (f32) and double (f64)
 * storage: copy in with rounding, run, copy the float values back. */
#define ORC_EMU_TYPED(S, SUF)                                                                              \
  int orc_emu_cholesky_##SUF(int p, S *A, int n, int64_t batch, int *info) {                                \
    const orc_rnd R = emu_rounding(p);                                                                     \
    if (!R || orc_check_dims(n, batch)) return ORC_EINVAL;                                                 \
    const int64_t nn = (int64_t)n * n;                                                                     \
    _Pragma("omp parallel for schedule(static) num_threads(orc_threads_for(batch))")                     \
    for (int64_t b = 0; b < batch; ++b) {                                                                  \
      float a[ORC_MAX_N * ORC_MAX_N];                                                                      \
      for (int64_t k = 0; k < nn; ++k) a[k] = R((float)A[b * nn + k]);                                    \
      const int r = emu_chol_rm(a, n, R);                                                                  \
      if (info) info[b] = r;                                                                               \
      for (int64_t k = 0; k < nn; ++k) A[b * nn + k] = (S)a[k];                                           \
    }                                                                                                      \
    return 0;                                                                                              \
  }                                                                                                        \
  int orc_emu_lu_##SUF(int p, S *A, int n, int64_t batch, int *piv, int *info) {                            \
    const orc_rnd R = emu_rounding(p);                                                                     \
    if (!R || orc_check_dims(n, batch) || !piv) return ORC_EINVAL;                                         \
    const int64_t nn = (int64_t)n * n;                                                                     \
    _Pragma("omp parallel for schedule(static) num_threads(orc_threads_for(batch))")                     \
    for (int64_t b = 0; b < batch; ++b) {                                                                  \
      float a[ORC_MAX_N * ORC_MAX_N];                                                                      \
      for (int64_t k = 0; k < nn; ++k) a[k] = R((float)A[b * nn + k]);                                    \
      const int r = emu_lu_rm(a, n, piv + b * n, R);                                                       \
      if (info) info[b] = r;                                                                               \
      for (int64_t k = 0; k < nn; ++k) A[b * nn + k] = (S)a[k];                                           \
    }                                                                                                      \
    return 0;                                                                                              \
  }                                                                                                        \
  int orc_emu_chol_solve_##SUF(int p, const S *L, int n, int64_t batch, S *B, int nrhs) {                   \
    const orc_rnd R = emu_rounding(p);                                                                     \
    if (!R || orc_check_dims(n, batch) || nrhs < 0 || nrhs > ORC_EMU_MAX_NRHS) return ORC_EINVAL;         \
    const int64_t nn = (int64_t)n * n, bn = (int64_t)n * nrhs;                                             \
    _Pragma("omp parallel for schedule(static) num_threads(orc_threads_for(batch))")                     \
    for (int64_t b = 0; b < batch; ++b) {                                                                  \
      float l[ORC_MAX_N * ORC_MAX_N], x[ORC_MAX_N * ORC_EMU_MAX_NRHS];                                     \
      for (int64_t k = 0; k < nn; ++k) l[k] = R((float)L[b * nn + k]);                                    \
      for (int64_t k = 0; k < bn; ++k) x[k] = R((float)B[b * bn + k]);                                    \
      emu_chol_solve_rm(l, n, x, nrhs, R);                                                                 \
      for (int64_t k = 0; k < bn; ++k) B[b * bn + k] = (S)x[k];                                           \
    }                                                                                                      \
    return 0;                                                                                              \
  }                                                                                                        \
  int orc_emu_lu_solve_##SUF(int p, const S *LU, const int *piv, int n, int64_t batch, S *B, int nrhs) {    \
    const orc_rnd R = emu_rounding(p);                                                                     \
    if (!R || orc_check_dims(n, batch) || !piv || nrhs < 0 || nrhs > ORC_EMU_MAX_NRHS) return ORC_EINVAL; \
    const int64_t nn = (int64_t)n * n, bn = (int64_t)n * nrhs;                                             \
    _Pragma("omp parallel for schedule(static) num_threads(orc_threads_for(batch))")                     \
    for (int64_t b = 0; b < batch; ++b) {                                                                  \
      float l[ORC_MAX_N * ORC_MAX_N], x[ORC_MAX_N * ORC_EMU_MAX_NRHS];                                     \
      for (int64_t k = 0; k < nn; ++k) l[k] = R((float)LU[b * nn + k]);                                   \
      for (int64_t k = 0; k < bn; ++k) x[k] = R((float)B[b * bn + k]);                                    \
      emu_lu_solve_rm(l, piv + b * n, n, x, nrhs, R);                                                      \
      for (int64_t k = 0; k < bn; ++k) B[b * bn + k] = (S)x[k];                                           \
    }                                                                                                      \
    return 0;                                                                                              \
  }

#define ORC_EMU_MAX_NRHS 16 /* oracle limit (stack buffers), not the reference's */
ORC_EMU_TYPED(float, f32)
ORC_EMU_TYPED(double, f64)
#undef ORC_EMU_TYPED

/* ------------------------------------------------ long-double ground truth */
typedef long double LD;

static LD ld_fro(const double *a, int cnt) {
  LD s = 0;
  for (int i = 0; i < cnt; ++i) s += (LD)a[i] * a[i];
  return sqrtl(s);
}

/* runtime.cpp:167-189: GEPP in long double on a copy, then back substitution */
static void ld_solve(const double *A, const double *b, int n, LD *x, LD *a) {
  for (int i = 0; i < n * n; ++i) a[i] = A[i];
  for (int i = 0; i < n; ++i) x[i] = b[i];
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (fabsl(a[i * n + k]) > fabsl(a[p * n + k])) p = i;
    if (p != k) {
      for (int j = 0; j < n; ++j) { LD t = a[k * n + j]; a[k * n + j] = a[p * n + j]; a[p * n + j] = t; }
      LD t = x[k]; x[k] = x[p]; x[p] = t;
    }
    for (int i = k + 1; i < n; ++i) {
      const LD f = a[i * n + k] / a[k * n + k];
      for (int j = k + 1; j < n; ++j) a[i * n + j] -= f * a[k * n + j];
      x[i] -= f * x[k];
    }
  }
  for (int i = n - 1; i >= 0; --i) {
    LD s = x[i];
    for (int j = i + 1; j < n; ++j) s -= a[i * n + j] * x[j];
    x[i] = s / a[i * n + i];
  }
}

/* runtime.cpp:194-207: ||A - L L^T||_F / ||A||_F using k <= min(i,j) of L */
void orc_chol_backward_error(const double *A, const double *L, int n, int64_t batch, double *out) {
#pragma omp parallel for schedule(dynamic) num_threads(orc_threads_for(batch))
  for (int64_t b = 0; b < batch; ++b) {
    const double *a = A + b * n * n, *l = L + b * n * n;
    LD num = 0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        LD s = 0;
        const int m = i < j ? i : j;
        for (int k = 0; k <= m; ++k) s += (LD)l[i * n + k] * l[j * n + k];
        const LD d = (LD)a[i * n + j] - s;
        num += d * d;
      }
    out[b] = (double)(sqrtl(num) / ld_fro(a, n * n));
  }
}

/* runtime.cpp:209-228: ||P A - L U||_F / ||A||_F, P from the sequential swaps */
void orc_lu_backward_error(const double *A, const double *LU, const int *piv, int n, int64_t batch, double *out) {
#pragma omp parallel for schedule(dynamic) num_threads(orc_threads_for(batch))
  for (int64_t b = 0; b < batch; ++b) {
    const double *a = A + b * n * n, *f = LU + b * n * n;
    const int *pv = piv + b * n;
    int perm[ORC_MAX_N];
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int k = 0; k < n; ++k) { int t = perm[k]; perm[k] = perm[pv[k]]; perm[pv[k]] = t; }
    LD num = 0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        LD s = 0;
        const int m = i < j ? i : j;
        for (int k = 0; k <= m; ++k) {
          const LD lik = (k == i) ? (LD)1 : (LD)f[i * n + k];
          s += lik * f[k * n + j];
        }
        const LD d = (LD)a[perm[i] * n + j] - s;
        num += d * d;
      }
    out[b] = (double)(sqrtl(num) / ld_fro(a, n * n));
  }
}

/* runtime.cpp:230-242: ||b - A x||_2 / (||A||_F ||x||_2 + ||b||_2) */
void orc_solve_backward_error(const double *A, const double *x, const double *bb, int n, int64_t batch, double *out) {
#pragma omp parallel for schedule(dynamic) num_threads(orc_threads_for(batch))
  for (int64_t b = 0; b < batch; ++b) {
    const double *a = A + b * n * n, *xv = x + b * n, *rhs = bb + b * n;
    LD num = 0;
    for (int i = 0; i < n; ++i) {
      LD s = rhs[i];
      for (int j = 0; j < n; ++j) s -= (LD)a[i * n + j] * xv[j];
      num += s * s;
    }
    const LD den = ld_fro(a, n * n) * ld_fro(xv, n) + ld_fro(rhs, n);
    out[b] = (double)(sqrtl(num) / den);
  }
}

/* runtime.cpp:244-250 */
void orc_reference_solve(const double *A, const double *b, int n, int64_t batch, double *x) {
#pragma omp parallel for schedule(dynamic) num_threads(orc_threads_for(batch))
  for (int64_t m = 0; m < batch; ++m) {
    LD xr[ORC_MAX_N], a[ORC_MAX_N * ORC_MAX_N];
    ld_solve(A + m * n * n, b + m * n, n, xr, a);
    for (int i = 0; i < n; ++i) x[m * n + i] = (double)xr[i];
  }
}

/* runtime.cpp:252-260: ||x - x_ref||_2 / ||x_ref||_2 against the long-double solve */
void orc_solve_forward_error(const double *A, const double *x, const double *b, int n, int64_t batch, double *out) {
#pragma omp parallel for schedule(dynamic) num_threads(orc_threads_for(batch))
  for (int64_t m = 0; m < batch; ++m) {
    LD xr[ORC_MAX_N], a[ORC_MAX_N * ORC_MAX_N];
    ld_solve(A + m * n * n, b + m * n, n, xr, a);
    LD num = 0, den = 0;
    for (int i = 0; i < n; ++i) {
      const LD d = (LD)x[m * n + i] - xr[i];
      num += d * d;
      den += xr[i] * xr[i];
    }
    out[m] = (double)sqrtl(num / den);
  }
}
