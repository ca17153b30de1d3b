/* tsdx_hip.h -- C ABI of the MI355X (gfx950) batched small-matrix solver
 * library (libtsdx_hip.so).
 *
 * Layout and semantics are the reference's (include/tinysolverdx/batched.hpp:1-6):
 * A is (batch, n, n) row-major with matrices contiguous, 1 <= n <= 64; the
 * Cholesky result is the lower factor with the upper triangle zeroed; LU is
 * packed L\U with 0-based sequential pivots piv (batch, n) (LAPACK ipiv - 1);
 * info (batch) is the LAPACK code (0 ok, k > 0: first failing pivot k), and
 * may be NULL; right-hand sides B / solutions X are (batch, n, nrhs) row-major.
 *
 * Factors, pivots and info are bit-identical to the reference's Auto CPU
 * backend (SIMD tile kernels for Cholesky and for LU up to n = 24, scalar LU
 * kernel from n = 25).  Solutions agree to rounding (the backward substitution
 * runs as a column sweep, a different summation order).
 *
 * Every function returns TSDX_HIP_OK, TSDX_HIP_EINVAL (bad size, batch, null
 * pointer: the reference throws std::invalid_argument) or TSDX_HIP_ERUNTIME
 * (no device, HIP failure: the reference throws std::runtime_error).  The
 * message of the last failure on the calling thread is tsdx_hip_last_error().
 *
 * Two families:
 *   tsdx_hip_<op>_<t>(..., stream)      device pointers, stream-ordered, no
 *                                       synchronisation, no allocation;
 *   tsdx_hip_<op>_batched_<t>(...)      host pointers, synchronous (stage to
 *                                       the device, run, copy back), the shape
 *                                       of the reference's CUDA backend.
 * <t> is f32 or f64.  stream is a hipStream_t (NULL = the null stream).
 */
#ifndef TSDX_HIP_H
#define TSDX_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TSDX_HIP_OK 0
#define TSDX_HIP_EINVAL (-1)
#define TSDX_HIP_ERUNTIME (-2)

/* 1 when a HIP device is visible (replaces cuda::available(),
 * include/tinysolverdx/cuda_backend.hpp:9 / src/cuda/cuda_backend.cu:150-153). */
int tsdx_hip_available(void);

/* Message of the last failed call on this thread ("" if none). */
const char *tsdx_hip_last_error(void);

/* Library version as major*10000 + minor*100 + patch. */
int tsdx_hip_version(void);

/* Device-memory helpers (no reference counterpart: the reference's API takes
 * host pointers only, README.md:99-101).  For callers that hold device arrays
 * without a GPU framework of their own, e.g. the Python module's device-array
 * path.  tsdx_hip_malloc allocates on the current device; tsdx_hip_memcpy is
 * synchronous and takes any pair of host / device pointers;
 * tsdx_hip_pointer_device gives the device of a device (or managed) pointer and
 * fails with TSDX_HIP_EINVAL for host memory. */
int tsdx_hip_malloc(void **ptr, size_t bytes);
int tsdx_hip_free(void *ptr);
int tsdx_hip_memcpy(void *dst, const void *src, size_t bytes);
int tsdx_hip_device_synchronize(void);
int tsdx_hip_get_device(int *device);
int tsdx_hip_pointer_device(const void *ptr, int *device);

/* ---------------------------------------------------------------------------
 * Device-pointer entry points (stream-ordered).
 * ------------------------------------------------------------------------ */

/* potrf: A := lower Cholesky factor.  Replaces the device work of
 * cuda::cholesky_batched (cuda_backend.hpp:10, cuda_backend.cu:155-170). */
int tsdx_hip_potrf_f32(float *A, int n, int64_t batch, int *info, void *stream);
int tsdx_hip_potrf_f64(double *A, int n, int64_t batch, int *info, void *stream);

/* getrf: A := packed L\U, piv (batch, n) must not be NULL.  Replaces the device
 * work of cuda::lu_batched (cuda_backend.hpp:11, cuda_backend.cu:172-188). */
int tsdx_hip_getrf_f32(float *A, int n, int64_t batch, int *piv, int *info, void *stream);
int tsdx_hip_getrf_f64(double *A, int n, int64_t batch, int *piv, int *info, void *stream);

/* potrs: B := (L L^T)^-1 B.  Device form of cholesky_solve_batched
 * (batched.hpp:248-257, which always runs on the CPU in the reference). */
int tsdx_hip_potrs_f32(const float *L, int n, int64_t batch, float *B, int nrhs, void *stream);
int tsdx_hip_potrs_f64(const double *L, int n, int64_t batch, double *B, int nrhs, void *stream);

/* getrs: B := A^-1 B from (LU, piv).  Device form of lu_solve_batched
 * (batched.hpp:259-268). */
int tsdx_hip_getrs_f32(const float *LU, const int *piv, int n, int64_t batch, float *B, int nrhs, void *stream);
int tsdx_hip_getrs_f64(const double *LU, const int *piv, int n, int64_t batch, double *B, int nrhs, void *stream);

/* posv: A := L, B := X in one launch (the reference composes two passes,
 * posv_batched, batched.hpp:271-275). */
int tsdx_hip_posv_f32(float *A, int n, int64_t batch, float *B, int nrhs, int *info, void *stream);
int tsdx_hip_posv_f64(double *A, int n, int64_t batch, double *B, int nrhs, int *info, void *stream);

/* gesv: A := L\U, piv, B := X in one launch (gesv_batched, batched.hpp:276-280). */
int tsdx_hip_gesv_f32(float *A, int n, int64_t batch, float *B, int nrhs, int *piv, int *info, void *stream);
int tsdx_hip_gesv_f64(double *A, int n, int64_t batch, double *B, int nrhs, int *piv, int *info, void *stream);

/* Solve-only posv / gesv: A and B are read, X is written, the factor is not
 * stored (piv may be NULL).  No reference counterpart: it is the plain path
 * with the algorithmic minimum of HBM traffic (read A and B, write X and info). */
int tsdx_hip_posv_solve_f32(const float *A, int n, int64_t batch, const float *B, float *X, int nrhs, int *info,
                            void *stream);
int tsdx_hip_posv_solve_f64(const double *A, int n, int64_t batch, const double *B, double *X, int nrhs, int *info,
                            void *stream);
int tsdx_hip_gesv_solve_f32(const float *A, int n, int64_t batch, const float *B, float *X, int nrhs, int *piv,
                            int *info, void *stream);
int tsdx_hip_gesv_solve_f64(const double *A, int n, int64_t batch, const double *B, double *X, int nrhs, int *piv,
                            int *info, void *stream);

/* ---------------------------------------------------------------------------
 * Mixed-precision posv with iterative refinement (float64 A, B, X).  No
 * reference counterpart: the reference has only the plain posv_batched
 * (batched.hpp:271-275).  Each matrix is factored in the low precision
 * `factor`, each right-hand side refined with float64 residuals until the
 * normwise backward error ||b - Ax|| / (||A||_F ||x|| + ||b||) <= tol, at most
 * max_iter corrections; matrices whose low-precision factorization fails or
 * whose refinement does not reach tol are solved again in float64 (the
 * fallback: for n > 32 a block Cholesky on the matrix cores with float64
 * refinement, itself backed by the plain float64 posv in the same wave; the
 * plain float64 posv for n <= 32), so every X meets tol unless A itself is
 * not SPD in float64 (then info > 0, as for tsdx_hip_posv_f64).
 *   factor: TSDX_HIP_FACTOR_FP32 / _BF16 / _FP16 (low precision + refinement),
 *           _FP64 (plain float64 solve for every matrix), _FP64_BLOCK (the
 *           float64 fallback tier for every matrix), _AUTO (FP32).
 *   stat:   (batch) or NULL: corrections used, or -1 for a matrix that took
 *           the float64 fallback.
 *   work:   device workspace of tsdx_hip_posv_ir_workspace_size(batch) bytes
 *           (the fallback list); calls sharing one workspace must not overlap.
 * ------------------------------------------------------------------------ */
#define TSDX_HIP_FACTOR_AUTO 0
#define TSDX_HIP_FACTOR_FP32 1
#define TSDX_HIP_FACTOR_BF16 2
#define TSDX_HIP_FACTOR_FP16 3
#define TSDX_HIP_FACTOR_FP64 4
#define TSDX_HIP_FACTOR_FP64_BLOCK 5

size_t tsdx_hip_posv_ir_workspace_size(int64_t batch);
int tsdx_hip_posv_ir_f64(const double *A, int n, int64_t batch, const double *B, double *X, int nrhs, int factor,
                         int max_iter, double tol, int *info, int *stat, void *work, size_t work_bytes, void *stream);

/* Allocation-free tsdx_hip_posv_ir_f64.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as tsdx_hip_posv_ir_f64 without the workspace: the
 * fallback list lives in a fixed device array of the library (16 MiB + 256
 * bytes on every GPU it runs on, held until the process ends), so batch may
 * not exceed TSDX_HIP_NOALLOC_MAX_BATCH (TSDX_HIP_EINVAL otherwise).  The call
 * allocates nothing.  All calls on one GPU share that array: calls that
 * overlap on one GPU (other streams, other threads) race and must not be
 * issued. */
#define TSDX_HIP_NOALLOC_MAX_BATCH (1 << 22)
int tsdx_hip_posv_ir_noalloc_f64(const double *A, int n, int64_t batch, const double *B, double *X, int nrhs,
                                 int factor, int max_iter, double tol, int *info, int *stat, void *stream);

/* ---------------------------------------------------------------------------
 * Mixed-precision gesv with iterative refinement (float64 A, B, X).  No
 * reference counterpart: the reference has only the plain gesv_batched
 * (batched.hpp:276-280, solve by lu_solve_rm, kernels.hpp:204-227).  Each
 * matrix is factored with partial pivoting in the low precision `factor`,
 * each right-hand side refined with float64 residuals until
 * ||b - Ax|| <= tol/2 (||A||_F ||x|| + ||b||) (the test of
 * tsdx_hip_posv_ir_f64), at most max_iter corrections; a matrix whose
 * low-precision factorization meets a zero pivot, or whose refinement stalls
 * or does not reach tol, is solved again by the plain float64 gesv, which
 * also gives its info (as tsdx_hip_gesv_solve_f64).  Converged matrices get
 * info 0.  Only X, info and stat are written.
 *   factor: TSDX_HIP_FACTOR_FP32 / _BF16 / _FP16 (low precision + refinement),
 *           _FP64 (plain float64 gesv for every matrix), _AUTO (FP32);
 *           _FP64_BLOCK is rejected (there is no block float64 LU tier).
 *   stat:   (batch) or NULL: corrections used (largest over the right-hand
 *           sides), or -1 for a matrix solved by the float64 gesv.
 *   work:   device workspace of tsdx_hip_gesv_ir_workspace_size(batch) bytes
 *           (the float64 list); calls sharing one workspace must not overlap.
 * ------------------------------------------------------------------------ */
size_t tsdx_hip_gesv_ir_workspace_size(int64_t batch);
int tsdx_hip_gesv_ir_f64(const double *A, int n, int64_t batch, const double *B, double *X, int nrhs, int factor,
                         int max_iter, double tol, int *info, int *stat, void *work, size_t work_bytes, void *stream);

/* ---------------------------------------------------------------------------
 * Learned precision router for posv (no reference counterpart).  A small
 * network picks, per matrix, the cheapest tier expected to reach the
 * tolerance: 0 = bf16 factor + float64 refinement, 1 = fp32 factor + float64
 * refinement, 2 = float64 direct.  Every answer is still judged by the float64
 * residual; refinement that does not converge falls back to float64.
 *   features  (count, 12) float: one pass over the lower triangle of A.
 *   params    TSDX_HIP_ROUTER_PARAMS floats: feature means (12) and scales
 *             (12), then W1 (64x12), b1, W2 (64x64), b2, W3 (3x64), b3,
 *             row-major [out][in]; ReLU hidden layers.
 * Training: tsdx_hip_router_grad gives the gradient of the mean softmax
 * cross-entropy (and the loss) over a batch of (features, label); the first
 * 24 gradient entries are zero.  tsdx_hip_router_adam applies one Adam step
 * (beta1 0.9, beta2 0.999, eps 1e-8; step counts from 1) to W1..b3.
 * ------------------------------------------------------------------------ */
#define TSDX_HIP_ROUTER_FEATURES 12
#define TSDX_HIP_ROUTER_TIERS 3
#define TSDX_HIP_ROUTER_PARAMS 5211

int tsdx_hip_router_features_f64(const double *A, int n, int64_t batch, float *feat, void *stream);
int tsdx_hip_router_forward(const float *params, const float *feat, int64_t count, float *logits, void *stream);
/* Inference as the routed posv runs it: the network on the matrix cores with
 * bfloat16 weights and activations and float accumulation (the feature
 * normalisation folded into the first layer), so its tiers are the routed
 * solve's choices for the same features; logits within bfloat16 rounding of
 * tsdx_hip_router_forward's.  tier (count) and / or logits (count, 3); work:
 * tsdx_hip_router_infer_workspace_size() bytes of device memory (16-byte
 * aligned) for the prepared weights. */
size_t tsdx_hip_router_infer_workspace_size(void);
int tsdx_hip_router_infer(const float *params, const float *feat, int64_t count, int *tier, float *logits, void *work,
                          size_t work_bytes, void *stream);
size_t tsdx_hip_router_grad_workspace_size(int64_t count);
int tsdx_hip_router_grad(const float *params, const float *feat, const int *label, int64_t count, float *grad,
                         float *loss, void *work, size_t work_bytes, void *stream);
int tsdx_hip_router_adam(float *params, const float *grad, float *m, float *v, int step, float lr, void *stream);

/* Routed posv (float64 A, B, X).  One kernel reads each matrix once, computes
 * its features, runs the network and solves in the chosen tier; tier-2
 * matrices and every refinement that does not converge then take the float64
 * solve.  stat as for tsdx_hip_posv_ir_f64; tier (batch) and feat (batch, 12)
 * may be NULL, else they receive the network's choice and the features (the
 * same definitions as tsdx_hip_router_features_f64).  Only the lower triangle
 * of A is read.  work: device workspace of
 * tsdx_hip_posv_routed_workspace_size(batch) bytes, 16-byte aligned (the
 * float64 and bfloat16 tier lists and the network prepared for the matrix
 * cores, rewritten on every call); calls sharing one workspace must not
 * overlap.  The network runs as in tsdx_hip_router_infer. */
size_t tsdx_hip_posv_routed_workspace_size(int64_t batch);
int tsdx_hip_posv_routed_f64(const double *A, int n, int64_t batch, const double *B, double *X, int nrhs,
                             const float *params, int max_iter, double tol, int *info, int *stat, int *tier,
                             float *feat, void *work, size_t work_bytes, void *stream);

/* ---------------------------------------------------------------------------
 * Ragged batches: matrices of different sizes (1..64) and kinds in one call.
 * No reference counterpart (the reference drivers take one n per call,
 * batched.hpp:121-244); each matrix gets exactly the result of the per-size
 * call of its kind (tsdx_hip_potrf / tsdx_hip_getrf), bit for bit.
 *   A          device buffer; matrix m (row-major n_m x n_m) starts at element
 *              offsets[m]; factored in place.
 *   sizes      (count) n_m;  kinds (count) TSDX_HIP_KIND_CHOLESKY or _LU.
 *   piv        pivots of LU matrix m at element piv_offsets[m] (n_m ints);
 *              may be NULL (then piv_offsets may be NULL).
 *   info       (count) LAPACK codes, may be NULL.
 * A plan reads sizes and kinds once (synchronising `stream`), buckets the
 * matrices by kind and size, and, when tune != 0, times the compiled variants
 * of each size class on synthetic matrices of the same sizes (the caller's
 * data is not touched) and keeps the fastest.  It captures the work - one
 * persistent kernel per size class fed by a work queue - into a HIP graph;
 * tsdx_hip_ragged_run replays it on `stream` (one graph launch).  The plan
 * keeps its own copy of sizes and kinds (later changes to the caller's arrays
 * do not reach it) and is bound to A, offsets, piv, piv_offsets and info,
 * whose addresses are captured: offsets and piv_offsets must not change after
 * planning.  Runs of one plan must not overlap (on one stream or several):
 * they share the plan's work counters.
 * ------------------------------------------------------------------------ */
#define TSDX_HIP_KIND_CHOLESKY 0
#define TSDX_HIP_KIND_LU 1

typedef struct tsdx_hip_ragged_plan tsdx_hip_ragged_plan;

int tsdx_hip_ragged_plan_create_f32(float *A, const int64_t *offsets, const int *sizes, const int *kinds, int64_t count,
                                    int *piv, const int64_t *piv_offsets, int *info, int tune,
                                    tsdx_hip_ragged_plan **plan, void *stream);
int tsdx_hip_ragged_plan_create_f64(double *A, const int64_t *offsets, const int *sizes, const int *kinds,
                                    int64_t count, int *piv, const int64_t *piv_offsets, int *info, int tune,
                                    tsdx_hip_ragged_plan **plan, void *stream);
int tsdx_hip_ragged_run(tsdx_hip_ragged_plan *plan, void *stream);
/* JSON text: per size class the matrices, wave tasks, chosen variant, grid and
 * the tuner's timings. */
const char *tsdx_hip_ragged_plan_report(const tsdx_hip_ragged_plan *plan);
int tsdx_hip_ragged_plan_destroy(tsdx_hip_ragged_plan *plan);

/* ---------------------------------------------------------------------------
 * Machine probes of the GPU (synchronous; they allocate and free their own
 * buffers).  They replace the reference's CPU probes measure_peak_gflops /
 * measure_bandwidth_gbs (include/tinysolverdx/tinysolverdx.hpp:31-32,
 * src/runtime.cpp:262-310), giving the measured ceilings of the roofline.
 *   bandwidth: TSDX_HIP_PROBE_TRIAD (c = a + 3b, 3 x 1 GiB, bytes counted 3x)
 *              or TSDX_HIP_PROBE_COPY (c = a, bytes counted 2x); GB/s.
 *   peak:      multiply-add loop on the vector ALUs, TSDX_HIP_PROBE_FP32 or
 *              _FP64; GFLOP/s (fp32 without packed instructions).
 *   mfma:      independent chains of one matrix-core instruction:
 *              _MFMA_FP32 (16x16x4 f32), _MFMA_FP64 (16x16x4 f64),
 *              _MFMA_BF16 / _MFMA_FP16 (16x16x32, float accumulation); TFLOP/s.
 * ------------------------------------------------------------------------ */
#define TSDX_HIP_PROBE_TRIAD 0
#define TSDX_HIP_PROBE_COPY 1
#define TSDX_HIP_PROBE_FP32 0
#define TSDX_HIP_PROBE_FP64 1
#define TSDX_HIP_PROBE_MFMA_FP32 0
#define TSDX_HIP_PROBE_MFMA_FP64 1
#define TSDX_HIP_PROBE_MFMA_BF16 2
#define TSDX_HIP_PROBE_MFMA_FP16 3

int tsdx_hip_measure_bandwidth_gbs(int kind, double *gbs);
int tsdx_hip_measure_peak_gflops(int precision, double *gflops);
int tsdx_hip_measure_mfma_tflops(int kind, double *tflops);

/* ---------------------------------------------------------------------------
 * Host-pointer entry points (synchronous).  Shapes of the reference's CUDA
 * backend (cuda_backend.hpp:10-11) plus the solve and driver calls of
 * batched.hpp:248-280.
 * ------------------------------------------------------------------------ */
int tsdx_hip_cholesky_batched_f32(float *A, int n, int64_t batch, int *info);
int tsdx_hip_cholesky_batched_f64(double *A, int n, int64_t batch, int *info);
int tsdx_hip_lu_batched_f32(float *A, int n, int64_t batch, int *piv, int *info);
int tsdx_hip_lu_batched_f64(double *A, int n, int64_t batch, int *piv, int *info);
int tsdx_hip_cholesky_solve_batched_f32(const float *L, int n, int64_t batch, float *B, int nrhs);
int tsdx_hip_cholesky_solve_batched_f64(const double *L, int n, int64_t batch, double *B, int nrhs);
int tsdx_hip_lu_solve_batched_f32(const float *LU, const int *piv, int n, int64_t batch, float *B, int nrhs);
int tsdx_hip_lu_solve_batched_f64(const double *LU, const int *piv, int n, int64_t batch, double *B, int nrhs);
int tsdx_hip_posv_batched_f32(float *A, int n, int64_t batch, float *B, int nrhs, int *info);
int tsdx_hip_posv_batched_f64(double *A, int n, int64_t batch, double *B, int nrhs, int *info);
int tsdx_hip_gesv_batched_f32(float *A, int n, int64_t batch, float *B, int nrhs, int *piv, int *info);
int tsdx_hip_gesv_batched_f64(double *A, int n, int64_t batch, double *B, int nrhs, int *piv, int *info);

/* ---------------------------------------------------------------------------
 * Emulated fp16 / bfloat16: the reference's runtime-precision calls with
 * Precision::FP16 / BF16 (runtime.cpp:72-140; Emu<F> arithmetic,
 * precision.hpp:71-97; Auto resolves to the scalar kernels cholesky_rm, lu_rm,
 * cholesky_solve_rm, lu_solve_rm, kernels.hpp:18-37, 91-119, 184-227).  The
 * caller's float32 / float64 array is read with every entry rounded to the
 * format (float64 through float first), each operation is computed in float
 * and rounded to the format, and the results are written back in the caller's
 * storage type: factors (a failed Cholesky leaves the entries the reference
 * leaves untouched as the rounded input), pivots, info and solutions are
 * bit-identical to the reference's.
 *   format: TSDX_HIP_EMU_FP16 or TSDX_HIP_EMU_BF16 (TSDX_HIP_EINVAL otherwise).
 * posv / gesv are the factor followed by the solve (posv_batched,
 * gesv_batched, batched.hpp:271-280).
 * ------------------------------------------------------------------------ */
#define TSDX_HIP_EMU_FP16 1
#define TSDX_HIP_EMU_BF16 2

/* stream-ordered, device pointers */
int tsdx_hip_potrf_emu_f32(float *A, int n, int64_t batch, int *info, int format, void *stream);
int tsdx_hip_potrf_emu_f64(double *A, int n, int64_t batch, int *info, int format, void *stream);
int tsdx_hip_getrf_emu_f32(float *A, int n, int64_t batch, int *piv, int *info, int format, void *stream);
int tsdx_hip_getrf_emu_f64(double *A, int n, int64_t batch, int *piv, int *info, int format, void *stream);
int tsdx_hip_potrs_emu_f32(const float *L, int n, int64_t batch, float *B, int nrhs, int format, void *stream);
int tsdx_hip_potrs_emu_f64(const double *L, int n, int64_t batch, double *B, int nrhs, int format, void *stream);
int tsdx_hip_getrs_emu_f32(const float *LU, const int *piv, int n, int64_t batch, float *B, int nrhs, int format,
                           void *stream);
int tsdx_hip_getrs_emu_f64(const double *LU, const int *piv, int n, int64_t batch, double *B, int nrhs, int format,
                           void *stream);

/* synchronous, host pointers */
int tsdx_hip_cholesky_emu_batched_f32(float *A, int n, int64_t batch, int *info, int format);
int tsdx_hip_cholesky_emu_batched_f64(double *A, int n, int64_t batch, int *info, int format);
int tsdx_hip_lu_emu_batched_f32(float *A, int n, int64_t batch, int *piv, int *info, int format);
int tsdx_hip_lu_emu_batched_f64(double *A, int n, int64_t batch, int *piv, int *info, int format);
int tsdx_hip_cholesky_solve_emu_batched_f32(const float *L, int n, int64_t batch, float *B, int nrhs, int format);
int tsdx_hip_cholesky_solve_emu_batched_f64(const double *L, int n, int64_t batch, double *B, int nrhs, int format);
int tsdx_hip_lu_solve_emu_batched_f32(const float *LU, const int *piv, int n, int64_t batch, float *B, int nrhs,
                                      int format);
int tsdx_hip_lu_solve_emu_batched_f64(const double *LU, const int *piv, int n, int64_t batch, double *B, int nrhs,
                                      int format);
int tsdx_hip_posv_emu_batched_f32(float *A, int n, int64_t batch, float *B, int nrhs, int *info, int format);
int tsdx_hip_posv_emu_batched_f64(double *A, int n, int64_t batch, double *B, int nrhs, int *info, int format);
int tsdx_hip_gesv_emu_batched_f32(float *A, int n, int64_t batch, float *B, int nrhs, int *piv, int *info,
                                  int format);
int tsdx_hip_gesv_emu_batched_f64(double *A, int n, int64_t batch, double *B, int nrhs, int *piv, int *info,
                                  int format);

#ifdef __cplusplus
}
#endif

#endif /* TSDX_HIP_H */
