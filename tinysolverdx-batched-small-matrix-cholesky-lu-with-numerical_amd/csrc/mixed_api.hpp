// Host-side declaration of the refinement-tier launcher (mixed_posv.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tsdx_hip {

struct IrArgs {
  const double* A;
  const double* B;
  double* X;
  int n;
  int64_t batch;
  int nrhs;
  int max_iter;
  double tol;
  int* info;    // (batch) or null
  int* stat;    // (batch) or null: corrections used, or -1 when sent to the float64 solve
  int* flist;   // (batch) matrices handed to the float64 solve
  int* fcount;  // (1) their number
  // The routed solve's list of matrices for the bfloat16 tier (filled by the
  // routed kernel, solved by a list launch of the refinement kernel): list
  // llist[0 .. *lcount).
  int* llist = nullptr;
  int* lcount = nullptr;
  // First matrix of this launch: workgroup w of the one-wave-per-matrix
  // kernels solves matrix m0 + w.  Large batches are split into launches of
  // at most kIrChunk workgroups (a grid must stay below 2^32 work items).
  int64_t m0 = 0;
};
constexpr int64_t kIrChunk = int64_t(1) << 24;

// low: 1 float, 2 bfloat16, 3 half (refinement), 4 plain float64 for all, 5
// the float64 fallback tier for all.
hipError_t launch_posv_ir(const IrArgs& a, int low, hipStream_t s);
// float64 posv of the matrices on a.flist[0 .. *a.fcount).
hipError_t launch_posv_gather(const IrArgs& a, hipStream_t s);

// Mixed-precision gesv of bucket NB (gesv_ir.hpp): refinement in low (1 float,
// 2 bfloat16, 3 half), then the float64 gesv of the list a.flist, whose count
// *a.fcount the caller has cleared.  ztile: zero-pivot rule of the float64
// gesv (divide by one, n <= 24).
template <int NB>
hipError_t launch_gesv_ir_nb(const IrArgs& a, int low, bool ztile, hipStream_t s);

}  // namespace tsdx_hip
