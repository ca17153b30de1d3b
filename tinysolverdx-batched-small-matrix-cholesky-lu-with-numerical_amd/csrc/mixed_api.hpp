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
  // Optional subset for the refinement kernel: when list is set, workgroup w
  // solves matrix list[w] for w < *lcount (the router's tier lists).
  const int* list = nullptr;
  const int* lcount = nullptr;
};

// low: 1 float, 2 bfloat16, 3 half (refinement), 4 plain float64 for all, 5
// the float64 fallback tier for all.
hipError_t launch_posv_ir(const IrArgs& a, int low, hipStream_t s);
// float64 posv of the matrices on a.flist[0 .. *a.fcount).
hipError_t launch_posv_gather(const IrArgs& a, hipStream_t s);

}  // namespace tsdx_hip
