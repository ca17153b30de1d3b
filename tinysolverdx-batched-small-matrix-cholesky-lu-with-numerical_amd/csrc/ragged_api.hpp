// Host-side declarations of the ragged-batch launchers (ragged_kernels.hpp),
// one explicit instantiation per (type, register class, variant) unit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tsdx_hip {

template <class T>
struct RaggedArgs;
struct RTask;

template <class T, int CLS, int VAR>
hipError_t launch_ragged(const RaggedArgs<T>& a, const RTask* tasks, int ntasks, int* counter, int grid, hipStream_t s);
// Workgroups of the variant resident per CU (occupancy), and its task capacity
// (matrices per wave task) for a bucket.
template <class T, int CLS, int VAR>
int ragged_blocks_per_cu();
template <class T, int CLS, int VAR>
int ragged_capacity(int code);
template <class T, int CLS, int VAR>
int ragged_wpb();

}  // namespace tsdx_hip
