// The DLPack exchange structures (the public DLPack C ABI, version 0.x
// "legacy" capsules named "dltensor"), declared here because the toolchain
// ships no dlpack.h.  Only what the device-array path reads and writes.
#pragma once

#include <cstdint>

namespace tsdx_py {

// Device types of the DLPack ABI used here.
constexpr int32_t kDLCUDA = 2;
constexpr int32_t kDLROCM = 10;
// Data type codes.
constexpr uint8_t kDLInt = 0;
constexpr uint8_t kDLFloat = 2;

struct DLDevice {
  int32_t device_type;
  int32_t device_id;
};

struct DLDataType {
  uint8_t code;
  uint8_t bits;
  uint16_t lanes;
};

struct DLTensor {
  void* data;
  DLDevice device;
  int32_t ndim;
  DLDataType dtype;
  int64_t* shape;
  int64_t* strides;  // in elements; NULL = compact row-major
  uint64_t byte_offset;
};

struct DLManagedTensor {
  DLTensor dl_tensor;
  void* manager_ctx;
  void (*deleter)(DLManagedTensor* self);
};

}  // namespace tsdx_py
