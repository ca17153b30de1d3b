// Precision and backend descriptors of the C++ host API (namespace tsdx), the
// counterpart of the reference's include/tinysolverdx/precision.hpp:32-33,
// :41-65, :118-126.  Backend keeps the reference's enumerator order and appends
// Hip, the backend this library implements.
#pragma once

#include <cstdint>

namespace tsdx {

enum class Precision { FP32, FP64, FP16, BF16 };
enum class Backend { Auto, Naive, Scalar, Simd, Cuda, Hip };

namespace detail {
// Nearest IEEE binary16 / bfloat16 value of x (ties to even, subnormals kept,
// overflow to infinity; NaN and infinities pass through unchanged).
float round_fp16(float x);
float round_bf16(float x);
}  // namespace detail

// Unit roundoff u = 2^-p of each format.
double unit_roundoff(Precision p);

template <class T>
struct is_native {
  static constexpr bool value = false;
};
template <>
struct is_native<float> {
  static constexpr bool value = true;
};
template <>
struct is_native<double> {
  static constexpr bool value = true;
};

}  // namespace tsdx
