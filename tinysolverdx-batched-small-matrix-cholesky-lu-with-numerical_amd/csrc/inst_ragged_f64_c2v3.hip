// Instantiates the ragged-batch persistent kernel for double, register class n <= 64, variant 3 (three waves per SIMD).
#include "ragged_launch_impl.hpp"
TSDX_INSTANTIATE_RAGGED(double, 2, 3)
