// Instantiates the ragged-batch persistent kernel for double, register class n <= 16, variant 0.
#include "ragged_launch_impl.hpp"
TSDX_INSTANTIATE_RAGGED(double, 0, 0)
