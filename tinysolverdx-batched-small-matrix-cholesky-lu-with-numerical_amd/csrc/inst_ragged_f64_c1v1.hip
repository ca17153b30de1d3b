// Instantiates the ragged-batch persistent kernel for double, register class n <= 32, variant 1.
#include "ragged_launch_impl.hpp"
TSDX_INSTANTIATE_RAGGED(double, 1, 1)
