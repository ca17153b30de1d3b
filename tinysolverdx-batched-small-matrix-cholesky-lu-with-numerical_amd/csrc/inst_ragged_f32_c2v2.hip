// Instantiates the ragged-batch persistent kernel for float, register class n <= 64, variant 2.
#include "ragged_launch_impl.hpp"
TSDX_INSTANTIATE_RAGGED(float, 2, 2)
