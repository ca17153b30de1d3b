// Instantiates the factor kernels and launchers for double, matrix sizes n <= 64.
#include "launch_impl.hpp"
TSDX_INSTANTIATE_FACTOR(double, 64)
