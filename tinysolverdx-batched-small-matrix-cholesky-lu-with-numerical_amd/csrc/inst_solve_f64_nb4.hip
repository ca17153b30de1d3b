// Instantiates the solve kernels and launchers for double, matrix sizes n <= 4.
#include "launch_impl.hpp"
TSDX_INSTANTIATE_SOLVE(double, 4)
