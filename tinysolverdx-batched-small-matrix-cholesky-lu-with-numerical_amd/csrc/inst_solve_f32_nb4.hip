// Instantiates the solve kernels and launchers for float, matrix sizes n <= 4.
#include "launch_impl.hpp"
TSDX_INSTANTIATE_SOLVE(float, 4)
