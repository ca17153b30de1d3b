// Instantiates the posv kernels and launchers for double, matrix sizes n <= 32.
#include "launch_impl.hpp"
TSDX_INSTANTIATE_POSV(double, 32)
