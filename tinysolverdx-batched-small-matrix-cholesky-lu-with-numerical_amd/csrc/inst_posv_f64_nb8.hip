// Instantiates the posv kernels and launchers for double, matrix sizes n <= 8.
#include "launch_impl.hpp"
TSDX_INSTANTIATE_POSV(double, 8)
