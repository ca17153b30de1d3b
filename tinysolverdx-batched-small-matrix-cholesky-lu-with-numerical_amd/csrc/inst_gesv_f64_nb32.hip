// Instantiates the gesv kernels and launchers for double, matrix sizes n <= 32.
#include "launch_impl.hpp"
TSDX_INSTANTIATE_GESV(double, 32)
