// Instantiates the gesv kernels and launchers for double, matrix sizes n <= 8.
#include "launch_impl.hpp"
TSDX_INSTANTIATE_GESV(double, 8)
