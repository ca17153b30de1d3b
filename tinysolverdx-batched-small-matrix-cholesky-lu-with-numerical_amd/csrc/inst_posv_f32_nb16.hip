// Instantiates the posv kernels and launchers for float, matrix sizes n <= 16.
#include "launch_impl.hpp"
TSDX_INSTANTIATE_POSV(float, 16)
