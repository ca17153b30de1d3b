// Instantiates the gesv kernels and launchers for float, matrix sizes n <= 4.
#include "launch_impl.hpp"
TSDX_INSTANTIATE_GESV(float, 4)
