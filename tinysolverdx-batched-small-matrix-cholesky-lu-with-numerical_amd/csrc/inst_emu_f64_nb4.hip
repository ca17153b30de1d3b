// Instantiates the emulated fp16 / bf16 kernels and launchers for double storage, n <= 4.
#include "emu_launch_impl.hpp"
TSDX_INSTANTIATE_EMU(double, 4)
