// Instantiates the emulated fp16 / bf16 kernels and launchers for float storage, n <= 32.
#include "emu_launch_impl.hpp"
TSDX_INSTANTIATE_EMU(float, 32)
