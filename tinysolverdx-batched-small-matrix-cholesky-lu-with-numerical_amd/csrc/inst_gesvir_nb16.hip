// Mixed-precision gesv launcher of bucket 16 (kernels in gesv_ir.hpp).
#include "gesv_ir.hpp"
template hipError_t tsdx_hip::launch_gesv_ir_nb<16>(const IrArgs&, int, bool, hipStream_t);
