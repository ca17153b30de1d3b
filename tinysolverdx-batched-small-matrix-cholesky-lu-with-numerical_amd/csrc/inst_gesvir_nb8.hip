// Mixed-precision gesv launcher of bucket 8 (kernels in gesv_ir.hpp).
#include "gesv_ir.hpp"
template hipError_t tsdx_hip::launch_gesv_ir_nb<8>(const IrArgs&, int, bool, hipStream_t);
