// Mixed-precision gesv launcher of bucket 4 (kernels in gesv_ir.hpp).
#include "gesv_ir.hpp"
template hipError_t tsdx_hip::launch_gesv_ir_nb<4>(const IrArgs&, int, bool, hipStream_t);
