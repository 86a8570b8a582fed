// Mixed-precision (FP32 segment arithmetic) instantiations of wfl_kernels.h.
#include "wfl_kernels.h"
#define WFL_R float
#define WFL_SUFFIX f32
#include "wfl_dispatch.inc"
