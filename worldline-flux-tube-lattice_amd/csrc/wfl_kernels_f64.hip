// FP64 instantiations of wfl_kernels.h.
#include "wfl_kernels.h"
#define WFL_R double
#define WFL_SUFFIX f64
#include "wfl_dispatch.inc"
