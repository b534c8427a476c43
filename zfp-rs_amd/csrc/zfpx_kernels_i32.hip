// Kernels of the field codec for int32_t fields (see zfpx_kernels.inc.h).
#include "zfpx_kernels.inc.h"

ZX_INSTANTIATE(int32_t, i32)
