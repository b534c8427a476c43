// Kernels of the field codec for float fields (see zfpx_kernels.inc.h).
#include "zfpx_kernels.inc.h"

ZX_INSTANTIATE(float, f32)
