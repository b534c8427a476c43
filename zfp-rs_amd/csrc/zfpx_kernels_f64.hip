// Kernels of the field codec for double fields (see zfpx_kernels.inc.h).
#include "zfpx_kernels.inc.h"

ZX_INSTANTIATE(double, f64)
