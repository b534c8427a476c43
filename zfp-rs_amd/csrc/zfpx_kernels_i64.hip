// Kernels of the field codec for int64_t fields (see zfpx_kernels.inc.h).
#include "zfpx_kernels.inc.h"

ZX_INSTANTIATE(int64_t, i64)
