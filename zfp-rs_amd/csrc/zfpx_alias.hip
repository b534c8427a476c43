// zfpx_alias.hip -- decoding into a field whose strides may alias.
//
// The blocks were decoded in parallel into a dense copy of the field; each
// element of the field then keeps the value with the largest ticket aimed at
// it (zfpx_alias.h: the serial decode's write order), found with an atomic
// max, and only that value stores.  The result equals the serial decode's
// whatever order the threads run in.
#include <hip/hip_runtime.h>

#include "zfpx_api.h"

namespace {

constexpr int kT = 256;

__global__ __launch_bounds__(kT) void k_alias_tickets(zx::AliasArgs a, unsigned long long *best)
{
  for (uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x; i < a.count;
       i += (uint64_t)gridDim.x * kT) {
    const zx::AliasSlot q = zx::alias_slot(a, i);
    atomicMax(best + q.addr, (unsigned long long)q.ticket);
  }
}

template <class W>
__global__ __launch_bounds__(kT) void k_alias_store(zx::AliasArgs a, const unsigned long long *best)
{
  const W *src = (const W *)a.dense;
  W *dst = (W *)a.lowest;
  for (uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x; i < a.count;
       i += (uint64_t)gridDim.x * kT) {
    const zx::AliasSlot q = zx::alias_slot(a, i);
    if (best[q.addr] == q.ticket) dst[q.addr] = src[i];
  }
}

}  // namespace

hipError_t zx_launch_alias_scatter(const zx::AliasArgs &a, unsigned long long *best, hipStream_t s)
{
  if (a.count == 0) return hipSuccess;
  if (a.esz != 4 && a.esz != 8) return hipErrorInvalidValue;
  hipError_t e = zx_launch_fill((uint64_t *)best, a.span, 0, s);
  if (e != hipSuccess) return e;
  uint64_t g = (a.count + kT - 1) / kT;
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(k_alias_tickets, dim3((uint32_t)g), dim3(kT), 0, s, a, best);
  if (a.esz == 4) hipLaunchKernelGGL(k_alias_store<uint32_t>, dim3((uint32_t)g), dim3(kT), 0, s, a, best);
  else hipLaunchKernelGGL(k_alias_store<uint64_t>, dim3((uint32_t)g), dim3(kT), 0, s, a, best);
  return hipGetLastError();
}
