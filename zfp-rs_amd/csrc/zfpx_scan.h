// zfpx_scan.h -- wave-level prefix sums and the decoupled look-back that
// the offset scan (zfpx_scan.hip) and the one-pass compress kernel share.
//
// A tile (a workgroup's share of the items) takes a ticket, so tiles start in
// ticket order, and finds the totals of all earlier tiles by a decoupled
// look-back (Merrill and Garland, "Single-pass parallel prefix scan with
// decoupled look-back", NVIDIA TR 2016): it publishes its own total at once,
// then one wave reads the statuses of the 64 tiles before it in one step,
// adds the totals down to the nearest tile whose inclusive prefix is known
// (found with a ballot) and publishes its own inclusive prefix.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zx {

// tile status words: the value in the low 62 bits, the kind in the top two
constexpr uint64_t kAgg = 1ull << 62;   // the tile's own total
constexpr uint64_t kIncl = 2ull << 62;  // the total of every tile up to and including it
constexpr uint64_t kValue = (1ull << 62) - 1;
// a look-back that waits this many rounds gives up and marks the scan
// failed; the tile then takes a prefix no stream reaches
constexpr uint32_t kSpinLimit = 1u << 24;
constexpr uint64_t kFailedPrefix = 1ull << 61;

// Inclusive prefix sum of v over the 64 lanes of the wave: Hillis-Steele
// within each 16-lane row (row_shr 1, 2, 4, 8; a lane with no source in its
// row adds 0), then row 0's last lane into row 1 and row 2's into row 3
// (row_bcast:15 on rows 1 and 3), then lane 31 into rows 2 and 3
// (row_bcast:31).
__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v)
{
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);
  return v;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v)
{
  for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

// Status words carry their own value and nothing else is read through them,
// so relaxed atomics at device scope suffice (an acquire or release here
// would invalidate or write back the caches at every access).
__device__ __forceinline__ uint64_t load_status(const uint64_t *p)
{
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_status(uint64_t *p, uint64_t v)
{
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The exclusive prefix of the tile with ticket `tile` whose own total is
// `total`, by one wave (all 64 lanes call it): `status` holds one word per
// tile, cleared before the launch; `failed` is set when the look-back gives
// up.  Tile 0's prefix is `start`.
__device__ __forceinline__ uint64_t lookback(uint64_t *status, uint64_t tile, uint64_t total,
                                             uint64_t start, uint64_t *failed)
{
  const uint32_t lane = threadIdx.x & 63;
  if (tile == 0) {
    if (lane == 0) store_status(&status[0], kIncl | ((start + total) & kValue));
    return start;
  }
  if (lane == 0) store_status(&status[tile], kAgg | total);
  uint64_t prefix = 0;
  int64_t j = (int64_t)tile - 1;
  uint32_t spins = 0;
  for (;;) {
    const int64_t k = j - (int64_t)lane;
    const uint64_t st = k >= 0 ? load_status(&status[k]) : kIncl;  // before tile 0: nothing
    const uint64_t incl = __ballot(st >= kIncl);
    const uint64_t none = __ballot(st < kAgg);
    const uint32_t first = incl ? (uint32_t)__builtin_ctzll(incl) : 64u;
    const uint64_t below = first < 64 ? (1ull << first) - 1 : ~0ull;
    if (none & below) {
      // a tile between this one and the nearest inclusive prefix has not
      // published yet: read the window again
      if (++spins > kSpinLimit) {
        if (lane == 0) *failed = 1;
        prefix = kFailedPrefix;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    prefix += wave_sum_u64(lane <= first ? (st & kValue) : 0ull);
    if (first < 64) break;
    j -= 64;
  }
  if (lane == 0) store_status(&status[tile], kIncl | ((prefix + total) & kValue));
  return prefix;
}

}  // namespace zx
