// zfpx_scan.hip -- exclusive scan of per-block bit lengths into stream offsets
// (variable-rate placement), the fill that zeroes a stream's word range, and
// the batch helpers.
//
// The scan is one launch (after a fill that clears its state): a tile of 256
// threads x 8 lengths takes a ticket, so tiles start in ticket order, scans
// its lengths within each wave with DPP row shifts and row broadcasts and
// across its four waves through LDS, and finds the lengths of all earlier
// tiles by a decoupled look-back (Merrill and Garland, "Single-pass parallel
// prefix scan with decoupled look-back", NVIDIA TR 2016): it publishes its own
// total at once, then one wave reads the statuses of the 64 tiles before it in
// one step, adds the totals down to the nearest tile whose inclusive prefix is
// known (found with a ballot) and publishes its own inclusive prefix.
#include <hip/hip_runtime.h>

#include "zfpx_api.h"

namespace {

constexpr int kT = 256;
constexpr int kItems = 8;
constexpr uint64_t kTile = (uint64_t)kT * kItems;

// tile status words: the value in the low 62 bits, the kind in the top two
constexpr uint64_t kAgg = 1ull << 62;   // the tile's own total
constexpr uint64_t kIncl = 2ull << 62;  // the total of every tile up to and including it
constexpr uint64_t kValue = (1ull << 62) - 1;
// a look-back that waits this many rounds gives up and marks the scan
// failed (state word 1); the tile then takes a prefix no stream reaches
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

}  // namespace

// The exclusive prefix of the tile with ticket `tile` whose own total is
// `total`, by one wave (all 64 lanes call it): `status` holds one word per
// tile, cleared before the launch; `failed` is set when the look-back gives
// up.  Tile 0's prefix is `start`.
__device__ uint64_t zx_lookback(uint64_t *status, uint64_t tile, uint64_t total, uint64_t start,
                                uint64_t *failed)
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

namespace {

// state[0]: ticket counter; state[1]: failure flag; state[2 + t]: tile t
__global__ __launch_bounds__(kT) void k_scan(const uint32_t *len, uint64_t n, uint64_t start,
                                             uint64_t *off, uint64_t *state)
{
  __shared__ uint32_t tile_s;
  __shared__ uint32_t wave_tot[kT / 64];
  __shared__ uint64_t prefix_s;
  if (threadIdx.x == 0) tile_s = (uint32_t)atomicAdd((unsigned long long *)&state[0], 1ull);
  __syncthreads();
  const uint64_t tile = tile_s;
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t base = tile * kTile + (uint64_t)threadIdx.x * kItems;
  uint32_t v[kItems];
  uint32_t s = 0;
  for (int i = 0; i < kItems; i++) {
    v[i] = base + i < n ? len[base + i] : 0u;
    s += v[i];
  }
  const uint32_t inc = wave_inclusive_sum(s);
  if (lane == 63) wave_tot[wave] = inc;
  __syncthreads();
  uint32_t before = 0, total = 0;
  for (int w = 0; w < kT / 64; w++) {
    before += w < (int)wave ? wave_tot[w] : 0u;
    total += wave_tot[w];
  }
  if (wave == 0) {
    const uint64_t p = zx_lookback(state + 2, tile, total, start, state + 1);
    if (lane == 0) prefix_s = p;
  }
  __syncthreads();
  uint64_t o = prefix_s + before + (inc - s);
  for (int i = 0; i < kItems; i++) {
    if (base + i < n) off[base + i] = o;
    o += v[i];
  }
  if (base < n && base + kItems >= n) off[n] = o;  // the thread holding the last length
}

__global__ void k_fill(uint64_t *w, uint64_t n, uint64_t first)
{
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) w[i] = i == 0 ? first : 0ull;
}

// ------------------------------------------------------------------ batches
// Per-array work is cut into chunks of kChunk items (words or blocks) spread
// over the whole grid: first[j] is array j's first chunk, and a workgroup
// finds its array by binary search.

constexpr uint64_t kChunk = 8192;

__device__ uint32_t chunk_job(const uint64_t *first, uint32_t n, uint64_t c)
{
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (first[mid] <= c) lo = mid;
    else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(kT) void k_batch_gather_index(const zx::BatchJob *jobs, const uint64_t *first,
                                                           uint32_t n, uint32_t *lengths)
{
  const uint32_t j = chunk_job(first, n, blockIdx.x);
  const zx::BatchJob &J = jobs[j];
  if (J.fixed || !J.index) return;
  const uint64_t b0 = (blockIdx.x - first[j]) * kChunk;
  const uint64_t b1 = b0 + kChunk < J.field.nblocks ? b0 + kChunk : J.field.nblocks;
  for (uint64_t b = b0 + threadIdx.x; b < b1; b += kT) lengths[J.lbase + b] = J.index[b];
}

__global__ void k_batch_ends(const zx::BatchJob *jobs, uint32_t n, const uint64_t *offsets)
{
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const zx::BatchJob &J = jobs[j];
  const uint64_t nb = J.field.nblocks;
  J.result[0] = J.fixed ? J.start + nb * (uint64_t)J.fixed
                        : J.start + (offsets[J.lbase + nb] - offsets[J.lbase]);
}

// zeroes an array's words through its last block's word, keeping the prefix
// bits of the first (after k_batch_ends)
__global__ __launch_bounds__(kT) void k_batch_fill(const zx::BatchJob *jobs, const uint64_t *first, uint32_t n)
{
  const uint32_t j = chunk_job(first, n, blockIdx.x);
  const zx::BatchJob &J = jobs[j];
  if (!J.fill) return;
  uint64_t nw = (J.result[0] + 63) / 64;
  if (nw > J.nwords) nw = J.nwords;
  const uint64_t w0 = (blockIdx.x - first[j]) * kChunk;
  const uint64_t w1 = w0 + kChunk < nw ? w0 + kChunk : nw;
  for (uint64_t i = w0 + threadIdx.x; i < w1; i += kT) J.words[i] = i == 0 ? J.prefix : 0ull;
}

__global__ __launch_bounds__(kT) void k_batch_keep_index(const zx::BatchJob *jobs, const uint64_t *first,
                                                         uint32_t n, const uint32_t *lengths)
{
  const uint32_t j = chunk_job(first, n, blockIdx.x);
  const zx::BatchJob &J = jobs[j];
  if (J.fixed || !J.index_out) return;
  const uint64_t b0 = (blockIdx.x - first[j]) * kChunk;
  const uint64_t b1 = b0 + kChunk < J.field.nblocks ? b0 + kChunk : J.field.nblocks;
  for (uint64_t b = b0 + threadIdx.x; b < b1; b += kT) J.index_out[b] = lengths[J.lbase + b];
}

}  // namespace

uint64_t zx_scan_tiles(uint64_t n) { return (n + kTile - 1) / kTile + 2; }

hipError_t zx_launch_scan(const uint32_t *lengths, uint64_t n, uint64_t start, uint64_t *offsets,
                          uint64_t *tiles, hipStream_t s)
{
  if (n == 0) {
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(1), 0, s, offsets, (uint64_t)1, start);
    return hipGetLastError();
  }
  const uint64_t nt = (n + kTile - 1) / kTile;
  hipError_t e = zx_launch_fill(tiles, nt + 2, 0, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_scan, dim3((uint32_t)nt), dim3(kT), 0, s, lengths, n, start, offsets, tiles);
  return hipGetLastError();
}

hipError_t zx_launch_fill(uint64_t *words, uint64_t n, uint64_t first, hipStream_t s)
{
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3((n + 255) / 256), dim3(256), 0, s, words, n, first);
  return hipGetLastError();
}

uint64_t zx_batch_chunks(uint64_t items) { return (items + kChunk - 1) / kChunk; }

hipError_t zx_batch_gather_index(const zx::BatchJob *jobs, const uint64_t *first, uint32_t n,
                                 uint64_t nchunks, uint32_t *lengths, hipStream_t s)
{
  if (nchunks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_batch_gather_index, dim3(nchunks), dim3(kT), 0, s, jobs, first, n, lengths);
  return hipGetLastError();
}

hipError_t zx_batch_ends(const zx::BatchJob *jobs, uint32_t n, const uint64_t *offsets, hipStream_t s)
{
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_batch_ends, dim3((n + 255) / 256), dim3(256), 0, s, jobs, n, offsets);
  return hipGetLastError();
}

hipError_t zx_batch_fill(const zx::BatchJob *jobs, const uint64_t *first, uint32_t n, uint64_t nchunks,
                         hipStream_t s)
{
  if (nchunks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_batch_fill, dim3(nchunks), dim3(kT), 0, s, jobs, first, n);
  return hipGetLastError();
}

hipError_t zx_batch_keep_index(const zx::BatchJob *jobs, const uint64_t *first, uint32_t n,
                               uint64_t nchunks, const uint32_t *lengths, hipStream_t s)
{
  if (nchunks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_batch_keep_index, dim3(nchunks), dim3(kT), 0, s, jobs, first, n, lengths);
  return hipGetLastError();
}
