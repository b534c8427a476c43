// zfpx_scan.hip -- exclusive scan of per-block bit lengths into stream offsets
// (variable-rate placement), and the fill that zeroes a stream's word range.
//
// Three passes: per-tile sums, one workgroup scanning the tile sums, then each
// tile rescans its lengths with the tile's base added.  A tile is 256 threads
// x 8 lengths.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "zfpx_api.h"

namespace {

constexpr int kT = 256;
constexpr int kItems = 8;
constexpr uint64_t kTile = (uint64_t)kT * kItems;

__global__ __launch_bounds__(kT) void k_tile_sums(const uint32_t *len, uint64_t n, uint64_t *tiles)
{
  typedef hipcub::BlockReduce<uint64_t, kT> Reduce;
  __shared__ typename Reduce::TempStorage tmp;
  const uint64_t base = (uint64_t)blockIdx.x * kTile;
  uint64_t s = 0;
  for (int i = 0; i < kItems; i++) {
    const uint64_t j = base + (uint64_t)i * kT + threadIdx.x;
    if (j < n) s += len[j];
  }
  const uint64_t t = Reduce(tmp).Sum(s);
  if (threadIdx.x == 0) tiles[blockIdx.x] = t;
}

__global__ __launch_bounds__(kT) void k_scan_tiles(uint64_t *tiles, uint64_t ntiles, uint64_t start)
{
  typedef hipcub::BlockScan<uint64_t, kT> Scan;
  __shared__ typename Scan::TempStorage tmp;
  __shared__ uint64_t carry;
  if (threadIdx.x == 0) carry = start;
  __syncthreads();
  for (uint64_t b = 0; b < ntiles; b += kT) {
    const uint64_t j = b + threadIdx.x;
    const uint64_t v = j < ntiles ? tiles[j] : 0;
    uint64_t ex, total;
    Scan(tmp).ExclusiveSum(v, ex, total);
    const uint64_t c = carry;
    if (j < ntiles) tiles[j] = c + ex;
    __syncthreads();
    if (threadIdx.x == 0) carry = c + total;
    __syncthreads();
  }
  if (threadIdx.x == 0) tiles[ntiles] = carry;
}

__global__ __launch_bounds__(kT) void k_apply(const uint32_t *len, uint64_t n, const uint64_t *tiles,
                                              uint64_t *off)
{
  typedef hipcub::BlockScan<uint64_t, kT> Scan;
  __shared__ typename Scan::TempStorage tmp;
  // thread t owns lengths base + t*kItems .. + kItems (blocked arrangement)
  const uint64_t base = (uint64_t)blockIdx.x * kTile + (uint64_t)threadIdx.x * kItems;
  uint64_t v[kItems];
  uint64_t s = 0;
  for (int i = 0; i < kItems; i++) {
    v[i] = base + i < n ? len[base + i] : 0;
    s += v[i];
  }
  uint64_t ex;
  Scan(tmp).ExclusiveSum(s, ex);
  uint64_t o = tiles[blockIdx.x] + ex;
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

uint64_t zx_scan_tiles(uint64_t n) { return (n + kTile - 1) / kTile + 1; }

hipError_t zx_launch_scan(const uint32_t *lengths, uint64_t n, uint64_t start, uint64_t *offsets,
                          uint64_t *tiles, hipStream_t s)
{
  if (n == 0) {
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(1), 0, s, offsets, (uint64_t)1, start);
    return hipGetLastError();
  }
  const uint64_t nt = (n + kTile - 1) / kTile;
  hipLaunchKernelGGL(k_tile_sums, dim3(nt), dim3(kT), 0, s, lengths, n, tiles);
  hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kT), 0, s, tiles, nt, start);
  hipLaunchKernelGGL(k_apply, dim3(nt), dim3(kT), 0, s, lengths, n, tiles, offsets);
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
