// zfpx_scan.hip -- exclusive scan of per-block bit lengths into stream offsets
// (variable-rate placement), the fill that zeroes a stream's word range, and
// the batch helpers.
//
// The scan is one launch (after a fill that clears its state): a tile of 256
// threads x 8 lengths takes a ticket, scans its lengths within each wave with
// DPP row shifts and row broadcasts and across its four waves through LDS,
// and finds the lengths of all earlier tiles by the decoupled look-back of
// zfpx_scan.h.
#include <hip/hip_runtime.h>

#include "zfpx_api.h"
#include "zfpx_fx.h"
#include "zfpx_scan.h"

namespace {

constexpr int kT = 256;
constexpr int kItems = 8;
constexpr uint64_t kTile = (uint64_t)kT * kItems;

using zx::lookback;
using zx::wave_inclusive_sum;

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
    const uint64_t p = lookback(state + 2, tile, total, start, state + 1);
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

__global__ __launch_bounds__(kT) void k_op_merge(uint64_t *words, uint64_t nwords, const uint64_t *state,
                                                 const uint64_t *edges, uint64_t groups, uint64_t start,
                                                 uint64_t prefix)
{
  const uint64_t g = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (g >= groups) return;
  zx::op_merge(words, nwords, state + 2, edges, g, groups, start, prefix);
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

hipError_t zx_launch_op_merge(uint64_t *words, uint64_t nwords, const uint64_t *state,
                              const uint64_t *edges, uint64_t groups, uint64_t start, uint64_t prefix,
                              hipStream_t s)
{
  if (groups == 0) return hipSuccess;
  hipLaunchKernelGGL(k_op_merge, dim3((groups + kT - 1) / kT), dim3(kT), 0, s, words, nwords, state, edges,
                     groups, start, prefix);
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
