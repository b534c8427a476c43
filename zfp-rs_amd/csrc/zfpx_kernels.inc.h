// zfpx_kernels.inc.h -- HIP kernels of the field codec, instantiated per
// scalar type by zfpx_kernels_{i32,i64,f32,f64}.hip.
//
// One thread codes one block.  Blocks are placed in the stream either at a
// fixed stride (fixed-rate: block b starts at start + b * bits, the
// reference's per-chunk seek, decompress.rs:205-330) or at offsets from an
// exclusive scan of per-block lengths (variable-rate).  Stream words a block
// owns outright are stored whole; the few words shared by neighbouring blocks
// are merged with an atomic OR into words the host zeroed first.
#pragma once

#include <hip/hip_runtime.h>

#include "zfpx_api.h"
#include "zfpx_field.h"

namespace zx {

constexpr int kThreads = 256;

// Rows of four contiguous values are loaded and stored as 16-byte vectors
// when the host has checked that every row of every whole block is aligned.
template <class S, int D>
__device__ __forceinline__ void gather_rows(const FieldDesc &f, const BlockPos &p,
                                            typename Ints<ST<S>::P>::U (&v)[Blk<D>::SIZE])
{
  typedef typename Ints<ST<S>::P>::U U;
  const U *base = (const U *)f.data + p.origin;
  ZX_UNROLL
  for (int r = 0; r < Blk<D>::SIZE / 4; r++) {
    int64_t off = 0;
    if (D >= 2) off += (int64_t)(r & 3) * f.s[1];
    if (D >= 3) off += (int64_t)(r >> 2) * f.s[2];
    if constexpr (sizeof(U) == 4) {
      const uint4 q = *(const uint4 *)(base + off);
      v[4 * r + 0] = q.x; v[4 * r + 1] = q.y; v[4 * r + 2] = q.z; v[4 * r + 3] = q.w;
    } else {
      const ulonglong2 a = *(const ulonglong2 *)(base + off);
      const ulonglong2 b = *(const ulonglong2 *)(base + off + 2);
      v[4 * r + 0] = a.x; v[4 * r + 1] = a.y; v[4 * r + 2] = b.x; v[4 * r + 3] = b.y;
    }
  }
}

template <class S, int D>
__device__ __forceinline__ void scatter_rows(const FieldDesc &f, const BlockPos &p,
                                             const typename Ints<ST<S>::P>::U (&v)[Blk<D>::SIZE])
{
  typedef typename Ints<ST<S>::P>::U U;
  U *base = (U *)f.data + p.origin;
  ZX_UNROLL
  for (int r = 0; r < Blk<D>::SIZE / 4; r++) {
    int64_t off = 0;
    if (D >= 2) off += (int64_t)(r & 3) * f.s[1];
    if (D >= 3) off += (int64_t)(r >> 2) * f.s[2];
    if constexpr (sizeof(U) == 4) {
      *(uint4 *)(base + off) = make_uint4(v[4 * r], v[4 * r + 1], v[4 * r + 2], v[4 * r + 3]);
    } else {
      ulonglong2 a, b;
      a.x = v[4 * r]; a.y = v[4 * r + 1]; b.x = v[4 * r + 2]; b.y = v[4 * r + 3];
      *(ulonglong2 *)(base + off) = a;
      *(ulonglong2 *)(base + off + 2) = b;
    }
  }
}

template <class S, int D>
__device__ __forceinline__ void load_block(const FieldDesc &f, uint64_t b, bool vec,
                                           typename Ints<ST<S>::P>::U (&v)[Blk<D>::SIZE])
{
  const BlockPos p = block_pos<D>(f, b);
  if (vec && p.full) gather_rows<S, D>(f, p, v);
  else gather<S, D>(f, p, v);
}

template <class S, int D>
__device__ __forceinline__ void store_block(const FieldDesc &f, uint64_t b, bool vec,
                                            const typename Ints<ST<S>::P>::U (&v)[Blk<D>::SIZE])
{
  const BlockPos p = block_pos<D>(f, b);
  if (vec && p.full) scatter_rows<S, D>(f, p, v);
  else scatter<S, D>(f, p, v);
}

// Encode every block; COUNT writes per-block bit lengths instead of bits.
template <class S, int D, bool REV, bool COUNT>
__global__ __launch_bounds__(kThreads) void k_encode(EncodeArgs a)
{
  const uint64_t b = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  if (b >= a.field.nblocks) return;
  typename Ints<ST<S>::P>::U v[Blk<D>::SIZE];
  load_block<S, D>(a.field, b, a.vec, v);
  if constexpr (COUNT) {
    CountSink cs{0};
    encode_values<S, D, REV>(cs, v, a.cfg);
    a.lengths[b] = (uint32_t)cs.n;
  } else {
    const uint64_t pos = a.offsets ? a.offsets[b] : a.start + b * (uint64_t)a.fixed;
    WordSink w;
    w.init(a.words, a.nwords, pos);
    encode_values<S, D, REV>(w, v, a.cfg);
    w.finish();
  }
}

template <class S, int D, bool REV>
__global__ __launch_bounds__(kThreads) void k_decode(DecodeArgs a)
{
  const uint64_t b = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  if (b >= a.field.nblocks) return;
  const uint64_t pos = a.offsets ? a.offsets[b] : a.start + b * (uint64_t)a.fixed;
  WordSource r;
  r.init(a.words, a.nwords, pos);
  typename Ints<ST<S>::P>::U v[Blk<D>::SIZE];
  decode_values<S, D, REV>(r, v, a.cfg);
  // Words this block loaded, as the reference's cursor counts them: those it
  // consumed bits of, plus the word a skip to min_bits lands inside.
  const uint64_t used = r.idx * 64 - r.bits - pos;
  uint64_t loaded = r.idx;
  if (used < a.cfg.minbits) {
    const uint64_t target = pos + a.cfg.minbits;
    if (target % 64) loaded = loaded > target / 64 + 1 ? loaded : target / 64 + 1;
  }
  atomicMax((unsigned long long *)a.maxword, (unsigned long long)loaded);
  store_block<S, D>(a.field, b, a.vec, v);
}

// Length walk of a stream without an index: one thread decodes the blocks in
// order, as the reference's serial decoder does, and records each block's
// length (what it read, raised to min_bits as skip_to does).
template <class S, int D, bool REV>
__global__ __launch_bounds__(64) void k_walk(DecodeArgs a, uint32_t *lengths)
{
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  uint64_t pos = a.start;
  for (uint64_t b = 0; b < a.field.nblocks; b++) {
    WordSource r;
    r.init(a.words, a.nwords, pos);
    typename Ints<ST<S>::P>::U v[Blk<D>::SIZE];
    decode_values<S, D, REV>(r, v, a.cfg);
    uint64_t used = r.idx * 64 - r.bits - pos;
    if (used < a.cfg.minbits) used = a.cfg.minbits;
    lengths[b] = (uint32_t)used;
    pos += used;
  }
}

template <class S, int D, bool REV>
static hipError_t launch_encode_t(const EncodeArgs &a, hipStream_t s)
{
  const uint64_t grid = (a.field.nblocks + kThreads - 1) / kThreads;
  if (grid == 0) return hipSuccess;
  if (a.lengths) hipLaunchKernelGGL((k_encode<S, D, REV, true>), dim3(grid), dim3(kThreads), 0, s, a);
  else hipLaunchKernelGGL((k_encode<S, D, REV, false>), dim3(grid), dim3(kThreads), 0, s, a);
  return hipGetLastError();
}

template <class S, int D, bool REV>
static hipError_t launch_decode_t(const DecodeArgs &a, uint32_t *walk_lengths, hipStream_t s)
{
  if (walk_lengths) {
    hipLaunchKernelGGL((k_walk<S, D, REV>), dim3(1), dim3(64), 0, s, a, walk_lengths);
    return hipGetLastError();
  }
  const uint64_t grid = (a.field.nblocks + kThreads - 1) / kThreads;
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL((k_decode<S, D, REV>), dim3(grid), dim3(kThreads), 0, s, a);
  return hipGetLastError();
}

template <class S>
static hipError_t launch_encode_type(int dims, bool rev, const EncodeArgs &a, hipStream_t s)
{
  switch (dims * 2 + (rev ? 1 : 0)) {
  case 2: return launch_encode_t<S, 1, false>(a, s);
  case 3: return launch_encode_t<S, 1, true>(a, s);
  case 4: return launch_encode_t<S, 2, false>(a, s);
  case 5: return launch_encode_t<S, 2, true>(a, s);
  case 6: return launch_encode_t<S, 3, false>(a, s);
  case 7: return launch_encode_t<S, 3, true>(a, s);
  default: return hipErrorInvalidValue;
  }
}

template <class S>
static hipError_t launch_decode_type(int dims, bool rev, const DecodeArgs &a, uint32_t *walk,
                                     hipStream_t s)
{
  switch (dims * 2 + (rev ? 1 : 0)) {
  case 2: return launch_decode_t<S, 1, false>(a, walk, s);
  case 3: return launch_decode_t<S, 1, true>(a, walk, s);
  case 4: return launch_decode_t<S, 2, false>(a, walk, s);
  case 5: return launch_decode_t<S, 2, true>(a, walk, s);
  case 6: return launch_decode_t<S, 3, false>(a, walk, s);
  case 7: return launch_decode_t<S, 3, true>(a, walk, s);
  default: return hipErrorInvalidValue;
  }
}

}  // namespace zx

#define ZX_INSTANTIATE(S, NAME)                                                              \
  hipError_t zx_launch_encode_##NAME(int dims, bool rev, const zx::EncodeArgs &a,           \
                                     hipStream_t s)                                        \
  {                                                                                        \
    return zx::launch_encode_type<S>(dims, rev, a, s);                                      \
  }                                                                                        \
  hipError_t zx_launch_decode_##NAME(int dims, bool rev, const zx::DecodeArgs &a,           \
                                     uint32_t *walk, hipStream_t s)                        \
  {                                                                                        \
    return zx::launch_decode_type<S>(dims, rev, a, walk, s);                                \
  }
