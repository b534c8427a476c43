// zfpx_kernels.inc.h -- HIP kernels of the field codec, instantiated per
// scalar type by zfpx_kernels_{i32,i64,f32,f64}.hip.
//
// One thread codes one block.  Blocks are placed in the stream either at a
// fixed stride (fixed-rate: block b starts at start + b * bits, the
// reference's per-chunk seek, decompress.rs:205-330) or at offsets from an
// exclusive scan of per-block lengths (variable-rate).
//
// Fixed-rate blocks of up to 64 words go through LDS: a workgroup's slice of
// the stream is one contiguous word range, read in (decode) or written out
// (encode) with coalesced accesses, while each thread works on its own words
// in LDS.  When blocks are whole words and the stream starts on a word, each
// thread owns its words outright (stride W + 1 with a private pad word);
// otherwise neighbours share boundary words and merge them with LDS atomic
// ORs, and the workgroup's two end words are merged into global memory.
// Other layouts use global memory directly (zfpx_block.h WordSink /
// WordSource), with the host zeroing the word range first.
#pragma once

#include <hip/hip_runtime.h>

#include "zfpx_api.h"
#include "zfpx_field.h"
#include "zfpx_fx.h"
#include "zfpx_scan.h"

namespace zx {

constexpr int kThreads = 256;

// Waves per SIMD the LDS kernels are held to: four for lossy 32-bit values
// (128 VGPRs); none for 64-bit values, whose 64 coefficients alone take 128,
// nor for reversible blocks, which would spill.
template <class S, bool REV> constexpr int kMinWaves = ST<S>::P == 32 && !REV ? 4 : 1;

// Waves per SIMD the decoders of 64-bit lossy blocks are held to: two (256
// registers; the 64 coefficients take 128 and a few values spill), where
// unconstrained they take about 270 and run one wave per SIMD, with nothing
// to hide the latency of their loads.
template <class S, bool REV> constexpr int kWideDecodeWaves = ST<S>::P == 64 && !REV ? 2 : 1;
template <class S, bool REV>
constexpr int kDecodeWaves = ST<S>::P == 32 ? kMinWaves<S, REV> : kWideDecodeWaves<S, REV>;
// 4-D blocks run one wave per SIMD: a 32-bit block's 256 values take 256 of
// the 512 registers (vector and accumulation registers) a lone wave may hold.
template <class S, int D, bool REV> constexpr int kFxDecodeWaves = D == 4 ? 1 : kDecodeWaves<S, REV>;

// Waves per SIMD the encoders are held to: 32-bit as kMinWaves; 64-bit lossy
// two (256 registers, a few values spilled), where unconstrained the
// double-precision ones take a few more and run one wave per SIMD (reversible
// double blocks would spill 60-150 bytes per lane under the cap).
template <class S, bool REV> constexpr int kWideEncodeWaves = ST<S>::P == 64 && !REV ? 2 : 1;
template <class S, bool REV>
constexpr int kEncodeWaves = ST<S>::P == 32 ? kMinWaves<S, REV> : kWideEncodeWaves<S, REV>;
template <class S, int D, bool REV> constexpr int kFxEncodeWaves = D == 4 ? 1 : kEncodeWaves<S, REV>;

// Fills the workgroup's expand8 table (zfpx_block.h); every kernel that
// encodes calls it first, with all of its threads.
__device__ __forceinline__ void expand8_fill()
{
  for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) zx_expand8_lds[i] = kExpand8Dev.v[i];
  __syncthreads();
}

// Fills the workgroup's compress8 table (zfpx_block.h), which the decoders of
// 2-D and 3-D blocks read (kDecodeTable); every kernel that decodes them
// calls it first, with all of its threads, before any thread leaves.  Other
// dimensionalities skip it.
template <int D>
__device__ __forceinline__ void compress8_load()
{
  if constexpr (kDecodeTable<D>)
    for (uint32_t i = threadIdx.x; i < 64; i += blockDim.x)
      ((uint32_t *)zx_compress8_lds)[i] = ((const uint32_t *)kCompress8Dev.v)[i];
}

// compress8_load and the barrier after it (2-D and 3-D only)
template <int D>
__device__ __forceinline__ void compress8_fill()
{
  if constexpr (kDecodeTable<D>) {
    compress8_load<D>();
    __syncthreads();
  }
}

// Rows of four contiguous values are loaded and stored as 16-byte vectors
// when the host has checked that every row of every whole block is aligned.
// The field is read or written once per call, so the accesses are
// nontemporal (streamed past the caches: 8 % faster encodes and 3 % faster
// decodes of 512^3 f32 at rate 8 than cached accesses).
typedef unsigned int zx_u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned long long zx_u64x2 __attribute__((ext_vector_type(2)));

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
    if (D >= 3) off += (int64_t)((r >> 2) & 3) * f.s[2];
    if (D >= 4) off += (int64_t)(r >> 4) * f.s[3];
    if constexpr (sizeof(U) == 4) {
      const zx_u32x4 q = __builtin_nontemporal_load((const zx_u32x4 *)(base + off));
      v[4 * r + 0] = q.x; v[4 * r + 1] = q.y; v[4 * r + 2] = q.z; v[4 * r + 3] = q.w;
    } else {
      const zx_u64x2 a = __builtin_nontemporal_load((const zx_u64x2 *)(base + off));
      const zx_u64x2 b = __builtin_nontemporal_load((const zx_u64x2 *)(base + off + 2));
      v[4 * r + 0] = a.x; v[4 * r + 1] = a.y; v[4 * r + 2] = b.x; v[4 * r + 3] = b.y;
    }
  }
  // An empty barrier: without it the compiler sinks these loads into a block
  // shared with the element loads of partial blocks, splitting each 16-byte
  // load into four with an address apiece.
  asm volatile("" ::: "memory");
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
    if (D >= 3) off += (int64_t)((r >> 2) & 3) * f.s[2];
    if (D >= 4) off += (int64_t)(r >> 4) * f.s[3];
    if constexpr (sizeof(U) == 4) {
      zx_u32x4 q;
      q.x = v[4 * r]; q.y = v[4 * r + 1]; q.z = v[4 * r + 2]; q.w = v[4 * r + 3];
      __builtin_nontemporal_store(q, (zx_u32x4 *)(base + off));
    } else {
      zx_u64x2 a, b;
      a.x = v[4 * r]; a.y = v[4 * r + 1]; b.x = v[4 * r + 2]; b.y = v[4 * r + 3];
      __builtin_nontemporal_store(a, (zx_u64x2 *)(base + off));
      __builtin_nontemporal_store(b, (zx_u64x2 *)(base + off + 2));
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

// ------------------------------------------------------------------ global path

// Encode every block; COUNT writes per-block bit lengths instead of bits.
template <class S, int D, bool REV, bool COUNT>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(D == 4 ? 1 : kWideEncodeWaves<S, REV>, 8))) void k_encode(EncodeArgs a)
{
  expand8_fill();
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
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(D == 4 ? 1 : kWideDecodeWaves<S, REV>, 8))) void k_decode(DecodeArgs a)
{
  compress8_fill<D>();
  const uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= (a.count ? a.count : a.field.nblocks)) return;
  const uint64_t b = a.first + i;
  const uint64_t pos = a.offsets ? a.offsets[b] : a.start + b * (uint64_t)a.fixed;
  WordSource r;
  r.init(GlobalWords{a.words, a.nwords}, pos);
  typename Ints<ST<S>::P>::U v[Blk<D>::SIZE];
  decode_values<S, D, REV>(r, v, a.cfg);
  // blocks are contiguous, so the last one loads the highest word
  if (a.maxword && b == a.field.nblocks - 1) *a.maxword = loaded_words(r, pos, a.cfg.minbits);
  store_block<S, D>(a.field, b, a.vec, v);
}

// ------------------------------------------------------------------ LDS path

// Fixed-rate encode through LDS (zfpx_fx.h): one workgroup per group of
// blockDim.x blocks.  32-bit values fit four waves per SIMD without spilling
// (the coder keeps packed chunk halves rather than the coefficients).
template <class S, int D, bool REV, bool OWN>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kFxEncodeWaves<S, D, REV>, 8))) void k_encode_fx(EncodeArgs a, FxPlan p)
{
  extern __shared__ uint64_t lds[];
  expand8_fill();
  const uint32_t t = threadIdx.x;
  const FxGroup q = fx_group(blockIdx.x, p.T, a.field.nblocks, a.start, a.fixed);
  if constexpr (!OWN) {
    for (uint32_t i = t; i < q.nw; i += p.T) fx_enc_clear(lds, p, i);
    __syncthreads();
  }
  if (t < q.nb) {
    typename Ints<ST<S>::P>::U v[Blk<D>::SIZE];
    load_block<S, D>(a.field, q.b0 + t, a.vec, v);
    fx_enc_block<S, D, REV, OWN>(lds, p, q, t, a.fixed, v, a.cfg);
  }
  __syncthreads();
  if constexpr (OWN) {
    // whole blocks of an even number of words from an even word: the words
    // go out in 16-byte pairs (a pair never spans two blocks' slots)
    const uint32_t W = a.fixed / 64;
    if (W % 2 == 0 && q.w0 % 2 == 0 && q.w0 + q.nw <= a.nwords) {
      for (uint32_t j = t; 2 * j < q.nw; j += p.T) {
        const uint32_t i = 2 * j;
        const uint32_t sl = i + i / W;
        zx_u64x2 pr;
        pr.x = lds[sl];
        pr.y = lds[sl + 1];
        __builtin_nontemporal_store(pr, (zx_u64x2 *)(a.words + q.w0 + i));
      }
      return;
    }
  }
  for (uint32_t i = t; i < q.nw; i += p.T)
    if (!fx_enc_store<OWN>(a.words, a.nwords, lds, p, q, i, a.fixed)) break;
}

// Fixed-rate decode through LDS: stage the group's words, then decode.  For
// 32-bit values held to four waves per SIMD (128 VGPRs, a few values spilled
// once per block): the group-test loop waits on LDS reads, and the extra wave
// hides more of that than the spills cost.
template <class S, int D, bool REV, bool FLAT>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kFxDecodeWaves<S, D, REV>, 8))) void k_decode_fx(DecodeArgs a, FxPlan p)
{
  extern __shared__ uint64_t lds[];
  const uint32_t t = threadIdx.x;
  const FxGroup q = fx_group(blockIdx.x, p.T, a.field.nblocks, a.start, a.fixed);
  fx_dec_stage_thread(lds, a.words, a.nwords, p, q, t);
  compress8_load<D>();
  __syncthreads();
  if (t >= q.nb) return;
  typename Ints<ST<S>::P>::U v[Blk<D>::SIZE];
  const uint64_t loaded = fx_dec_block<S, D, REV, FLAT>(lds, p, q, t, a.fixed, v, a.cfg);
  // blocks are contiguous, so the last one loads the highest word
  // block indices fit in 32 bits (the host caps nblocks below 2^32)
  const uint32_t b = (uint32_t)q.b0 + t;
  if (a.maxword && b == (uint32_t)(a.field.nblocks - 1)) *a.maxword = loaded;
  store_block<S, D>(a.field, b, a.vec, v);
}

// Variable-rate decode through LDS: a workgroup whose blocks span at most
// p.cap words (by the offsets) stages them, its loads batched, and decodes
// from LDS; the groups over the capacity are left to k_decode_dense.
template <class S, int D, bool REV>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kDecodeWaves<S, REV>, 8))) void k_decode_vx(DecodeArgs a, FxPlan p)
{
  extern __shared__ uint64_t lds[];
  const uint32_t t = threadIdx.x;
  const FxGroup q = vx_group(blockIdx.x, p.T, a.field.nblocks, a.offsets);
  if (q.nw > p.cap) return;
  fx_dec_stage_thread(lds, a.words, a.nwords, p, q, t);
  compress8_load<D>();
  __syncthreads();
  if (t >= q.nb) return;
  const uint64_t b = q.b0 + t;
  typename Ints<ST<S>::P>::U v[Blk<D>::SIZE];
  const uint64_t loaded = vx_dec_block<S, D, REV>(lds, p, q, a.offsets[b], v, a.cfg);
  if (a.maxword && b == a.field.nblocks - 1) *a.maxword = loaded;
  store_block<S, D>(a.field, b, a.vec, v);
}

// The groups k_decode_vx leaves: each block decoded from HBM at its offset,
// as k_decode does; a workgroup tests blockDim.x groups at once (one per
// thread) and then decodes the dense ones among them.
template <class S, int D, bool REV>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kWideDecodeWaves<S, REV>, 8))) void k_decode_dense(DecodeArgs a, FxPlan p, uint64_t ngroups)
{
  __shared__ uint32_t dense[kThreads];
  __shared__ uint32_t ndense;
  const uint32_t t = threadIdx.x;
  if (t == 0) ndense = 0;
  __syncthreads();
  const uint64_t g0 = (uint64_t)blockIdx.x * blockDim.x;
  if (g0 + t < ngroups && vx_group(g0 + t, p.T, a.field.nblocks, a.offsets).nw > p.cap)
    dense[atomicAdd(&ndense, 1u)] = t;
  compress8_load<D>();
  __syncthreads();
  const uint32_t n = ndense;
  for (uint32_t k = 0; k < n; k++) {
    const FxGroup q = vx_group(g0 + dense[k], p.T, a.field.nblocks, a.offsets);
    for (uint32_t j = t; j < q.nb; j += blockDim.x) {
      const uint64_t b = q.b0 + j;
      const uint64_t pos = a.offsets[b];
      WordSource r;
      r.init(GlobalWords{a.words, a.nwords}, pos);
      typename Ints<ST<S>::P>::U v[Blk<D>::SIZE];
      decode_values<S, D, REV>(r, v, a.cfg);
      if (a.maxword && b == a.field.nblocks - 1) *a.maxword = loaded_words(r, pos, a.cfg.minbits);
      store_block<S, D>(a.field, b, a.vec, v);
    }
  }
}

// Variable-rate write pass through LDS: a workgroup whose blocks touch at most
// p.cap words encodes them into LDS at their scanned offsets and stores the
// range with coalesced stores (atomic ORs for the two words it may share with
// its neighbours).  Denser workgroups are left to k_encode_dense.  The choice
// is uniform over the workgroup, so the barriers are too.
template <class S, int D, bool REV>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kEncodeWaves<S, REV>, 8))) void k_encode_vx(EncodeArgs a, FxPlan p)
{
  extern __shared__ uint64_t lds[];
  const uint32_t t = threadIdx.x;
  const FxGroup q = vx_group(blockIdx.x, p.T, a.field.nblocks, a.offsets);
  if (q.nw > p.cap) return;
  expand8_fill();
  for (uint32_t i = t; i < q.nw; i += p.T) fx_enc_clear(lds, p, i);
  __syncthreads();
  if (t < q.nb) {
    typename Ints<ST<S>::P>::U v[Blk<D>::SIZE];
    load_block<S, D>(a.field, q.b0 + t, a.vec, v);
    fx_enc_at<S, D, REV>(lds, p, (uint32_t)(a.offsets[q.b0 + t] - 64 * q.w0), v, a.cfg);
  }
  __syncthreads();
  for (uint32_t i = t; i < q.nw; i += p.T)
    if (!fx_enc_store<false>(a.words, a.nwords, lds, p, q, i, 0)) break;
}

// The groups k_encode_vx leaves: each block written to HBM at its offset, as
// k_encode does.  Shared words are merged with atomic ORs, so the two launches
// may run in either order.  Most streams have few or no dense groups, so a
// workgroup first tests blockDim.x groups at once (one per thread) and then
// codes the dense ones among them.
template <class S, int D, bool REV>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kWideEncodeWaves<S, REV>, 8))) void k_encode_dense(EncodeArgs a, FxPlan p, uint64_t ngroups)
{
  __shared__ uint32_t dense[kThreads];
  __shared__ uint32_t ndense;
  const uint32_t t = threadIdx.x;
  if (t == 0) ndense = 0;
  __syncthreads();
  const uint64_t g0 = (uint64_t)blockIdx.x * p.T;
  if (g0 + t < ngroups && vx_group(g0 + t, p.T, a.field.nblocks, a.offsets).nw > p.cap)
    dense[atomicAdd(&ndense, 1u)] = t;
  __syncthreads();
  const uint32_t n = ndense;
  if (n == 0) return;
  expand8_fill();
  for (uint32_t k = 0; k < n; k++) {
    const FxGroup q = vx_group(g0 + dense[k], p.T, a.field.nblocks, a.offsets);
    if (t < q.nb) {
      typename Ints<ST<S>::P>::U v[Blk<D>::SIZE];
      load_block<S, D>(a.field, q.b0 + t, a.vec, v);
      WordSink w;
      w.init(a.words, a.nwords, a.offsets[q.b0 + t]);
      encode_values<S, D, REV>(w, v, a.cfg);
      w.finish();
    }
  }
}

// One-pass variable-rate compress (zfpx_fx.h, op_*): each workgroup takes a
// ticket for its group of blocks, prepares and counts its blocks, scans the
// counts, looks back for the bits of the groups before it, then codes its
// blocks from the same registers into LDS (or, for a group over the LDS
// capacity, straight to HBM) and stores its words; the shared edge words go
// to a_op_edges for k_op_merge.  The staged/direct choice is uniform over the
// workgroup, so the barriers are too.
template <class S, int D, bool REV>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kEncodeWaves<S, REV>, 8))) void k_encode_1p(EncodeArgs a, OnePlan p)
{
  extern __shared__ uint64_t lds[];
  __shared__ uint32_t group_s;
  __shared__ uint32_t wave_tot[kThreads / 64];
  __shared__ uint64_t prefix_s;
  __shared__ uint64_t edge_s[2];
  const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
  if (t == 0) {
    group_s = (uint32_t)atomicAdd((unsigned long long *)&a.op_state[0], 1ull);
    edge_s[0] = 0;
    edge_s[1] = 0;
  }
  expand8_fill();
  const uint64_t g = group_s;
  const uint64_t b = g * p.T + t;
  const bool live = b < a.field.nblocks;
  PreparedBlock<S, D> pb;
  uint32_t len = 0;
  if (live) {
    typename Ints<ST<S>::P>::U v[Blk<D>::SIZE];
    load_block<S, D>(a.field, b, a.vec, v);
    prepare_block<S, D, REV>(v, a.cfg, pb);
    CountSink cs{0};
    len = code_block<S, D>(cs, pb, a.cfg);
    if (a.lengths) a.lengths[b] = len;
  }
  const uint32_t inc = wave_inclusive_sum(len);
  if (lane == 63) wave_tot[wave] = inc;
  __syncthreads();
  uint32_t before = 0, total = 0;
  for (uint32_t w = 0; w < p.T / 64; w++) {
    before += w < wave ? wave_tot[w] : 0u;
    total += wave_tot[w];
  }
  if (wave == 0) {
    const uint64_t pre = lookback(a.op_state + 2, g, total, a.start, a.op_state + 1);
    if (lane == 0) prefix_s = pre;
  }
  __syncthreads();
  const OneRange r = op_range(prefix_s, total);
  const uint32_t off = before + (inc - len);  // the block's first bit within the group
  if (r.nw <= p.cap) {
    for (uint32_t i = t; i < r.nw; i += p.T) lds[fx_slot(i, p.skew)] = 0;
    __syncthreads();
    if (live) op_stage<S, D>(lds, p, (uint32_t)(r.b0 & 63) + off, pb, a.cfg);
    __syncthreads();
    for (uint32_t i = t; i < r.nw; i += p.T) op_store(a.words, a.nwords, lds, p, r, i, edge_s);
  } else {
    for (uint32_t i = t; i < r.nw; i += p.T) op_clear(a.words, a.nwords, r, i);
    __threadfence();
    __syncthreads();
    if (live) op_direct<S, D>(a.words, a.nwords, r, r.b0 + off, pb, a.cfg, edge_s);
  }
  __syncthreads();
  if (t == 0) {
    a.op_edges[2 * g] = edge_s[0];
    a.op_edges[2 * g + 1] = edge_s[1];
  }
}

// ------------------------------------------------------------------ batches

// The array holding a launch's block g: the last j with firsts[j] <= g.
__device__ __forceinline__ uint32_t job_of(const uint64_t *firsts, uint32_t njobs, uint64_t g)
{
  uint32_t lo = 0, hi = njobs;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (firsts[mid] <= g) lo = mid;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ uint64_t batch_pos(const BatchArgs &a, const BatchJob &J, uint64_t b)
{
  return J.fixed ? J.start + b * (uint64_t)J.fixed : J.start + (a.offsets[J.lbase + b] - a.offsets[J.lbase]);
}

// Every block of every array of one class; COUNT writes bit lengths.
template <class S, int D, bool REV, bool COUNT>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kWideEncodeWaves<S, REV>, 8))) void k_batch_encode(BatchArgs a)
{
  expand8_fill();
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.nblocks) return;
  const uint32_t j = job_of(a.firsts, a.njobs, g);
  const BatchJob &J = a.jobs[j];
  const uint64_t b = g - a.firsts[j];
  typename Ints<ST<S>::P>::U v[Blk<D>::SIZE];
  load_block<S, D>(J.field, b, J.vec, v);
  if constexpr (COUNT) {
    CountSink cs{0};
    encode_values<S, D, REV>(cs, v, J.cfg);
    a.lengths[J.lbase + b] = (uint32_t)cs.n;
  } else {
    WordSink w;
    w.init(J.words, J.nwords, batch_pos(a, J, b));
    encode_values<S, D, REV>(w, v, J.cfg);
    w.finish();
  }
}

template <class S, int D, bool REV>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kWideDecodeWaves<S, REV>, 8))) void k_batch_decode(BatchArgs a)
{
  compress8_fill<D>();
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.nblocks) return;
  const uint32_t j = job_of(a.firsts, a.njobs, g);
  const BatchJob &J = a.jobs[j];
  const uint64_t b = g - a.firsts[j];
  const uint64_t pos = batch_pos(a, J, b);
  WordSource r;
  r.init(GlobalWords{J.words, J.nwords}, pos);
  typename Ints<ST<S>::P>::U v[Blk<D>::SIZE];
  decode_values<S, D, REV>(r, v, J.cfg);
  if (b == J.field.nblocks - 1) J.result[1] = loaded_words(r, pos, J.cfg.minbits);
  store_block<S, D>(J.field, b, J.vec, v);
}

// Index-less variable-rate arrays: one thread per array reads its blocks'
// lengths in order (skip_values, the length reader the host runs for a single
// stream).
template <class S, int D, bool REV>
__global__ __launch_bounds__(64) void k_batch_walk(BatchArgs a)
{
  const uint32_t j = blockIdx.x * 64 + threadIdx.x;
  if (j >= a.njobs) return;
  const BatchJob &J = a.jobs[j];
  if (J.fixed || J.index) return;
  uint64_t pos = J.start;
  for (uint64_t b = 0; b < J.field.nblocks; b++) {
    WordSource r;
    r.init(GlobalWords{J.words, J.nwords}, pos);
    const uint32_t used = skip_values<S, D, REV>(r, J.cfg);
    a.lengths[J.lbase + b] = used;
    pos += used;
  }
}

// ------------------------------------------------------------------ launchers

// The global-memory kernels: every block coded from and to HBM at its own
// position (variable-rate length and write passes, decodes from offsets,
// blocks too large to stage).  4-D blocks (256 values) have no batch
// launches and no one-pass compress.
template <class S, int D, bool REV>
static hipError_t launch_global_encode_t(const EncodeArgs &a, hipStream_t s)
{
  const uint64_t grid = (a.field.nblocks + kThreads - 1) / kThreads;
  if (a.lengths) hipLaunchKernelGGL((k_encode<S, D, REV, true>), dim3(grid), dim3(kThreads), 0, s, a);
  else hipLaunchKernelGGL((k_encode<S, D, REV, false>), dim3(grid), dim3(kThreads), 0, s, a);
  return hipGetLastError();
}

template <class S, int D, bool REV>
static hipError_t launch_encode_t(const EncodeArgs &a, hipStream_t s)
{
  if (a.field.nblocks == 0) return hipSuccess;
  // 4-D blocks take the staged fixed-rate encode in its whole-block layout
  // only and the global-memory kernels otherwise: every 4-D kernel holds a
  // 256-value codec, and each one instantiated adds minutes to the build
  FxPlan p;
  if (!a.op_state && !a.lengths && !a.offsets && fx_plan(a.fixed, a.start, false, a.slice, p) &&
      (D != 4 || p.own)) {
    const uint64_t grid = (a.field.nblocks + p.T - 1) / p.T;
    const size_t bytes = (size_t)p.lds_words * 8;
    if (p.own) hipLaunchKernelGGL((k_encode_fx<S, D, REV, true>), dim3(grid), dim3(p.T), bytes, s, a, p);
    else if constexpr (D != 4) hipLaunchKernelGGL((k_encode_fx<S, D, REV, false>), dim3(grid), dim3(p.T), bytes, s, a, p);
    return hipGetLastError();
  }
  if constexpr (D == 4) {
    // (4-D fields keep the length pass, scan and write pass)
    if (a.op_state) return hipErrorNotSupported;
  } else {
    if (!a.op_state && !a.lengths && a.offsets && a.vbits) {
      vx_plan(a.vbits, a.field.nblocks, a.slice, a.vcap, p);
      const uint64_t grid = (a.field.nblocks + p.T - 1) / p.T;
      hipLaunchKernelGGL((k_encode_vx<S, D, REV>), dim3(grid), dim3(p.T), (size_t)p.lds_words * 8, s, a, p);
      hipLaunchKernelGGL((k_encode_dense<S, D, REV>), dim3((grid + p.T - 1) / p.T), dim3(p.T), 0, s, a, p, grid);
      return hipGetLastError();
    }
    if (a.op_state) {
      const OnePlan p = op_plan(block_bound(a.cfg, ST<S>::P, Blk<D>::SIZE), 0, a.op_T, a.vcap);
      const uint64_t grid = (a.field.nblocks + p.T - 1) / p.T;
      hipLaunchKernelGGL((k_encode_1p<S, D, REV>), dim3(grid), dim3(p.T), (size_t)p.lds_words * 8, s, a, p);
      return hipGetLastError();
    }
  }
  return launch_global_encode_t<S, D, REV>(a, s);
}

template <class S, int D, bool REV>
static hipError_t launch_decode_t(const DecodeArgs &a, hipStream_t s)
{
  if (a.field.nblocks == 0) return hipSuccess;
  // 4-D blocks: the staged fixed-rate decode of whole-word blocks only (as
  // the encode above)
  FxPlan p;
  if (!a.offsets && fx_plan(a.fixed, a.start, true, a.slice, p) && (D != 4 || p.flat)) {
    const uint64_t grid = (a.field.nblocks + p.T - 1) / p.T;
    if (p.flat) hipLaunchKernelGGL((k_decode_fx<S, D, REV, true>), dim3(grid), dim3(p.T), (size_t)p.lds_words * 8, s, a, p);
    else if constexpr (D != 4) hipLaunchKernelGGL((k_decode_fx<S, D, REV, false>), dim3(grid), dim3(p.T), (size_t)p.lds_words * 8, s, a, p);
    return hipGetLastError();
  }
  if constexpr (D != 4) {
    if (a.offsets && a.vbits && !a.count) {
      // one wave per workgroup, as the fixed-rate decode
      vx_plan(a.vbits, a.field.nblocks, a.slice ? a.slice : 64u, 0, p);
      const uint64_t groups = (a.field.nblocks + p.T - 1) / p.T;
      hipLaunchKernelGGL((k_decode_vx<S, D, REV>), dim3(groups), dim3(p.T), (size_t)p.lds_words * 8, s, a, p);
      hipLaunchKernelGGL((k_decode_dense<S, D, REV>), dim3((groups + kThreads - 1) / kThreads), dim3(kThreads), 0,
                         s, a, p, groups);
      return hipGetLastError();
    }
  }
  const uint64_t grid = ((a.count ? a.count : a.field.nblocks) + kThreads - 1) / kThreads;
  hipLaunchKernelGGL((k_decode<S, D, REV>), dim3(grid), dim3(kThreads), 0, s, a);
  return hipGetLastError();
}

template <class S, int D, bool REV>
static hipError_t launch_batch_encode_t(const BatchArgs &a, bool count, hipStream_t s)
{
  if constexpr (D == 4) {
    return hipErrorNotSupported;
  } else {
    if (a.nblocks == 0) return hipSuccess;
    const uint32_t T = a.threads ? a.threads : kThreads;
    const uint64_t grid = (a.nblocks + T - 1) / T;
    if (count) hipLaunchKernelGGL((k_batch_encode<S, D, REV, true>), dim3(grid), dim3(T), 0, s, a);
    else hipLaunchKernelGGL((k_batch_encode<S, D, REV, false>), dim3(grid), dim3(T), 0, s, a);
    return hipGetLastError();
  }
}

template <class S, int D, bool REV>
static hipError_t launch_batch_decode_t(const BatchArgs &a, bool walk, hipStream_t s)
{
  if constexpr (D == 4) {
    return hipErrorNotSupported;
  } else {
    if (walk) {
      if (a.njobs == 0) return hipSuccess;
      hipLaunchKernelGGL((k_batch_walk<S, D, REV>), dim3((a.njobs + 63) / 64), dim3(64), 0, s, a);
      return hipGetLastError();
    }
    if (a.nblocks == 0) return hipSuccess;
    const uint32_t T = a.threads ? a.threads : kThreads;
    const uint64_t grid = (a.nblocks + T - 1) / T;
    hipLaunchKernelGGL((k_batch_decode<S, D, REV>), dim3(grid), dim3(T), 0, s, a);
    return hipGetLastError();
  }
}

}  // namespace zx
