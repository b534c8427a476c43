// zfpx_api.h -- launch interface between the C ABI (zfp_capi.cpp) and the
// kernels (zfpx_kernels_*.hip, zfpx_scan.hip, zfpx_alias.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "zfpx_alias.h"
#include "zfpx_field.h"

namespace zx {

struct EncodeArgs {
  FieldDesc field;
  Cfg cfg;
  uint64_t *words;          // stream words (device)
  uint64_t nwords;          // capacity in words; words past it are dropped
  uint64_t start;           // bit of block 0 (fixed placement)
  uint32_t fixed;           // bits per block (fixed placement)
  const uint64_t *offsets;  // per-block start bits (variable placement), or null
  uint32_t *lengths;        // non-null: length pass, write bit counts here
  bool vec;                 // rows of whole blocks may be moved as 16-byte vectors
  uint32_t slice;           // blocks per workgroup of the LDS path, or 0 for its default
  uint64_t vbits;           // variable placement: bits of all blocks (offsets[nblocks] -
                            // offsets[0]), sizing the LDS path; 0 keeps to global memory
  uint32_t vcap;            // variable placement: words a workgroup may stage, or 0 for
                            // the default (from vbits; one pass: the largest group)
  // one-pass variable-rate compress (op_* in zfpx_fx.h), when op_state is set;
  // `lengths` then receives each block's bit count (the offset index), or is null
  uint64_t *op_state;       // [0] ticket counter, [1] failure flag, [2 + g] group g's
                            // look-back status; cleared before the launch
  uint64_t *op_edges;       // two words per group: its shared first and last words
  uint32_t op_T;            // blocks per workgroup, or 0 for the default
};

struct DecodeArgs {
  FieldDesc field;
  Cfg cfg;
  const uint64_t *words;
  uint64_t nwords;
  uint64_t start;
  uint32_t fixed;
  const uint64_t *offsets;
  uint64_t *maxword;        // one past the highest word loaded (written by the last block), or null
  bool vec;
  uint32_t slice;           // blocks per workgroup of the LDS path, or 0 for its default
  uint64_t first;           // variable placement: decode blocks [first, first + count) only,
  uint64_t count;           // or every block when count is 0
  uint64_t vbits;           // variable placement through offsets: bits of all blocks, which
                            // sizes the LDS staging (0: unknown, decode from HBM)
};

// One array of a batch (zfp_hip_compress_batch / zfp_hip_decompress_batch).
struct BatchJob {
  FieldDesc field;
  Cfg cfg;
  uint64_t *words;          // stream words from the array's first word
  uint64_t nwords;          // capacity from there: writes past it drop, reads give zero
  uint64_t start;           // bit of block 0 within words[0]
  uint64_t prefix;          // encode: bits of words[0] below `start`, kept by the fill
  uint64_t lbase;           // variable placement: block 0's slot in the batch's lengths
  const uint32_t *index;    // decode: the stream's offset index, or null (walk the stream)
  uint32_t *index_out;      // encode: where to keep the offset index, or null
  uint64_t *result;         // [0] end bit of the last block; [1] decode: one past the
                            // highest word loaded
  uint32_t fixed;           // bits per block (fixed placement), or 0
  uint32_t fill;            // encode: zero the words first (blocks share words)
  bool vec;                 // rows of whole blocks may be moved as 16-byte vectors
};

// One launch over the arrays of one class (same scalar type, dimensionality,
// reversible or not, fixed or variable placement).
struct BatchArgs {
  const BatchJob *jobs;
  const uint64_t *firsts;   // firsts[j]: index of jobs[j]'s block 0 among the launch's blocks
  uint32_t njobs;
  uint64_t nblocks;
  uint32_t *lengths;        // batch-wide per-block bit lengths (variable placement)
  const uint64_t *offsets;  // batch-wide exclusive scan of `lengths`
  uint32_t threads;         // workgroup size (blocks per workgroup): 64, 128 or 256
};

}  // namespace zx

// zfpx_alias.hip: ticket pass and store pass; `best` holds `span` words of
// scratch.
hipError_t zx_launch_alias_scatter(const zx::AliasArgs &a, unsigned long long *best, hipStream_t s);

// kernel launchers, one pair per (type, dimensionality, reversible):
// zx_{enc,dec}_<f32|f64|i32|i64>_<1|2|3|4>_<0|1> (zfpx_kernels.hip); the batch
// launchers of 4-D return hipErrorNotSupported.  zx_walk_* is the host's
// length reader (walk_offsets, zfpx_field.h) over `count` blocks of host words.
#define ZX_DECLARE1(T, D, R)                                                              \
  hipError_t zx_enc_##T##_##D##_##R(const zx::EncodeArgs &a, hipStream_t s);              \
  hipError_t zx_dec_##T##_##D##_##R(const zx::DecodeArgs &a, hipStream_t s);              \
  hipError_t zx_benc_##T##_##D##_##R(const zx::BatchArgs &a, bool count, hipStream_t s);  \
  hipError_t zx_bdec_##T##_##D##_##R(const zx::BatchArgs &a, bool walk, hipStream_t s);   \
  uint64_t zx_walk_##T##_##D##_##R(const uint64_t *words, uint64_t nwords, uint64_t pos,   \
                                   uint64_t count, const zx::Cfg &c, uint64_t *offsets);
#define ZX_DECLARE(T)                                                                     \
  ZX_DECLARE1(T, 1, 0) ZX_DECLARE1(T, 1, 1) ZX_DECLARE1(T, 2, 0) ZX_DECLARE1(T, 2, 1)     \
  ZX_DECLARE1(T, 3, 0) ZX_DECLARE1(T, 3, 1) ZX_DECLARE1(T, 4, 0) ZX_DECLARE1(T, 4, 1)
ZX_DECLARE(f32)
ZX_DECLARE(f64)
ZX_DECLARE(i32)
ZX_DECLARE(i64)
#undef ZX_DECLARE
#undef ZX_DECLARE1

// zfpx_scan.hip
// offsets[i] = start + sum(lengths[0..i)), offsets[n] = end; `tiles` holds
// zx_scan_tiles(n) words of scratch, whose word 1 is nonzero after a scan
// whose look-back gave up (its offsets are then past any stream).
uint64_t zx_scan_tiles(uint64_t n);
hipError_t zx_launch_scan(const uint32_t *lengths, uint64_t n, uint64_t start, uint64_t *offsets,
                          uint64_t *tiles, hipStream_t s);
// words[0] = first, words[1..n) = 0
hipError_t zx_launch_fill(uint64_t *words, uint64_t n, uint64_t first, hipStream_t s);
// The second launch of the one-pass compress: every stream word shared by two
// groups (or by the first group and the caller's `prefix` bits below `start`)
// from the groups' edges; `state` is the one-pass state after its launch.
hipError_t zx_launch_op_merge(uint64_t *words, uint64_t nwords, const uint64_t *state,
                              const uint64_t *edges, uint64_t groups, uint64_t start, uint64_t prefix,
                              hipStream_t s);


// Batch helpers over all `n` arrays of a batch.  Per-array work is cut into
// chunks of zx_batch_chunks(items) workgroups; first[j] is array j's first
// chunk (an exclusive prefix over the arrays) and nchunks the total.
uint64_t zx_batch_chunks(uint64_t items);
// copy each variable-rate array's offset index into the batch-wide lengths
// (chunks over blocks);
hipError_t zx_batch_gather_index(const zx::BatchJob *jobs, const uint64_t *first, uint32_t n,
                                 uint64_t nchunks, uint32_t *lengths, hipStream_t s);
// write each array's end bit to result[0];
hipError_t zx_batch_ends(const zx::BatchJob *jobs, uint32_t n, const uint64_t *offsets, hipStream_t s);
// zero the words of the arrays with `fill` set, keeping the prefix bits of the
// first (chunks over each array's word capacity; after zx_batch_ends);
hipError_t zx_batch_fill(const zx::BatchJob *jobs, const uint64_t *first, uint32_t n, uint64_t nchunks,
                         hipStream_t s);
// copy each variable-rate array's lengths to its index_out (chunks over blocks).
hipError_t zx_batch_keep_index(const zx::BatchJob *jobs, const uint64_t *first, uint32_t n,
                               uint64_t nchunks, const uint32_t *lengths, hipStream_t s);
