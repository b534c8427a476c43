// zfpx_api.h -- launch interface between the C ABI (zfp_capi.cpp) and the
// kernels (zfpx_kernels_*.hip, zfpx_scan.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

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
};

struct DecodeArgs {
  FieldDesc field;
  Cfg cfg;
  const uint64_t *words;
  uint64_t nwords;
  uint64_t start;
  uint32_t fixed;
  const uint64_t *offsets;
  uint64_t *maxword;        // one past the highest word loaded (written by the last block)
  bool vec;
};

}  // namespace zx

// kernel launchers, one pair per (type, dimensionality, reversible):
// zx_{enc,dec}_<f32|f64|i32|i64>_<1|2|3>_<0|1> (zfpx_kernels.hip)
#define ZX_DECLARE1(T, D, R)                                                              \
  hipError_t zx_enc_##T##_##D##_##R(const zx::EncodeArgs &a, hipStream_t s);              \
  hipError_t zx_dec_##T##_##D##_##R(const zx::DecodeArgs &a, uint32_t *walk, hipStream_t s);
#define ZX_DECLARE(T)                                                                     \
  ZX_DECLARE1(T, 1, 0) ZX_DECLARE1(T, 1, 1) ZX_DECLARE1(T, 2, 0) ZX_DECLARE1(T, 2, 1)     \
  ZX_DECLARE1(T, 3, 0) ZX_DECLARE1(T, 3, 1)
ZX_DECLARE(f32)
ZX_DECLARE(f64)
ZX_DECLARE(i32)
ZX_DECLARE(i64)
#undef ZX_DECLARE
#undef ZX_DECLARE1

// zfpx_scan.hip
// offsets[i] = start + sum(lengths[0..i)), offsets[n] = end; `tiles` holds
// zx_scan_tiles(n) words of scratch.
uint64_t zx_scan_tiles(uint64_t n);
hipError_t zx_launch_scan(const uint32_t *lengths, uint64_t n, uint64_t start, uint64_t *offsets,
                          uint64_t *tiles, hipStream_t s);
// words[0] = first, words[1..n) = 0
hipError_t zx_launch_fill(uint64_t *words, uint64_t n, uint64_t first, hipStream_t s);
