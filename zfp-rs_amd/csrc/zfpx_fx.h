// zfpx_fx.h -- the fixed-rate LDS path, written as per-thread phases that the
// kernels run between workgroup barriers and the CPU test build runs one
// thread after another.
//
// A workgroup codes T consecutive blocks.  At a fixed rate its blocks occupy
// one contiguous bit range of the stream (block b at start + b * fixed, the
// reference's per-chunk seek, src/compress/pipeline.rs and
// src/decompress/pipeline.rs fixed-rate paths), so the workgroup moves its
// whole word range between HBM and LDS with coalesced accesses while every
// thread codes its own block against LDS.
//
// Encode layouts:
//  - OWN: blocks are whole words and block 0 starts on a word, so thread t
//    owns words [t*W, t*W + W) outright, kept at LDS stride W + 1 (the extra
//    word is the thread's private spill slot for branch-free candidate
//    stores, and it staggers the lanes across LDS banks);
//  - shared: blocks straddle words; the range is kept contiguous with a pad
//    word every 2^skew words, zeroed first, and neighbours merge the words
//    they share with LDS atomic ORs.  The range's first and last words may be
//    shared with the neighbouring workgroups and are merged into HBM with
//    atomic ORs (the host zeroes the stream's range beforehand); all other
//    words are stored whole.
// Decode stages the range (plus two words the reader looks ahead into) in the
// contiguous layout, each pad slot holding a copy of the word after it so that
// a word and its successor are always adjacent; words at or past the stream's
// capacity read as zero.
#pragma once

#include "zfpx_field.h"

namespace zx {

struct FxPlan {
  uint32_t T;          // blocks (threads) per workgroup
  uint32_t skew;       // contiguous layouts: word i at LDS slot i + (i >> skew)
  bool own;            // encode layout OWN (see above)
  bool flat;           // decode: blocks are whole words from a word boundary and
                       // 2^skew is a multiple of their word count, so no pad slot
                       // splits a block (readers address their words directly)
  uint32_t lds_words;  // LDS words per workgroup
  uint32_t cap;        // variable-rate: the most words a staged group may touch
};

ZX_HD uint32_t fx_log2(uint32_t v)
{
  uint32_t s = 0;
  while ((2u << s) <= v) s++;
  return s;
}

// Stream words are read (decode staging) or written (encode) once per call:
// nontemporal accesses on the device, plain ones in the CPU build.
ZX_HD uint64_t fx_load_once(const uint64_t *p)
{
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
}
ZX_HD void fx_store_once(uint64_t *p, uint64_t v)
{
#if defined(__HIP_DEVICE_COMPILE__)
  __builtin_nontemporal_store(v, p);
#else
  *p = v;
#endif
}

// LDS a workgroup of the path may use.
constexpr uint64_t kFxMaxLdsBytes = 64 * 1024;

// Plan of the LDS path for blocks of `fixed` bits from bit `start`; false when
// blocks are too large for it.  `T` (at most 256) overrides the workgroup size
// when nonzero; false when that workgroup's words would not fit in LDS.
ZX_HD bool fx_plan(uint32_t fixed, uint64_t start, bool decode, uint32_t T, FxPlan &p)
{
  if (fixed == 0 || fixed > 64 * 64 || T > 256) return false;
  const uint32_t W = (fixed + 63) / 64;
  // decode: one wave per workgroup (its waves share nothing but the staging,
  // and a one-wave workgroup takes any free wave slot, which keeps more waves
  // resident: 2-3 % faster than 256 on 512^3 f32 at rate 8); encode: 256
  // blocks, or fewer for large blocks
  p.T = T ? T : decode ? 64u : W <= 16 ? 256u : W <= 32 ? 128u : 64u;
  p.own = !decode && fixed % 64 == 0 && start % 64 == 0;
  const uint32_t s = fx_log2(W);
  p.skew = p.own ? 31u : s < 3 ? 3u : s;
  p.flat = decode && fixed % 64 == 0 && start % 64 == 0 && ((1u << p.skew) % W) == 0;
  p.cap = 0;
  if (p.own) {
    p.lds_words = p.T * (W + 1);
  } else {
    const uint32_t words = (uint32_t)(((uint64_t)p.T * fixed + 63) / 64) + 3;
    p.lds_words = words + (words >> p.skew) + 1;
  }
  return (uint64_t)p.lds_words * 8 <= kFxMaxLdsBytes;
}

// One workgroup's share of the stream.
struct FxGroup {
  uint64_t b0;  // first block
  uint32_t nb;  // blocks
  uint64_t w0;  // first word touched
  uint32_t r0;  // bit of block b0 within word w0
  uint32_t nw;  // words touched
  bool first_shared, last_shared;
};

ZX_HD FxGroup fx_group(uint64_t g, uint32_t T, uint64_t nblocks, uint64_t start, uint32_t fixed)
{
  FxGroup q;
  q.b0 = g * T;
  q.nb = nblocks - q.b0 < T ? (uint32_t)(nblocks - q.b0) : T;
  const uint64_t bit0 = start + q.b0 * fixed;
  const uint64_t bit_end = bit0 + (uint64_t)q.nb * fixed;
  q.w0 = bit0 >> 6;
  q.r0 = (uint32_t)(bit0 & 63);
  q.nw = (uint32_t)(((bit_end + 63) >> 6) - q.w0);
  q.first_shared = q.r0 != 0;
  q.last_shared = (bit_end & 63) != 0;
  return q;
}

ZX_HD uint32_t fx_slot(uint32_t i, uint32_t skew) { return i + (i >> skew); }

// The variable-rate write pass takes the same encode phases, with each block
// placed by the scan of the block lengths (offsets[b] is block b's first bit
// and offsets[nblocks] the end) in the shared layout with a pad word every 8.
// A workgroup stages its range when it touches at most `cap` words; a denser
// range is written block by block to HBM.  The LDS a launch reserves bounds
// how many workgroups fit on a CU, so the capacity follows the stream's mean
// block length: room for 1.5 mean groups, at most 64 KiB.  `T` (at most 256)
// overrides the workgroup size and `cap` the capacity when nonzero.  Decodes
// through an offset index stage their groups the same way (k_decode_vx).
ZX_HD void vx_plan(uint64_t bits, uint64_t nblocks, uint32_t T, uint32_t cap, FxPlan &p)
{
  const uint64_t mean = nblocks ? (bits + nblocks - 1) / nblocks : 0;
  const uint64_t W = (mean + 63) / 64;
  p.T = T ? T : W <= 16 ? 256u : W <= 32 ? 128u : 64u;
  p.skew = 3;
  p.own = false;
  p.flat = false;
  // words 0 .. cap + 1 staged, and slots up to word cap + 2 (a reader near
  // the end loads three words: Source<StagedWords>::read_peek)
  const uint64_t most = (kFxMaxLdsBytes / 8 - 1) * 8 / 9 - 3;
  const uint64_t want = cap ? cap : (mean * p.T * 3 / 2 + 63) / 64 + 16;
  p.cap = (uint32_t)(want < most ? want : most);
  p.lds_words = fx_slot(p.cap + 2, p.skew) + 1;
}

ZX_HD FxGroup vx_group(uint64_t g, uint32_t T, uint64_t nblocks, const uint64_t *offsets)
{
  FxGroup q;
  q.b0 = g * T;
  q.nb = nblocks - q.b0 < T ? (uint32_t)(nblocks - q.b0) : T;
  const uint64_t bit0 = offsets[q.b0];
  const uint64_t bit_end = offsets[q.b0 + q.nb];
  q.w0 = bit0 >> 6;
  q.r0 = (uint32_t)(bit0 & 63);
  const uint64_t nw = ((bit_end + 63) >> 6) - q.w0;
  q.nw = nw > 0xffffffffull ? 0xffffffffu : (uint32_t)nw;
  q.first_shared = q.r0 != 0;
  q.last_shared = (bit_end & 63) != 0;
  return q;
}

// Words a block loaded, as the reference's cursor counts them (one past the
// highest): those it consumed bits of, plus the word a skip to min_bits lands
// inside (decompress.rs:57-74, bitstream/core.rs:265-284).  Positions are
// relative to the reader's word 0.
template <class Src>
ZX_HD uint64_t loaded_words(const Src &r, uint64_t pos, uint32_t minbits)
{
  const uint64_t used = r.pos() - pos;
  uint64_t loaded = r.words_loaded();
  if (used < minbits) {
    const uint64_t target = pos + minbits;
    if (target % 64) loaded = loaded > target / 64 + 1 ? loaded : target / 64 + 1;
  }
  return loaded;
}

// ------------------------------------------------------------------ encode

// phase 1 (shared layout only): zero word slot i
ZX_HD void fx_enc_clear(uint64_t *lds, const FxPlan &p, uint32_t i) { lds[fx_slot(i, p.skew)] = 0; }

// phase 2 (shared layout): a block's values v encoded from bit `rel` of the
// group's first word
template <class S, int D, bool REV>
ZX_HD void fx_enc_at(uint64_t *lds, const FxPlan &p, uint32_t rel,
                     const typename Ints<ST<S>::P>::U (&v)[Blk<D>::SIZE], const Cfg &c)
{
  LdsSink<true> w;
  w.init(lds, p.skew, rel);
  encode_values<S, D, REV>(w, v, c);
  w.finish();
}

// phase 2: thread t encodes its block's values v
template <class S, int D, bool REV, bool OWN>
ZX_HD void fx_enc_block(uint64_t *lds, const FxPlan &p, const FxGroup &q, uint32_t t, uint32_t fixed,
                        const typename Ints<ST<S>::P>::U (&v)[Blk<D>::SIZE], const Cfg &c)
{
  if constexpr (OWN) {
    LdsSink<false> w;
    w.init(lds + (size_t)t * (fixed / 64 + 1), 31, 0);
    encode_values<S, D, REV>(w, v, c);
    w.finish();
  } else {
    fx_enc_at<S, D, REV>(lds, p, q.r0 + t * fixed, v, c);
  }
}

// phase 3: word i of the group to the stream; false past the capacity
template <bool OWN>
ZX_HD bool fx_enc_store(uint64_t *words, uint64_t nwords, const uint64_t *lds, const FxPlan &p,
                        const FxGroup &q, uint32_t i, uint32_t fixed)
{
  const uint64_t gw = q.w0 + i;
  if (gw >= nwords) return false;  // the reference drops writes past the end
  if constexpr (OWN) {
    const uint32_t W = fixed / 64;
    fx_store_once(words + gw, lds[i + i / W]);
  } else {
    const uint64_t v = lds[fx_slot(i, p.skew)];
    if ((i == 0 && q.first_shared) || (i == q.nw - 1 && q.last_shared)) WordSink::merge(words + gw, v);
    else fx_store_once(words + gw, v);
  }
  return true;
}

// ------------------------------------------------------------------ decode

// staged word i in its slot, and in the pad slot before it when it has one
// (StagedWords)
ZX_HD void fx_dec_put(uint64_t *lds, const FxPlan &p, uint32_t i, uint64_t v)
{
  const uint32_t s = fx_slot(i, p.skew);
  lds[s] = v;
  if (i != 0 && (i & ((1u << p.skew) - 1)) == 0) lds[s - 1] = v;
}

// Words a thread stages per batch: all of a batch's loads are issued before
// the first is used, so their latencies overlap.
constexpr uint32_t kFxBatch = 16;

// phase 1 for thread t of T: stage words t, t + T, ... below nw + 2 (the
// reader looks up to two words ahead), a batch of loads issued before the
// batch is written
ZX_HD void fx_dec_stage_thread(uint64_t *lds, const uint64_t *words, uint64_t nwords, const FxPlan &p,
                               const FxGroup &q, uint32_t t)
{
  const uint32_t total = q.nw + 2;
  for (uint32_t base = 0; base < total; base += kFxBatch * p.T) {
    uint64_t v[kFxBatch];
    ZX_UNROLL
    for (uint32_t k = 0; k < kFxBatch; k++) {
      const uint32_t i = base + t + k * p.T;
      const uint64_t gw = q.w0 + i;
      v[k] = i < total && gw < nwords ? fx_load_once(words + gw) : 0ull;
    }
    ZX_UNROLL
    for (uint32_t k = 0; k < kFxBatch; k++) {
      const uint32_t i = base + t + k * p.T;
      if (i < total) fx_dec_put(lds, p, i, v[k]);
    }
  }
}

// phase 2: thread t decodes its block into v; returns one past the highest
// stream word the reference's reader would have loaded for it
template <class S, int D, bool REV, bool FLAT>
ZX_HD uint64_t fx_dec_block(const uint64_t *lds, const FxPlan &p, const FxGroup &q, uint32_t t,
                            uint32_t fixed, typename Ints<ST<S>::P>::U (&v)[Blk<D>::SIZE],
                            const Cfg &c)
{
  const uint64_t pos = q.r0 + (uint64_t)t * fixed;
  if constexpr (FLAT) {
    // positions relative to the block's first word (q.r0 is 0)
    const uint32_t i0 = (uint32_t)(pos >> 6);
    Source<StagedWords<true>> r;
    r.init(StagedWords<true>{lds + fx_slot(i0, p.skew), p.skew}, 0);
    decode_values<S, D, REV>(r, v, c);
    return q.w0 + i0 + loaded_words(r, 0, c.minbits);
  } else {
    Source<StagedWords<false>> r;
    r.init(StagedWords<false>{lds, p.skew}, pos);
    decode_values<S, D, REV>(r, v, c);
    return q.w0 + loaded_words(r, pos, c.minbits);
  }
}

// Variable-rate decode from a staged group (vx_group's range, staged with
// fx_dec_stage_thread): the block at stream bit `pos` of the words the
// group's w0 counts from; returns one past the highest word its reader
// would have loaded.
template <class S, int D, bool REV>
ZX_HD uint64_t vx_dec_block(const uint64_t *lds, const FxPlan &p, const FxGroup &q, uint64_t pos,
                            typename Ints<ST<S>::P>::U (&v)[Blk<D>::SIZE], const Cfg &c)
{
  const uint64_t rel = pos - 64 * q.w0;
  Source<StagedWords<false>> r;
  r.init(StagedWords<false>{lds, p.skew}, rel);
  decode_values<S, D, REV>(r, v, c);
  return q.w0 + loaded_words(r, rel, c.minbits);
}

// ------------------------------------------------------------------ one pass

// One-pass variable-rate compress (the single-array path for 1-3-D fields):
// a workgroup of T consecutive blocks prepares each block once
// (prepare_block), counts its bits, scans the counts across the workgroup,
// finds the bits of all earlier workgroups by a decoupled look-back, and then
// codes its blocks from the same registers: into LDS at their offsets when
// the group's word range fits `cap` words, straight to HBM otherwise (the
// group's words zeroed first, blocks' shared words merged with atomic ORs).
// Words wholly inside a group's range are stored by that group; the two
// words it may share with its neighbours (its first, when it starts inside a
// word, and its last, when it ends inside one) go to `edges`, and a second
// launch (op_merge) ORs each shared word together from the two groups' edges.
// No stream word is written twice and none needs clearing beforehand.  This
// restates the reference's chunked compress (compress.rs:132-236: chunks
// coded apart, then concatenated by bit-shifted appends) with the offsets
// found on the device instead of by a serial append.
struct OnePlan {
  uint32_t T;          // blocks per workgroup
  uint32_t cap;        // words a workgroup may stage in LDS
  uint32_t skew;       // LDS slot of word i: i + (i >> skew)
  uint32_t lds_words;  // LDS words per workgroup
};

// Most bits one block of `size` values can take: min_bits, or the header
// (at most 19 bits: a reversible float's), every coded plane's verbatim bits
// and one test (planes <= min(max_prec, P)), and at most three bits per
// coefficient over all planes (its zeros scanned, its value bit and its
// test), capped by the budget.
ZX_HD uint32_t block_bound(const Cfg &c, int P, int size)
{
  const uint64_t planes = c.maxprec < (uint32_t)P ? c.maxprec : (uint32_t)P;
  const uint64_t coded = 19 + planes * (uint64_t)(size + 1) + 3ull * size;
  const uint64_t budget = c.maxbits > 19 ? c.maxbits : 19;
  const uint64_t most = coded < budget ? coded : budget;
  return (uint32_t)(most > c.minbits ? most : c.minbits);
}

// LDS a one-pass workgroup may stage: more costs more than the groups that
// spill past it (measured on the 384^3 f64 accuracy workload).
constexpr uint64_t kOpRoomWords = (24 * 1024 / 8 - 2) * 8 / 9;

// The plan for blocks of at most `bound` bits, `mean` bits on average when a
// previous call on the stream tells (0 when not): the largest workgroup of
// 256, 128 or 64 blocks whose expected range (1.5 mean blocks, or the
// largest blocks) fits the room, 64 when none does, staging that range or
// the room, whichever is less.  `T` and `cap` override when nonzero.
ZX_HD OnePlan op_plan(uint32_t bound, uint32_t mean, uint32_t T, uint32_t cap)
{
  OnePlan p;
  p.skew = 3;
  uint64_t e = mean ? (uint64_t)mean * 3 / 2 + 1 : bound;
  if (e > bound) e = bound;
  p.T = T;
  if (!p.T) {
    p.T = 64;
    for (uint32_t t = 256; t > 64; t >>= 1)
      if (((uint64_t)t * e + 63) / 64 + 2 <= kOpRoomWords) {
        p.T = t;
        break;
      }
  }
  const uint64_t want = cap ? cap : ((uint64_t)p.T * e + 63) / 64 + 2;
  p.cap = (uint32_t)(want < kOpRoomWords ? want : kOpRoomWords);
  p.lds_words = fx_slot(p.cap, p.skew) + 1;
  return p;
}

// The range of a group: [b0, b1) in stream bits, its word count and shared edges.
struct OneRange {
  uint64_t b0, b1;
  uint64_t w0;      // first word
  uint32_t nw;      // words touched
  bool head, tail;  // first / last word shared with a neighbour (or the stream's prefix)
};

ZX_HD OneRange op_range(uint64_t b0, uint64_t total)
{
  OneRange r;
  r.b0 = b0;
  r.b1 = b0 + total;
  r.w0 = b0 >> 6;
  const uint64_t nw = ((r.b1 + 63) >> 6) - r.w0;
  r.nw = nw > 0xffffffffull ? 0xffffffffu : (uint32_t)nw;
  r.head = (b0 & 63) != 0;
  r.tail = (r.b1 & 63) != 0;
  return r;
}

// staged path, phase 2: a prepared block at bit `rel` of the group's first word
template <class S, int D>
ZX_HD void op_stage(uint64_t *lds, const OnePlan &p, uint32_t rel, const PreparedBlock<S, D> &pb, const Cfg &c)
{
  LdsSink<true> w;
  w.init(lds, p.skew, rel);
  code_block<S, D>(w, pb, c);
  w.finish();
}

// staged path, phase 3: word i of the group to the stream, or to the
// group's edges (edges[0]: first word, edges[1]: last word) when shared
ZX_HD void op_store(uint64_t *words, uint64_t nwords, const uint64_t *lds, const OnePlan &p,
                    const OneRange &r, uint32_t i, uint64_t *edges)
{
  const uint64_t v = lds[fx_slot(i, p.skew)];
  const bool first = i == 0 && r.head, last = i == r.nw - 1 && r.tail;
  if (first) edges[0] = v;
  if (last) edges[1] = v;
  if (!first && !last && r.w0 + i < nwords) fx_store_once(words + r.w0 + i, v);
}

// dense path, phase 2: word i of the group cleared where blocks may share it
ZX_HD void op_clear(uint64_t *words, uint64_t nwords, const OneRange &r, uint32_t i)
{
  const bool first = i == 0 && r.head, last = i == r.nw - 1 && r.tail;
  if (!first && !last && r.w0 + i < nwords) words[r.w0 + i] = 0;
}

// dense path, phase 3: a prepared block at stream bit `pos`
template <class S, int D>
ZX_HD void op_direct(uint64_t *words, uint64_t nwords, const OneRange &r, uint64_t pos,
                     const PreparedBlock<S, D> &pb, const Cfg &c, uint64_t *edge)
{
  EdgeSink w;
  w.init(words, nwords, pos, r.head ? r.w0 : ~0ull, r.tail ? r.w0 + r.nw - 1 : ~0ull, edge);
  code_block<S, D>(w, pb, c);
  w.finish();
}

// op_merge for group g of `groups`: end[g] is the bit after its last block
// in the low 62 bits (the scan's inclusive status word), the first group
// starting at `start`; edges[2g] and edges[2g + 1] hold its shared first and
// last words.  A word that groups share is written by the last group that
// starts inside it: its own first word ORed with the bits every earlier group
// left in that word (their last words, back to a group that began before the
// word or at its first bit) and, when the walk reaches the first group, the
// caller's `prefix` bits below `start`.  A last group that ends inside a word
// it did not start inside writes that word alone.  (Groups of 64 blocks or
// more hold at least 64 bits, so the walk takes one step; smaller groups
// may share a word several ways.)
ZX_HD void op_merge(uint64_t *words, uint64_t nwords, const uint64_t *end, const uint64_t *edges,
                    uint64_t g, uint64_t groups, uint64_t start, uint64_t prefix)
{
  constexpr uint64_t M = (1ull << 62) - 1;
  const uint64_t b0 = g ? end[g - 1] & M : start;
  const uint64_t b1 = end[g] & M;
  const uint64_t w = b0 >> 6;
  if ((b0 & 63) && !(g + 1 < groups && (b1 >> 6) == w)) {
    uint64_t v = edges[2 * g];
    for (uint64_t h = g;;) {
      if (h == 0) {
        v |= prefix;
        break;
      }
      h--;
      // group h ends at the bit where h + 1 starts, inside word w
      v |= edges[2 * h + 1];
      const uint64_t h0 = h ? end[h - 1] & M : start;
      if (h0 <= 64 * w) break;  // it began before the word or at its first bit
    }
    if (w < nwords) words[w] = v;
  }
  const uint64_t last = (b1 - 1) >> 6;
  if (g == groups - 1 && (b1 & 63) && !((b0 & 63) && last == w) && last < nwords) words[last] = edges[2 * g + 1];
}

}  // namespace zx
