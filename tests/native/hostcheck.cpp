// hostcheck.cpp -- builds the GPU block codec (zfp-rs_amd/csrc/zfpx_*.h) for
// the CPU and runs it with the kernels' placement logic, so the test suite can
// compare the kernels' own code with the oracle without a GPU.  Test
// infrastructure only.
#include <stdlib.h>
#include <vector>

#include "zfpx_alias.h"
#include "zfpx_fx.h"

using namespace zx;

struct HcCfg {
  uint32_t minbits, maxbits, maxprec;
  int32_t minexp, rounding, tight;
};

static FieldDesc make_field(int dims, void *data, const int64_t n[4], const int64_t s[4])
{
  FieldDesc f;
  f.data = data;
  for (int a = 0; a < 4; a++) {
    f.n[a] = a < dims ? n[a] : 1;
    f.s[a] = a < dims ? s[a] : 0;
  }
  set_block_grid(f);
  return f;
}

template <class S, int D, bool REV>
static uint64_t compress_t(const FieldDesc &f, const Cfg &c, uint64_t *words, uint64_t nwords,
                           uint64_t start, uint32_t fixed, uint32_t *lens)
{
  if (fixed) {
    for (uint64_t b = 0; b < f.nblocks; b++) {
      WordSink w;
      w.init(words, nwords, start + b * fixed);
      uint32_t bits = encode_field_block<S, D, REV>(f, b, w, c);
      w.finish();
      if (bits != fixed) abort();
    }
    return start + f.nblocks * fixed;
  }
  for (uint64_t b = 0; b < f.nblocks; b++) {
    CountSink cs{0};
    encode_field_block<S, D, REV>(f, b, cs, c);
    lens[b] = (uint32_t)cs.n;
  }
  uint64_t off = start;
  for (uint64_t b = 0; b < f.nblocks; b++) {
    WordSink w;
    w.init(words, nwords, off);
    uint32_t bits = encode_field_block<S, D, REV>(f, b, w, c);
    w.finish();
    if (bits != lens[b]) abort();
    off += lens[b];
  }
  return off;
}

template <class S, int D, bool REV>
static void decompress_t(const FieldDesc &f, const Cfg &c, const uint64_t *words, uint64_t nwords,
                         uint64_t start, uint32_t fixed, const uint32_t *lens)
{
  uint64_t off = start;
  for (uint64_t b = 0; b < f.nblocks; b++) {
    WordSource r;
    r.init(GlobalWords{words, nwords}, off);
    decode_field_block<S, D, REV>(f, b, r, c);
    off += fixed ? fixed : lens[b];
  }
}

// The fixed-rate LDS path of the kernels (zfpx_fx.h), one workgroup of T
// threads at a time, each phase run for every thread before the next.
template <class S, int D, bool REV>
static uint64_t compress_fx(const FieldDesc &f, const Cfg &c, uint64_t *words, uint64_t nwords,
                            uint64_t start, uint32_t fixed, const FxPlan &p)
{
  typedef typename Ints<ST<S>::P>::U U;
  std::vector<uint64_t> lds(p.lds_words);
  const uint64_t groups = (f.nblocks + p.T - 1) / p.T;
  for (uint64_t g = 0; g < groups; g++) {
    const FxGroup q = fx_group(g, p.T, f.nblocks, start, fixed);
    for (auto &x : lds) x = 0x5555aaaa5555aaaaull;  // LDS starts undefined
    if (!p.own)
      for (uint32_t i = 0; i < q.nw; i++) fx_enc_clear(lds.data(), p, i);
    for (uint32_t t = 0; t < q.nb; t++) {
      U v[Blk<D>::SIZE];
      gather<S, D>(f, block_pos<D>(f, q.b0 + t), v);
      if (p.own) fx_enc_block<S, D, REV, true>(lds.data(), p, q, t, fixed, v, c);
      else fx_enc_block<S, D, REV, false>(lds.data(), p, q, t, fixed, v, c);
    }
    for (uint32_t i = 0; i < q.nw; i++) {
      const bool in = p.own ? fx_enc_store<true>(words, nwords, lds.data(), p, q, i, fixed)
                            : fx_enc_store<false>(words, nwords, lds.data(), p, q, i, fixed);
      if (!in) break;
    }
  }
  return start + f.nblocks * fixed;
}

// LDS starts undefined: each group's copy is filled with different words
// first, so that a read of a slot the group did not stage shows as a
// mismatch with the oracle
static void lds_garbage(std::vector<uint64_t> &lds, uint64_t g)
{
  uint64_t z = 0x9e3779b97f4a7c15ull * (g + 1);
  for (auto &x : lds) {
    z ^= z << 13;
    z ^= z >> 7;
    z ^= z << 17;
    x = z;
  }
}

template <class S, int D, bool REV>
static uint64_t decompress_fx(const FieldDesc &f, const Cfg &c, const uint64_t *words,
                              uint64_t nwords, uint64_t start, uint32_t fixed, const FxPlan &p)
{
  typedef typename Ints<ST<S>::P>::U U;
  std::vector<uint64_t> lds(p.lds_words);
  const uint64_t groups = (f.nblocks + p.T - 1) / p.T;
  uint64_t maxword = 0;
  for (uint64_t g = 0; g < groups; g++) {
    const FxGroup q = fx_group(g, p.T, f.nblocks, start, fixed);
    lds_garbage(lds, g);
    // the kernel's staging phase, thread by thread
    for (uint32_t t = 0; t < p.T; t++) fx_dec_stage_thread(lds.data(), words, nwords, p, q, t);
    for (uint32_t t = 0; t < q.nb; t++) {
      U v[Blk<D>::SIZE];
      const uint64_t loaded = p.flat ? fx_dec_block<S, D, REV, true>(lds.data(), p, q, t, fixed, v, c)
                                     : fx_dec_block<S, D, REV, false>(lds.data(), p, q, t, fixed, v, c);
      if (q.b0 + t == f.nblocks - 1) maxword = loaded;
      scatter<S, D>(f, block_pos<D>(f, q.b0 + t), v);
    }
  }
  return maxword;
}

// The variable-rate decode through LDS (k_decode_vx, then k_decode_dense for
// the groups over the capacity), one workgroup and one thread at a time.
template <class S, int D, bool REV>
static uint64_t decompress_vx(const FieldDesc &f, const Cfg &c, const uint64_t *words,
                              uint64_t nwords, uint64_t start, const uint32_t *lens, uint32_t T,
                              uint32_t cap)
{
  typedef typename Ints<ST<S>::P>::U U;
  std::vector<uint64_t> offs(f.nblocks + 1);
  offs[0] = start;
  for (uint64_t b = 0; b < f.nblocks; b++) offs[b + 1] = offs[b] + lens[b];
  FxPlan p;
  vx_plan(offs[f.nblocks] - start, f.nblocks, T, cap, p);
  std::vector<uint64_t> lds(p.lds_words);
  const uint64_t groups = (f.nblocks + p.T - 1) / p.T;
  uint64_t maxword = 0;
  for (uint64_t g = 0; g < groups; g++) {
    const FxGroup q = vx_group(g, p.T, f.nblocks, offs.data());
    lds_garbage(lds, g);
    const bool staged = q.nw <= p.cap;
    if (staged)
      for (uint32_t t = 0; t < p.T; t++) fx_dec_stage_thread(lds.data(), words, nwords, p, q, t);
    for (uint32_t t = 0; t < q.nb; t++) {
      U v[Blk<D>::SIZE];
      const uint64_t b = q.b0 + t;
      uint64_t loaded;
      if (staged) {
        loaded = vx_dec_block<S, D, REV>(lds.data(), p, q, offs[b], v, c);
      } else {
        WordSource r;
        r.init(GlobalWords{words, nwords}, offs[b]);
        decode_values<S, D, REV>(r, v, c);
        loaded = loaded_words(r, offs[b], c.minbits);
      }
      if (b == f.nblocks - 1) maxword = loaded;
      scatter<S, D>(f, block_pos<D>(f, b), v);
    }
  }
  return maxword;
}

// The variable-rate LDS path: the length pass, the scan, then the write pass
// one workgroup at a time, staged through LDS or, for a workgroup over the
// capacity, block by block (the kernels' second launch; its blocks' bits are
// disjoint from the staged workgroups', so the order does not matter).
template <class S, int D, bool REV>
static uint64_t compress_vx(const FieldDesc &f, const Cfg &c, uint64_t *words, uint64_t nwords,
                            uint64_t start, uint32_t *lens, uint32_t T, uint32_t cap)
{
  typedef typename Ints<ST<S>::P>::U U;
  std::vector<uint64_t> offs(f.nblocks + 1);
  offs[0] = start;
  for (uint64_t b = 0; b < f.nblocks; b++) {
    CountSink cs{0};
    encode_field_block<S, D, REV>(f, b, cs, c);
    lens[b] = (uint32_t)cs.n;
    offs[b + 1] = offs[b] + lens[b];
  }
  FxPlan p;
  vx_plan(offs[f.nblocks] - start, f.nblocks, T, cap, p);
  std::vector<uint64_t> lds(p.lds_words);
  const uint64_t groups = (f.nblocks + p.T - 1) / p.T;
  for (uint64_t g = 0; g < groups; g++) {
    const FxGroup q = vx_group(g, p.T, f.nblocks, offs.data());
    const bool staged = q.nw <= p.cap;
    for (auto &x : lds) x = 0x5555aaaa5555aaaaull;  // LDS starts undefined
    if (staged)
      for (uint32_t i = 0; i < q.nw; i++) fx_enc_clear(lds.data(), p, i);
    for (uint32_t t = 0; t < q.nb; t++) {
      U v[Blk<D>::SIZE];
      gather<S, D>(f, block_pos<D>(f, q.b0 + t), v);
      const uint64_t pos = offs[q.b0 + t];
      if (staged) {
        fx_enc_at<S, D, REV>(lds.data(), p, (uint32_t)(pos - 64 * q.w0), v, c);
      } else {
        WordSink w;
        w.init(words, nwords, pos);
        encode_values<S, D, REV>(w, v, c);
        w.finish();
      }
    }
    if (staged)
      for (uint32_t i = 0; i < q.nw; i++)
        if (!fx_enc_store<false>(words, nwords, lds.data(), p, q, i, 0)) break;
  }
  return offs[f.nblocks];
}

// The one-pass compress (zfpx_fx.h op_*): each workgroup's phases one thread
// after another, the workgroups in ticket order (so each group's look-back
// prefix is the running end), then the merge of the shared edge words.
// Stream words are not cleared first: the path writes every word of the
// stream's range itself, and the caller's bits below `start` come from the
// word that holds them, as the cursor's buffer holds them.
template <class S, int D, bool REV>
static uint64_t compress_1p(const FieldDesc &f, const Cfg &c, uint64_t *words, uint64_t nwords,
                            uint64_t start, uint32_t *lens, uint32_t T, uint32_t cap)
{
  typedef typename Ints<ST<S>::P>::U U;
  const OnePlan p = op_plan(block_bound(c, ST<S>::P, Blk<D>::SIZE), 0, T, cap);
  const uint64_t prefix = (start & 63) ? words[start >> 6] & lowmask((uint32_t)(start & 63)) : 0;
  const uint64_t groups = (f.nblocks + p.T - 1) / p.T;
  std::vector<uint64_t> lds(p.lds_words), edges(2 * groups), ends(groups);
  std::vector<PreparedBlock<S, D>> pb(p.T);
  std::vector<uint32_t> len(p.T);
  uint64_t pos = start;
  for (uint64_t g = 0; g < groups; g++) {
    uint32_t total = 0;
    for (uint32_t t = 0; t < p.T; t++) {
      const uint64_t b = g * p.T + t;
      len[t] = 0;
      if (b >= f.nblocks) continue;
      U v[Blk<D>::SIZE];
      gather<S, D>(f, block_pos<D>(f, b), v);
      prepare_block<S, D, REV>(v, c, pb[t]);
      CountSink cs{0};
      len[t] = code_block<S, D>(cs, pb[t], c);
      lens[b] = len[t];
      total += len[t];
    }
    const OneRange r = op_range(pos, total);
    uint64_t edge[2] = {0, 0};
    uint32_t off = 0;
    if (r.nw <= p.cap) {
      for (auto &x : lds) x = 0x5555aaaa5555aaaaull;  // LDS starts undefined
      for (uint32_t i = 0; i < r.nw; i++) lds[fx_slot(i, p.skew)] = 0;
      for (uint32_t t = 0; t < p.T && g * p.T + t < f.nblocks; t++) {
        op_stage<S, D>(lds.data(), p, (uint32_t)(r.b0 & 63) + off, pb[t], c);
        off += len[t];
      }
      for (uint32_t i = 0; i < r.nw; i++) op_store(words, nwords, lds.data(), p, r, i, edge);
    } else {
      for (uint32_t i = 0; i < r.nw; i++) op_clear(words, nwords, r, i);
      for (uint32_t t = 0; t < p.T && g * p.T + t < f.nblocks; t++) {
        op_direct<S, D>(words, nwords, r, r.b0 + off, pb[t], c, edge);
        off += len[t];
      }
    }
    edges[2 * g] = edge[0];
    edges[2 * g + 1] = edge[1];
    pos = r.b1;
    ends[g] = pos;
  }
  for (uint64_t g = 0; g < groups; g++)
    op_merge(words, nwords, ends.data(), edges.data(), g, groups, start, prefix);
  return pos;
}

template <class S, bool REV>
static uint64_t comp1p_d(int dims, const FieldDesc &f, const Cfg &c, uint64_t *w, uint64_t nw,
                         uint64_t st, uint32_t *l, uint32_t T, uint32_t cap)
{
  switch (dims) {
  case 1: return compress_1p<S, 1, REV>(f, c, w, nw, st, l, T, cap);
  case 2: return compress_1p<S, 2, REV>(f, c, w, nw, st, l, T, cap);
  default: return compress_1p<S, 3, REV>(f, c, w, nw, st, l, T, cap);
  }
}

template <class S, bool REV>
static uint64_t comp_d(int dims, const FieldDesc &f, const Cfg &c, uint64_t *w, uint64_t nw,
                       uint64_t st, uint32_t fx, uint32_t *l, uint32_t T, uint32_t cap)
{
  FxPlan p;
  if (T && !fx && dims < 4) {
    switch (dims) {
    case 1: return compress_vx<S, 1, REV>(f, c, w, nw, st, l, T, cap);
    case 2: return compress_vx<S, 2, REV>(f, c, w, nw, st, l, T, cap);
    default: return compress_vx<S, 3, REV>(f, c, w, nw, st, l, T, cap);
    }
  }
  // (4-D: the whole-block layout only, as the kernels' launcher)
  if (T && fx_plan(fx, st, false, T, p) && (dims < 4 || p.own)) {
    switch (dims) {
    case 1: return compress_fx<S, 1, REV>(f, c, w, nw, st, fx, p);
    case 2: return compress_fx<S, 2, REV>(f, c, w, nw, st, fx, p);
    case 3: return compress_fx<S, 3, REV>(f, c, w, nw, st, fx, p);
    default: return compress_fx<S, 4, REV>(f, c, w, nw, st, fx, p);
    }
  }
  switch (dims) {
  case 1: return compress_t<S, 1, REV>(f, c, w, nw, st, fx, l);
  case 2: return compress_t<S, 2, REV>(f, c, w, nw, st, fx, l);
  case 3: return compress_t<S, 3, REV>(f, c, w, nw, st, fx, l);
  default: return compress_t<S, 4, REV>(f, c, w, nw, st, fx, l);
  }
}

template <class S, bool REV>
static uint64_t decomp_d(int dims, const FieldDesc &f, const Cfg &c, const uint64_t *w, uint64_t nw,
                         uint64_t st, uint32_t fx, const uint32_t *l, uint32_t T)
{
  FxPlan p;
  if (T && fx_plan(fx, st, true, T, p) && (dims < 4 || p.flat)) {
    switch (dims) {
    case 1: return decompress_fx<S, 1, REV>(f, c, w, nw, st, fx, p);
    case 2: return decompress_fx<S, 2, REV>(f, c, w, nw, st, fx, p);
    case 3: return decompress_fx<S, 3, REV>(f, c, w, nw, st, fx, p);
    default: return decompress_fx<S, 4, REV>(f, c, w, nw, st, fx, p);
    }
  }
  if (T && !fx && l && dims < 4) {
    // odd T: a capacity of the mean group's words, so that some groups are
    // staged and some left to the dense pass
    uint64_t bits = 0;
    for (uint64_t b = 0; b < f.nblocks; b++) bits += l[b];
    const uint32_t cap = T % 2 ? (uint32_t)((bits * T / (f.nblocks ? f.nblocks : 1) + 63) / 64) + 1 : 0u;
    switch (dims) {
    case 1: return decompress_vx<S, 1, REV>(f, c, w, nw, st, l, T, cap);
    case 2: return decompress_vx<S, 2, REV>(f, c, w, nw, st, l, T, cap);
    default: return decompress_vx<S, 3, REV>(f, c, w, nw, st, l, T, cap);
    }
  }
  switch (dims) {
  case 1: decompress_t<S, 1, REV>(f, c, w, nw, st, fx, l); break;
  case 2: decompress_t<S, 2, REV>(f, c, w, nw, st, fx, l); break;
  case 3: decompress_t<S, 3, REV>(f, c, w, nw, st, fx, l); break;
  default: decompress_t<S, 4, REV>(f, c, w, nw, st, fx, l); break;
  }
  return 0;
}

template <class S, bool REV>
static uint64_t walk_d(int dims, const Cfg &c, const uint64_t *w, uint64_t nw, uint64_t st,
                       uint64_t count, uint64_t *offs)
{
  switch (dims) {
  case 1: return walk_offsets<S, 1, REV>(w, nw, st, count, c, offs);
  case 2: return walk_offsets<S, 2, REV>(w, nw, st, count, c, offs);
  case 3: return walk_offsets<S, 3, REV>(w, nw, st, count, c, offs);
  default: return walk_offsets<S, 4, REV>(w, nw, st, count, c, offs);
  }
}

extern "C" {

// The one-pass variable-rate compress of a 1-3-D field; T and cap override
// the plan's workgroup size and LDS capacity when nonzero.
uint64_t hc_compress_1p(int type, int dims, int rev, void *data, const int64_t *n, const int64_t *s,
                        const HcCfg *hc, uint64_t *words, uint64_t nwords, uint64_t start_bit,
                        uint32_t *lengths, uint32_t T, uint32_t cap)
{
  FieldDesc f = make_field(dims, data, n, s);
  Cfg c{hc->minbits, hc->maxbits, hc->maxprec, hc->minexp, hc->rounding, hc->tight};
  switch (type * 2 + (rev ? 1 : 0)) {
  case 2: return comp1p_d<int32_t, false>(dims, f, c, words, nwords, start_bit, lengths, T, cap);
  case 3: return comp1p_d<int32_t, true>(dims, f, c, words, nwords, start_bit, lengths, T, cap);
  case 4: return comp1p_d<int64_t, false>(dims, f, c, words, nwords, start_bit, lengths, T, cap);
  case 5: return comp1p_d<int64_t, true>(dims, f, c, words, nwords, start_bit, lengths, T, cap);
  case 6: return comp1p_d<float, false>(dims, f, c, words, nwords, start_bit, lengths, T, cap);
  case 7: return comp1p_d<float, true>(dims, f, c, words, nwords, start_bit, lengths, T, cap);
  case 8: return comp1p_d<double, false>(dims, f, c, words, nwords, start_bit, lengths, T, cap);
  default: return comp1p_d<double, true>(dims, f, c, words, nwords, start_bit, lengths, T, cap);
  }
}

// The length reader of index-less decodes: offsets[0..count) and the end bit.
uint64_t hc_walk(int type, int dims, int rev, const HcCfg *hc, const uint64_t *words,
                 uint64_t nwords, uint64_t start_bit, uint64_t count, uint64_t *offsets)
{
  Cfg c{hc->minbits, hc->maxbits, hc->maxprec, hc->minexp, hc->rounding, hc->tight};
  switch (type * 2 + (rev ? 1 : 0)) {
  case 2: return walk_d<int32_t, false>(dims, c, words, nwords, start_bit, count, offsets);
  case 3: return walk_d<int32_t, true>(dims, c, words, nwords, start_bit, count, offsets);
  case 4: return walk_d<int64_t, false>(dims, c, words, nwords, start_bit, count, offsets);
  case 5: return walk_d<int64_t, true>(dims, c, words, nwords, start_bit, count, offsets);
  case 6: return walk_d<float, false>(dims, c, words, nwords, start_bit, count, offsets);
  case 7: return walk_d<float, true>(dims, c, words, nwords, start_bit, count, offsets);
  case 8: return walk_d<double, false>(dims, c, words, nwords, start_bit, count, offsets);
  default: return walk_d<double, true>(dims, c, words, nwords, start_bit, count, offsets);
  }
}

uint64_t hc_compress(int type, int dims, int rev, void *data, const int64_t *n, const int64_t *s,
                     const HcCfg *hc, uint64_t *words, uint64_t nwords, uint64_t start_bit,
                     uint32_t fixed_bits, uint32_t *lengths, uint32_t lds_T,
                     uint32_t lds_cap)
{
  FieldDesc f = make_field(dims, data, n, s);
  Cfg c{hc->minbits, hc->maxbits, hc->maxprec, hc->minexp, hc->rounding, hc->tight};
  switch (type * 2 + (rev ? 1 : 0)) {
  case 2: return comp_d<int32_t, false>(dims, f, c, words, nwords, start_bit, fixed_bits, lengths, lds_T, lds_cap);
  case 3: return comp_d<int32_t, true>(dims, f, c, words, nwords, start_bit, fixed_bits, lengths, lds_T, lds_cap);
  case 4: return comp_d<int64_t, false>(dims, f, c, words, nwords, start_bit, fixed_bits, lengths, lds_T, lds_cap);
  case 5: return comp_d<int64_t, true>(dims, f, c, words, nwords, start_bit, fixed_bits, lengths, lds_T, lds_cap);
  case 6: return comp_d<float, false>(dims, f, c, words, nwords, start_bit, fixed_bits, lengths, lds_T, lds_cap);
  case 7: return comp_d<float, true>(dims, f, c, words, nwords, start_bit, fixed_bits, lengths, lds_T, lds_cap);
  case 8: return comp_d<double, false>(dims, f, c, words, nwords, start_bit, fixed_bits, lengths, lds_T, lds_cap);
  default: return comp_d<double, true>(dims, f, c, words, nwords, start_bit, fixed_bits, lengths, lds_T, lds_cap);
  }
}

uint64_t hc_decompress(int type, int dims, int rev, void *data, const int64_t *n, const int64_t *s,
                   const HcCfg *hc, const uint64_t *words, uint64_t nwords, uint64_t start_bit,
                   uint32_t fixed_bits, const uint32_t *lengths, uint32_t lds_T)
{
  FieldDesc f = make_field(dims, data, n, s);
  Cfg c{hc->minbits, hc->maxbits, hc->maxprec, hc->minexp, hc->rounding, hc->tight};
  switch (type * 2 + (rev ? 1 : 0)) {
  case 2: return decomp_d<int32_t, false>(dims, f, c, words, nwords, start_bit, fixed_bits, lengths, lds_T);
  case 3: return decomp_d<int32_t, true>(dims, f, c, words, nwords, start_bit, fixed_bits, lengths, lds_T);
  case 4: return decomp_d<int64_t, false>(dims, f, c, words, nwords, start_bit, fixed_bits, lengths, lds_T);
  case 5: return decomp_d<int64_t, true>(dims, f, c, words, nwords, start_bit, fixed_bits, lengths, lds_T);
  case 6: return decomp_d<float, false>(dims, f, c, words, nwords, start_bit, fixed_bits, lengths, lds_T);
  case 7: return decomp_d<float, true>(dims, f, c, words, nwords, start_bit, fixed_bits, lengths, lds_T);
  case 8: return decomp_d<double, false>(dims, f, c, words, nwords, start_bit, fixed_bits, lengths, lds_T);
  default: return decomp_d<double, true>(dims, f, c, words, nwords, start_bit, fixed_bits, lengths, lds_T);
  }
}

// The alias scatter of zfpx_alias.hip run on the CPU, its two passes in the
// opposite thread order to the natural one (the result must not depend on it).
int hc_alias_scatter(const void *dense, void *lowest, const int64_t *n, const int64_t *s,
                     int64_t imin, int dims, uint32_t esz)
{
  AliasArgs a{};
  a.dense = dense;
  a.lowest = lowest;
  a.count = 1;
  int64_t lo = 0, hi = 0;
  for (int k = 0; k < 4; k++) {
    a.n[k] = k < dims ? n[k] : 1;
    a.s[k] = k < dims ? s[k] : 0;
    if (k < dims) {
      a.count *= (uint64_t)n[k];
      const int64_t e = s[k] * (n[k] - 1);
      if (e < 0) lo += e;
      else hi += e;
    }
  }
  a.imin = imin;
  a.span = (uint64_t)(hi - lo + 1);
  a.dims = dims;
  a.esz = esz;
  if (lo != imin) return -1;
  std::vector<uint64_t> best(a.span, 0);
  for (uint64_t i = a.count; i-- > 0;) {
    const AliasSlot q = alias_slot(a, i);
    if ((uint64_t)q.addr >= a.span) return -2;
    if (q.ticket > best[q.addr]) best[q.addr] = q.ticket;
  }
  for (uint64_t i = a.count; i-- > 0;) {
    const AliasSlot q = alias_slot(a, i);
    if (best[q.addr] != q.ticket) continue;
    if (esz == 4) ((uint32_t *)lowest)[q.addr] = ((const uint32_t *)dense)[i];
    else ((uint64_t *)lowest)[q.addr] = ((const uint64_t *)dense)[i];
  }
  return 0;
}

// n / d through the block grid's invariant-divisor path, for count values
void hc_divide(const uint32_t *n, uint32_t d, uint32_t *q, uint64_t count)
{
  const Divisor v = make_divisor(d);
  for (uint64_t i = 0; i < count; i++) q[i] = divide(n[i], v);
}

}  // extern "C"
