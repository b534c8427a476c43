// hostcheck.cpp -- builds the GPU block codec (zfp-rs_amd/csrc/zfpx_*.h) for
// the CPU and runs it with the kernels' placement logic, so the test suite can
// compare the kernels' own code with the oracle without a GPU.  Test
// infrastructure only.
#include <stdlib.h>
#include <vector>

#include "zfpx_field.h"

using namespace zx;

struct HcCfg {
  uint32_t minbits, maxbits, maxprec;
  int32_t minexp, rounding, tight;
};

static FieldDesc make_field(int dims, void *data, const int64_t n[3], const int64_t s[3])
{
  FieldDesc f;
  f.data = data;
  f.nblocks = 1;
  for (int a = 0; a < 3; a++) {
    f.n[a] = a < dims ? n[a] : 1;
    f.s[a] = a < dims ? s[a] : 0;
    f.nb[a] = (uint32_t)((f.n[a] + 3) / 4);
    f.nblocks *= f.nb[a];
  }
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
    r.init(words, nwords, off);
    decode_field_block<S, D, REV>(f, b, r, c);
    off += fixed ? fixed : lens[b];
  }
}

template <class S, bool REV>
static uint64_t comp_d(int dims, const FieldDesc &f, const Cfg &c, uint64_t *w, uint64_t nw,
                       uint64_t st, uint32_t fx, uint32_t *l)
{
  switch (dims) {
  case 1: return compress_t<S, 1, REV>(f, c, w, nw, st, fx, l);
  case 2: return compress_t<S, 2, REV>(f, c, w, nw, st, fx, l);
  default: return compress_t<S, 3, REV>(f, c, w, nw, st, fx, l);
  }
}

template <class S, bool REV>
static void decomp_d(int dims, const FieldDesc &f, const Cfg &c, const uint64_t *w, uint64_t nw,
                     uint64_t st, uint32_t fx, const uint32_t *l)
{
  switch (dims) {
  case 1: decompress_t<S, 1, REV>(f, c, w, nw, st, fx, l); break;
  case 2: decompress_t<S, 2, REV>(f, c, w, nw, st, fx, l); break;
  default: decompress_t<S, 3, REV>(f, c, w, nw, st, fx, l); break;
  }
}

extern "C" {

uint64_t hc_compress(int type, int dims, int rev, void *data, const int64_t *n, const int64_t *s,
                     const HcCfg *hc, uint64_t *words, uint64_t nwords, uint64_t start_bit,
                     uint32_t fixed_bits, uint32_t *lengths)
{
  FieldDesc f = make_field(dims, data, n, s);
  Cfg c{hc->minbits, hc->maxbits, hc->maxprec, hc->minexp, hc->rounding, hc->tight};
  switch (type * 2 + (rev ? 1 : 0)) {
  case 2: return comp_d<int32_t, false>(dims, f, c, words, nwords, start_bit, fixed_bits, lengths);
  case 3: return comp_d<int32_t, true>(dims, f, c, words, nwords, start_bit, fixed_bits, lengths);
  case 4: return comp_d<int64_t, false>(dims, f, c, words, nwords, start_bit, fixed_bits, lengths);
  case 5: return comp_d<int64_t, true>(dims, f, c, words, nwords, start_bit, fixed_bits, lengths);
  case 6: return comp_d<float, false>(dims, f, c, words, nwords, start_bit, fixed_bits, lengths);
  case 7: return comp_d<float, true>(dims, f, c, words, nwords, start_bit, fixed_bits, lengths);
  case 8: return comp_d<double, false>(dims, f, c, words, nwords, start_bit, fixed_bits, lengths);
  default: return comp_d<double, true>(dims, f, c, words, nwords, start_bit, fixed_bits, lengths);
  }
}

void hc_decompress(int type, int dims, int rev, void *data, const int64_t *n, const int64_t *s,
                   const HcCfg *hc, const uint64_t *words, uint64_t nwords, uint64_t start_bit,
                   uint32_t fixed_bits, const uint32_t *lengths)
{
  FieldDesc f = make_field(dims, data, n, s);
  Cfg c{hc->minbits, hc->maxbits, hc->maxprec, hc->minexp, hc->rounding, hc->tight};
  switch (type * 2 + (rev ? 1 : 0)) {
  case 2: decomp_d<int32_t, false>(dims, f, c, words, nwords, start_bit, fixed_bits, lengths); break;
  case 3: decomp_d<int32_t, true>(dims, f, c, words, nwords, start_bit, fixed_bits, lengths); break;
  case 4: decomp_d<int64_t, false>(dims, f, c, words, nwords, start_bit, fixed_bits, lengths); break;
  case 5: decomp_d<int64_t, true>(dims, f, c, words, nwords, start_bit, fixed_bits, lengths); break;
  case 6: decomp_d<float, false>(dims, f, c, words, nwords, start_bit, fixed_bits, lengths); break;
  case 7: decomp_d<float, true>(dims, f, c, words, nwords, start_bit, fixed_bits, lengths); break;
  case 8: decomp_d<double, false>(dims, f, c, words, nwords, start_bit, fixed_bits, lengths); break;
  default: decomp_d<double, true>(dims, f, c, words, nwords, start_bit, fixed_bits, lengths); break;
  }
}

}  // extern "C"
