// zfp_capi.cpp -- the C ABI (include/zfp.h, include/zfp_hip.h) over the HIP
// field codec.
//
// Mirrors zfp-rs-ffi (zfp-rs-ffi/src/{stream,field,bitstream}.rs): the same
// entry points, argument meaning and error behaviour (0 / null / zfp_false on
// failure, never a crash for bad parameters).  The bit-stream cursor keeps
// the reference's exact semantics (src/bitstream/core.rs), so header I/O and
// user bit I/O interleave with GPU compression exactly as with zfp-rs.
//
// zfp_compress / zfp_decompress run every block on the GPU; there is no CPU
// codec in this library.  Host-resident fields and streams are staged through
// device memory; device-resident ones are coded in place.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <mutex>
#include <atomic>
#include <vector>

#include "../../include/zfp.h"
#include "../../include/zfp_hip.h"
#include "../../include/zfp_block.h"
#include "zfpx_api.h"
#include "zfpx_fx.h"

extern "C" {
const size_t stream_word_bits = 64;
const uint32_t zfp_codec_version = ZFP_CODEC;
const uint32_t zfp_library_version = 10001;  // 1.0.1 as major*10000 + minor*100 + patch
}

// ===================================================================== memory

namespace {

bool is_device_ptr(const void *p)
{
  if (!p) return false;
  hipPointerAttribute_t a;
  hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

// Grow-only device buffers shared by the codec calls of this process.
// The reason the calling thread's last call failed inside the library
// (zfp_hip_last_error): the source line and the HIP error or reservation
// that failed.
thread_local char g_last_error[96];

void note_failure(int line, const char *what)
{
  snprintf(g_last_error, sizeof g_last_error, "zfp_capi.cpp:%d: %s", line, what);
}

// Injected reservation failures for tests (zfp_hip_fail_after): counts the
// reservations of workspace, staging and pinned memory from the moment it is
// armed, and makes the n-th fail as an out-of-memory would, once.
std::atomic<long> g_fail_after{-1};

bool injected_failure()
{
  long n = g_fail_after.load();
  while (n >= 0 && !g_fail_after.compare_exchange_weak(n, n - 1)) {
  }
  return n == 0;
}

struct Workspace {
  std::mutex mu;
  void *field = nullptr; size_t field_cap = 0;
  void *words = nullptr; size_t words_cap = 0;
  void *lens = nullptr; size_t lens_cap = 0;
  void *offs = nullptr; size_t offs_cap = 0;
  void *tiles = nullptr; size_t tiles_cap = 0;
  void *dense = nullptr; size_t dense_cap = 0;  // decode of an aliasing field
  void *best = nullptr; size_t best_cap = 0;    // its per-element tickets
  void *one = nullptr;  // one pinned host word for reading back a scan total
  // pinned host memory of the length reader (index-less variable-rate decode):
  // a copy of a device stream's words, and the block offsets it finds
  void *hwords = nullptr; size_t hwords_cap = 0;
  void *hoffs = nullptr; size_t hoffs_cap = 0;
  int device = -1;

  static bool grow(void *&p, size_t &cap, size_t need)
  {
    if (injected_failure()) {
      note_failure(__LINE__, "injected reservation failure");
      return false;
    }
    if (need <= cap) return true;
    const size_t n = std::max(need, cap + cap / 2);
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    if (hipMalloc(&p, n) != hipSuccess) {
      note_failure(__LINE__, "workspace hipMalloc");
      (void)hipGetLastError();
      p = nullptr;
      return false;
    }
    cap = n;
    return true;
  }
};

Workspace &ws()
{
  static Workspace w;
  return w;
}

}  // namespace

// ===================================================================== bit stream

// A bit stream over 64-bit words in host or device memory, with the cursor of
// src/bitstream/core.rs:7-18.
struct bitstream {
  uint64_t *words;
  size_t nwords;
  bool owned;
  bool device;
  size_t word_pos;
  uint64_t buffer;
  uint32_t bits;
  bool overflowed;
  bool overread;
  size_t seek_word;
  // offset index of the last variable-rate compression (device memory)
  uint32_t *index;
  size_t index_n;
  size_t index_cap;
  uint64_t index_start;
  uint64_t index_key;
  uint64_t index_bits;  // bits of the indexed blocks (sizes the decode's LDS staging; 0: unknown)
  // the batch call that last took this stream (duplicates within one call
  // go to the single-array calls)
  uint64_t batch_epoch;
  // mean block length in bits of the last one-pass compress through this
  // stream (0: none yet), which sizes the next one's LDS staging
  uint32_t mean_bits;

  uint64_t get(size_t i) const
  {
    if (i >= nwords) return 0;
    if (!device) return words[i];
    uint64_t v = 0;
    (void)hipMemcpy(&v, words + i, 8, hipMemcpyDeviceToHost);
    return v;
  }
  void set(size_t i, uint64_t v)
  {
    if (!device) words[i] = v;
    else (void)hipMemcpy(words + i, &v, 8, hipMemcpyHostToDevice);
  }
  // read_word_raw / write_word_raw (core.rs:151-170)
  uint64_t load()
  {
    uint64_t w = get(word_pos);
    word_pos++;
    return w;
  }
  void store(uint64_t v)
  {
    if (word_pos < nwords) set(word_pos, v);
    else overflowed = true;
    word_pos++;
  }
  // read_bits_impl (core.rs:192-229)
  uint64_t read_bits(uint32_t n)
  {
    if (n > 64) n = 64;
    uint64_t value = buffer;
    if (bits < n) {
      uint64_t w = load();
      buffer = w;
      value += w << bits;
      bits = bits + 64 - n;
      if (bits == 0) {
        buffer = 0;
      } else {
        buffer >>= 64 - bits;
        if (n < 64) value &= (2ull << (n - 1)) - 1;
      }
    } else {
      bits -= n;
      buffer = n < 64 ? buffer >> n : 0;
      if (n < 64) value &= (1ull << n) - 1;
    }
    return value;
  }
  uint32_t read_bit()
  {
    if (bits == 0) {
      buffer = load();
      bits = 64;
    }
    bits--;
    uint32_t b = (uint32_t)(buffer & 1);
    buffer >>= 1;
    return b;
  }
  // write_bits_impl (core.rs:296-326)
  uint64_t write_bits(uint64_t value, uint32_t n)
  {
    if (n > 64) n = 64;
    uint32_t old = bits;
    buffer += value << bits;
    bits += n;
    if (bits >= 64) {
      uint64_t full = buffer;
      bits -= 64;
      store(full);
      buffer = (value >> 1) >> (63 - old);
    }
    if (bits < 64) buffer &= (1ull << bits) - 1;
    return n < 64 ? value >> n : 0;
  }
  void write_bit(uint32_t bit)
  {
    buffer += (uint64_t)bit << bits;
    bits++;
    if (bits == 64) {
      store(buffer);
      buffer = 0;
      bits = 0;
    }
  }
  bool loaded_since_seek() const { return word_pos > std::max(nwords, seek_word); }
  // seek_read_impl (core.rs:265-284)
  void rseek(uint64_t off)
  {
    uint32_t n = (uint32_t)(off % 64);
    overread = overread || loaded_since_seek();
    word_pos = (size_t)(off / 64);
    seek_word = word_pos;
    if (n) {
      buffer = load() >> n;
      bits = 64 - n;
    } else {
      buffer = 0;
      bits = 0;
    }
  }
  // seek_write_impl (core.rs:354-375)
  void wseek(uint64_t off)
  {
    uint32_t n = (uint32_t)(off % 64);
    word_pos = (size_t)(off / 64);
    overflowed = false;
    if (n) {
      buffer = word_pos < nwords ? get(word_pos) & ((1ull << n) - 1) : 0;
      bits = n;
    } else {
      buffer = 0;
      bits = 0;
    }
  }
  uint64_t wtell() const { return (uint64_t)word_pos * 64 + bits; }
  uint64_t rtell() const { return (uint64_t)word_pos * 64 - bits; }
  void skip(uint64_t n) { rseek(rtell() + n); }
  // pad_impl (core.rs:460-478)
  // pad_impl (core.rs:460-478): the count saturates, and words past the end
  // are skipped at once
  void pad(uint64_t n)
  {
    uint64_t b = n > UINT64_MAX - bits ? UINT64_MAX : (uint64_t)bits + n;
    while (b >= 64) {
      if (word_pos >= nwords) {
        const size_t words = (size_t)(b / 64);
        word_pos = words > SIZE_MAX - word_pos ? SIZE_MAX : word_pos + words;
        buffer = 0;
        overflowed = true;
        b %= 64;
        break;
      }
      store(buffer);
      buffer = 0;
      b -= 64;
    }
    bits = (uint32_t)b;
  }
  uint32_t align()
  {
    uint32_t b = bits;
    if (b) skip(b);
    return b;
  }
  uint32_t flush()
  {
    uint32_t p = (64 - bits) % 64;
    if (p) pad(p);
    return p;
  }
  void rewind()
  {
    word_pos = 0;
    buffer = 0;
    bits = 0;
    overflowed = false;
    overread = false;
    seek_word = 0;
  }
  bool overread_now() const { return overread || loaded_since_seek(); }
  void reset_overread()
  {
    overread = false;
    seek_word = bits > 0 ? (word_pos ? word_pos - 1 : 0) : word_pos;
  }
};

namespace {

bitstream *new_stream(uint64_t *words, size_t nwords, bool owned, bool device)
{
  bitstream *s = (bitstream *)calloc(1, sizeof(bitstream));
  if (!s) return nullptr;
  s->words = words;
  s->nwords = nwords;
  s->owned = owned;
  s->device = device;
  return s;
}

void drop_index(bitstream *s)
{
  if (s->index) (void)hipFree(s->index);
  s->index = nullptr;
  s->index_n = s->index_cap = 0;
  s->index_bits = 0;
}

}  // namespace

extern "C" {

// stream_open (zfp-rs-ffi/src/bitstream.rs:168-198): a null buffer allocates,
// an aligned one is borrowed, an unaligned host buffer is copied.
bitstream *stream_open(void *buffer, size_t bytes)
{
  if (!buffer) {
    size_t n = (bytes + 7) / 8;
    uint64_t *w = (uint64_t *)calloc(n ? n : 1, 8);
    if (!w) return nullptr;
    return new_stream(w, n, true, false);
  }
  const bool dev = is_device_ptr(buffer);
  if (((uintptr_t)buffer % 8) != 0) {
    if (dev) return nullptr;
    size_t n = (bytes + 7) / 8;
    uint64_t *w = (uint64_t *)calloc(n ? n : 1, 8);
    if (!w) return nullptr;
    memcpy(w, buffer, bytes);
    return new_stream(w, n, true, false);
  }
  return new_stream((uint64_t *)buffer, bytes / 8, false, dev);
}

void stream_close(bitstream *s)
{
  if (!s) return;
  drop_index(s);
  if (s->owned) free(s->words);
  free(s);
}

// deep copy of the committed words (C stream_clone)
bitstream *stream_clone(const bitstream *s)
{
  if (!s) return nullptr;
  size_t n = std::min(s->word_pos, s->nwords);
  uint64_t *w = (uint64_t *)calloc(n ? n : 1, 8);
  if (!w) return nullptr;
  if (n) {
    if (s->device) (void)hipMemcpy(w, s->words, n * 8, hipMemcpyDeviceToHost);
    else memcpy(w, s->words, n * 8);
  }
  return new_stream(w, n, true, false);
}

bitstream_count stream_alignment(void) { return 8; }
void *stream_data(const bitstream *s) { return s ? (void *)s->words : nullptr; }
size_t stream_size(const bitstream *s) { return s ? s->word_pos * 8 : 0; }
size_t stream_capacity(const bitstream *s) { return s ? s->nwords * 8 : 0; }
size_t stream_stride_block(const bitstream *) { return 8; }
ptrdiff_t stream_stride_delta(const bitstream *) { return -8; }
uint stream_read_bit(bitstream *s) { return s ? s->read_bit() : 0; }
uint stream_write_bit(bitstream *s, uint bit)
{
  if (!s) return 0;
  s->write_bit(bit & 1u);
  return bit;
}
uint64 stream_read_bits(bitstream *s, bitstream_count n)
{
  return s ? s->read_bits(n > 64 ? 64u : (uint32_t)n) : 0;
}
uint64 stream_write_bits(bitstream *s, uint64 value, bitstream_count n)
{
  return s ? s->write_bits(value, n > 64 ? 64u : (uint32_t)n) : 0;
}
bitstream_offset stream_rtell(const bitstream *s) { return s ? s->rtell() : 0; }
bitstream_offset stream_wtell(const bitstream *s) { return s ? s->wtell() : 0; }
void stream_rewind(bitstream *s)
{
  if (s) s->rewind();
}
void stream_rseek(bitstream *s, bitstream_offset o)
{
  if (s) s->rseek(o);
}
void stream_wseek(bitstream *s, bitstream_offset o)
{
  if (s) s->wseek(o);
}
void stream_skip(bitstream *s, bitstream_size n)
{
  if (s) s->skip(n);
}
void stream_pad(bitstream *s, bitstream_size n)
{
  if (s) s->pad(n);
}
bitstream_count stream_align(bitstream *s) { return s ? s->align() : 0; }
bitstream_count stream_flush(bitstream *s) { return s ? s->flush() : 0; }

// copy_impl (core.rs:490-522)
void stream_copy(bitstream *dst, bitstream *src, bitstream_size n)
{
  if (!dst || !src) return;
  uint64_t rem = n;
  while (rem > 64) {
    bool src_done = src->word_pos >= src->nwords && src->buffer == 0;
    bool dst_done = dst->word_pos >= dst->nwords;
    if (src_done && dst_done) {
      uint64_t skipped = (rem - 1) / 64;
      src->word_pos += (size_t)skipped;
      dst->word_pos += (size_t)skipped;
      dst->buffer = 0;
      dst->overflowed = true;
      rem -= skipped * 64;
      break;
    }
    dst->write_bits(src->read_bits(64), 64);
    rem -= 64;
  }
  if (rem > 0) dst->write_bits(src->read_bits((uint32_t)rem), (uint32_t)rem);
}

}  // extern "C"

// ===================================================================== config

namespace {

const uint32_t kMinBits = ZFP_MIN_BITS, kMaxBits = ZFP_MAX_BITS, kMaxPrec = ZFP_MAX_PREC;
const int32_t kMinExp = ZFP_MIN_EXP;

uint32_t header_bits(zfp_type t) { return t == zfp_type_float ? 9 : t == zfp_type_double ? 12 : 0; }
uint32_t type_prec(zfp_type t) { return (t == zfp_type_int32 || t == zfp_type_float) ? 32 : 64; }
size_t type_size(zfp_type t)
{
  return (t == zfp_type_int32 || t == zfp_type_float) ? 4 : (t == zfp_type_int64 || t == zfp_type_double) ? 8 : 0;
}
bool valid_type(zfp_type t) { return t >= zfp_type_int32 && t <= zfp_type_double; }

bool valid_params(uint32_t minbits, uint32_t maxbits, uint32_t maxprec)
{
  return minbits <= maxbits && maxprec > 0 && maxprec <= 64;
}

void set_params(zfp_stream *z, uint32_t a, uint32_t b, uint32_t c, int32_t d)
{
  z->minbits = a;
  z->maxbits = b;
  z->maxprec = c;
  z->minexp = d;
}

// ZfpConfig::mode (config.rs:469-527)
zfp_mode mode_of(uint32_t minb, uint32_t maxb, uint32_t maxp, int32_t mine)
{
  if (!valid_params(minb, maxb, maxp)) return zfp_mode_null;
  if (minb == kMinBits && maxb == kMaxBits && maxp == kMaxPrec && mine == kMinExp) return zfp_mode_expert;
  if (minb == maxb && maxb >= 1 && maxb <= kMaxBits && maxp >= kMaxPrec && mine == kMinExp)
    return zfp_mode_fixed_rate;
  if (minb <= kMinBits && maxb >= kMaxBits && maxp >= 1 && mine == kMinExp) return zfp_mode_fixed_precision;
  if (minb <= kMinBits && maxb >= kMaxBits && maxp >= kMaxPrec && mine >= kMinExp)
    return zfp_mode_fixed_accuracy;
  if (minb <= kMinBits && maxb >= kMaxBits && maxp >= kMaxPrec && mine < kMinExp) return zfp_mode_reversible;
  return zfp_mode_expert;
}
zfp_mode mode_of(const zfp_stream *z) { return mode_of(z->minbits, z->maxbits, z->maxprec, z->minexp); }

// encode_expert_mode (config.rs:178-198)
uint64_t expert_mode_bits(uint32_t minb, uint32_t maxb, uint32_t maxp, int32_t mine)
{
  uint64_t a = std::min(std::max(minb, 1u), 0x8000u) - 1;
  uint64_t b = std::min(std::max(maxb, 1u), 0x8000u) - 1;
  uint64_t c = std::min(std::max(maxp, 1u), 0x80u) - 1;
  int64_t e = (int64_t)mine + 16495;
  if (e < 0) e = 0;
  if (e > 0x7fff) e = 0x7fff;
  uint64_t m = 0;
  m = (m << 15) | (uint64_t)e;
  m = (m << 7) | c;
  m = (m << 15) | b;
  m = (m << 15) | a;
  m = (m << 12) | 0xfffu;
  return m;
}

// mode_bits (config.rs:570-600)
uint64_t mode_bits(uint32_t minb, uint32_t maxb, uint32_t maxp, int32_t mine)
{
  switch (mode_of(minb, maxb, maxp, mine)) {
  case zfp_mode_fixed_rate:
    if (maxb <= 2048) return maxb - 1;
    break;
  case zfp_mode_fixed_precision:
    if (maxp <= 128) return (uint64_t)(maxp - 1) + 2048;
    break;
  case zfp_mode_reversible: return 2048 + 128;
  case zfp_mode_fixed_accuracy:
    if (mine <= 843) return (uint64_t)(mine - kMinExp) + (2048 + 128 + 1);
    break;
  default: break;
  }
  return expert_mode_bits(minb, maxb, maxp, mine);
}

// from_mode_bits (config.rs:215-247); false when the parameters are invalid
bool params_from_mode(uint64_t m, uint32_t &a, uint32_t &b, uint32_t &c, int32_t &d)
{
  if (m <= ZFP_MODE_SHORT_MAX) {
    if (m < 2048) { a = b = (uint32_t)m + 1; c = kMaxPrec; d = kMinExp; }
    else if (m < 2048 + 128) { a = kMinBits; b = kMaxBits; c = (uint32_t)m + 1 - 2048; d = kMinExp; }
    else if (m == 2048 + 128) { a = kMinBits; b = kMaxBits; c = kMaxPrec; d = kMinExp - 1; }
    else { a = kMinBits; b = kMaxBits; c = kMaxPrec; d = (int32_t)m + kMinExp - (2048 + 128 + 1); }
  } else {
    uint64_t x = m >> 12;
    a = (uint32_t)(x & 0x7fff) + 1; x >>= 15;
    b = (uint32_t)(x & 0x7fff) + 1; x >>= 15;
    c = (uint32_t)(x & 0x7f) + 1; x >>= 7;
    d = (int32_t)(x & 0x7fff) - 16495;
  }
  return valid_params(a, b, c);
}

// rate_bits (config.rs:85-102); 0 on failure
bool rate_bits(double rate, zfp_type t, uint32_t dims, bool align, uint32_t &out)
{
  double values = (double)(1u << (2 * dims));
  double r = floor(values * rate + 0.5);
  if (!(r >= 0.0 && r <= 4294967295.0)) return false;
  uint32_t b = (uint32_t)r;
  b = std::max(b, header_bits(t));
  if (align) {
    uint64_t x = ((uint64_t)b + 63) / 64 * 64;
    if (x > 0xffffffffull) return false;
    b = (uint32_t)x;
  }
  out = b;
  return true;
}

// block_bits (config.rs:670-686)
uint32_t block_bits(const zx::Cfg &c, zfp_type t, int dims)
{
  uint32_t values = 1u << (2 * dims);
  uint32_t header;
  if (c.minexp < kMinExp)
    header = t == zfp_type_int32 ? 5 : t == zfp_type_int64 ? 6 : t == zfp_type_float ? 15 : 19;
  else
    header = header_bits(t);
  uint64_t most = (uint64_t)header + values - 1 + (uint64_t)values * std::min(c.maxprec, type_prec(t));
  most = std::min<uint64_t>(most, c.maxbits);
  most = std::max<uint64_t>(most, header);
  most = std::max<uint64_t>(most, c.minbits);
  return (uint32_t)most;
}

// Library-owned execution parameters; they begin as zfp_exec_params_omp does.
struct ExecParams {
  uint threads;
  uint chunk_size;
  int32_t rounding;  // zx::ROUND_*
  int32_t tight;
  hipStream_t hip;
  // kernel timing (zfp_stream_set_timing): events around the main kernel
  int32_t timing;
  hipEvent_t ev[4];
  float kernel_ms;   // the block-coding kernel of the last call
  float call_ms;     // every kernel of the last call
};

// Records timing events on `st` when the stream asked for timing.
struct Timer {
  ExecParams *p;
  hipStream_t st;
  bool on;
  Timer(ExecParams *pp, hipStream_t s) : p(pp), st(s), on(pp && pp->timing)
  {
    if (on && !p->ev[0]) {
      for (int i = 0; i < 4; i++)
        if (hipEventCreate(&p->ev[i]) != hipSuccess) { (void)hipGetLastError(); on = false; }
    }
  }
  void mark(int i)
  {
    if (on) (void)hipEventRecord(p->ev[i], st);
  }
  // after the stream was synchronized
  void collect()
  {
    if (!on) return;
    float a = 0, b = 0;
    if (hipEventElapsedTime(&a, p->ev[1], p->ev[2]) != hipSuccess) { (void)hipGetLastError(); a = -1; }
    if (hipEventElapsedTime(&b, p->ev[0], p->ev[3]) != hipSuccess) { (void)hipGetLastError(); b = -1; }
    p->kernel_ms = a;
    p->call_ms = b;
  }
};

ExecParams *params(zfp_stream *z, bool create)
{
  if (!z->exec.params && create) z->exec.params = calloc(1, sizeof(ExecParams));
  return (ExecParams *)z->exec.params;
}
const ExecParams *params(const zfp_stream *z) { return (const ExecParams *)z->exec.params; }

zx::Cfg cfg_of(const zfp_stream *z)
{
  const ExecParams *p = params(z);
  return zx::Cfg{z->minbits, z->maxbits, z->maxprec, z->minexp, p ? p->rounding : 0,
                 p ? p->tight : 0};
}

hipStream_t hip_stream_of(const zfp_stream *z)
{
  const ExecParams *p = params(z);
  return p ? p->hip : nullptr;
}

// ===================================================================== fields

int dimensionality(const zfp_field *f)
{
  return f->ny == 0 ? 1 : f->nz == 0 ? 2 : f->nw == 0 ? 3 : 4;
}

struct Plan {
  zfp_type type;
  int dims;
  int64_t n[4];
  int64_t s[4];
  int64_t imin, imax;  // element span relative to element [0,0,0,0]
  size_t esz;
  uint64_t nblocks;    // saturates at UINT64_MAX, as the reference's num_blocks
  uint64_t nelems;     // saturates at UINT64_MAX, as num_elements
  bool fits;           // logical_shape_fits (field.rs:566-569)
  bool may_alias;      // strides_may_alias (field_plan.rs:92-118)
  uint32_t nb[3];
};

// Whether two index tuples of the field may address one element: the active
// axes of two or more values, sorted by |stride|, must each step past the
// reach of the smaller ones (field_plan.rs:92-118; true means "cannot be
// ruled out").
bool strides_may_alias(const Plan &p)
{
  uint64_t st[4], dm[4];
  int k = 0;
  for (int a = 0; a < p.dims; a++) {
    if (p.n[a] >= 2) {
      st[k] = p.s[a] < 0 ? 0 - (uint64_t)p.s[a] : (uint64_t)p.s[a];
      dm[k] = (uint64_t)p.n[a];
      k++;
    }
  }
  for (int i = 1; i < k; i++)  // insertion sort by (stride, extent)
    for (int j = i; j > 0 && (st[j] < st[j - 1] || (st[j] == st[j - 1] && dm[j] < dm[j - 1])); j--) {
      std::swap(st[j], st[j - 1]);
      std::swap(dm[j], dm[j - 1]);
    }
  uint64_t reach = 1;
  for (int i = 0; i < k; i++) {
    if (st[i] < reach) return true;
    const unsigned __int128 r = (unsigned __int128)st[i] * (dm[i] - 1) + reach;
    reach = r > UINT64_MAX ? UINT64_MAX : (uint64_t)r;
  }
  return false;
}

// The field as the reference's FFI takes it (field_to_rust, zfp-rs-ffi/src/
// field.rs:43-86): checked_index_span with natural strides for zero entries
// (field.rs:621-647) and a span of at most isize::MAX bytes; false where that
// conversion fails.  The block grid (field_plan.rs:41-77), the element and
// block counts (saturating), logical_shape_fits and strides_may_alias come
// with it.
bool make_plan(const zfp_field *f, Plan &p)
{
  if (!f || !valid_type(f->type)) return false;
  p.type = f->type;
  p.esz = type_size(f->type);
  p.dims = dimensionality(f);
  const size_t dims[4] = {f->nx, f->ny, f->nz, f->nw};
  const ptrdiff_t strides[4] = {f->sx, f->sy, f->sz, f->sw};
  __int128 imin = 0, imax = 0;
  unsigned __int128 below = 1, elems = 1, blocks = 1;  // saturate at 2^64
  const unsigned __int128 kSat = (unsigned __int128)UINT64_MAX + 1;
  for (int a = 0; a < 4; a++) {
    p.n[a] = 1;
    p.s[a] = 0;
  }
  for (int a = 0; a < p.dims; a++) {
    const uint64_t d = (uint64_t)dims[a];
    const unsigned __int128 natural = below;
    below = below * d < kSat ? below * d : kSat;
    elems = elems * d < kSat ? elems * d : kSat;
    blocks = blocks * ((d + 3) / 4) < kSat ? blocks * ((d + 3) / 4) : kSat;
    int64_t st = strides[a];
    if (d == 0) {
      p.n[a] = 0;
      p.s[a] = st != 0 ? st : (natural <= (unsigned __int128)INT64_MAX ? (int64_t)natural : INT64_MAX);
      continue;
    }
    if (d > (uint64_t)INT64_MAX) return false;
    if (st == 0) {
      if (natural > (unsigned __int128)INT64_MAX) return false;
      st = (int64_t)natural;
    }
    p.n[a] = (int64_t)d;
    p.s[a] = st;
    const __int128 ext = (__int128)st * (__int128)(d - 1);
    if (ext > INT64_MAX || ext < INT64_MIN) return false;
    if (ext < 0) imin += ext;
    else imax += ext;
    if (imin < INT64_MIN || imax > INT64_MAX) return false;
  }
  const __int128 span = imax - imin + 1;
  if (span > INT64_MAX || span * (__int128)p.esz > INT64_MAX) return false;
  p.imin = (int64_t)imin;
  p.imax = (int64_t)imax;
  p.nelems = elems >= kSat ? UINT64_MAX : (uint64_t)elems;
  p.nblocks = blocks >= kSat ? UINT64_MAX : (uint64_t)blocks;
  p.fits = elems <= (unsigned __int128)INT64_MAX && blocks < kSat;
  for (int a = 0; a < 3; a++) p.nb[a] = (uint32_t)((p.n[a] + 3) / 4);
  p.may_alias = strides_may_alias(p);
  return true;
}

// Blocks the codec's kernels can index: their grid arithmetic is 32-bit.
constexpr uint64_t kMaxCodecBlocks = 0xffffffffull;

// make_plan for zfp_compress / zfp_decompress and the batched calls: the
// reference also rejects a shape that does not fit (validate_layout,
// field.rs:84-106, ShapeTooLarge), and this library a field of more than
// kMaxCodecBlocks blocks.
bool codec_plan(const zfp_field *f, Plan &p)
{
  return make_plan(f, p) && p.fits && p.nblocks <= kMaxCodecBlocks;
}

bool contiguous(const Plan &p)
{
  int64_t nat = 1;
  for (int a = 0; a < p.dims; a++) {
    if (p.s[a] != nat) return false;
    nat *= p.n[a];
  }
  return true;
}

}  // namespace

// ===================================================================== codec

namespace {

typedef hipError_t (*EncFn)(const zx::EncodeArgs &, hipStream_t);
typedef hipError_t (*DecFn)(const zx::DecodeArgs &, hipStream_t);

#define ZX_ROW(T) {zx_enc_##T##_1_0, zx_enc_##T##_1_1, zx_enc_##T##_2_0, zx_enc_##T##_2_1, \
                   zx_enc_##T##_3_0, zx_enc_##T##_3_1, zx_enc_##T##_4_0, zx_enc_##T##_4_1}
const EncFn kEnc[4][8] = {ZX_ROW(i32), ZX_ROW(i64), ZX_ROW(f32), ZX_ROW(f64)};
#undef ZX_ROW
#define ZX_ROW(T) {zx_dec_##T##_1_0, zx_dec_##T##_1_1, zx_dec_##T##_2_0, zx_dec_##T##_2_1, \
                   zx_dec_##T##_3_0, zx_dec_##T##_3_1, zx_dec_##T##_4_0, zx_dec_##T##_4_1}
const DecFn kDec[4][8] = {ZX_ROW(i32), ZX_ROW(i64), ZX_ROW(f32), ZX_ROW(f64)};
#undef ZX_ROW

typedef uint64_t (*WalkFn)(const uint64_t *, uint64_t, uint64_t, uint64_t, const zx::Cfg &, uint64_t *);
#define ZX_ROW(T) {zx_walk_##T##_1_0, zx_walk_##T##_1_1, zx_walk_##T##_2_0, zx_walk_##T##_2_1, \
                   zx_walk_##T##_3_0, zx_walk_##T##_3_1, zx_walk_##T##_4_0, zx_walk_##T##_4_1}
const WalkFn kWalk[4][8] = {ZX_ROW(i32), ZX_ROW(i64), ZX_ROW(f32), ZX_ROW(f64)};
#undef ZX_ROW

hipError_t launch_encode(zfp_type t, int dims, bool rev, const zx::EncodeArgs &a, hipStream_t s)
{
  if (t < zfp_type_int32 || t > zfp_type_double || dims < 1 || dims > 4) return hipErrorInvalidValue;
  return kEnc[t - zfp_type_int32][2 * (dims - 1) + (rev ? 1 : 0)](a, s);
}

hipError_t launch_decode(zfp_type t, int dims, bool rev, const zx::DecodeArgs &a, hipStream_t s)
{
  if (t < zfp_type_int32 || t > zfp_type_double || dims < 1 || dims > 4) return hipErrorInvalidValue;
  return kDec[t - zfp_type_int32][2 * (dims - 1) + (rev ? 1 : 0)](a, s);
}

uint64_t cfg_key(const zx::Cfg &c, const Plan &p)
{
  uint64_t h = 1469598103934665603ull;
  const uint64_t v[] = {c.minbits, c.maxbits, c.maxprec, (uint64_t)(uint32_t)c.minexp,
                        (uint64_t)c.rounding, (uint64_t)c.tight, (uint64_t)p.type,
                        (uint64_t)p.dims, p.nblocks};
  for (uint64_t x : v) h = (h ^ x) * 1099511628211ull;
  return h;
}

zx::FieldDesc field_desc(const Plan &p, void *origin)
{
  zx::FieldDesc f;
  f.data = origin;
  for (int a = 0; a < 4; a++) {
    f.n[a] = p.n[a];
    f.s[a] = p.s[a];
  }
  zx::set_block_grid(f);
  return f;
}

bool vec_ok(const Plan &p, const void *origin)
{
  if (p.s[0] != 1) return false;
  if (p.dims >= 2 && (p.s[1] % 4) != 0) return false;
  if (p.dims >= 3 && (p.s[2] % 4) != 0) return false;
  if (p.dims >= 4 && (p.s[3] % 4) != 0) return false;
  return ((uintptr_t)origin % 16) == 0;
}

bool fixed_layout(const zx::Cfg &c, const Plan &p)
{
  return c.minbits == c.maxbits && block_bits(c, p.type, p.dims) == c.maxbits && c.maxbits > 0;
}

#define ZCHECK(x)                                        \
  do {                                                   \
    const hipError_t zcheck_e = (x);                     \
    if (zcheck_e != hipSuccess) {                        \
      note_failure(__LINE__, hipGetErrorName(zcheck_e)); \
      (void)hipGetLastError();                           \
      return false;                                      \
    }                                                    \
  } while (0)

// Device pointer to element [0,0,0] of the field: the caller's own, or the
// span staged into the workspace (copied in when `copy_in`).
bool stage_field(Workspace &W, const zfp_field *f, const Plan &p, bool copy_in, hipStream_t st,
                 void *&dorigin, bool &staged)
{
  staged = !is_device_ptr(f->data);
  if (!staged) {
    dorigin = f->data;
    return true;
  }
  const size_t span = (size_t)(p.imax - p.imin + 1) * p.esz;
  if (!Workspace::grow(W.field, W.field_cap, span)) return false;
  const char *lo = (const char *)f->data + p.imin * (int64_t)p.esz;
  if (copy_in) ZCHECK(hipMemcpyAsync(W.field, lo, span, hipMemcpyHostToDevice, st));
  dorigin = (char *)W.field - p.imin * (int64_t)p.esz;
  return true;
}

// Device scratch of a variable-rate compress supplied by the caller in place
// of the workspace's (zfp_hip_compress_noalloc): per-block lengths, offsets
// (blocks + 1; the one-pass state and edges use the same words) and scan
// tiles, for at most `blocks` blocks.
struct Scratch {
  uint32_t *lens;
  uint64_t *offs;
  uint64_t *tiles;
  uint64_t blocks;
  uint64_t offs_words;
};

// Keeps the block lengths `lens` (device) as the stream's offset index; with
// `no_alloc`, only where the stream's index buffer already has room (else the
// old index is marked stale).  A failed allocation leaves no index.
void keep_offset_index(bitstream *s, const uint32_t *lens, const Plan &p, const zx::Cfg &c, uint64_t start,
                       bool no_alloc, hipStream_t st)
{
  if (no_alloc && s->index_cap < p.nblocks) {
    s->index_n = 0;
    return;
  }
  if (s->index_cap < p.nblocks) {
    drop_index(s);
    if (hipMalloc((void **)&s->index, p.nblocks * 4) != hipSuccess) {
      (void)hipGetLastError();
      s->index = nullptr;
      return;
    }
    s->index_cap = p.nblocks;
  }
  if (hipMemcpyAsync(s->index, lens, p.nblocks * 4, hipMemcpyDeviceToDevice, st) != hipSuccess) {
    (void)hipGetLastError();
    s->index_n = 0;
    return;
  }
  s->index_n = p.nblocks;
  s->index_start = start;
  s->index_key = cfg_key(c, p);
}

// Encode every block of the field at the cursor of `s`; leaves the cursor
// where compress + finish would (compress.rs:22-53).  Returns false on a
// device failure (nothing consistent was written).  `end_bit` receives the
// bit after the last block; `keep_index` keeps a variable-rate stream's
// offset index beside it.  With `sc`, field and stream words must be device
// memory and nothing is allocated: the scratch is `sc`'s, the index is kept
// only where the stream's index buffer already holds it (else the old index
// is marked stale), and the end offset is read back into host stack memory.
bool gpu_compress(bitstream *s, const zfp_field *f, const Plan &p, const zx::Cfg &c, hipStream_t st,
                  ExecParams *xp, bool keep_index = true, uint64_t *end_bit = nullptr,
                  const Scratch *sc = nullptr)
{
  Workspace &W = ws();
  std::lock_guard<std::mutex> lock(W.mu);
  if (sc && (!s->device || !is_device_ptr(f->data) || p.nblocks > sc->blocks)) return false;
  const uint64_t start = s->wtell();
  const uint64_t prefix = s->bits ? s->buffer : 0;
  const uint64_t w0 = start / 64;
  const bool rev = c.minexp < kMinExp;
  void *dorigin = nullptr;
  bool staged = false;
  Timer tm(xp, st);
  if (!stage_field(W, f, p, true, st, dorigin, staged)) return false;
  tm.mark(0);
  zx::EncodeArgs a{};
  a.field = field_desc(p, dorigin);
  a.cfg = c;
  a.vec = vec_ok(p, dorigin);
  a.offsets = nullptr;
  a.lengths = nullptr;
  a.start = start - 64 * w0;
  const bool fixed = fixed_layout(c, p);
  uint64_t end_rel;
  uint64_t *offs = nullptr;
  uint32_t *lens = nullptr;
  // variable-rate 1-3-D fields take the one-pass kernel
  const int size = 1 << (2 * p.dims);
  const int tbits = p.type == zfp_type_int32 || p.type == zfp_type_float ? 32 : 64;
  if (!fixed && p.dims <= 3) {
    // workgroup size and staging from the stream's last mean block length
    const zx::OnePlan op = zx::op_plan(zx::block_bound(c, tbits, size), s->mean_bits, 0, 0);
    const uint64_t groups = (p.nblocks + op.T - 1) / op.T;
    uint64_t *state, *edges;
    if (sc) {
      if (2 + 3 * groups > sc->offs_words) return false;
      lens = sc->lens;
      state = sc->offs;
      edges = sc->offs + 2 + groups;
    } else {
      if (!Workspace::grow(W.lens, W.lens_cap, p.nblocks * 4)) return false;
      if (!Workspace::grow(W.tiles, W.tiles_cap, (2 + groups) * 8)) return false;
      if (!Workspace::grow(W.offs, W.offs_cap, 2 * groups * 8)) return false;
      if (!W.one) ZCHECK(hipHostMalloc(&W.one, 32, hipHostMallocDefault));
      lens = (uint32_t *)W.lens;
      state = (uint64_t *)W.tiles;
      edges = (uint64_t *)W.offs;
    }
    const uint64_t cap_words = s->nwords > w0 ? s->nwords - w0 : 0;
    uint64_t *dwords;
    uint64_t nwords = cap_words;
    if (s->device) {
      dwords = s->words + w0;
    } else {
      // staging for the words the stream can hold, at most every block's largest code
      const uint64_t bound = a.start + p.nblocks * (uint64_t)zx::block_bound(c, tbits, size);
      nwords = std::min<uint64_t>(cap_words, (bound + 63) / 64);
      if (!Workspace::grow(W.words, W.words_cap, std::max<uint64_t>(nwords, 1) * 8)) return false;
      dwords = (uint64_t *)W.words;
    }
    ZCHECK(zx_launch_fill(state, 2 + groups, 0, st));
    a.fixed = 0;
    a.words = dwords;
    a.nwords = nwords;
    a.lengths = lens;
    a.op_state = state;
    a.op_edges = edges;
    a.op_T = op.T;
    a.vcap = op.cap;
    tm.mark(1);
    ZCHECK(launch_encode(p.type, p.dims, rev, a, st));
    ZCHECK(zx_launch_op_merge(dwords, nwords, state, edges, groups, a.start, prefix, st));
    tm.mark(2);
    uint64_t end_host[2] = {0, 0};
    uint64_t *rb = sc ? end_host : (uint64_t *)W.one;
    ZCHECK(hipMemcpyAsync(rb, state + 2 + groups - 1, 8, hipMemcpyDeviceToHost, st));
    ZCHECK(hipMemcpyAsync(rb + 1, state + 1, 8, hipMemcpyDeviceToHost, st));
    if (keep_index) keep_offset_index(s, lens, p, c, start, sc != nullptr, st);
    tm.mark(3);
    ZCHECK(hipStreamSynchronize(st));
    tm.collect();
    if (rb[1]) return false;  // a look-back gave up
    end_rel = rb[0] & ((1ull << 62) - 1);
    s->mean_bits = (uint32_t)std::min<uint64_t>((end_rel - a.start + p.nblocks - 1) / p.nblocks, 1u << 20);
    if (keep_index) s->index_bits = end_rel - a.start;
    const uint64_t w1 = (64 * w0 + end_rel + 63) / 64;  // words through the flush
    const uint64_t live = std::min<uint64_t>(w1 - w0, nwords);
    if (!s->device && live) {
      ZCHECK(hipMemcpyAsync(s->words + w0, dwords, live * 8, hipMemcpyDeviceToHost, st));
      ZCHECK(hipStreamSynchronize(st));
    }
    if (end_bit) *end_bit = 64 * w0 + end_rel;
    if (w1 > s->nwords) s->overflowed = true;
    s->word_pos = (size_t)w1;
    s->bits = 0;
    s->buffer = 0;
    return true;
  }
  if (fixed) {
    a.fixed = c.maxbits;
    end_rel = a.start + p.nblocks * (uint64_t)c.maxbits;
  } else {
    a.fixed = 0;
    uint64_t *tiles;
    if (sc) {
      lens = sc->lens;
      offs = sc->offs;
      tiles = sc->tiles;
    } else {
      if (!Workspace::grow(W.lens, W.lens_cap, p.nblocks * 4)) return false;
      if (!Workspace::grow(W.offs, W.offs_cap, (p.nblocks + 1) * 8)) return false;
      if (!Workspace::grow(W.tiles, W.tiles_cap, zx_scan_tiles(p.nblocks) * 8)) return false;
      if (!W.one) ZCHECK(hipHostMalloc(&W.one, 32, hipHostMallocDefault));
      lens = (uint32_t *)W.lens;
      offs = (uint64_t *)W.offs;
      tiles = (uint64_t *)W.tiles;
    }
    a.lengths = lens;
    a.words = nullptr;
    a.nwords = 0;
    ZCHECK(launch_encode(p.type, p.dims, rev, a, st));
    ZCHECK(zx_launch_scan(lens, p.nblocks, a.start, offs, tiles, st));
    uint64_t end_host[2] = {0, 0};
    uint64_t *rb = sc ? end_host : (uint64_t *)W.one;
    ZCHECK(hipMemcpyAsync(rb, offs + p.nblocks, 8, hipMemcpyDeviceToHost, st));
    ZCHECK(hipMemcpyAsync(rb + 1, tiles + 1, 8, hipMemcpyDeviceToHost, st));
    ZCHECK(hipStreamSynchronize(st));
    if (rb[1]) return false;  // the scan's look-back gave up
    end_rel = rb[0];
    a.lengths = nullptr;
    a.offsets = offs;
    a.vbits = end_rel - a.start;
  }
  const uint64_t w1 = (64 * w0 + end_rel + 63) / 64;  // words through the flush
  const uint64_t cap_words = s->nwords > w0 ? s->nwords - w0 : 0;
  const uint64_t live = std::min<uint64_t>(w1, s->nwords) > w0 ? std::min<uint64_t>(w1, s->nwords) - w0 : 0;
  uint64_t *dwords;
  if (s->device) {
    dwords = s->words + w0;
  } else {
    if (!Workspace::grow(W.words, W.words_cap, std::max<uint64_t>(live, 1) * 8)) return false;
    dwords = (uint64_t *)W.words;
  }
  a.words = dwords;
  a.nwords = std::min(cap_words, live);
  const bool aligned = a.start == 0 && fixed && (c.maxbits % 64) == 0;
  if (!aligned && live) ZCHECK(zx_launch_fill(dwords, live, prefix, st));
  tm.mark(1);
  if (live) ZCHECK(launch_encode(p.type, p.dims, rev, a, st));
  tm.mark(2);
  if (!s->device && live) {
    ZCHECK(hipMemcpyAsync(s->words + w0, dwords, live * 8, hipMemcpyDeviceToHost, st));
  }
  if (!fixed && keep_index) {
    keep_offset_index(s, lens, p, c, start, sc != nullptr, st);
    s->index_bits = a.vbits;
  }
  tm.mark(3);
  ZCHECK(hipStreamSynchronize(st));
  tm.collect();
  if (end_bit) *end_bit = 64 * w0 + end_rel;
  // the cursor after the last block, then flush (core.rs:536-542)
  if (w1 > s->nwords) s->overflowed = true;
  s->word_pos = (size_t)w1;
  s->bits = 0;
  s->buffer = 0;
  return true;
}

bool grow_host(void *&p, size_t &cap, size_t need)
{
  if (injected_failure()) {
    note_failure(__LINE__, "injected reservation failure");
    return false;
  }
  if (need <= cap) return true;
  const size_t n = std::max(need, cap + cap / 2);
  if (p) (void)hipHostFree(p);
  p = nullptr;
  cap = 0;
  if (hipHostMalloc(&p, n, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    p = nullptr;
    return false;
  }
  cap = n;
  return true;
}

// Blocks per run of the length reader.
constexpr uint64_t kWalkRun = 65536;

// Decode of a variable-rate stream without an offset index, pipelined as the
// reference's index-less decoder is (decompress/pipeline.rs:99-229, a serial
// plane reader feeding parallel reconstruction): the host reads block
// lengths in order, a run of kWalkRun blocks at a time, and each run's
// offsets are copied up and its blocks decoded on the GPU while the host
// reads the next run.  `a` is set up for the whole field; `need` words of the
// stream from word w0 on are readable.  end_rel is the end bit of the last
// block, relative to word w0.
bool walk_decode(Workspace &W, const bitstream *s, const Plan &p, zx::DecodeArgs a, bool rev,
                 uint64_t w0, uint64_t need, hipStream_t st, uint64_t &end_rel)
{
  const uint64_t *hw = s->words + w0;
  if (s->device) {
    if (!grow_host(W.hwords, W.hwords_cap, std::max<uint64_t>(need, 1) * 8)) return false;
    if (need) ZCHECK(hipMemcpyAsync(W.hwords, s->words + w0, need * 8, hipMemcpyDeviceToHost, st));
    ZCHECK(hipStreamSynchronize(st));
    hw = (const uint64_t *)W.hwords;
  }
  if (!grow_host(W.hoffs, W.hoffs_cap, (p.nblocks + 1) * 8)) return false;
  uint64_t *ho = (uint64_t *)W.hoffs;
  uint64_t *doffs = (uint64_t *)W.offs;
  const WalkFn walk = kWalk[p.type - zfp_type_int32][2 * (p.dims - 1) + (rev ? 1 : 0)];
  uint64_t pos = a.start;
  a.offsets = doffs;
  for (uint64_t b0 = 0; b0 < p.nblocks; b0 += kWalkRun) {
    const uint64_t nb = std::min<uint64_t>(kWalkRun, p.nblocks - b0);
    pos = walk(hw, need, pos, nb, a.cfg, ho + b0);
    const bool last = b0 + nb == p.nblocks;
    if (last) ho[p.nblocks] = pos;
    ZCHECK(hipMemcpyAsync(doffs + b0, ho + b0, (nb + (last ? 1 : 0)) * 8, hipMemcpyHostToDevice, st));
    a.first = b0;
    a.count = nb;
    ZCHECK(launch_decode(p.type, p.dims, rev, a, st));
  }
  end_rel = pos;
  return true;
}

// Decode every block of the field from the cursor of `s`; leaves the cursor
// and overread state where decompress + finish would (decompress.rs:27-74).
bool gpu_decompress(bitstream *s, zfp_field *f, const Plan &p, const zx::Cfg &c, hipStream_t st,
                    ExecParams *xp, uint64_t &end_out)
{
  Workspace &W = ws();
  std::lock_guard<std::mutex> lock(W.mu);
  const uint64_t start = s->rtell();
  const uint64_t w0 = start / 64;
  const bool rev = c.minexp < kMinExp;
  s->reset_overread();
  const size_t seek_word0 = s->seek_word;
  void *dorigin = nullptr;
  bool staged = false;
  Timer tm(xp, st);
  if (!stage_field(W, f, p, !contiguous(p), st, dorigin, staged)) return false;
  // A field whose strides may alias is decoded into a dense copy first; the
  // alias scatter then leaves each element the serial decode's last value.
  Plan q = p;
  void *korigin = dorigin;
  if (p.may_alias) {
    int64_t nat = 1;
    for (int a = 0; a < p.dims; a++) {
      q.s[a] = nat;
      nat *= p.n[a];
    }
    q.imin = 0;
    q.imax = nat - 1;
    if (!Workspace::grow(W.dense, W.dense_cap, (size_t)p.nelems * p.esz)) return false;
    if (!Workspace::grow(W.best, W.best_cap, (size_t)(p.imax - p.imin + 1) * 8)) return false;
    korigin = W.dense;
  }
  tm.mark(0);
  zx::DecodeArgs a{};
  a.field = field_desc(q, korigin);
  a.cfg = c;
  a.vec = vec_ok(q, korigin);
  a.start = start - 64 * w0;
  a.offsets = nullptr;
  if (!Workspace::grow(W.tiles, W.tiles_cap, zx_scan_tiles(p.nblocks) * 8 + 8)) return false;
  // the last tiles word collects the highest word any block loaded
  uint64_t *maxword = (uint64_t *)W.tiles + zx_scan_tiles(p.nblocks);
  a.maxword = maxword;
  const bool fixed = fixed_layout(c, p);
  // device words from w0 on
  const uint64_t avail = s->nwords > w0 ? s->nwords - w0 : 0;
  uint64_t need = avail;
  // A fixed-rate decode loads no word past the one holding its last bit, so
  // when the stream holds that word nothing can be overread and the highest
  // loaded word is not needed.
  bool whole = false;
  if (fixed) {
    const uint64_t end_rel = a.start + p.nblocks * (uint64_t)c.maxbits;
    need = std::min<uint64_t>(avail, (end_rel + 63) / 64);
    whole = avail >= (end_rel + 63) / 64;
    if (whole) a.maxword = nullptr;
  }
  const uint64_t *dwords;
  if (s->device) {
    dwords = s->words + w0;
  } else {
    if (!Workspace::grow(W.words, W.words_cap, std::max<uint64_t>(need, 1) * 8)) return false;
    if (need) ZCHECK(hipMemcpyAsync(W.words, s->words + w0, need * 8, hipMemcpyHostToDevice, st));
    dwords = (const uint64_t *)W.words;
  }
  a.words = dwords;
  a.nwords = need;
  uint64_t end_rel = 0;
  bool walked = false;  // no offset index: the host reads the block lengths
  if (!W.one) ZCHECK(hipHostMalloc(&W.one, 32, hipHostMallocDefault));
  ((uint64_t *)W.one)[2] = 0;
  if (fixed) {
    a.fixed = c.maxbits;
    end_rel = a.start + p.nblocks * (uint64_t)c.maxbits;
  } else {
    a.fixed = 0;
    if (!Workspace::grow(W.offs, W.offs_cap, (p.nblocks + 1) * 8)) return false;
    if (s->index && s->index_n == p.nblocks && s->index_start == start &&
        (s->index_key == 0 || s->index_key == cfg_key(c, p))) {
      uint64_t *offs = (uint64_t *)W.offs;
      ZCHECK(zx_launch_scan(s->index, p.nblocks, a.start, offs, (uint64_t *)W.tiles, st));
      ZCHECK(hipMemcpyAsync(W.one, offs + p.nblocks, 8, hipMemcpyDeviceToHost, st));
      ZCHECK(hipMemcpyAsync((uint64_t *)W.one + 2, (uint64_t *)W.tiles + 1, 8, hipMemcpyDeviceToHost, st));
      a.offsets = offs;
      a.vbits = s->index_bits;
    } else {
      walked = true;
    }
  }
  if (!whole) ZCHECK(zx_launch_fill(maxword, 1, 0, st));
  tm.mark(1);
  if (walked) {
    if (!walk_decode(W, s, p, a, rev, w0, need, st, end_rel)) return false;
  } else {
    ZCHECK(launch_decode(p.type, p.dims, rev, a, st));
  }
  if (p.may_alias) {
    zx::AliasArgs aa{};
    aa.dense = W.dense;
    aa.lowest = (char *)dorigin + p.imin * (int64_t)p.esz;
    for (int k = 0; k < 4; k++) {
      aa.n[k] = p.n[k];
      aa.s[k] = p.s[k];
    }
    aa.imin = p.imin;
    aa.count = p.nelems;
    aa.span = (uint64_t)(p.imax - p.imin + 1);
    aa.dims = p.dims;
    aa.esz = (uint32_t)p.esz;
    ZCHECK(zx_launch_alias_scatter(aa, (unsigned long long *)W.best, st));
  }
  tm.mark(2);
  if (!whole) ZCHECK(hipMemcpyAsync((uint64_t *)W.one + 1, maxword, 8, hipMemcpyDeviceToHost, st));
  if (staged) {
    const size_t span = (size_t)(p.imax - p.imin + 1) * p.esz;
    char *lo = (char *)f->data + p.imin * (int64_t)p.esz;
    ZCHECK(hipMemcpyAsync(lo, W.field, span, hipMemcpyDeviceToHost, st));
  }
  tm.mark(3);
  ZCHECK(hipStreamSynchronize(st));
  tm.collect();
  if (!fixed && ((uint64_t *)W.one)[2]) return false;  // the scan's look-back gave up
  if (!fixed && !walked) end_rel = *(uint64_t *)W.one;
  const uint64_t end = 64 * w0 + end_rel;
  end_out = end;
  // park + align: the cursor at the word boundary after the last block.  A
  // load past the end (not a skip over missing padding) is a truncation.
  const uint64_t w1 = (end + 63) / 64;
  // one past the highest loaded word
  const uint64_t loaded_end = whole ? w1 : w0 + ((uint64_t *)W.one)[1];
  const bool over = loaded_end > std::max<uint64_t>(s->nwords, seek_word0);
  s->word_pos = (size_t)w1;
  s->bits = 0;
  s->buffer = 0;
  s->seek_word = s->word_pos;
  s->overread = over;
  return true;
}

// ---------------------------------------------------------------- single blocks

// The plan of one block of `dims` axes with `len` values (1..4) and `strides`
// per axis, taken as given (a zero stride repeats a value, as the reference's
// pointer walk does); false for a length outside 1..4 (check_lengths,
// src/codec/block/strided.rs:392-405) or a span that overflows.
bool block_plan(zfp_type t, int dims, const size_t len[4], const ptrdiff_t str[4], Plan &p)
{
  if (!valid_type(t) || dims < 1 || dims > 4) return false;
  p.type = t;
  p.esz = type_size(t);
  p.dims = dims;
  p.imin = p.imax = 0;
  p.nblocks = 1;
  for (int a = 0; a < 4; a++) {
    p.n[a] = 1;
    p.s[a] = 0;
  }
  for (int a = 0; a < dims; a++) {
    if (len[a] < 1 || len[a] > 4) return false;
    const __int128 ext = (__int128)str[a] * (int64_t)(len[a] - 1);
    if (ext > INT64_MAX / 4 || ext < INT64_MIN / 4) return false;
    p.n[a] = (int64_t)len[a];
    p.s[a] = str[a];
    if (ext < 0) p.imin += (int64_t)ext;
    else p.imax += (int64_t)ext;
  }
  for (int a = 0; a < 3; a++) p.nb[a] = 1;
  p.nelems = (uint64_t)(p.n[0] * p.n[1] * p.n[2] * p.n[3]);
  p.fits = true;
  // one thread writes the block's values in the reference's order, so
  // strides that repeat an element leave its last value without the dense
  // copy and scatter of an aliasing field
  p.may_alias = false;
  return true;
}

// One block coded on the GPU at the stream's cursor, as the block-level
// entry points of zfp-rs-ffi do (zfp-rs-ffi/src/block.rs, over
// src/codec/block/strided.rs:420-700): the same coder as a one-block field,
// with the cursor left right after the block (the writer holding the block's
// last partial word, unflushed; the reader unaligned).  Returns the block's
// bits, or 0 for a null block, a length outside 1..4 or a device failure.
size_t block_code(bool encode, zfp_stream *z, const void *block, zfp_type t, int dims,
                  const size_t len[4], const ptrdiff_t str[4])
{
  if (!z || !z->stream || !block) return 0;
  Plan p{};
  if (!block_plan(t, dims, len, str, p)) return 0;
  if (((uintptr_t)((const char *)block + p.imin * (int64_t)p.esz)) % p.esz) return 0;
  zfp_field f{};
  f.type = t;
  f.data = const_cast<void *>(block);
  bitstream *s = z->stream;
  const zx::Cfg c = cfg_of(z);
  uint64_t end = 0;
  if (encode) {
    const uint64_t start = s->wtell();
    const bool was_over = s->overflowed;
    if (!gpu_compress(s, &f, p, c, hip_stream_of(z), params(z, false), false, &end)) return 0;
    // a whole word past the capacity was dropped
    const bool over = was_over || end / 64 > s->nwords;
    s->wseek(end);
    s->overflowed = over;
    return (size_t)(end - start);
  }
  const uint64_t start = s->rtell();
  if (!gpu_decompress(s, &f, p, c, hip_stream_of(z), params(z, false), end)) return 0;
  const bool over = s->overread;
  s->rseek(end);
  s->overread = over;
  return (size_t)(end - start);
}

size_t block_whole(bool encode, zfp_stream *z, const void *b, zfp_type t, int dims)
{
  const size_t len[4] = {4, 4, 4, 4};
  const ptrdiff_t str[4] = {1, 4, 16, 64};
  return block_code(encode, z, b, t, dims, len, str);
}

size_t block_strided(bool encode, zfp_stream *z, const void *b, zfp_type t, int dims, ptrdiff_t sx,
                     ptrdiff_t sy, ptrdiff_t sz, ptrdiff_t sw)
{
  const size_t len[4] = {4, 4, 4, 4};
  const ptrdiff_t str[4] = {sx, sy, sz, sw};
  return block_code(encode, z, b, t, dims, len, str);
}

size_t block_partial(bool encode, zfp_stream *z, const void *b, zfp_type t, int dims, size_t nx,
                     size_t ny, size_t nz, size_t nw, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz,
                     ptrdiff_t sw)
{
  const size_t len[4] = {nx, ny, nz, nw};
  const ptrdiff_t str[4] = {sx, sy, sz, sw};
  return block_code(encode, z, b, t, dims, len, str);
}

}  // namespace

// ===================================================================== batches

namespace {

typedef hipError_t (*BEncFn)(const zx::BatchArgs &, bool, hipStream_t);
typedef hipError_t (*BDecFn)(const zx::BatchArgs &, bool, hipStream_t);

#define ZX_ROW(T) {zx_benc_##T##_1_0, zx_benc_##T##_1_1, zx_benc_##T##_2_0, zx_benc_##T##_2_1, \
                   zx_benc_##T##_3_0, zx_benc_##T##_3_1}
const BEncFn kBEnc[4][6] = {ZX_ROW(i32), ZX_ROW(i64), ZX_ROW(f32), ZX_ROW(f64)};
#undef ZX_ROW
#define ZX_ROW(T) {zx_bdec_##T##_1_0, zx_bdec_##T##_1_1, zx_bdec_##T##_2_0, zx_bdec_##T##_2_1, \
                   zx_bdec_##T##_3_0, zx_bdec_##T##_3_1}
const BDecFn kBDec[4][6] = {ZX_ROW(i32), ZX_ROW(i64), ZX_ROW(f32), ZX_ROW(f64)};
#undef ZX_ROW

// Fixed-rate arrays of at least this many blocks get launches of their own
// (the LDS-staged kernels of the single-array path) inside a batch; smaller
// ones share their class's launch.
constexpr uint64_t kSoloBlocks = 1 << 14;
constexpr int kSoloClass = 1 << 20;

// One array of a batch, on the host side.
struct BatchItem {
  size_t input;      // position in the caller's arrays
  int cls;           // ((type, dims, rev) variant) * 2 + variable placement, or
                     // kSoloClass + variant for an array launched on its own
  Plan plan;
  zx::BatchJob job;
  uint64_t w0;       // first stream word of the array
  uint64_t start;    // cursor bit at the call
  size_t seek_word0; // decode: seek word at the call
};

// Grow-only pinned and device buffers of the batch entry points.
// Side streams that a batch's independent class launches are spread over
// (forked from and joined back into the caller's stream by events), so that
// the many small launches of a mixed batch run beside each other and beside
// the large ones.  Default priority, non-blocking.  Three, with the caller's
// stream four: HIP's default of four hardware queues per process (seven
// measured no faster for planned batches and slower for batch calls).
constexpr int kLanes = 3;
struct Lanes {
  hipStream_t s[kLanes] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t join[kLanes] = {};
  bool ready = false;

  bool init()
  {
    if (ready) return true;
    for (int i = 0; i < kLanes; i++)
      if (hipStreamCreateWithFlags(&s[i], hipStreamNonBlocking) != hipSuccess) return fail();
    if (hipEventCreateWithFlags(&fork, hipEventDisableTiming) != hipSuccess) return fail();
    for (int i = 0; i < kLanes; i++)
      if (hipEventCreateWithFlags(&join[i], hipEventDisableTiming) != hipSuccess) return fail();
    ready = true;
    return true;
  }
  bool fail()
  {
    (void)hipGetLastError();
    release();
    return false;
  }
  void release()
  {
    for (int i = 0; i < kLanes; i++) {
      if (s[i]) (void)hipStreamDestroy(s[i]);
      if (join[i]) (void)hipEventDestroy(join[i]);
      s[i] = nullptr;
      join[i] = nullptr;
    }
    if (fork) (void)hipEventDestroy(fork);
    fork = nullptr;
    ready = false;
  }
};

struct BatchBuffers {
  void *hbuf = nullptr; size_t hcap = 0;  // pinned: job table + firsts, then results
  void *dbuf = nullptr; size_t dcap = 0;  // device copy of the same
  Lanes lanes;                             // for the synchronous batch calls
};

BatchBuffers &batch_buffers()
{
  static BatchBuffers b;
  return b;
}

bool grow_pinned(void *&p, size_t &cap, size_t need)
{
  if (injected_failure()) {
    note_failure(__LINE__, "injected reservation failure");
    return false;
  }
  if (need <= cap) return true;
  const size_t n = std::max(need, cap + cap / 2);
  if (p) (void)hipHostFree(p);
  p = nullptr;
  cap = 0;
  if (hipHostMalloc(&p, n, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    p = nullptr;
    return false;
  }
  cap = n;
  return true;
}

// Checks an array as zfp_compress / zfp_decompress do; false when the call
// would fail before touching the stream.
bool batch_plan(zfp_stream *z, const zfp_field *f, bool decode, Plan &p)
{
  if (!z || !f || !z->stream || !f->data) return false;
  if (decode && !valid_type(f->type)) return false;
  if (!codec_plan(f, p)) return false;
  if (((uintptr_t)((const char *)f->data + p.imin * (int64_t)p.esz)) % p.esz) return false;
  return true;
}

// A batch as planned on the host: the arrays the batched launches take,
// ordered by class, with their chunk and table layout.
struct BatchSet {
  bool encode = true;
  std::vector<BatchItem> items;
  std::vector<size_t> later;  // arrays for the single-array path, in input order
  std::vector<uint64_t> firsts, fill_first, blk_first;
  uint64_t var_total = 0, fill_chunks = 0, blk_chunks = 0;
  size_t jbytes = 0, fbytes = 0, rbytes = 0;
};

// Device side of a batch: pinned tables and results, their device copy, the
// variable-rate scratch, and the tiling of the launches.
struct BatchDev {
  void *hbuf = nullptr, *dbuf = nullptr;
  void *lens = nullptr, *offs = nullptr, *tiles = nullptr;
  uint32_t threads = 256;  // blocks per workgroup of the class launches
  uint32_t slice = 0;      // blocks per workgroup of arrays launched alone (0: default)
  const Lanes *lanes = nullptr;  // side streams for the class launches, or none
};

// Plans the arrays; sizes[] is zeroed.  Arrays the batch does not take
// (host-resident, a stream met before, no blocks, 4-D, a decode into
// aliasing strides) go to `later`; invalid arrays are skipped (size 0).
void batch_prepare(bool encode, zfp_stream *const *zs, zfp_field *const *fs, size_t count,
                   size_t *sizes, uint flags, BatchSet &S)
{
  S.encode = encode;
  static std::atomic<uint64_t> epochs{0};
  const uint64_t epoch = ++epochs;
  std::vector<BatchItem> items;
  items.reserve(count);
  for (size_t i = 0; i < count; i++) {
    sizes[i] = 0;
    zfp_stream *z = zs[i];
    zfp_field *f = fs[i];
    Plan p;
    if (!batch_plan(z, f, !encode, p)) continue;
    bitstream *s = z->stream;
    if (s->batch_epoch == epoch) {
      S.later.push_back(i);
      continue;
    }
    s->batch_epoch = epoch;
    // nothing to code (the calls' own handling), 4-D blocks, which the batch
    // kernels are not built for, or a decode into strides that may alias,
    // which needs the single-array call's ordered scatter; the stream's later
    // arrays follow
    if (p.nblocks == 0 || p.dims > 3 || (!encode && p.may_alias)) {
      S.later.push_back(i);
      continue;
    }
    const bool dev = s->device && ((flags & ZFP_HIP_BATCH_DEVICE_FIELDS) || is_device_ptr(f->data));
    if (!dev) {
      S.later.push_back(i);
      continue;
    }
    BatchItem it;
    it.input = i;
    it.plan = p;
    const zx::Cfg c = cfg_of(z);
    const bool rev = c.minexp < kMinExp;
    const bool fixed = fixed_layout(c, p);
    if (encode) {
      it.start = s->wtell();
    } else {
      // the seek word reset_overread will leave (batch_commit resets it)
      it.seek_word0 = s->bits > 0 ? (s->word_pos ? s->word_pos - 1 : 0) : s->word_pos;
      it.start = s->rtell();
    }
    it.w0 = it.start / 64;
    zx::BatchJob &J = it.job;
    memset(&J, 0, sizeof(J));
    J.field = field_desc(p, f->data);
    J.cfg = c;
    J.vec = vec_ok(p, f->data);
    J.words = s->words + it.w0;
    J.nwords = s->nwords > it.w0 ? s->nwords - it.w0 : 0;
    J.start = it.start - 64 * it.w0;
    J.prefix = encode && s->bits ? s->buffer : 0;
    J.fixed = fixed ? c.maxbits : 0;
    J.fill = !(fixed && J.start == 0 && (c.maxbits % 64) == 0);
    if (!fixed && !encode && s->index && s->index_n == p.nblocks && s->index_start == it.start &&
        (s->index_key == 0 || s->index_key == cfg_key(c, p))) {
      J.index = s->index;
    }
    const int variant = ((int)p.type - (int)zfp_type_int32) * 6 + (p.dims - 1) * 2 + (rev ? 1 : 0);
    it.cls = fixed && p.nblocks >= kSoloBlocks ? kSoloClass + variant : variant * 2 + (fixed ? 0 : 1);
    items.push_back(it);
  }
  if (items.empty()) return;
  // items grouped by class, in input order within a class: a counting sort
  // over the class ranks (48 shared classes, then 24 of arrays launched
  // alone), each item moved once
  constexpr int kRanks = 48 + 24;
  auto rank = [](int cls) { return cls >= kSoloClass ? 48 + (cls - kSoloClass) : cls; };
  size_t at[kRanks + 1] = {};
  for (const BatchItem &it : items) at[rank(it.cls) + 1]++;
  for (int r = 0; r < kRanks; r++) at[r + 1] += at[r];
  S.items.resize(items.size());
  for (BatchItem &it : items) S.items[at[rank(it.cls)]++] = it;
  const size_t n = S.items.size();
  S.firsts.assign(n, 0);
  S.fill_first.assign(n, 0);
  S.blk_first.assign(n, 0);
  for (size_t k = 0; k < n; k++) {
    if (!S.items[k].job.fixed) {
      S.items[k].job.lbase = S.var_total;
      S.var_total += S.items[k].plan.nblocks;
    }
  }
  for (size_t k = 0; k < n;) {
    size_t e = k;
    uint64_t acc = 0;
    while (e < n && S.items[e].cls == S.items[k].cls) {
      S.firsts[e] = acc;
      acc += S.items[e].plan.nblocks;
      e++;
    }
    k = e;
  }
  // chunks of the per-array helpers: word fills and index copies
  for (size_t k = 0; k < n; k++) {
    const zx::BatchJob &J = S.items[k].job;
    S.fill_first[k] = S.fill_chunks;
    S.blk_first[k] = S.blk_chunks;
    if (encode && J.fill) {
      uint64_t words = J.nwords;
      if (J.fixed) words = std::min<uint64_t>(words, (J.start + S.items[k].plan.nblocks * J.fixed + 63) / 64);
      S.fill_chunks += zx_batch_chunks(words);
    }
    if (!J.fixed) S.blk_chunks += zx_batch_chunks(S.items[k].plan.nblocks);
  }
  S.jbytes = (n * sizeof(zx::BatchJob) + 15) / 16 * 16;
  S.fbytes = n * 8 * 3;  // firsts, fill_first, blk_first
  S.rbytes = n * 16;
}

// Bytes of the pinned and device table buffers, and of the scratch.
size_t batch_host_bytes(const BatchSet &S) { return std::max(S.jbytes + S.fbytes, S.rbytes); }
size_t batch_dev_bytes(const BatchSet &S) { return S.jbytes + S.fbytes + S.rbytes; }

// Writes the job table and chunk tables into D.hbuf (pinned); each job's
// results go to its slot after the tables in D.dbuf.
void batch_tables(BatchSet &S, const BatchDev &D)
{
  const size_t n = S.items.size();
  uint64_t *dres = (uint64_t *)((char *)D.dbuf + S.jbytes + S.fbytes);
  zx::BatchJob *hjobs = (zx::BatchJob *)D.hbuf;
  for (size_t k = 0; k < n; k++) {
    S.items[k].job.result = dres + 2 * k;
    hjobs[k] = S.items[k].job;
  }
  memcpy((char *)D.hbuf + S.jbytes, S.firsts.data(), n * 8);
  memcpy((char *)D.hbuf + S.jbytes + n * 8, S.fill_first.data(), n * 8);
  memcpy((char *)D.hbuf + S.jbytes + n * 16, S.blk_first.data(), n * 8);
}

// Records the batch's device work on `st`: the table upload (when `upload`),
// length passes and scan of variable-rate arrays, end bits, word fills, the
// coding launches, index copies, and the results' download into D.hbuf.  No
// host synchronisation and no memsets, so the sequence can be captured.
bool batch_enqueue(const BatchSet &S, const BatchDev &D, bool upload, hipStream_t st)
{
  const bool encode = S.encode;
  const size_t n = S.items.size();
  const zx::BatchJob *djobs = (const zx::BatchJob *)D.dbuf;
  const uint64_t *dfirsts = (const uint64_t *)((char *)D.dbuf + S.jbytes);
  const uint64_t *dfill_first = dfirsts + n;
  const uint64_t *dblk_first = dfirsts + 2 * n;
  uint64_t *dres = (uint64_t *)((char *)D.dbuf + S.jbytes + S.fbytes);
  uint32_t *lens = (uint32_t *)D.lens;
  uint64_t *offs = (uint64_t *)D.offs;
  const auto &items = S.items;
  const Lanes *L = D.lanes && D.lanes->ready ? D.lanes : nullptr;
  // every variable-rate array of items [k, e) has its offset index
  auto indexed = [&](size_t k, size_t e) {
    for (size_t q = k; q < e; q++)
      if (!items[q].job.fixed && !items[q].job.index) return false;
    return true;
  };
  auto launches = [&](bool var_only, bool count_pass, auto &lane) -> bool {
    for (size_t k = 0; k < n;) {
      size_t e = k;
      uint64_t nb = 0;
      while (e < n && items[e].cls == items[k].cls) nb += items[e++].plan.nblocks;
      const int cls = items[k].cls;
      if (cls >= kSoloClass) {
        if (!var_only) {
          // arrays launched on their own, through the single-array kernels
          for (size_t q = k; q < e; q++) {
            const zx::BatchJob &J = items[q].job;
            const int v = cls - kSoloClass;
            const zfp_type t = (zfp_type)(zfp_type_int32 + v / 6);
            const int dims = (v % 6) / 2 + 1;
            const bool rev = v & 1;
            hipError_t r;
            if (encode) {
              zx::EncodeArgs ea{};
              ea.field = J.field;
              ea.cfg = J.cfg;
              ea.words = J.words;
              ea.nwords = J.nwords;
              ea.start = J.start;
              ea.fixed = J.fixed;
              ea.vec = J.vec;
              ea.slice = D.slice;
              r = launch_encode(t, dims, rev, ea, lane());
            } else {
              zx::DecodeArgs da{};
              da.field = J.field;
              da.cfg = J.cfg;
              da.words = J.words;
              da.nwords = J.nwords;
              da.start = J.start;
              da.fixed = J.fixed;
              da.maxword = dres + 2 * q + 1;
              da.vec = J.vec;
              da.slice = D.slice;
              r = launch_decode(t, dims, rev, da, lane());
            }
            if (r != hipSuccess) {
              (void)hipGetLastError();
              return false;
            }
          }
        }
      } else if ((!var_only || (cls & 1)) && !(var_only && !encode && indexed(k, e))) {
        // (a decode's walk is launched only for a class with an array that
        // has no offset index)
        zx::BatchArgs a{};
        a.jobs = djobs + k;
        a.firsts = dfirsts + k;
        a.njobs = (uint32_t)(e - k);
        a.nblocks = nb;
        a.lengths = lens;
        a.offsets = offs;
        a.threads = D.threads;
        const int v = cls >> 1;
        const hipStream_t ls = lane();
        const hipError_t r = encode ? kBEnc[v / 6][v % 6](a, count_pass, ls) : kBDec[v / 6][v % 6](a, count_pass, ls);
        if (r != hipSuccess) {
          (void)hipGetLastError();
          return false;
        }
      }
      k = e;
    }
    return true;
  };
  // launch i of a phase goes to stream i mod (kLanes + 1): the caller's, then
  // the side streams, which start after the work before the phase and are
  // joined back before the work after it (also when a launch fails)
  auto run = [&](bool var_only, bool count_pass) -> bool {
    const bool forked = L && hipEventRecord(L->fork, st) == hipSuccess;
    uint32_t ready = 0;
    while (forked && ready < (uint32_t)kLanes && hipStreamWaitEvent(L->s[ready], L->fork, 0) == hipSuccess)
      ready++;
    uint32_t launch = 0;
    auto lane = [&]() -> hipStream_t {
      const uint32_t i = launch++ % (kLanes + 1);
      return i && i <= ready ? L->s[i - 1] : st;
    };
    const bool good = launches(var_only, count_pass, lane);
    bool joined = true;
    for (uint32_t i = 0; i < ready; i++)
      joined = hipEventRecord(L->join[i], L->s[i]) == hipSuccess &&
               hipStreamWaitEvent(st, L->join[i], 0) == hipSuccess && joined;
    return good && joined && (!L || forked);
  };
  bool good = true;
  if (upload) good = hipMemcpyAsync(D.dbuf, D.hbuf, S.jbytes + S.fbytes, hipMemcpyHostToDevice, st) == hipSuccess;
  if (good && S.var_total) {
    if (encode) good = run(true, true);
    else
      good = zx_batch_gather_index(djobs, dblk_first, (uint32_t)n, S.blk_chunks, lens, st) == hipSuccess &&
             run(true, true);
    good = good && zx_launch_scan(lens, S.var_total, 0, offs, (uint64_t *)D.tiles, st) == hipSuccess;
  }
  good = good && zx_batch_ends(djobs, (uint32_t)n, offs, st) == hipSuccess;
  if (good && encode) good = zx_batch_fill(djobs, dfill_first, (uint32_t)n, S.fill_chunks, st) == hipSuccess;
  good = good && run(false, false);
  if (good && encode && S.var_total)
    good = zx_batch_keep_index(djobs, dblk_first, (uint32_t)n, S.blk_chunks, lens, st) == hipSuccess;
  good = good && hipMemcpyAsync(D.hbuf, dres, S.rbytes, hipMemcpyDeviceToHost, st) == hipSuccess;
  if (!good) (void)hipGetLastError();
  return good;
}

// After the device work: each array's cursor (and index record) from the
// downloaded results, and its size; returns the arrays that succeeded.
size_t batch_finish(const BatchSet &S, const BatchDev &D, zfp_stream *const *zs, size_t *sizes)
{
  size_t ok = 0;
  const uint64_t *res = (const uint64_t *)D.hbuf;
  for (size_t k = 0; k < S.items.size(); k++) {
    const BatchItem &it = S.items[k];
    zfp_stream *z = zs[it.input];
    bitstream *s = z->stream;
    const uint64_t end = 64 * it.w0 + res[2 * k];
    const uint64_t w1 = (end + 63) / 64;
    if (S.encode) {
      // the cursor after the last block, then flush (core.rs:536-542)
      if (w1 > s->nwords) s->overflowed = true;
      s->word_pos = (size_t)w1;
      s->bits = 0;
      s->buffer = 0;
      if (!it.job.fixed && s->index) {
        s->index_n = it.plan.nblocks;
        s->index_start = it.start;
        s->index_key = cfg_key(cfg_of(z), it.plan);
        s->index_bits = end - it.start;
      }
      if (!s->overflowed) {
        sizes[it.input] = s->word_pos * 8;
        ok++;
      }
    } else {
      const uint64_t loaded_end = it.w0 + res[2 * k + 1];
      s->word_pos = (size_t)w1;
      s->bits = 0;
      s->buffer = 0;
      s->seek_word = s->word_pos;
      s->overread = loaded_end > std::max<uint64_t>(s->nwords, it.seek_word0);
      if (!s->overread_now()) {
        sizes[it.input] = s->word_pos * 8;
        ok++;
      }
    }
  }
  return ok;
}

// What a planned batch changes in its streams before its device work, once
// the batch is accepted (batch_prepare changes nothing a caller sees): a decode's overread
// state is reset, as zfp_decompress does; a variable-rate encode gets room for
// its offset index beside the stream, as zfp_compress keeps it.
void batch_commit(zfp_stream *const *zs, BatchSet &S)
{
  for (BatchItem &it : S.items) {
    bitstream *s = zs[it.input]->stream;
    if (!S.encode) {
      s->reset_overread();
    } else if (!it.job.fixed) {
      if (s->index_cap < it.plan.nblocks) {
        drop_index(s);
        if (hipMalloc((void **)&s->index, it.plan.nblocks * 4) != hipSuccess) {
          (void)hipGetLastError();
          s->index = nullptr;
        } else {
          s->index_cap = it.plan.nblocks;
        }
      }
      it.job.index_out = s->index;
    }
  }
}

size_t batch_run(bool encode, zfp_stream *const *zs, zfp_field *const *fs, size_t count, size_t *sizes,
                 uint flags)
{
  if (!zs || !fs || !sizes) return 0;
  size_t ok = 0;
  BatchSet S;
  batch_prepare(encode, zs, fs, count, sizes, flags, S);
  batch_commit(zs, S);
  if (!S.items.empty()) {
    const hipStream_t st = hip_stream_of(zs[S.items[0].input]);
    Workspace &W = ws();
    BatchBuffers &B = batch_buffers();
    std::lock_guard<std::mutex> lock(W.mu);
    bool good = grow_pinned(B.hbuf, B.hcap, batch_host_bytes(S)) &&
                Workspace::grow(B.dbuf, B.dcap, batch_dev_bytes(S));
    if (good && S.var_total) {
      good = Workspace::grow(W.lens, W.lens_cap, S.var_total * 4) &&
             Workspace::grow(W.offs, W.offs_cap, (S.var_total + 1) * 8) &&
             Workspace::grow(W.tiles, W.tiles_cap, zx_scan_tiles(S.var_total) * 8);
    }
    BatchDev D;
    D.hbuf = B.hbuf;
    D.dbuf = B.dbuf;
    D.lens = W.lens;
    D.offs = W.offs;
    D.tiles = W.tiles;
    if (B.lanes.init()) D.lanes = &B.lanes;
    if (good) {
      batch_tables(S, D);
      good = batch_enqueue(S, D, true, st) && hipStreamSynchronize(st) == hipSuccess;
      if (!good) (void)hipGetLastError();
    }
    if (good) ok += batch_finish(S, D, zs, sizes);
  }
  std::sort(S.later.begin(), S.later.end());
  for (size_t i : S.later) {
    sizes[i] = encode ? zfp_compress(zs[i], fs[i]) : zfp_decompress(zs[i], fs[i]);
    if (sizes[i]) ok++;
  }
  return ok;
}

// Cursor state of a stream that a planned batch restores before each run.
struct SavedCursor {
  size_t word_pos;
  uint64_t buffer;
  uint32_t bits;
  size_t seek_word;
};

SavedCursor save_cursor(const bitstream *s) { return {s->word_pos, s->buffer, s->bits, s->seek_word}; }

void restore_cursor(bitstream *s, const SavedCursor &c)
{
  s->word_pos = c.word_pos;
  s->buffer = c.buffer;
  s->bits = c.bits;
  s->seek_word = c.seek_word;
}

bool same_field(const zfp_field &a, const zfp_field &b)
{
  return a.type == b.type && a.nx == b.nx && a.ny == b.ny && a.nz == b.nz && a.nw == b.nw &&
         a.sx == b.sx && a.sy == b.sy && a.sz == b.sz && a.sw == b.sw && a.data == b.data;
}

bool same_cfg(const zx::Cfg &a, const zx::Cfg &b)
{
  return a.minbits == b.minbits && a.maxbits == b.maxbits && a.maxprec == b.maxprec &&
         a.minexp == b.minexp && a.rounding == b.rounding && a.tight == b.tight;
}

// Buffers of an asynchronous submission (pinned tables and results, their
// device copy, the variable-rate scratch), reused across submissions: a
// submission takes a set from the pool and its wait returns it.
struct SubmitBuffers {
  void *hbuf = nullptr; size_t hcap = 0;
  void *dbuf = nullptr; size_t dcap = 0;
  void *lens = nullptr; size_t lens_cap = 0;
  void *offs = nullptr; size_t offs_cap = 0;
  void *tiles = nullptr; size_t tiles_cap = 0;
};

struct SubmitPool {
  std::mutex mu;
  std::vector<SubmitBuffers *> free;
};

SubmitPool &submit_pool()
{
  static SubmitPool p;
  return p;
}

}  // namespace

// A batch planned once and replayed from a captured HIP graph
// (zfp_hip_batch_plan), or submitted once without waiting
// (zfp_hip_batch_submit).
struct zfp_hip_batch {
  BatchSet S;
  BatchDev D;
  std::vector<zfp_stream *> zs;
  std::vector<zfp_field> fields;       // the fields as planned
  std::vector<SavedCursor> cursors;    // per input: the cursor at plan time
  std::vector<bitstream> streams;      // per input: the stream as planned (words, index)
  hipStream_t capture = nullptr;       // private stream the graph was captured on
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double tuned_ms = 0;                 // the chosen tiling's time at plan time
  SubmitBuffers *pooled = nullptr;     // a submission's buffers (returned to the pool)
  hipEvent_t done = nullptr;           // a submission's completion
  Lanes lanes;                         // a plan's side streams

  ~zfp_hip_batch()
  {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    lanes.release();
    if (capture) (void)hipStreamDestroy(capture);
    if (done) (void)hipEventDestroy(done);
    if (pooled) {
      SubmitPool &pool = submit_pool();
      std::lock_guard<std::mutex> lock(pool.mu);
      pool.free.push_back(pooled);
      return;
    }
    if (D.hbuf) (void)hipHostFree(D.hbuf);
    for (void *p : {D.dbuf, D.lens, D.offs, D.tiles})
      if (p) (void)hipFree(p);
  }
};

namespace {

bool dev_alloc(void *&p, size_t bytes)
{
  if (hipMalloc(&p, std::max<size_t>(bytes, 8)) != hipSuccess) {
    (void)hipGetLastError();
    p = nullptr;
    return false;
  }
  return true;
}

// Plans a batch whose every array the batched launches take (else null,
// cursors left as they were) and writes its tables into its pinned buffer:
// buffers owned by the batch, or (pooled) a set from the submission pool.
std::unique_ptr<zfp_hip_batch> make_batch(bool encode, zfp_stream *const *streams,
                                          zfp_field *const *fields, size_t count, uint flags,
                                          bool pooled)
{
  if (!streams || !fields || count == 0) return nullptr;
  std::unique_ptr<zfp_hip_batch> P(new zfp_hip_batch);
  std::vector<SavedCursor> before(count);
  for (size_t i = 0; i < count; i++) {
    if (!streams[i] || !streams[i]->stream || !fields[i]) return nullptr;
    before[i] = save_cursor(streams[i]->stream);
  }
  std::vector<size_t> sizes(count);
  batch_prepare(encode, streams, fields, count, sizes.data(), flags, P->S);
  // every array must be one the batched launches take
  if (!P->S.later.empty() || P->S.items.size() != count) {
    for (size_t i = 0; i < count; i++) restore_cursor(streams[i]->stream, before[i]);
    return nullptr;
  }
  batch_commit(streams, P->S);
  P->zs.assign(streams, streams + count);
  for (size_t i = 0; i < count; i++) {
    P->fields.push_back(*fields[i]);
    P->cursors.push_back(before[i]);
    P->streams.push_back(*streams[i]->stream);
  }
  BatchSet &S = P->S;
  BatchDev &D = P->D;
  if (pooled) {
    SubmitPool &pool = submit_pool();
    SubmitBuffers *b = nullptr;
    {
      std::lock_guard<std::mutex> lock(pool.mu);
      if (!pool.free.empty()) {
        b = pool.free.back();
        pool.free.pop_back();
      }
    }
    if (!b) b = new SubmitBuffers;
    P->pooled = b;  // back to the pool with the batch, whatever happens next
    bool good = grow_pinned(b->hbuf, b->hcap, batch_host_bytes(S)) &&
                Workspace::grow(b->dbuf, b->dcap, batch_dev_bytes(S));
    if (good && S.var_total)
      good = Workspace::grow(b->lens, b->lens_cap, S.var_total * 4) &&
             Workspace::grow(b->offs, b->offs_cap, (S.var_total + 1) * 8) &&
             Workspace::grow(b->tiles, b->tiles_cap, zx_scan_tiles(S.var_total) * 8);
    if (!good) return nullptr;
    D.hbuf = b->hbuf;
    D.dbuf = b->dbuf;
    D.lens = b->lens;
    D.offs = b->offs;
    D.tiles = b->tiles;
  } else {
    if (hipHostMalloc(&D.hbuf, batch_host_bytes(S), hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      D.hbuf = nullptr;
      return nullptr;
    }
    if (!dev_alloc(D.dbuf, batch_dev_bytes(S))) return nullptr;
    if (S.var_total && !(dev_alloc(D.lens, S.var_total * 4) && dev_alloc(D.offs, (S.var_total + 1) * 8) &&
                         dev_alloc(D.tiles, zx_scan_tiles(S.var_total) * 8)))
      return nullptr;
  }
  batch_tables(S, D);
  return P;
}

// Time of one run of the batch's device work with the current tiling
// (best of `reps`), or a negative value when a launch fails.
double time_batch(zfp_hip_batch &P, hipStream_t st, int reps)
{
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess) return -1;
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    return -1;
  }
  double best = -1;
  for (int r = 0; r < reps; r++) {
    float ms = 0;
    const bool good = hipEventRecord(a, st) == hipSuccess && batch_enqueue(P.S, P.D, false, st) &&
                      hipEventRecord(b, st) == hipSuccess && hipEventSynchronize(b) == hipSuccess &&
                      hipEventElapsedTime(&ms, a, b) == hipSuccess;
    if (!good) {
      (void)hipGetLastError();
      best = -1;
      break;
    }
    if (best < 0 || ms < best) best = ms;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return best;
}

// Slices of the LDS path every array launched alone can take.
bool slice_fits(const BatchSet &S, uint32_t T)
{
  for (const BatchItem &it : S.items) {
    if (it.cls < kSoloClass) continue;
    zx::FxPlan p;
    if (!zx::fx_plan(it.job.fixed, it.job.start, !S.encode, T, p)) return false;
  }
  return true;
}

}  // namespace

// ===================================================================== C ABI

extern "C" {

size_t zfp_type_size(zfp_type t) { return type_size(t); }

zfp_config zfp_config_none(void)
{
  zfp_config c;
  memset(&c, 0, sizeof(c));
  c.mode = zfp_mode_null;
  return c;
}
zfp_config zfp_config_rate(double rate, zfp_bool align)
{
  zfp_config c = zfp_config_none();
  c.mode = zfp_mode_fixed_rate;
  c.arg.rate = (align == zfp_false && rate >= 0.0) ? -rate : rate;  // negative: unaligned
  return c;
}
zfp_config zfp_config_precision(uint32_t precision)
{
  zfp_config c = zfp_config_none();
  c.mode = zfp_mode_fixed_precision;
  c.arg.precision = precision;
  return c;
}
zfp_config zfp_config_accuracy(double tolerance)
{
  zfp_config c = zfp_config_none();
  c.mode = zfp_mode_fixed_accuracy;
  c.arg.tolerance = tolerance;
  return c;
}
zfp_config zfp_config_reversible(void)
{
  zfp_config c = zfp_config_none();
  c.mode = zfp_mode_reversible;
  return c;
}
zfp_config zfp_config_expert(uint32_t minbits, uint32_t maxbits, uint32_t maxprec, int32_t minexp)
{
  zfp_config c = zfp_config_none();
  c.mode = zfp_mode_expert;
  c.arg.expert.minbits = minbits;
  c.arg.expert.maxbits = maxbits;
  c.arg.expert.maxprec = maxprec;
  c.arg.expert.minexp = minexp;
  return c;
}

zfp_stream *zfp_stream_open(bitstream *bs)
{
  zfp_stream *z = (zfp_stream *)calloc(1, sizeof(zfp_stream));
  if (!z) return nullptr;
  set_params(z, kMinBits, kMaxBits, kMaxPrec, kMinExp);
  z->stream = bs;
  z->exec.policy = zfp_exec_serial;
  z->exec.params = nullptr;
  return z;
}

void zfp_stream_close(zfp_stream *z)
{
  if (!z) return;
  if (ExecParams *p = params(z, false))
    for (int i = 0; i < 4; i++)
      if (p->ev[i]) (void)hipEventDestroy(p->ev[i]);
  free(z->exec.params);
  free(z);
}

bitstream *zfp_stream_bit_stream(const zfp_stream *z) { return z ? z->stream : nullptr; }
zfp_mode zfp_stream_compression_mode(const zfp_stream *z) { return z ? mode_of(z) : zfp_mode_null; }

double zfp_stream_rate(const zfp_stream *z, uint dims)
{
  if (!z || dims < 1 || dims > 4 || mode_of(z) != zfp_mode_fixed_rate) return 0.0;
  return (double)z->maxbits / (double)(1u << (2 * dims));
}
uint zfp_stream_precision(const zfp_stream *z)
{
  return z && mode_of(z) == zfp_mode_fixed_precision ? z->maxprec : 0;
}
double zfp_stream_accuracy(const zfp_stream *z)
{
  return z && mode_of(z) == zfp_mode_fixed_accuracy ? ldexp(1.0, z->minexp) : 0.0;
}
uint64 zfp_stream_mode(const zfp_stream *z)
{
  return z ? mode_bits(z->minbits, z->maxbits, z->maxprec, z->minexp) : 0;
}
void zfp_stream_params(const zfp_stream *z, uint *minbits, uint *maxbits, uint *maxprec, int32_t *minexp)
{
  if (!z || !minbits || !maxbits || !maxprec || !minexp) return;
  *minbits = z->minbits;
  *maxbits = z->maxbits;
  *maxprec = z->maxprec;
  *minexp = z->minexp;
}
size_t zfp_stream_compressed_size(const zfp_stream *z) { return z && z->stream ? z->stream->word_pos * 8 : 0; }

// zfp_block_maximum_size (zfp-rs-ffi/src/stream.rs:178-206)
size_t zfp_block_maximum_size(zfp_type ty, uint dims, zfp_bool reversible)
{
  static const size_t sizes[2][4][4] = {
      {{131, 527, 2111, 8447}, {259, 1039, 4159, 16639}, {140, 536, 2120, 8456}, {271, 1051, 4171, 16651}},
      {{136, 532, 2116, 8452}, {265, 1045, 4165, 16645}, {146, 542, 2126, 8462}, {278, 1058, 4178, 16658}},
  };
  if (!valid_type(ty) || dims < 1 || dims > 4) return 0;
  return sizes[reversible != zfp_false][ty - 1][dims - 1];
}

// maximum_size (config.rs:642-658)
size_t zfp_stream_maximum_size(const zfp_stream *z, const zfp_field *f)
{
  if (!z || !f || !valid_type(f->type)) return 0;
  size_t d[4] = {f->nx, f->ny, f->nz, f->nw};
  for (int i = 0; i < 4; i++)
    if (d[i] == 0) {
      for (int j = i; j < 4; j++) d[j] = 0;
      break;
    }
  if (d[0] == 0) return 0;
  int dims = d[1] == 0 ? 1 : d[2] == 0 ? 2 : d[3] == 0 ? 3 : 4;
  unsigned __int128 blocks = 1;
  for (int i = 0; i < dims; i++) blocks *= (d[i] + 3) / 4;
  zx::Cfg c = cfg_of(z);
  unsigned __int128 bits = blocks * block_bits(c, f->type, dims) + 148;
  bits = (bits + 63) / 64 * 64;
  if (bits / 8 > (unsigned __int128)SIZE_MAX) return 0;
  return (size_t)(bits / 8);
}

// field metadata (field.rs:678-735)
static bool field_metadata(const zfp_field *f, uint64_t &meta)
{
  const size_t d[4] = {f->nx, f->ny, f->nz, f->nw};
  if (d[0] == 0) return false;
  bool zero = false;
  for (int i = 1; i < 4; i++) {
    if (d[i] == 0) zero = true;
    else if (zero) return false;
  }
  int dims = dimensionality(f);
  int width = 48 / dims;
  uint64_t maxv = (1ull << width) - 1;
  uint64_t m = 0;
  int shift = 0;
  for (int i = 0; i < dims; i++) {
    uint64_t e = (uint64_t)d[i] - 1;
    if (e > maxv) return false;
    m |= e << shift;
    shift += width;
  }
  m = (m << 2) | (uint64_t)(dims - 1);
  m = (m << 2) | (uint64_t)(f->type - 1);
  meta = m;
  return true;
}

static bool metadata_to_field(uint64_t meta, zfp_field *f)
{
  if (meta >> 52) return false;
  zfp_type t = (zfp_type)((meta & 3) + 1);
  uint64_t m = meta >> 2;
  int dims = (int)(m & 3) + 1;
  m >>= 2;
  int width = 48 / dims;
  uint64_t mask = (1ull << width) - 1;
  size_t d[4] = {0, 0, 0, 0};
  for (int i = 0; i < dims; i++) {
    d[i] = (size_t)((m & mask) + 1);
    m >>= width;
  }
  f->type = t;
  f->nx = d[0]; f->ny = d[1]; f->nz = d[2]; f->nw = d[3];
  f->sx = f->sy = f->sz = f->sw = 0;
  return true;
}

// zfp_write_header (zfp-rs-ffi/src/stream.rs:230-276)
size_t zfp_write_header(zfp_stream *z, const zfp_field *f, uint mask)
{
  if (!z || !f || !z->stream || !valid_type(f->type)) return 0;
  Plan p;
  if (!make_plan(f, p)) return 0;
  bitstream *s = z->stream;
  size_t bits = 0;
  if (mask & ZFP_HEADER_MAGIC) {
    s->write_bits('z', 8);
    s->write_bits('f', 8);
    s->write_bits('p', 8);
    s->write_bits(ZFP_CODEC, 8);
    bits += ZFP_MAGIC_BITS;
  }
  if (mask & ZFP_HEADER_META) {
    uint64_t meta;
    if (!field_metadata(f, meta)) return 0;
    s->write_bits(meta, ZFP_META_BITS);
    bits += ZFP_META_BITS;
  }
  if (mask & ZFP_HEADER_MODE) {
    uint64_t m = mode_bits(z->minbits, z->maxbits, z->maxprec, z->minexp);
    uint32_t size = m > ZFP_MODE_SHORT_MAX ? ZFP_MODE_LONG_BITS : ZFP_MODE_SHORT_BITS;
    s->write_bits(m, size);
    bits += size;
  }
  return bits;
}

// zfp_read_header (header.rs:132-173)
size_t zfp_read_header(zfp_stream *z, zfp_field *f, uint mask)
{
  if (!z || !f || !z->stream) return 0;
  bitstream *s = z->stream;
  size_t bits = 0;
  if (mask & ZFP_HEADER_MAGIC) {
    if (s->read_bits(8) != 'z' || s->read_bits(8) != 'f' || s->read_bits(8) != 'p' ||
        s->read_bits(8) != ZFP_CODEC)
      return 0;
    bits += ZFP_MAGIC_BITS;
  }
  zfp_field tmp = *f;
  bool have_meta = false;
  if (mask & ZFP_HEADER_META) {
    uint64_t meta = s->read_bits(ZFP_META_BITS);
    if (!metadata_to_field(meta, &tmp)) return 0;
    have_meta = true;
    bits += ZFP_META_BITS;
  }
  uint32_t a = 0, b = 0, c = 0;
  int32_t d = 0;
  bool have_mode = false;
  if (mask & ZFP_HEADER_MODE) {
    uint64_t m = s->read_bits(ZFP_MODE_SHORT_BITS);
    bits += ZFP_MODE_SHORT_BITS;
    if (m > ZFP_MODE_SHORT_MAX) {
      m |= s->read_bits(ZFP_MODE_LONG_BITS - ZFP_MODE_SHORT_BITS) << ZFP_MODE_SHORT_BITS;
      bits += ZFP_MODE_LONG_BITS - ZFP_MODE_SHORT_BITS;
    }
    if (!params_from_mode(m, a, b, c, d)) return 0;
    have_mode = true;
  }
  if (have_meta) {
    f->type = tmp.type;
    f->nx = tmp.nx; f->ny = tmp.ny; f->nz = tmp.nz; f->nw = tmp.nw;
    f->sx = f->sy = f->sz = f->sw = 0;
  }
  if (have_mode) set_params(z, a, b, c, d);
  return bits;
}

// zfp_compress (zfp-rs-ffi/src/stream.rs:321-342): bytes of stream after the
// flush, or 0 on any failure.
size_t zfp_compress(zfp_stream *z, const zfp_field *f)
{
  if (!z || !f || !z->stream || !f->data) return 0;
  Plan p;
  if (!codec_plan(f, p)) return 0;
  if (((uintptr_t)((const char *)f->data + p.imin * (int64_t)p.esz)) % p.esz) return 0;
  bitstream *s = z->stream;
  const zx::Cfg c = cfg_of(z);
  if (p.nblocks > 0) {
    if (!gpu_compress(s, f, p, c, hip_stream_of(z), params(z, false))) return 0;
  } else {
    s->flush();
  }
  if (s->overflowed) return 0;
  return s->word_pos * 8;
}

// ---------------------------------------------------------------- allocation-free compress

// Scratch of zfp_hip_compress_noalloc: one array per GPU, held until the
// process ends.  Layout (64-bit words): lengths (4 B per block), offsets
// (blocks + 1), scan tiles.  Nothing in it has to start at zero: the length
// pass writes every length, the scan every offset and tile.
constexpr uint64_t kNaBlocks = ZFP_HIP_NOALLOC_MAX_BLOCKS;
constexpr uint64_t kNaLensWords = kNaBlocks / 2;
constexpr uint64_t kNaOffsWords = kNaBlocks + 32;
constexpr uint64_t kNaTileWords = 1024;
static __device__ __align__(256) uint64_t scratch[kNaLensWords + kNaOffsWords + kNaTileWords];

// The current device's address of `scratch`, looked up once per device with
// that device current.  Host code must use this address, never the array's
// name (which on the host is a stand-in, not a device address).
static uint64_t *noalloc_scratch()
{
  static std::mutex mu;
  static uint64_t *addr[64] = {};
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
    (void)hipGetLastError();
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void *ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
      (void)hipGetLastError();
      return nullptr;
    }
    addr[dev] = (uint64_t *)ptr;
  }
  return addr[dev];
}

// zfp_compress without allocating (no counterpart in the reference; for
// callers that cannot allocate): device-resident field and stream words, at
// most ZFP_HIP_NOALLOC_MAX_BLOCKS blocks, scratch from the static array.
size_t zfp_hip_compress_noalloc(zfp_stream *z, const zfp_field *f)
{
  if (!z || !f || !z->stream || !f->data) return 0;
  Plan p;
  if (!codec_plan(f, p)) return 0;
  if (p.nblocks > kNaBlocks || zx_scan_tiles(p.nblocks) > kNaTileWords) return 0;
  if (((uintptr_t)((const char *)f->data + p.imin * (int64_t)p.esz)) % p.esz) return 0;
  bitstream *s = z->stream;
  if (!s->device || !is_device_ptr(f->data)) return 0;
  const zx::Cfg c = cfg_of(z);
  if (p.nblocks > 0) {
    uint64_t *base = noalloc_scratch();
    if (!base) return 0;
    const Scratch sc{(uint32_t *)base, base + kNaLensWords, base + kNaLensWords + kNaOffsWords,
                     kNaBlocks, kNaOffsWords};
    if (!gpu_compress(s, f, p, c, hip_stream_of(z), nullptr, true, nullptr, &sc)) return 0;
  } else {
    s->flush();
  }
  if (s->overflowed) return 0;
  return s->word_pos * 8;
}

uint64 zfp_hip_noalloc_max_blocks(void) { return kNaBlocks; }

// zfp_decompress (zfp-rs-ffi/src/stream.rs:345-367): bytes read through the
// aligned end of the last block, or 0 on any failure (truncation included).
size_t zfp_decompress(zfp_stream *z, zfp_field *f)
{
  if (!z || !f || !z->stream || !f->data || !valid_type(f->type)) return 0;
  Plan p;
  if (!codec_plan(f, p)) return 0;
  if (((uintptr_t)((const char *)f->data + p.imin * (int64_t)p.esz)) % p.esz) return 0;
  bitstream *s = z->stream;
  const zx::Cfg c = cfg_of(z);
  if (p.nblocks > 0) {
    uint64_t end;
    if (!gpu_decompress(s, f, p, c, hip_stream_of(z), params(z, false), end)) return 0;
  } else {
    s->reset_overread();
    s->align();
  }
  if (s->overread_now()) return 0;
  return s->word_pos * 8;
}

// Block-level entry points (zfp-rs-ffi/src/block.rs): T the C scalar type,
// N its name in the symbols, ZT its zfp_type.
#define ZX_BLOCK_API(T, N, ZT)                                                                     \
  size_t zfp_encode_block_##N##_1(zfp_stream *z, const T *b) { return block_whole(true, z, b, ZT, 1); } \
  size_t zfp_encode_block_##N##_2(zfp_stream *z, const T *b) { return block_whole(true, z, b, ZT, 2); } \
  size_t zfp_encode_block_##N##_3(zfp_stream *z, const T *b) { return block_whole(true, z, b, ZT, 3); } \
  size_t zfp_encode_block_##N##_4(zfp_stream *z, const T *b) { return block_whole(true, z, b, ZT, 4); } \
  size_t zfp_decode_block_##N##_1(zfp_stream *z, T *b) { return block_whole(false, z, b, ZT, 1); }      \
  size_t zfp_decode_block_##N##_2(zfp_stream *z, T *b) { return block_whole(false, z, b, ZT, 2); }      \
  size_t zfp_decode_block_##N##_3(zfp_stream *z, T *b) { return block_whole(false, z, b, ZT, 3); }      \
  size_t zfp_decode_block_##N##_4(zfp_stream *z, T *b) { return block_whole(false, z, b, ZT, 4); }      \
  size_t zfp_encode_block_strided_##N##_1(zfp_stream *z, const T *b, ptrdiff_t sx)                      \
  {                                                                                                    \
    return block_strided(true, z, b, ZT, 1, sx, 0, 0, 0);                                              \
  }                                                                                                    \
  size_t zfp_encode_block_strided_##N##_2(zfp_stream *z, const T *b, ptrdiff_t sx, ptrdiff_t sy)        \
  {                                                                                                    \
    return block_strided(true, z, b, ZT, 2, sx, sy, 0, 0);                                             \
  }                                                                                                    \
  size_t zfp_encode_block_strided_##N##_3(zfp_stream *z, const T *b, ptrdiff_t sx, ptrdiff_t sy,        \
                                          ptrdiff_t sz)                                                \
  {                                                                                                    \
    return block_strided(true, z, b, ZT, 3, sx, sy, sz, 0);                                            \
  }                                                                                                    \
  size_t zfp_encode_block_strided_##N##_4(zfp_stream *z, const T *b, ptrdiff_t sx, ptrdiff_t sy,        \
                                          ptrdiff_t sz, ptrdiff_t sw)                                  \
  {                                                                                                    \
    return block_strided(true, z, b, ZT, 4, sx, sy, sz, sw);                                           \
  }                                                                                                    \
  size_t zfp_decode_block_strided_##N##_1(zfp_stream *z, T *b, ptrdiff_t sx)                            \
  {                                                                                                    \
    return block_strided(false, z, b, ZT, 1, sx, 0, 0, 0);                                             \
  }                                                                                                    \
  size_t zfp_decode_block_strided_##N##_2(zfp_stream *z, T *b, ptrdiff_t sx, ptrdiff_t sy)              \
  {                                                                                                    \
    return block_strided(false, z, b, ZT, 2, sx, sy, 0, 0);                                            \
  }                                                                                                    \
  size_t zfp_decode_block_strided_##N##_3(zfp_stream *z, T *b, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz) \
  {                                                                                                    \
    return block_strided(false, z, b, ZT, 3, sx, sy, sz, 0);                                           \
  }                                                                                                    \
  size_t zfp_decode_block_strided_##N##_4(zfp_stream *z, T *b, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz, \
                                          ptrdiff_t sw)                                                \
  {                                                                                                    \
    return block_strided(false, z, b, ZT, 4, sx, sy, sz, sw);                                          \
  }                                                                                                    \
  size_t zfp_encode_partial_block_strided_##N##_1(zfp_stream *z, const T *b, size_t nx, ptrdiff_t sx)   \
  {                                                                                                    \
    return block_partial(true, z, b, ZT, 1, nx, 1, 1, 1, sx, 0, 0, 0);                                 \
  }                                                                                                    \
  size_t zfp_encode_partial_block_strided_##N##_2(zfp_stream *z, const T *b, size_t nx, size_t ny,      \
                                                  ptrdiff_t sx, ptrdiff_t sy)                          \
  {                                                                                                    \
    return block_partial(true, z, b, ZT, 2, nx, ny, 1, 1, sx, sy, 0, 0);                               \
  }                                                                                                    \
  size_t zfp_encode_partial_block_strided_##N##_3(zfp_stream *z, const T *b, size_t nx, size_t ny,      \
                                                  size_t nz, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz) \
  {                                                                                                    \
    return block_partial(true, z, b, ZT, 3, nx, ny, nz, 1, sx, sy, sz, 0);                             \
  }                                                                                                    \
  size_t zfp_encode_partial_block_strided_##N##_4(zfp_stream *z, const T *b, size_t nx, size_t ny,      \
                                                  size_t nz, size_t nw, ptrdiff_t sx, ptrdiff_t sy,    \
                                                  ptrdiff_t sz, ptrdiff_t sw)                          \
  {                                                                                                    \
    return block_partial(true, z, b, ZT, 4, nx, ny, nz, nw, sx, sy, sz, sw);                           \
  }                                                                                                    \
  size_t zfp_decode_partial_block_strided_##N##_1(zfp_stream *z, T *b, size_t nx, ptrdiff_t sx)         \
  {                                                                                                    \
    return block_partial(false, z, b, ZT, 1, nx, 1, 1, 1, sx, 0, 0, 0);                                \
  }                                                                                                    \
  size_t zfp_decode_partial_block_strided_##N##_2(zfp_stream *z, T *b, size_t nx, size_t ny,            \
                                                  ptrdiff_t sx, ptrdiff_t sy)                          \
  {                                                                                                    \
    return block_partial(false, z, b, ZT, 2, nx, ny, 1, 1, sx, sy, 0, 0);                              \
  }                                                                                                    \
  size_t zfp_decode_partial_block_strided_##N##_3(zfp_stream *z, T *b, size_t nx, size_t ny, size_t nz, \
                                                  ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz)            \
  {                                                                                                    \
    return block_partial(false, z, b, ZT, 3, nx, ny, nz, 1, sx, sy, sz, 0);                            \
  }                                                                                                    \
  size_t zfp_decode_partial_block_strided_##N##_4(zfp_stream *z, T *b, size_t nx, size_t ny, size_t nz, \
                                                  size_t nw, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz, \
                                                  ptrdiff_t sw)                                        \
  {                                                                                                    \
    return block_partial(false, z, b, ZT, 4, nx, ny, nz, nw, sx, sy, sz, sw);                          \
  }

ZX_BLOCK_API(int32, int32, zfp_type_int32)
ZX_BLOCK_API(int64, int64, zfp_type_int64)
ZX_BLOCK_API(float, float, zfp_type_float)
ZX_BLOCK_API(double, double, zfp_type_double)
#undef ZX_BLOCK_API

void zfp_stream_rewind(zfp_stream *z)
{
  if (z && z->stream) z->stream->rewind();
}
void zfp_stream_set_bit_stream(zfp_stream *z, bitstream *bs)
{
  if (z) z->stream = bs;
}
void zfp_stream_set_reversible(zfp_stream *z)
{
  if (z) set_params(z, kMinBits, kMaxBits, kMaxPrec, kMinExp - 1);
}

// zfp_stream_set_rate (zfp-rs-ffi/src/stream.rs:393-423)
double zfp_stream_set_rate(zfp_stream *z, double rate, zfp_type ty, uint dims, zfp_bool align)
{
  if (!z || dims < 1 || dims > 4 || !valid_type(ty)) return 0.0;
  uint32_t bits;
  if (!rate_bits(rate, ty, dims, align == zfp_true, bits)) return 0.0;
  set_params(z, bits, bits, kMaxPrec, kMinExp);
  return (double)bits / (double)(1u << (2 * dims));
}

uint zfp_stream_set_precision(zfp_stream *z, uint precision)
{
  if (!z) return 0;
  uint32_t p = precision > 0 ? std::min(precision, kMaxPrec) : kMaxPrec;
  set_params(z, kMinBits, kMaxBits, p, kMinExp);
  return p;
}

double zfp_stream_set_accuracy(zfp_stream *z, double tolerance)
{
  if (!z) return 0.0;
  int32_t emin = kMinExp;
  if (tolerance > 0.0) {
    int e;
    frexp(tolerance, &e);
    if (isinf(tolerance)) e = 0;  // frexp's exponent of infinity is taken as zero
    emin = e - 1;
  }
  set_params(z, kMinBits, kMaxBits, kMaxPrec, emin);
  return tolerance > 0.0 ? ldexp(1.0, emin) : 0.0;
}

zfp_mode zfp_stream_set_mode(zfp_stream *z, uint64 mode)
{
  if (!z) return zfp_mode_null;
  uint32_t a, b, c;
  int32_t d;
  if (!params_from_mode(mode, a, b, c, d)) return zfp_mode_null;
  set_params(z, a, b, c, d);
  return mode_of(z);
}

zfp_bool zfp_stream_set_params(zfp_stream *z, uint minbits, uint maxbits, uint maxprec, int32_t minexp)
{
  if (!z || !valid_params(minbits, maxbits, maxprec)) return zfp_false;
  set_params(z, minbits, maxbits, maxprec, minexp);
  return zfp_true;
}

zfp_exec_policy zfp_stream_execution(const zfp_stream *z) { return z ? z->exec.policy : zfp_exec_serial; }
uint zfp_stream_omp_threads(const zfp_stream *z) { return z && params(z) ? params(z)->threads : 0; }
uint zfp_stream_omp_chunk_size(const zfp_stream *z) { return z && params(z) ? params(z)->chunk_size : 0; }

zfp_bool zfp_stream_set_execution(zfp_stream *z, zfp_exec_policy policy)
{
  if (!z) return zfp_false;
  switch (policy) {
  case zfp_exec_serial:
    if (ExecParams *p = params(z, false)) p->threads = p->chunk_size = 0;
    z->exec.policy = policy;
    return zfp_true;
  case zfp_exec_omp:
  case zfp_exec_hip:
    z->exec.policy = policy;
    return zfp_true;
  default:
    return zfp_false;
  }
}

zfp_bool zfp_stream_set_omp_threads(zfp_stream *z, uint threads)
{
  if (!z) return zfp_false;
  ExecParams *p = params(z, true);
  if (!p) return zfp_false;
  p->threads = threads;
  p->chunk_size = 0;
  z->exec.policy = zfp_exec_omp;
  return zfp_true;
}

zfp_bool zfp_stream_set_omp_chunk_size(zfp_stream *z, uint chunk_size)
{
  if (!z) return zfp_false;
  ExecParams *p = params(z, true);
  if (!p) return zfp_false;
  p->chunk_size = chunk_size;
  z->exec.policy = zfp_exec_omp;
  return zfp_true;
}

size_t zfp_stream_flush(zfp_stream *z) { return z && z->stream ? z->stream->flush() : 0; }
size_t zfp_stream_align(zfp_stream *z) { return z && z->stream ? z->stream->align() : 0; }

// ---- fields (zfp-rs-ffi/src/field.rs)

static zfp_field *alloc_field(void *data, zfp_type t, size_t nx, size_t ny, size_t nz, size_t nw)
{
  zfp_field *f = (zfp_field *)calloc(1, sizeof(zfp_field));
  if (!f) return nullptr;
  f->type = t;
  f->nx = nx; f->ny = ny; f->nz = nz; f->nw = nw;
  f->data = data;
  return f;
}
zfp_field *zfp_field_alloc(void) { return alloc_field(nullptr, zfp_type_none, 0, 0, 0, 0); }
zfp_field *zfp_field_1d(void *d, zfp_type t, size_t nx) { return alloc_field(d, t, nx, 0, 0, 0); }
zfp_field *zfp_field_2d(void *d, zfp_type t, size_t nx, size_t ny) { return alloc_field(d, t, nx, ny, 0, 0); }
zfp_field *zfp_field_3d(void *d, zfp_type t, size_t nx, size_t ny, size_t nz)
{
  return alloc_field(d, t, nx, ny, nz, 0);
}
zfp_field *zfp_field_4d(void *d, zfp_type t, size_t nx, size_t ny, size_t nz, size_t nw)
{
  return alloc_field(d, t, nx, ny, nz, nw);
}
void zfp_field_free(zfp_field *f) { free(f); }
void *zfp_field_pointer(const zfp_field *f) { return f ? f->data : nullptr; }
void *zfp_field_begin(const zfp_field *f)
{
  Plan p;
  if (!f || !f->data || !make_plan(f, p)) return nullptr;
  return (char *)f->data + p.imin * (int64_t)p.esz;
}
zfp_type zfp_field_type(const zfp_field *f) { return f ? f->type : zfp_type_none; }
uint zfp_field_precision(const zfp_field *f)
{
  return f && valid_type(f->type) ? type_prec(f->type) : 0;
}
uint zfp_field_dimensionality(const zfp_field *f)
{
  Plan p;
  return f && make_plan(f, p) ? (uint)p.dims : 0;
}
size_t zfp_field_size(const zfp_field *f, size_t *dims)
{
  if (!f) return 0;
  if (dims) {
    dims[0] = f->nx; dims[1] = f->ny; dims[2] = f->nz; dims[3] = f->nw;
  }
  Plan p;
  if (!make_plan(f, p)) return 0;
  return (size_t)p.nelems;
}
size_t zfp_field_size_bytes(const zfp_field *f)
{
  Plan p;
  if (!f || !make_plan(f, p)) return 0;
  return (size_t)(p.imax - p.imin + 1) * p.esz;
}
size_t zfp_field_blocks(const zfp_field *f)
{
  Plan p;
  return f && make_plan(f, p) ? (size_t)p.nblocks : 0;
}
zfp_bool zfp_field_is_contiguous(const zfp_field *f)
{
  Plan p;
  if (!f || !make_plan(f, p)) return zfp_false;
  return contiguous(p) ? zfp_true : zfp_false;
}
zfp_bool zfp_hip_field_may_alias(const zfp_field *f)
{
  Plan p;
  return f && make_plan(f, p) && p.may_alias ? zfp_true : zfp_false;
}
zfp_bool zfp_field_stride(const zfp_field *f, ptrdiff_t *strides)
{
  if (!f) return zfp_false;
  if (strides) {
    strides[0] = f->sx; strides[1] = f->sy; strides[2] = f->sz; strides[3] = f->sw;
  }
  return zfp_field_is_contiguous(f) ? zfp_false : zfp_true;
}
uint64 zfp_field_metadata(const zfp_field *f)
{
  if (!f) return 0;
  uint64_t m;
  if (!valid_type(f->type) || !field_metadata(f, m)) return UINT64_MAX;
  return m;
}
void zfp_field_set_pointer(zfp_field *f, void *ptr)
{
  if (f) f->data = ptr;
}
zfp_type zfp_field_set_type(zfp_field *f, zfp_type t)
{
  if (!f) return zfp_type_int32;
  if (!valid_type(t)) return zfp_type_none;
  f->type = t;
  return t;
}
void zfp_field_set_size_1d(zfp_field *f, size_t nx)
{
  if (f) { f->nx = nx; f->ny = f->nz = f->nw = 0; }
}
void zfp_field_set_size_2d(zfp_field *f, size_t nx, size_t ny)
{
  if (f) { f->nx = nx; f->ny = ny; f->nz = f->nw = 0; }
}
void zfp_field_set_size_3d(zfp_field *f, size_t nx, size_t ny, size_t nz)
{
  if (f) { f->nx = nx; f->ny = ny; f->nz = nz; f->nw = 0; }
}
void zfp_field_set_size_4d(zfp_field *f, size_t nx, size_t ny, size_t nz, size_t nw)
{
  if (f) { f->nx = nx; f->ny = ny; f->nz = nz; f->nw = nw; }
}
void zfp_field_set_stride_1d(zfp_field *f, ptrdiff_t sx)
{
  if (f) { f->sx = sx; f->sy = f->sz = f->sw = 0; }
}
void zfp_field_set_stride_2d(zfp_field *f, ptrdiff_t sx, ptrdiff_t sy)
{
  if (f) { f->sx = sx; f->sy = sy; f->sz = f->sw = 0; }
}
void zfp_field_set_stride_3d(zfp_field *f, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz)
{
  if (f) { f->sx = sx; f->sy = sy; f->sz = sz; f->sw = 0; }
}
void zfp_field_set_stride_4d(zfp_field *f, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz, ptrdiff_t sw)
{
  if (f) { f->sx = sx; f->sy = sy; f->sz = sz; f->sw = sw; }
}
zfp_bool zfp_field_set_metadata(zfp_field *f, uint64 meta)
{
  if (!f) return zfp_false;
  return metadata_to_field(meta, f) ? zfp_true : zfp_false;
}

// ---- extensions (include/zfp_hip.h)

zfp_bool zfp_stream_set_rounding(zfp_stream *z, int rounding, zfp_bool tight)
{
  if (!z) return zfp_false;
  int r;
  if (rounding == ZFP_ROUND_NEVER) r = zx::ROUND_NEVER;
  else if (rounding == ZFP_ROUND_FIRST) r = zx::ROUND_FIRST;
  else if (rounding == ZFP_ROUND_LAST) r = zx::ROUND_LAST;
  else return zfp_false;
  ExecParams *p = params(z, true);
  if (!p) return zfp_false;
  p->rounding = r;
  p->tight = tight ? 1 : 0;
  return zfp_true;
}

int zfp_stream_rounding(const zfp_stream *z, zfp_bool *tight)
{
  const ExecParams *p = z ? params(z) : nullptr;
  if (tight) *tight = p && p->tight ? zfp_true : zfp_false;
  if (!p) return ZFP_ROUND_NEVER;
  return p->rounding == zx::ROUND_FIRST ? ZFP_ROUND_FIRST
         : p->rounding == zx::ROUND_LAST ? ZFP_ROUND_LAST : ZFP_ROUND_NEVER;
}

zfp_bool zfp_stream_set_hip_stream(zfp_stream *z, void *hs)
{
  if (!z) return zfp_false;
  ExecParams *p = params(z, true);
  if (!p) return zfp_false;
  p->hip = (hipStream_t)hs;
  return zfp_true;
}

size_t zfp_stream_index_export(const bitstream *s, uint32_t *lengths, size_t capacity, uint64 *start)
{
  if (!s || !s->index || !s->index_n) return 0;
  if (start) *start = s->index_start;
  if (lengths) {
    size_t n = std::min(capacity, s->index_n);
    if (hipMemcpy(lengths, s->index, n * 4, hipMemcpyDeviceToHost) != hipSuccess) {
      (void)hipGetLastError();
      return 0;
    }
  }
  return s->index_n;
}

zfp_bool zfp_stream_index_import(bitstream *s, const uint32_t *lengths, size_t count, uint64 start)
{
  if (!s || !lengths || !count) return zfp_false;
  if (s->index_cap < count) {
    drop_index(s);
    if (hipMalloc((void **)&s->index, count * 4) != hipSuccess) {
      (void)hipGetLastError();
      s->index = nullptr;
      return zfp_false;
    }
    s->index_cap = count;
  }
  if (hipMemcpy(s->index, lengths, count * 4, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipGetLastError();
    return zfp_false;
  }
  s->index_n = count;
  s->index_start = start;
  s->index_key = 0;  // an imported index matches on start and block count
  uint64_t bits = 0;
  for (size_t i = 0; i < count; i++) bits += lengths[i];
  s->index_bits = bits;
  return zfp_true;
}

zfp_bool zfp_stream_set_timing(zfp_stream *z, zfp_bool on)
{
  if (!z) return zfp_false;
  ExecParams *p = params(z, true);
  if (!p) return zfp_false;
  p->timing = on ? 1 : 0;
  return zfp_true;
}

double zfp_stream_kernel_ms(const zfp_stream *z, double *call_ms)
{
  const ExecParams *p = z ? params(z) : nullptr;
  if (call_ms) *call_ms = p && p->timing ? p->call_ms : -1.0;
  return p && p->timing ? p->kernel_ms : -1.0;
}

// Batched codec: count independent arrays coded with a few launches per phase
// (one per class of arrays sharing scalar type, dimensionality, reversible
// mode and fixed or variable placement) instead of several per array.  Each
// array's stream, cursor and return value are those of zfp_compress /
// zfp_decompress on it; arrays the batch cannot take (host-resident data or
// words, a stream used twice) are coded by those calls after the batch, in
// input order.  Returns the number of arrays coded without error.
size_t zfp_hip_compress_batch(zfp_stream *const *streams, const zfp_field *const *fields, size_t count,
                              size_t *sizes, uint flags)
{
  return batch_run(true, streams, (zfp_field *const *)fields, count, sizes, flags);
}

size_t zfp_hip_decompress_batch(zfp_stream *const *streams, zfp_field *const *fields, size_t count,
                                size_t *sizes, uint flags)
{
  return batch_run(false, streams, fields, count, sizes, flags);
}

zfp_hip_batch *zfp_hip_batch_plan(int encode, zfp_stream *const *streams, zfp_field *const *fields,
                                  size_t count, uint flags)
{
  std::unique_ptr<zfp_hip_batch> P = make_batch(encode != 0, streams, fields, count, flags, false);
  if (!P) return nullptr;
  BatchSet &S = P->S;
  BatchDev &D = P->D;
  if (P->lanes.init()) D.lanes = &P->lanes;
  const hipStream_t st = hip_stream_of(streams[S.items[0].input]);
  if (hipMemcpyAsync(D.dbuf, D.hbuf, S.jbytes + S.fbytes, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  // The tuner: blocks per workgroup of the class launches, then the slice of
  // arrays launched alone, each the fastest of a few timed runs of the whole
  // batch (runs that write what a run writes).
  const bool tune = !(flags & ZFP_HIP_BATCH_NO_TUNE);
  double best = -1;
  if (tune) {
    uint32_t pick = D.threads;
    for (uint32_t T : {64u, 128u, 256u}) {
      D.threads = T;
      const double ms = time_batch(*P, st, 2);
      if (ms < 0) return nullptr;
      if (best < 0 || ms < best) {
        best = ms;
        pick = T;
      }
    }
    D.threads = pick;
    bool solo = false;
    for (const BatchItem &it : S.items) solo = solo || it.cls >= kSoloClass;
    if (solo) {
      uint32_t spick = 0;
      for (uint32_t T : {64u, 128u, 256u}) {
        if (!slice_fits(S, T)) continue;
        D.slice = T;
        const double ms = time_batch(*P, st, 2);
        if (ms < 0) return nullptr;
        if (ms < best) {
          best = ms;
          spick = T;
        }
      }
      D.slice = spick;
    }
    // the tuning runs wrote each variable-rate stream's words and offset
    // index as a run does: keep the index's description, as a run would, so
    // that a decompress planned before the first run reads through it
    if (S.encode) {
      for (const BatchItem &it : S.items) {
        bitstream *s = streams[it.input]->stream;
        if (it.job.fixed || !it.job.index_out || s->index != it.job.index_out) continue;
        s->index_n = it.plan.nblocks;
        s->index_start = it.start;
        s->index_key = cfg_key(it.job.cfg, it.plan);
      }
    }
  }
  P->tuned_ms = best;
  // capture the device work once, on a private stream
  if (hipStreamCreateWithFlags(&P->capture, hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    P->capture = nullptr;
    return nullptr;
  }
  if (hipStreamBeginCapture(P->capture, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  const bool rec = batch_enqueue(S, D, false, P->capture);
  const hipError_t e = hipStreamEndCapture(P->capture, &P->graph);
  if (!rec || e != hipSuccess || !P->graph) {
    (void)hipGetLastError();
    return nullptr;
  }
  if (hipGraphInstantiate(&P->exec, P->graph, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGetLastError();
    P->exec = nullptr;
    return nullptr;
  }
  for (size_t i = 0; i < count; i++) restore_cursor(streams[i]->stream, P->cursors[i]);
  return P.release();
}

size_t zfp_hip_batch_run(zfp_hip_batch *P, size_t *sizes)
{
  if (!P || !sizes || !P->exec) return 0;
  const size_t count = P->zs.size();
  for (size_t i = 0; i < count; i++) sizes[i] = 0;
  // the arrays, streams and configurations must be the ones planned
  for (const BatchItem &it : P->S.items) {
    const zfp_stream *z = P->zs[it.input];
    const bitstream *s = z->stream;
    const bitstream &was = P->streams[it.input];
    if (!s || s->words != was.words || s->nwords != was.nwords) return 0;
    if (!same_cfg(cfg_of(z), it.job.cfg)) return 0;
    if (it.job.index_out && (s->index != it.job.index_out || s->index_cap < it.plan.nblocks)) return 0;
    if (it.job.index && s->index != it.job.index) return 0;
  }
  for (const BatchItem &it : P->S.items) {
    bitstream *s = P->zs[it.input]->stream;
    restore_cursor(s, P->cursors[it.input]);
    if (!P->S.encode) s->reset_overread();
  }
  const hipStream_t st = hip_stream_of(P->zs[P->S.items[0].input]);
  if (hipGraphLaunch(P->exec, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return batch_finish(P->S, P->D, P->zs.data(), sizes);
}

size_t zfp_hip_batch_count(const zfp_hip_batch *P) { return P ? P->zs.size() : 0; }

double zfp_hip_batch_tiling(const zfp_hip_batch *P, uint *threads, uint *slice)
{
  if (!P) return -1;
  if (threads) *threads = P->D.threads;
  if (slice) *slice = P->D.slice;
  return P->tuned_ms;
}

void zfp_hip_batch_free(zfp_hip_batch *P) { delete P; }

zfp_hip_batch *zfp_hip_batch_submit(int encode, zfp_stream *const *streams, zfp_field *const *fields,
                                    size_t count, uint flags, void *hip_stream)
{
  std::unique_ptr<zfp_hip_batch> P = make_batch(encode != 0, streams, fields, count, flags, true);
  if (!P) return nullptr;
  const hipStream_t st = hip_stream ? (hipStream_t)hip_stream : hip_stream_of(streams[P->S.items[0].input]);
  const bool good =
      hipEventCreateWithFlags(&P->done, hipEventDisableTiming) == hipSuccess &&
      hipMemcpyAsync(P->D.dbuf, P->D.hbuf, P->S.jbytes + P->S.fbytes, hipMemcpyHostToDevice, st) == hipSuccess &&
      batch_enqueue(P->S, P->D, false, st) && hipEventRecord(P->done, st) == hipSuccess;
  if (!good) {
    (void)hipGetLastError();
    if (P->done) (void)hipEventSynchronize(P->done);
    (void)hipStreamSynchronize(st);
    for (size_t i = 0; i < count; i++) restore_cursor(streams[i]->stream, P->cursors[i]);
    return nullptr;
  }
  return P.release();
}

zfp_bool zfp_hip_batch_ready(const zfp_hip_batch *P)
{
  if (!P || !P->done) return zfp_false;
  const hipError_t e = hipEventQuery(P->done);
  if (e == hipSuccess) return zfp_true;
  if (e != hipErrorNotReady) (void)hipGetLastError();
  return zfp_false;
}

size_t zfp_hip_batch_wait(zfp_hip_batch *P, size_t *sizes)
{
  if (!P || !P->done) return 0;  // (a planned batch is not a submission)
  std::unique_ptr<zfp_hip_batch> hold(P);
  const size_t count = P->zs.size();
  std::vector<size_t> scratch;
  if (!sizes) {
    scratch.resize(count);
    sizes = scratch.data();
  }
  for (size_t i = 0; i < count; i++) sizes[i] = 0;
  if (hipEventSynchronize(P->done) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return batch_finish(P->S, P->D, P->zs.data(), sizes);
}

const char *zfp_hip_last_error(void) { return g_last_error; }

long zfp_hip_fail_after(long n)
{
  return g_fail_after.exchange(n < 0 ? -1 : n);
}

int zfp_hip_device(void)
{
  int d = -1;
  if (hipGetDevice(&d) != hipSuccess) {
    (void)hipGetLastError();
    return -1;
  }
  return d;
}

}  // extern "C"
