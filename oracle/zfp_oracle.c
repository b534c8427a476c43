/*
 * zfp_oracle.c -- CPU restatement of zfp-rs's codec (test oracle and CPU
 * baseline only; see zfp_oracle.h for the scope and the reference files each
 * part follows).  Built by oracle/Makefile into oracle/liboracle.so.
 */
#include "zfp_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ======================= small helpers ======================= */

static inline int orc_clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }
static inline int orc_ctz64(uint64_t x) { return x ? __builtin_ctzll(x) : 64; }
static inline uint32_t orc_ctz64_or64(uint64_t x) { return x ? (uint32_t)__builtin_ctzll(x) : 64u; }
static inline uint64_t orc_low_mask(uint32_t n) { return n < 64 ? (1ull << n) - 1 : ~0ull; }
static inline uint32_t orc_sat_sub(uint32_t a, uint32_t b) { return a > b ? a - b : 0; }

/* with_maxbits (encode/core.rs:183-185): (maxprec + 1) * size > maxbits + 1 */
static inline int orc_with_maxbits(uint32_t maxbits, uint32_t maxprec, uint32_t size)
{
  return ((uint64_t)maxprec + 1) * (uint64_t)size > (uint64_t)maxbits + 1;
}

/* In-place transpose of the 64x64 bit matrix a[row] bit col (the recursive
 * quadrant swap of Hacker's Delight 7-3, as bitplane.rs:75-92 uses). */
static void orc_transpose64(uint64_t a[64])
{
  static const uint64_t masks[6] = {
    0x5555555555555555ull, 0x3333333333333333ull, 0x0f0f0f0f0f0f0f0full,
    0x00ff00ff00ff00ffull, 0x0000ffff0000ffffull, 0x00000000ffffffffull,
  };
  for (int s = 5; s >= 0; s--) {
    const int j = 1 << s;
    const uint64_t m = masks[s];
    for (int k = 0; k < 64; k += 2 * j)
      for (int i = k; i < k + j; i++) {
        uint64_t t = ((a[i] >> j) ^ a[i + j]) & m;
        a[i] ^= t << j;
        a[i + j] ^= t;
      }
  }
}

/* ======================= bit stream (bitstream/core.rs) ======================= */

void orc_stream_init(orc_stream *s, uint64_t *words, size_t nwords)
{
  memset(s, 0, sizeof(*s));
  s->words = words;
  s->nwords = nwords;
}

/* read_word_raw (core.rs:151-157): zero past the end */
static uint64_t read_word_raw(orc_stream *s)
{
  uint64_t w = s->word_pos < s->nwords ? s->words[s->word_pos] : 0;
  s->word_pos++;
  return w;
}

/* write_word_raw (core.rs:159-170): drop past the end and flag it */
static void write_word_raw(orc_stream *s, uint64_t v)
{
  if (s->word_pos < s->nwords)
    s->words[s->word_pos] = v;
  else
    s->overflowed = 1;
  s->word_pos++;
}

/* read_bits_impl (core.rs:192-229) */
uint64_t orc_read_bits(orc_stream *s, uint32_t n)
{
  if (n > 64) n = 64;
  uint64_t value = s->buffer;
  if (s->bits < n) {
    uint64_t w = read_word_raw(s);
    s->buffer = w;
    value += w << s->bits;
    s->bits += 64;
    s->bits -= n;
    if (s->bits == 0) {
      s->buffer = 0;
    } else {
      s->buffer >>= 64 - s->bits;
      if (n < 64) value &= (2ull << (n - 1)) - 1;
    }
  } else {
    s->bits -= n;
    s->buffer = n < 64 ? s->buffer >> n : 0;
    if (n < 64) value &= (1ull << n) - 1;
  }
  return value;
}

/* read_bit_impl (core.rs:235-247) */
uint32_t orc_read_bit(orc_stream *s)
{
  if (s->bits == 0) {
    s->buffer = read_word_raw(s);
    s->bits = 64;
  }
  s->bits--;
  uint32_t bit = (uint32_t)(s->buffer & 1u);
  s->buffer >>= 1;
  return bit;
}

/* loaded_since_seek (core.rs:384-386) */
static int loaded_since_seek(const orc_stream *s)
{
  size_t m = s->nwords > s->seek_word ? s->nwords : s->seek_word;
  return s->word_pos > m;
}

/* seek_read_impl (core.rs:265-284) */
void orc_rseek(orc_stream *s, uint64_t offset)
{
  uint32_t n = (uint32_t)(offset % 64);
  s->overread |= loaded_since_seek(s);
  s->word_pos = (size_t)(offset / 64);
  s->seek_word = s->word_pos;
  if (n) {
    uint64_t w = read_word_raw(s);
    s->buffer = w >> n;
    s->bits = 64 - n;
  } else {
    s->buffer = 0;
    s->bits = 0;
  }
}

/* write_bits_impl (core.rs:296-326) */
uint64_t orc_write_bits(orc_stream *s, uint64_t value, uint32_t n)
{
  if (n > 64) n = 64;
  uint32_t bits_old = s->bits;
  s->buffer += value << s->bits;
  s->bits += n;
  if (s->bits >= 64) {
    uint32_t after = s->bits - 64;
    uint64_t buf = s->buffer;
    s->bits = after;
    write_word_raw(s, buf);
    /* (value >> 1) >> (n - 1 - after) == value >> (64 - bits_old) */
    s->buffer = (value >> 1) >> (63 - bits_old);
  }
  if (s->bits < 64) s->buffer &= (1ull << s->bits) - 1;
  return n < 64 ? value >> n : 0;
}

/* write_bit_impl (core.rs:332-347) */
uint32_t orc_write_bit(orc_stream *s, uint32_t bit)
{
  s->buffer += (uint64_t)bit << s->bits;
  s->bits++;
  if (s->bits == 64) {
    write_word_raw(s, s->buffer);
    s->buffer = 0;
    s->bits = 0;
  }
  return bit;
}

/* seek_write_impl (core.rs:354-375) */
void orc_wseek(orc_stream *s, uint64_t offset)
{
  uint32_t n = (uint32_t)(offset % 64);
  s->word_pos = (size_t)(offset / 64);
  s->overflowed = 0;
  if (n) {
    s->buffer = s->word_pos < s->nwords ? s->words[s->word_pos] & ((1ull << n) - 1) : 0;
    s->bits = n;
  } else {
    s->buffer = 0;
    s->bits = 0;
  }
}

uint64_t orc_wtell(const orc_stream *s) { return (uint64_t)s->word_pos * 64u + s->bits; }
uint64_t orc_rtell(const orc_stream *s) { return (uint64_t)s->word_pos * 64u - s->bits; }

/* skip_impl (core.rs:439-443) */
void orc_skip(orc_stream *s, uint64_t n) { orc_rseek(s, orc_rtell(s) + n); }

/* pad_impl (core.rs:460-478) */
void orc_pad(orc_stream *s, uint64_t n)
{
  /* saturating, as pad_impl (core.rs:460-478) */
  uint64_t bits = n > UINT64_MAX - s->bits ? UINT64_MAX : (uint64_t)s->bits + n;
  while (bits >= 64) {
    if (s->word_pos >= s->nwords) {
      const size_t words = (size_t)(bits / 64);
      s->word_pos = words > SIZE_MAX - s->word_pos ? SIZE_MAX : s->word_pos + words;
      s->buffer = 0;
      s->overflowed = 1;
      bits %= 64;
      break;
    }
    write_word_raw(s, s->buffer);
    s->buffer = 0;
    bits -= 64;
  }
  s->bits = (uint32_t)bits;
}

/* align_impl (core.rs:524-530) */
uint32_t orc_align(orc_stream *s)
{
  uint32_t b = s->bits;
  if (b) orc_skip(s, b);
  return b;
}

/* flush_impl (core.rs:536-542) */
uint32_t orc_flush(orc_stream *s)
{
  uint32_t pad = (64 - s->bits) % 64;
  if (pad) orc_pad(s, pad);
  return pad;
}

void orc_rewind(orc_stream *s)
{
  s->word_pos = 0; s->buffer = 0; s->bits = 0;
  s->overflowed = 0; s->overread = 0; s->seek_word = 0;
}

int orc_overread(const orc_stream *s) { return s->overread || loaded_since_seek(s); }

/* reset_overread (core.rs:396-404) */
void orc_reset_overread(orc_stream *s)
{
  s->overread = 0;
  s->seek_word = s->bits > 0 ? (s->word_pos ? s->word_pos - 1 : 0) : s->word_pos;
}

/* BitWriter::put / put_zeros (core.rs:671-694): value has no bits above n */
static inline void orc_put(orc_stream *s, uint64_t value, uint32_t n)
{
  if (n == 0) return;
  s->buffer |= value << s->bits;
  uint32_t total = s->bits + n;
  if (total >= 64) {
    uint32_t spilled = s->bits;
    write_word_raw(s, s->buffer);
    s->bits = total - 64;
    s->buffer = (value >> 1) >> (63 - spilled);
  } else {
    s->bits = total;
  }
}

static void orc_put_zeros(orc_stream *s, uint32_t n)
{
  while (n > 64) { orc_put(s, 0, 64); n -= 64; }
  orc_put(s, 0, n);
}

/* BitReader::peek / consume / read (core.rs:584-614) */
static inline uint64_t orc_peek(const orc_stream *s)
{
  uint64_t next = s->word_pos < s->nwords ? s->words[s->word_pos] : 0;
  return s->buffer | (s->bits < 64 ? next << s->bits : 0);
}

static inline void orc_consume(orc_stream *s, uint32_t n)
{
  if (n <= s->bits) {
    s->buffer = n < 64 ? s->buffer >> n : 0;
    s->bits -= n;
  } else {
    uint64_t w = s->word_pos < s->nwords ? s->words[s->word_pos] : 0;
    s->word_pos++;
    s->bits = s->bits + 64 - n;
    s->buffer = (w >> 1) >> (63 - s->bits);
  }
}

static inline uint64_t orc_read_low(orc_stream *s, uint32_t n)
{
  uint64_t v = orc_peek(s);
  orc_consume(s, n);
  return n < 64 ? v & ((1ull << n) - 1) : v;
}

/* Bit planes of up to 256 coefficients as four words, coefficient i at bit
 * i % 64 of word i / 64 (bitplane.rs:161-218). */
/* the low m bits of plane x, word by word */
static void orc_put_plane(orc_stream *s, const uint64_t x[4], uint32_t m)
{
  for (int c = 0; m != 0; c++) {
    const uint32_t t = m < 64 ? m : 64;
    orc_put(s, x[c] & orc_low_mask(t), t);
    m -= t;
  }
}

/* read m bits into the low bits of x; the rest of x is zero */
static void orc_read_plane(orc_stream *s, uint64_t x[4], uint32_t m)
{
  for (int c = 0; c < 4; c++) {
    const uint32_t t = m < 64 ? m : 64;
    x[c] = t ? orc_read_low(s, t) : 0;
    m -= t;
  }
}

/* the lowest one-bit of x at or above bit n, or size when there is none */
static uint32_t orc_next_one(const uint64_t x[4], uint32_t n, uint32_t size)
{
  for (uint32_t c = n / 64; c * 64 < size; c++) {
    const uint64_t v = c == n / 64 ? x[c] >> (n % 64) << (n % 64) : x[c];
    if (v) return c * 64 + (uint32_t)orc_ctz64(v);
  }
  return size;
}

/* pad_to / skip_to (encode/core.rs:147-172) */
static uint32_t orc_pad_to(orc_stream *s, uint32_t bits, uint32_t minbits)
{
  if (bits < minbits) { orc_pad(s, minbits - bits); return minbits; }
  return bits;
}

static uint32_t orc_skip_to(orc_stream *s, uint32_t bits, uint32_t minbits)
{
  if (bits < minbits) { orc_skip(s, minbits - bits); return minbits; }
  return bits;
}

/* ======================= permutations (encode/core.rs:12-39) ======================= */

static const uint8_t PERM1[4] = {0, 1, 2, 3};
static const uint8_t PERM2[16] = {0, 1, 4, 5, 2, 8, 6, 9, 3, 12, 10, 7, 13, 11, 14, 15};
static const uint8_t PERM3[64] = {
  0, 1, 4, 16, 20, 17, 5, 2, 8, 32, 21, 6, 18, 24, 9, 33, 36, 3, 12, 48, 22, 25, 37, 40, 34, 10,
  7, 19, 28, 13, 49, 52, 41, 38, 26, 23, 29, 53, 11, 35, 44, 14, 50, 56, 42, 27, 39, 45, 30, 54,
  57, 60, 51, 15, 43, 46, 58, 61, 55, 31, 62, 59, 47, 63,
};

/* index(i,j,k,l) = i + 4 j + 16 k + 64 l (encode/core.rs:26-39) */
static const uint8_t PERM4[256] = {
  0, 1, 4, 16, 64, 5, 80, 17, 68, 65, 20, 2, 8, 32, 128, 84, 81, 69, 21, 6, 18, 66, 24, 72, 9,
  96, 33, 36, 129, 132, 144, 3, 12, 48, 192, 85, 82, 70, 22, 73, 25, 88, 37, 100, 97, 148, 145,
  133, 10, 160, 34, 136, 130, 40, 7, 19, 67, 28, 76, 13, 112, 49, 52, 193, 196, 208, 86, 89, 101,
  149, 161, 137, 41, 134, 38, 164, 26, 152, 146, 104, 98, 74, 83, 71, 23, 77, 29, 92, 53, 116,
  113, 212, 209, 197, 11, 35, 131, 44, 140, 14, 176, 50, 56, 194, 200, 224, 90, 165, 102, 153,
  150, 105, 168, 162, 138, 42, 87, 93, 117, 213, 27, 75, 99, 39, 135, 147, 108, 45, 141, 156, 30,
  78, 177, 180, 54, 114, 120, 57, 198, 210, 216, 201, 225, 228, 15, 240, 51, 204, 195, 60, 169,
  166, 154, 106, 91, 103, 151, 109, 157, 94, 181, 118, 121, 214, 217, 229, 163, 139, 43, 142, 46,
  172, 58, 184, 178, 232, 226, 202, 241, 205, 61, 199, 55, 244, 31, 220, 211, 124, 115, 79, 170,
  167, 155, 107, 158, 110, 173, 122, 185, 182, 233, 230, 218, 95, 245, 119, 221, 215, 125, 242,
  206, 62, 203, 59, 248, 47, 236, 227, 188, 179, 143, 171, 174, 186, 234, 246, 222, 126, 219,
  123, 249, 111, 237, 231, 189, 183, 159, 252, 243, 207, 63, 175, 250, 187, 238, 235, 190, 253,
  247, 223, 127, 254, 251, 239, 191, 255,
};

static const uint8_t *orc_perm(int size)
{
  return size == 4 ? PERM1 : size == 16 ? PERM2 : size == 64 ? PERM3 : PERM4;
}

/* ======================= templated codec ======================= */

#define W 32
#define TINT int32_t
#define TUINT uint32_t
#define TFLOAT float
#define NBMASK 0xaaaaaaaau
#define EBITS 8
#define EBIAS 127
#define PBITS 5
#define MINCAST (-97)
#define ORC_LDEXP ldexpf
#define FN(x) x##_32
#include "codec_tmpl.h"
#undef W
#undef TINT
#undef TUINT
#undef TFLOAT
#undef NBMASK
#undef EBITS
#undef EBIAS
#undef PBITS
#undef MINCAST
#undef ORC_LDEXP
#undef FN

#define W 64
#define TINT int64_t
#define TUINT uint64_t
#define TFLOAT double
#define NBMASK 0xaaaaaaaaaaaaaaaaull
#define EBITS 11
#define EBIAS 1023
#define PBITS 6
#define MINCAST (-961)
#define ORC_LDEXP ldexp
#define FN(x) x##_64
#include "codec_tmpl.h"

/* ======================= config helpers (config.rs) ======================= */

static uint32_t header_bits(int type)
{
  return type == ORC_FLOAT ? 9u : type == ORC_DOUBLE ? 12u : 0u;
}

int orc_is_reversible(const orc_config *c) { return c->minexp < -1074; }

/* rate_bits (config.rs:85-102) */
uint32_t orc_rate_bits(double rate, int type, int dims, int align)
{
  double values = (double)(1u << (2 * dims));
  double rounded = floor(values * rate + 0.5);
  if (!(rounded >= 0.0 && rounded <= 4294967295.0)) return 0;
  uint32_t bits = (uint32_t)rounded;
  uint32_t hb = header_bits(type);
  if (bits < hb) bits = hb;
  if (align) {
    uint64_t b = ((uint64_t)bits + 63) / 64 * 64;
    if (b > 0xffffffffull) return 0;
    bits = (uint32_t)b;
  }
  return bits;
}

/* block_bits (config.rs:670-686) */
uint32_t orc_block_bits(const orc_config *c, int type, int dims)
{
  uint32_t values = 1u << (2 * dims);
  uint32_t header;
  if (orc_is_reversible(c)) {
    header = type == ORC_INT32 ? 5 : type == ORC_INT64 ? 6 : type == ORC_FLOAT ? 15 : 19;
  } else {
    header = header_bits(type);
  }
  uint32_t prec = (type == ORC_INT32 || type == ORC_FLOAT) ? 32 : 64;
  uint32_t mp = c->maxprec < prec ? c->maxprec : prec;
  uint32_t most = header + values - 1 + values * mp;
  if (most > c->maxbits) most = c->maxbits;
  if (most < header) most = header;
  if (most < c->minbits) most = c->minbits;
  return most;
}

static int field_dims(const orc_field *f)
{
  if (f->nx == 0) return 0;
  if (f->ny == 0) return 1;
  if (f->nz == 0) return 2;
  return 3;
}

size_t orc_num_blocks(const orc_field *f)
{
  int d = field_dims(f);
  if (d == 0) return 0;
  size_t n = (f->nx + 3) / 4;
  if (d >= 2) n *= (f->ny + 3) / 4;
  if (d >= 3) n *= (f->nz + 3) / 4;
  return n;
}

/* maximum_size (config.rs:642-658) */
size_t orc_maximum_size(const orc_config *c, int type, const size_t dims[3])
{
  orc_field f = {type, dims[0], dims[1], dims[2], 0, 0, 0, NULL};
  int d = field_dims(&f);
  if (d == 0) return 0;
  uint64_t total = (uint64_t)orc_num_blocks(&f) * orc_block_bits(c, type, d) + 148;
  total = (total + 63) / 64 * 64;
  return (size_t)(total / 8);
}

/* ======================= block gather / scatter ======================= */

/* C pad_block as a closed form (strided.rs:716-751): position i of a run of
 * four with n real values reads from i, from 0 (i == 3), else from n - 1. */
static inline size_t pad_src(size_t i, size_t n) { return i < n ? i : (i == 3 ? 0 : n - 1); }

static void gather(void *out, const void *data, int type, int dims, const ptrdiff_t s[4],
                   const size_t n[4])
{
  const size_t esz = (type == ORC_INT32 || type == ORC_FLOAT) ? 4 : 8;
  const char *base = (const char *)data;
  int size = 1 << (2 * dims);
  for (int i = 0; i < size; i++) {
    size_t x = (size_t)i & 3, y = ((size_t)i >> 2) & 3, z = ((size_t)i >> 4) & 3;
    size_t w = ((size_t)i >> 6) & 3;
    ptrdiff_t off = (ptrdiff_t)pad_src(x, n[0]) * s[0];
    if (dims >= 2) off += (ptrdiff_t)pad_src(y, n[1]) * s[1];
    if (dims >= 3) off += (ptrdiff_t)pad_src(z, n[2]) * s[2];
    if (dims >= 4) off += (ptrdiff_t)pad_src(w, n[3]) * s[3];
    memcpy((char *)out + (size_t)i * esz, base + off * (ptrdiff_t)esz, esz);
  }
}

static void scatter(const void *in, void *data, int type, int dims, const ptrdiff_t s[4],
                    const size_t n[4])
{
  const size_t esz = (type == ORC_INT32 || type == ORC_FLOAT) ? 4 : 8;
  char *base = (char *)data;
  int size = 1 << (2 * dims);
  for (int i = 0; i < size; i++) {
    size_t x = (size_t)i & 3, y = ((size_t)i >> 2) & 3, z = ((size_t)i >> 4) & 3;
    size_t w = ((size_t)i >> 6) & 3;
    if (x >= n[0]) continue;
    if (dims >= 2 && y >= n[1]) continue;
    if (dims >= 3 && z >= n[2]) continue;
    if (dims >= 4 && w >= n[3]) continue;
    ptrdiff_t off = (ptrdiff_t)x * s[0];
    if (dims >= 2) off += (ptrdiff_t)y * s[1];
    if (dims >= 3) off += (ptrdiff_t)z * s[2];
    if (dims >= 4) off += (ptrdiff_t)w * s[3];
    memcpy(base + off * (ptrdiff_t)esz, (const char *)in + (size_t)i * esz, esz);
  }
}

/* ======================= block entry points ======================= */

uint32_t orc_encode_block(orc_stream *s, const orc_config *c, int type, int dims,
                          const void *data, const ptrdiff_t strides[4], const size_t lengths[4])
{
  uint64_t blk[256];
  int size = 1 << (2 * dims);
  uint64_t before = orc_wtell(s);
  gather(blk, data, type, dims, strides, lengths);
  int rev = orc_is_reversible(c);
  switch (type) {
  case ORC_INT32:
    if (rev) rev_encode_int_32(s, (const int32_t *)blk, size, c);
    else encode_int_block_32(s, (const int32_t *)blk, size, c->minbits, c->maxbits, c->maxprec,
                             c->rounding);
    break;
  case ORC_INT64:
    if (rev) rev_encode_int_64(s, (const int64_t *)blk, size, c);
    else encode_int_block_64(s, (const int64_t *)blk, size, c->minbits, c->maxbits, c->maxprec,
                             c->rounding);
    break;
  case ORC_FLOAT:
    if (rev) rev_encode_float_32(s, (const float *)blk, size, c);
    else encode_float_block_32(s, (const float *)blk, size, dims, c);
    break;
  case ORC_DOUBLE:
    if (rev) rev_encode_float_64(s, (const double *)blk, size, c);
    else encode_float_block_64(s, (const double *)blk, size, dims, c);
    break;
  }
  return (uint32_t)(orc_wtell(s) - before);
}

uint32_t orc_decode_block(orc_stream *s, const orc_config *c, int type, int dims, void *data,
                          const ptrdiff_t strides[4], const size_t lengths[4])
{
  uint64_t blk[256];
  int size = 1 << (2 * dims);
  uint64_t before = orc_rtell(s);
  int rev = orc_is_reversible(c);
  switch (type) {
  case ORC_INT32:
    if (rev) rev_decode_int_32(s, (int32_t *)blk, size, c);
    else decode_int_block_32(s, (int32_t *)blk, size, c->minbits, c->maxbits, c->maxprec,
                             c->rounding);
    break;
  case ORC_INT64:
    if (rev) rev_decode_int_64(s, (int64_t *)blk, size, c);
    else decode_int_block_64(s, (int64_t *)blk, size, c->minbits, c->maxbits, c->maxprec,
                             c->rounding);
    break;
  case ORC_FLOAT:
    if (rev) rev_decode_float_32(s, (float *)blk, size, c);
    else decode_float_block_32(s, (float *)blk, size, dims, c);
    break;
  case ORC_DOUBLE:
    if (rev) rev_decode_float_64(s, (double *)blk, size, c);
    else decode_float_block_64(s, (double *)blk, size, dims, c);
    break;
  }
  scatter(blk, data, type, dims, strides, lengths);
  return (uint32_t)(orc_rtell(s) - before);
}

/* ======================= field drivers ======================= */

typedef struct plan {
  int dims, type;
  size_t n[3], nb[3], nblocks;
  ptrdiff_t s[4]; /* s[3] unused: fields are 1-D to 3-D */
  size_t esz;
} plan;

static int make_plan(const orc_field *f, plan *p)
{
  p->dims = field_dims(f);
  if (p->dims == 0 || !f->data) return 0;
  if (f->type < ORC_INT32 || f->type > ORC_DOUBLE) return 0;
  p->type = f->type;
  p->esz = (f->type == ORC_INT32 || f->type == ORC_FLOAT) ? 4 : 8;
  p->n[0] = f->nx;
  p->n[1] = p->dims >= 2 ? f->ny : 1;
  p->n[2] = p->dims >= 3 ? f->nz : 1;
  /* effective strides (field.rs:580-590): 0 selects the natural stride */
  p->s[0] = f->sx ? f->sx : 1;
  p->s[1] = f->sy ? f->sy : (ptrdiff_t)p->n[0];
  p->s[2] = f->sz ? f->sz : (ptrdiff_t)(p->n[0] * p->n[1]);
  for (int a = 0; a < 3; a++) p->nb[a] = (p->n[a] + 3) / 4;
  p->nblocks = p->nb[0] * p->nb[1] * p->nb[2];
  return 1;
}

/* block geometry (field_plan.rs:128-200), relative to data = element [0,0,0] */
static const char *block_origin(const plan *p, const orc_field *f, size_t b, size_t len[4])
{
  size_t ix = b % p->nb[0], iy = (b / p->nb[0]) % p->nb[1], iz = b / (p->nb[0] * p->nb[1]);
  size_t c[3] = {ix, iy, iz};
  ptrdiff_t off = 0;
  for (int a = 0; a < 3; a++) {
    size_t o = 4 * c[a];
    size_t rem = p->n[a] - o;
    len[a] = rem < 4 ? rem : 4;
    off += (ptrdiff_t)o * p->s[a];
  }
  return (const char *)f->data + off * (ptrdiff_t)p->esz;
}

static void compress_range(orc_stream *s, const orc_config *c, const orc_field *f,
                           const plan *p, size_t b0, size_t b1)
{
  orc_config cc = *c;
  for (size_t b = b0; b < b1; b++) {
    size_t len[4] = {1, 1, 1, 1};
    const char *o = block_origin(p, f, b, len);
    orc_encode_block(s, &cc, p->type, p->dims, o, p->s, len);
  }
}

static void decompress_range(orc_stream *s, const orc_config *c, const orc_field *f,
                             const plan *p, size_t b0, size_t b1)
{
  for (size_t b = b0; b < b1; b++) {
    size_t len[4] = {1, 1, 1, 1};
    char *o = (char *)block_origin(p, f, b, len);
    orc_decode_block(s, c, p->type, p->dims, o, p->s, len);
  }
}

/* compress + finish (compress.rs:22-53) */
size_t orc_compress(orc_stream *s, const orc_config *c, const orc_field *f, int *err)
{
  plan p;
  *err = 0;
  if (!make_plan(f, &p)) { *err = 1; return 0; }
  compress_range(s, c, f, &p, 0, p.nblocks);
  orc_flush(s);
  if (s->overflowed) { *err = 2; return 0; }
  return s->word_pos * 8;
}

/* decompress + finish (decompress.rs:27-74) */
size_t orc_decompress(orc_stream *s, const orc_config *c, const orc_field *f, int *err)
{
  plan p;
  *err = 0;
  if (!make_plan(f, &p)) { *err = 1; return 0; }
  orc_reset_overread(s);
  decompress_range(s, c, f, &p, 0, p.nblocks);
  orc_align(s);
  size_t required = s->word_pos * 8;
  if (orc_overread(s)) { *err = 3; return 0; }
  return required;
}

int orc_block_offsets(orc_stream *s, const orc_config *c, const orc_field *f, uint64_t *offsets)
{
  plan p;
  if (!make_plan(f, &p)) return 0;
  /* decode into a scratch copy of one block only */
  for (size_t b = 0; b < p.nblocks; b++) {
    uint64_t blk[64];
    ptrdiff_t unit[4] = {1, 4, 16, 64};
    size_t full[4] = {4, 4, 4, 4};
    offsets[b] = orc_rtell(s);
    orc_decode_block(s, c, p.type, p.dims, blk, unit, full);
  }
  offsets[p.nblocks] = orc_rtell(s);
  return 1;
}

/* ---------- parallel drivers (compress.rs:132-236, decompress.rs:205-330) ---------- */

/* chunk_ranges (execution.rs:119-148) with chunk_size 0: one chunk per thread */
static size_t chunk_start(size_t blocks, size_t chunks, size_t i)
{
  return (size_t)(((unsigned __int128)blocks * i) / chunks);
}

typedef struct cjob {
  const orc_config *c;
  const orc_field *f;
  const plan *p;
  size_t b0, b1;
  uint64_t *words;
  size_t nwords;
  uint64_t bits;
  uint64_t start_bit;
  int overread;
} cjob;

static void *compress_worker(void *arg)
{
  cjob *j = (cjob *)arg;
  orc_stream s;
  orc_stream_init(&s, j->words, j->nwords);
  compress_range(&s, j->c, j->f, j->p, j->b0, j->b1);
  j->bits = orc_wtell(&s);
  orc_flush(&s);
  return NULL;
}

size_t orc_compress_threads(orc_stream *s, const orc_config *c, const orc_field *f, int threads,
                            int *err)
{
  plan p;
  *err = 0;
  if (!make_plan(f, &p)) { *err = 1; return 0; }
  if (threads <= 1 || p.nblocks < 2) return orc_compress(s, c, f, err);
  size_t chunks = (size_t)threads < p.nblocks ? (size_t)threads : p.nblocks;
  cjob *jobs = calloc(chunks, sizeof(cjob));
  pthread_t *tids = calloc(chunks, sizeof(pthread_t));
  uint32_t bb = orc_block_bits(c, p.type, p.dims);
  for (size_t i = 0; i < chunks; i++) {
    jobs[i].c = c; jobs[i].f = f; jobs[i].p = &p;
    jobs[i].b0 = chunk_start(p.nblocks, chunks, i);
    jobs[i].b1 = chunk_start(p.nblocks, chunks, i + 1);
    jobs[i].nwords = (size_t)(((uint64_t)(jobs[i].b1 - jobs[i].b0) * bb) / 64 + 1);
    jobs[i].words = calloc(jobs[i].nwords, 8);
    pthread_create(&tids[i], NULL, compress_worker, &jobs[i]);
  }
  for (size_t i = 0; i < chunks; i++) pthread_join(tids[i], NULL);
  /* append_bits (compress.rs:205-216) */
  for (size_t i = 0; i < chunks; i++) {
    uint64_t full = jobs[i].bits / 64;
    for (uint64_t w = 0; w < full; w++) orc_put(s, jobs[i].words[w], 64);
    uint32_t rest = (uint32_t)(jobs[i].bits % 64);
    if (rest) orc_put(s, jobs[i].words[full] & ((1ull << rest) - 1), rest);
    free(jobs[i].words);
  }
  free(jobs);
  free(tids);
  orc_flush(s);
  if (s->overflowed) { *err = 2; return 0; }
  return s->word_pos * 8;
}

static void *decompress_worker(void *arg)
{
  cjob *j = (cjob *)arg;
  orc_stream s;
  orc_stream_init(&s, j->words, j->nwords);
  orc_rseek(&s, j->start_bit);
  decompress_range(&s, j->c, j->f, j->p, j->b0, j->b1);
  j->overread = orc_overread(&s);
  return NULL;
}

size_t orc_decompress_threads(orc_stream *s, const orc_config *c, const orc_field *f, int threads,
                              int *err)
{
  plan p;
  *err = 0;
  if (!make_plan(f, &p)) { *err = 1; return 0; }
  int fixed = c->minbits == c->maxbits && c->maxbits >= 1 && c->maxbits <= 16658 &&
              c->maxprec >= 64 && c->minexp == -1074;
  if (threads <= 1 || p.nblocks < 2 || !fixed ||
      orc_block_bits(c, p.type, p.dims) != c->maxbits)
    return orc_decompress(s, c, f, err);
  uint64_t start = orc_rtell(s);
  uint64_t end = start + (uint64_t)p.nblocks * c->maxbits;
  size_t chunks = (size_t)threads < p.nblocks ? (size_t)threads : p.nblocks;
  cjob *jobs = calloc(chunks, sizeof(cjob));
  pthread_t *tids = calloc(chunks, sizeof(pthread_t));
  for (size_t i = 0; i < chunks; i++) {
    jobs[i].c = c; jobs[i].f = f; jobs[i].p = &p;
    jobs[i].b0 = chunk_start(p.nblocks, chunks, i);
    jobs[i].b1 = chunk_start(p.nblocks, chunks, i + 1);
    jobs[i].words = s->words;
    jobs[i].nwords = s->nwords;
    jobs[i].start_bit = start + (uint64_t)jobs[i].b0 * c->maxbits;
    pthread_create(&tids[i], NULL, decompress_worker, &jobs[i]);
  }
  int over = 0;
  for (size_t i = 0; i < chunks; i++) {
    pthread_join(tids[i], NULL);
    over |= jobs[i].overread;
  }
  free(jobs);
  free(tids);
  /* park (decompress.rs:190-196), then finish */
  orc_reset_overread(s);
  orc_rseek(s, end);
  if (over) s->overread = 1;
  orc_align(s);
  size_t required = s->word_pos * 8;
  if (orc_overread(s)) { *err = 3; return 0; }
  return required;
}

/* ======================= test hooks ======================= */
/* The per-block primitives, exported for the closed-form checks of the test
 * suite (exponent vs frexp, exact casts, transform matrices, rounding
 * biases, plane transpose).  size is the block's value count (4, 16, 64). */

int orc_exponent_block_f32(const float *f, int size) { return exponent_block_32(f, size); }
int orc_exponent_block_f64(const double *f, int size) { return exponent_block_64(f, size); }
void orc_fwd_cast_f32(int32_t *i, const float *f, int size, int emax) { fwd_cast_32(i, f, size, emax); }
void orc_fwd_cast_f64(int64_t *i, const double *f, int size, int emax) { fwd_cast_64(i, f, size, emax); }
void orc_inv_cast_f32(const int32_t *i, float *f, int size, int emax) { inv_cast_32(i, f, size, emax); }
void orc_inv_cast_f64(const int64_t *i, double *f, int size, int emax) { inv_cast_64(i, f, size, emax); }
int64_t orc_truncate_f64(double v) { return truncate_64(v); }
uint32_t orc_precision_f(int emax, uint32_t maxprec, int32_t minexp, int dims, int tight)
{
  return precision_32(emax, maxprec, minexp, dims, tight);
}

/* which: 0 fwd_xform, 1 inv_xform, 2 rev_fwd_xform, 3 rev_inv_xform */
void orc_xform_i32(int32_t *p, int size, int which)
{
  switch (which) {
  case 0: fwd_xform_32(p, size, fwd_lift_32); break;
  case 1: inv_xform_32(p, size, inv_lift_32); break;
  case 2: fwd_xform_32(p, size, rev_fwd_lift_32); break;
  default: inv_xform_32(p, size, rev_inv_lift_32); break;
  }
}
void orc_xform_i64(int64_t *p, int size, int which)
{
  switch (which) {
  case 0: fwd_xform_64(p, size, fwd_lift_64); break;
  case 1: inv_xform_64(p, size, inv_lift_64); break;
  case 2: fwd_xform_64(p, size, rev_fwd_lift_64); break;
  default: inv_xform_64(p, size, rev_inv_lift_64); break;
  }
}

void orc_fwd_round_i32(int32_t *b, int size, uint32_t maxprec) { fwd_round_32(b, size, maxprec); }
void orc_fwd_round_i64(int64_t *b, int size, uint32_t maxprec) { fwd_round_64(b, size, maxprec); }
void orc_inv_round_u32(uint32_t *d, int size, uint32_t m, uint32_t prec) { inv_round_32(d, size, m, prec); }
void orc_inv_round_u64(uint64_t *d, int size, uint32_t m, uint32_t prec) { inv_round_64(d, size, m, prec); }

/* planes[k] bit i = bit k of u[i] (u given as 64-bit values) */
void orc_to_planes(const uint64_t *u, int size, uint64_t planes[64])
{
  uint64_t p4[64][4];
  to_planes_64(u, size, p4);
  for (int k = 0; k < 64; k++) planes[k] = p4[k][0];
}
