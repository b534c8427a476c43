/*
 * codec_tmpl.h -- the oracle's block codec, instantiated once per integer
 * width by zfp_oracle.c (test infrastructure only; see zfp_oracle.h).
 *
 * Template parameters (macros set by the includer):
 *   W        32 or 64 (intprec P of the block)
 *   TINT     int32_t / int64_t      TUINT  uint32_t / uint64_t
 *   TFLOAT   float / double         NBMASK 0xaaaa... of width W
 *   EBITS    8 / 11                 EBIAS  127 / 1023
 *   PBITS    5 / 6                  MINCAST -97 / -961
 *   FN(x)    x##_32 / x##_64
 */

/* ---------------- lifting (src/codec/transform.rs:74-142) ---------------- */

static inline TINT FN(wadd)(TINT a, TINT b) { return (TINT)((TUINT)a + (TUINT)b); }
static inline TINT FN(wsub)(TINT a, TINT b) { return (TINT)((TUINT)a - (TUINT)b); }
/* arithmetic shift right by one (gcc shifts signed values arithmetically) */
static inline TINT FN(half)(TINT a) { return a >> 1; }

/* transform.rs:74-86 */
static void FN(fwd_lift)(TINT *p, int s)
{
  TINT x = p[0], y = p[s], z = p[2 * s], w = p[3 * s];
  x = FN(half)(FN(wadd)(x, w));
  w = FN(wsub)(w, x);
  z = FN(half)(FN(wadd)(z, y));
  y = FN(wsub)(y, z);
  x = FN(half)(FN(wadd)(x, z));
  z = FN(wsub)(z, x);
  w = FN(half)(FN(wadd)(w, y));
  y = FN(wsub)(y, w);
  w = FN(wadd)(w, FN(half)(y));
  y = FN(wsub)(y, FN(half)(w));
  p[0] = x; p[s] = y; p[2 * s] = z; p[3 * s] = w;
}

/* transform.rs:98-110 */
static void FN(inv_lift)(TINT *p, int s)
{
  TINT x = p[0], y = p[s], z = p[2 * s], w = p[3 * s];
  y = FN(wadd)(y, FN(half)(w));
  w = FN(wsub)(w, FN(half)(y));
  y = FN(wadd)(y, w);
  w = FN(wsub)(w, FN(wsub)(y, w));
  z = FN(wadd)(z, x);
  x = FN(wsub)(x, FN(wsub)(z, x));
  y = FN(wadd)(y, z);
  z = FN(wsub)(z, FN(wsub)(y, z));
  w = FN(wadd)(w, x);
  x = FN(wsub)(x, FN(wsub)(w, x));
  p[0] = x; p[s] = y; p[2 * s] = z; p[3 * s] = w;
}

/* transform.rs:118-126 */
static void FN(rev_fwd_lift)(TINT *p, int s)
{
  TINT x = p[0], y = p[s], z = p[2 * s], w = p[3 * s];
  w = FN(wsub)(w, z);
  z = FN(wsub)(z, y);
  y = FN(wsub)(y, x);
  w = FN(wsub)(w, z);
  z = FN(wsub)(z, y);
  w = FN(wsub)(w, z);
  p[s] = y; p[2 * s] = z; p[3 * s] = w;
}

/* transform.rs:134-142 */
static void FN(rev_inv_lift)(TINT *p, int s)
{
  TINT x = p[0], y = p[s], z = p[2 * s], w = p[3 * s];
  w = FN(wadd)(w, z);
  z = FN(wadd)(z, y);
  w = FN(wadd)(w, z);
  y = FN(wadd)(y, x);
  z = FN(wadd)(z, y);
  w = FN(wadd)(w, z);
  p[s] = y; p[2 * s] = z; p[3 * s] = w;
}

/* Lift every line along the axis of stride s (transform.rs:151-170). */
static void FN(lift_axis)(TINT *p, int size, int s, void (*lift)(TINT *, int))
{
  for (int group = 0; group < size; group += 4 * s)
    for (int i = group; i < group + s; i++)
      lift(p + i, s);
}

/* forward: x, then y, then z, then w (transform.rs:173-216) */
static void FN(fwd_xform)(TINT *p, int size, void (*lift)(TINT *, int))
{
  FN(lift_axis)(p, size, 1, lift);
  if (size >= 16) FN(lift_axis)(p, size, 4, lift);
  if (size >= 64) FN(lift_axis)(p, size, 16, lift);
  if (size >= 256) FN(lift_axis)(p, size, 64, lift);
}

/* inverse: w, then z, then y, then x (transform.rs:193-216) */
static void FN(inv_xform)(TINT *p, int size, void (*lift)(TINT *, int))
{
  if (size >= 256) FN(lift_axis)(p, size, 64, lift);
  if (size >= 64) FN(lift_axis)(p, size, 16, lift);
  if (size >= 16) FN(lift_axis)(p, size, 4, lift);
  FN(lift_axis)(p, size, 1, lift);
}

/* ---------------- negabinary, order, rounding ---------------- */

/* encode/core.rs:82-91 */
static inline TUINT FN(int2uint)(TINT x) { return ((TUINT)x + NBMASK) ^ NBMASK; }
/* decode/core.rs:25-33 */
static inline TINT FN(uint2int)(TUINT x) { return (TINT)((x ^ NBMASK) - NBMASK); }

/* encode/core.rs:98-109 */
static void FN(fwd_order)(TUINT *u, const TINT *iblock, const uint8_t *perm, int size)
{
  for (int j = 0; j < size; j++) u[j] = FN(int2uint)(iblock[perm[j]]);
}

/* decode/core.rs:40-51 */
static void FN(inv_order)(const TUINT *u, TINT *iblock, const uint8_t *perm, int size)
{
  for (int j = 0; j < size; j++) iblock[perm[j]] = FN(uint2int)(u[j]);
}

/* ROUND_FIRST bias (encode/core.rs:53-74); NBMASK is unsigned, shifts logical */
static void FN(fwd_round)(TINT *iblock, int size, uint32_t maxprec)
{
  if (maxprec < W) {
    TINT bias = (TINT)((NBMASK >> 2) >> maxprec);
    if (maxprec & 1u) {
      for (int i = 0; i < size; i++) iblock[i] = FN(wadd)(iblock[i], bias);
    } else {
      for (int i = 0; i < size; i++) iblock[i] = FN(wsub)(iblock[i], bias);
    }
  }
}

/* ROUND_LAST bias (decode/core.rs:62-82) */
static void FN(inv_round)(TUINT *data, int size, uint32_t m, uint32_t prec)
{
  if (prec < W - 1) {
    uint32_t mm = m < (uint32_t)size ? m : (uint32_t)size;
    TUINT hi = (NBMASK >> 2) >> prec;
    TUINT lo = (NBMASK >> 1) >> prec;
    for (uint32_t i = 0; i < mm; i++) data[i] += hi;
    for (int i = (int)mm; i < size; i++) data[i] += lo;
  }
}

/* ---------------- bit-plane coder (src/codec/bitplane.rs) ---------------- */

/* top(): planes up to and including the highest one-bit (bitplane.rs:271-273) */
static uint32_t FN(top)(const TUINT *u, int size)
{
  TUINT acc = 0;
  for (int i = 0; i < size; i++) acc |= u[i];
  return acc ? (uint32_t)W - (uint32_t)orc_clz64((uint64_t)acc) + (64u - W) : 0u;
}

/* planes[k][c] bit i = bit k of coefficient 64 c + i (a 4-D block's 256
 * coefficients span four words, bitplane.rs:161-218) */
static void FN(to_planes)(const TUINT *u, int size, uint64_t planes[64][4])
{
  for (int c = 0; c < (size + 63) / 64; c++) {
    uint64_t a[64];
    for (int i = 0; i < 64; i++) a[i] = 64 * c + i < size ? (uint64_t)u[64 * c + i] : 0;
    orc_transpose64(a);
    for (int k = 0; k < 64; k++) planes[k][c] = a[k];
  }
}

static void FN(from_planes)(const uint64_t planes[64][4], int size, TUINT *u)
{
  for (int c = 0; c < (size + 63) / 64; c++) {
    uint64_t a[64];
    for (int k = 0; k < 64; k++) a[k] = planes[k][c];
    orc_transpose64(a);
    for (int i = 0; i < 64 && 64 * c + i < size; i++) u[64 * c + i] = (TUINT)a[i];
  }
}

/* encode_ints + encode_planes (bitplane.rs:509-610) */
static uint32_t FN(encode_ints)(orc_stream *s, uint32_t maxbits, uint32_t maxprec,
                                const TUINT *u, int size)
{
  const uint32_t budget = orc_with_maxbits(maxbits, maxprec, (uint32_t)size) ? maxbits : UINT32_MAX;
  const uint32_t kmin = W > maxprec ? W - maxprec : 0;
  const uint32_t usize_ = (uint32_t)size;
  uint64_t planes[64][4];
  FN(to_planes)(u, size, planes);
  uint32_t top = FN(top)(u, size);
  uint32_t bits = budget;
  uint32_t n = 0;
  uint32_t empty = W - (top > kmin ? top : kmin);
  if (empty > bits) empty = bits;
  orc_put_zeros(s, empty);
  bits -= empty;
  uint32_t k = W - empty;
  while (bits != 0 && k > kmin) {
    if (n == usize_) {
      while (bits > usize_ && k > kmin) {
        k--;
        orc_put_plane(s, planes[k], usize_);
        bits -= usize_;
      }
      if (k == kmin) break;
    }
    k--;
    const uint64_t *x = planes[k];
    uint32_t m = n < bits ? n : bits;
    bits -= m;
    orc_put_plane(s, x, m);
    while (bits != 0 && n < usize_) {
      const uint32_t p = orc_next_one(x, n, usize_);
      if (p == usize_) {
        orc_put(s, 0, 1);
        bits -= 1;
        break;
      }
      uint32_t zeros = p - n;
      int last = p == usize_ - 1;
      uint32_t len = zeros + 2 - (uint32_t)last;
      uint32_t take = len < bits ? len : bits;
      if (take <= 64) {
        uint64_t code = (last || zeros >= 63) ? 1u : (1u | (2ull << zeros));
        orc_put(s, code & orc_low_mask(take), take);
      } else {
        orc_put(s, 1, 1);
        orc_put_zeros(s, zeros < take - 1 ? zeros : take - 1);
        if (take == len && !last) orc_put(s, 1, 1);
      }
      bits -= take;
      n = p + 1;
    }
  }
  return budget - bits;
}

/* decode_ints + decode_planes + round_state (bitplane.rs:626-790).
 * Returns bits read; *zero_out is the "skip the inverse transform" flag. */
static uint32_t FN(decode_ints)(orc_stream *s, uint32_t maxbits, uint32_t maxprec,
                                int round_last, TUINT *u, int size, int *zero_out)
{
  const int constrained = orc_with_maxbits(maxbits, maxprec, (uint32_t)size);
  const uint32_t budget = constrained ? maxbits : UINT32_MAX;
  const uint32_t kmin = W > maxprec ? W - maxprec : 0;
  const uint32_t usize_ = (uint32_t)size;
  uint64_t planes[64][4];
  memset(planes, 0, sizeof(planes));
  uint32_t bits = budget, n = 0, m = 0, k = W, low = W;
  while (bits != 0) {
    m = 0;
    if (k <= kmin) { k = k - 1u; break; }
    if (n == 0) {
      uint32_t empty = orc_ctz64_or64(orc_peek(s));
      if (empty > k - kmin) empty = k - kmin;
      if (empty > bits) empty = bits;
      if (empty != 0) {
        orc_consume(s, empty);
        bits -= empty;
        k -= empty;
        m = usize_;
        continue;
      }
    } else if (n == usize_) {
      while (bits > usize_ && k > kmin) {
        k--;
        orc_read_plane(s, planes[k], usize_);
        bits -= usize_;
        low = k;
      }
      if (k == kmin) { k = k - 1u; break; }
    }
    k--;
    m = n < bits ? n : bits;
    bits -= m;
    uint64_t x[4];
    orc_read_plane(s, x, m);
    while (bits != 0 && n < usize_) {
      uint64_t ahead = orc_peek(s);
      bits -= 1;
      if ((ahead & 1u) == 0) {
        orc_consume(s, 1);
        m = usize_;
        break;
      }
      uint32_t limit = usize_ - 1 - n;
      if (limit > bits) limit = bits;
      uint32_t zeros = orc_ctz64_or64(ahead >> 1);
      uint32_t lim63 = limit < 63 ? limit : 63;
      if (zeros < lim63) {
        orc_consume(s, zeros + 2);
        bits -= zeros + 1;
        n += zeros;
      } else if (limit <= 63) {
        orc_consume(s, limit + 1);
        bits -= limit;
        n += limit;
      } else {
        /* runs longer than the lookahead only occur for 4-D blocks */
        orc_consume(s, 64);
        bits -= 63;
        n += 63;
        uint32_t lim = limit - 63;
        while (lim != 0) {
          uint32_t z = orc_ctz64_or64(orc_peek(s));
          uint32_t span = lim < 64 ? lim : 64;
          if (z < span) { orc_consume(s, z + 1); bits -= z + 1; n += z; break; }
          orc_consume(s, span); bits -= span; n += span; lim -= span;
        }
      }
      x[n >> 6] |= 1ull << (n & 63);
      n += 1;
      m = n;
    }
    memcpy(planes[k], x, sizeof(x));
    low = k;
  }
  if (n == 0) low = W;
  uint32_t prec = (uint32_t)W - k; /* wrapping, as the reference's u32 */
  int zero = low == W;
  if (zero) {
    for (int i = 0; i < size; i++) u[i] = 0;
  } else {
    FN(from_planes)(planes, size, u);
  }
  if (round_last) {
    uint32_t rm = m, rp = prec;
    if (!constrained) { rm = 0; rp = (uint32_t)W - (kmin - 1u); }
    FN(inv_round)(u, size, rm, rp);
  }
  *zero_out = zero && !round_last;
  return budget - bits;
}

/* ---------------- integer block (encode/integer.rs:50-93, decode/core.rs:97-131) --- */

static uint32_t FN(encode_int_block)(orc_stream *s, const TINT *iblock_in, int size,
                                     uint32_t minbits, uint32_t maxbits, uint32_t maxprec,
                                     int rounding)
{
  TINT block[256];
  TUINT u[256];
  memcpy(block, iblock_in, sizeof(TINT) * (size_t)size);
  FN(fwd_xform)(block, size, FN(fwd_lift));
  if (rounding == ORC_ROUND_FIRST) FN(fwd_round)(block, size, maxprec);
  FN(fwd_order)(u, block, orc_perm(size), size);
  uint32_t bits = FN(encode_ints)(s, maxbits, maxprec, u, size);
  return orc_pad_to(s, bits, minbits);
}

static uint32_t FN(decode_int_block)(orc_stream *s, TINT *iblock, int size, uint32_t minbits,
                                     uint32_t maxbits, uint32_t maxprec, int rounding)
{
  TUINT u[256];
  int zero;
  uint32_t bits = FN(decode_ints)(s, maxbits, maxprec, rounding == ORC_ROUND_LAST, u, size, &zero);
  orc_skip_to(s, bits, minbits);
  for (int i = 0; i < size; i++) iblock[i] = 0;
  if (!zero) {
    FN(inv_order)(u, iblock, orc_perm(size), size);
    FN(inv_xform)(iblock, size, FN(inv_lift));
  }
  return bits < minbits ? minbits : bits;
}

/* ---------------- float block (encode/core.rs:251-423, encode/float.rs, decode/core.rs) -- */

/* frexp exponent of max |x| over non-NaN values (encode/core.rs:280-307) */
static int FN(exponent_block)(const TFLOAT *f, int size)
{
  const TUINT absmask = ~((TUINT)1 << (W - 1));
  const TUINT inf = (TUINT)(((TUINT)1 << EBITS) - 1) << (W - 1 - EBITS);
  TUINT max = 0;
  for (int i = 0; i < size; i++) {
    TUINT b;
    memcpy(&b, &f[i], sizeof(b));
    b &= absmask;
    if (b > inf) b = 0; /* NaN */
    if (b > max) max = b;
  }
  if (max == 0) return -EBIAS;
  uint32_t biased = (uint32_t)(max >> (W - 1 - EBITS));
  if (biased == (1u << EBITS) - 1) return 0;
  return (int)biased - (EBIAS - 1);
}

/* exact power of two as TFLOAT: ldexp(1, e) with round-to-nearest-even underflow */
static TFLOAT FN(pow2)(int e)
{
  return (TFLOAT)ORC_LDEXP((TFLOAT)1, e);
}

/* x86-64 truncation: out of range or NaN -> minimum integer (encode/core.rs:391-423) */
static TINT FN(truncate)(TFLOAT v)
{
  const TFLOAT lim = (TFLOAT)ORC_LDEXP((TFLOAT)1, W - 1);
  if (v >= -lim && v < lim) return (TINT)v;
  return (TINT)((TUINT)1 << (W - 1));
}

/* fwd_cast with the exact two-step path for tiny blocks (encode/core.rs:323-389) */
static void FN(fwd_cast)(TINT *iblock, const TFLOAT *f, int size, int emax)
{
  if (emax < MINCAST) {
    const TFLOAT two64 = (TFLOAT)18446744073709551616.0;
    TFLOAT sc = FN(pow2)((W - 2) - 64 - emax);
    for (int i = 0; i < size; i++) {
      volatile TFLOAT t = f[i] * two64;
      iblock[i] = FN(truncate)(t * sc);
    }
    return;
  }
  TFLOAT sc = FN(pow2)((W - 2) - emax);
  for (int i = 0; i < size; i++) {
    volatile TFLOAT t = sc * f[i];
    iblock[i] = FN(truncate)(t);
  }
}

/* inv_cast (decode/core.rs:162-175) */
static void FN(inv_cast)(const TINT *iblock, TFLOAT *f, int size, int emax)
{
  TFLOAT sc = FN(pow2)(emax - (W - 2));
  for (int i = 0; i < size; i++) {
    volatile TFLOAT t = (TFLOAT)iblock[i];
    f[i] = sc * t;
  }
}

/* precision_f (encode/core.rs:231-244) */
static uint32_t FN(precision)(int emax, uint32_t maxprec, int32_t minexp, int dims, int tight)
{
  int64_t raw = (int64_t)emax - (int64_t)minexp + 2 * (int64_t)dims + (tight ? 1 : 2);
  if (raw < 0) raw = 0;
  if (raw > (int64_t)maxprec) raw = maxprec;
  return (uint32_t)raw;
}

/* encode_float_block / encode_double_block (encode/float.rs:22-113) */
static uint32_t FN(encode_float_block)(orc_stream *s, const TFLOAT *f, int size, int dims,
                                       const orc_config *c)
{
  int emax = FN(exponent_block)(f, size);
  uint32_t prec = FN(precision)(emax, c->maxprec, c->minexp, dims, c->tight);
  uint32_t e = prec ? (uint32_t)(emax + EBIAS) : 0u;
  if (e) {
    const uint32_t hb = EBITS + 1;
    uint32_t bmin = orc_sat_sub(c->minbits, hb), bmax = orc_sat_sub(c->maxbits, hb);
    orc_write_bits(s, 2 * (uint64_t)e + 1, hb);
    TINT iblock[256];
    FN(fwd_cast)(iblock, f, size, emax);
    return hb + FN(encode_int_block)(s, iblock, size, bmin, bmax, prec, c->rounding);
  }
  orc_write_bit(s, 0);
  return orc_pad_to(s, 1, c->minbits);
}

/* decode_float_block! (decode/core.rs:184-260) */
static uint32_t FN(decode_float_block)(orc_stream *s, TFLOAT *f, int size, int dims,
                                       const orc_config *c)
{
  for (int i = 0; i < size; i++) f[i] = 0;
  if (orc_read_bit(s)) {
    const uint32_t hb = 1 + EBITS;
    int emax = (int)orc_read_bits(s, EBITS) - EBIAS;
    uint32_t prec = FN(precision)(emax, c->maxprec, c->minexp, dims, c->tight);
    uint32_t bmin = orc_sat_sub(c->minbits, hb), bmax = orc_sat_sub(c->maxbits, hb);
    TINT iblock[256];
    uint32_t b = FN(decode_int_block)(s, iblock, size, bmin, bmax, prec, c->rounding);
    FN(inv_cast)(iblock, f, size, emax);
    return hb + b;
  }
  return orc_skip_to(s, 1, c->minbits);
}

/* ---------------- reversible (encode/reversible.rs, decode/reversible.rs) --------------- */

static uint32_t FN(rev_precision)(const TUINT *u, int size)
{
  TUINT acc = 0;
  for (int i = 0; i < size; i++) acc |= u[i];
  return acc ? (uint32_t)W - (uint32_t)orc_ctz64((uint64_t)acc) : 0u;
}

/* rev_encode_int_block (encode/reversible.rs:88-120); iblock already rev-transformed */
static uint32_t FN(rev_encode_ints)(orc_stream *s, const TINT *iblock, int size,
                                    uint32_t maxbits, uint32_t maxprec)
{
  TUINT u[256];
  FN(fwd_order)(u, iblock, orc_perm(size), size);
  uint32_t prec = FN(rev_precision)(u, size);
  if (prec > maxprec) prec = maxprec;
  if (prec < 1) prec = 1;
  orc_write_bits(s, (uint64_t)(prec - 1), PBITS);
  uint32_t mb = orc_sat_sub(maxbits, PBITS);
  return PBITS + FN(encode_ints)(s, mb, prec, u, size);
}

/* rev_decode_int_block (decode/reversible.rs:50-80) */
static uint32_t FN(rev_decode_ints)(orc_stream *s, TINT *iblock, int size, uint32_t maxbits)
{
  uint32_t prec = (uint32_t)orc_read_bits(s, PBITS) + 1;
  uint32_t mb = orc_sat_sub(maxbits, PBITS);
  TUINT u[256];
  int zero;
  uint32_t ub = FN(decode_ints)(s, mb, prec, 0, u, size, &zero);
  FN(inv_order)(u, iblock, orc_perm(size), size);
  return PBITS + ub;
}

/* integer reversible encode (encode/reversible.rs:286-300) */
static uint32_t FN(rev_encode_int)(orc_stream *s, const TINT *in, int size, const orc_config *c)
{
  TINT block[256];
  memcpy(block, in, sizeof(TINT) * (size_t)size);
  FN(fwd_xform)(block, size, FN(rev_fwd_lift));
  uint32_t bits = FN(rev_encode_ints)(s, block, size, c->maxbits, c->maxprec);
  return orc_pad_to(s, bits, c->minbits);
}

static uint32_t FN(rev_decode_int)(orc_stream *s, TINT *out, int size, const orc_config *c)
{
  uint32_t bits = FN(rev_decode_ints)(s, out, size, c->maxbits);
  bits = orc_skip_to(s, bits, c->minbits);
  FN(inv_xform)(out, size, FN(rev_inv_lift));
  return bits;
}

/* reversible float encode (encode/reversible.rs:172-240) */
static uint32_t FN(rev_encode_float)(orc_stream *s, const TFLOAT *f, int size, const orc_config *c)
{
  int emax = FN(exponent_block)(f, size);
  uint32_t e = (uint32_t)(emax + EBIAS);
  int tiny = e != 0 && emax < MINCAST;
  TINT iblock[256];
  for (int i = 0; i < size; i++) iblock[i] = 0;
  if (e != 0 && !tiny) FN(fwd_cast)(iblock, f, size, emax);
  int reversible = 0;
  if (!tiny) {
    TFLOAT sc = FN(pow2)(emax - (W - 2));
    reversible = 1;
    for (int i = 0; i < size; i++) {
      volatile TFLOAT t = (TFLOAT)iblock[i];
      TFLOAT r = sc * t;
      TUINT rb, fb;
      memcpy(&rb, &r, sizeof(rb));
      memcpy(&fb, &f[i], sizeof(fb));
      if (rb != fb) { reversible = 0; break; }
    }
  }
  uint32_t bits;
  if (reversible) {
    if (e != 0) {
      orc_write_bits(s, 1, 2); /* 0b01 */
      orc_write_bits(s, e, EBITS);
      const uint32_t hb = 2 + EBITS;
      FN(fwd_xform)(iblock, size, FN(rev_fwd_lift));
      bits = hb + FN(rev_encode_ints)(s, iblock, size, orc_sat_sub(c->maxbits, hb), c->maxprec);
    } else {
      orc_write_bit(s, 0);
      bits = 1;
    }
  } else {
    TINT tc[256];
    for (int i = 0; i < size; i++) {
      TUINT b;
      memcpy(&b, &f[i], sizeof(b));
      if (b >> (W - 1)) b ^= (TUINT)(~((TUINT)1 << (W - 1)));
      tc[i] = (TINT)b;
    }
    orc_write_bits(s, 3, 2); /* 0b11 */
    FN(fwd_xform)(tc, size, FN(rev_fwd_lift));
    bits = 2 + FN(rev_encode_ints)(s, tc, size, orc_sat_sub(c->maxbits, 2), c->maxprec);
  }
  return orc_pad_to(s, bits, c->minbits);
}

/* reversible float decode (decode/reversible.rs:138-175) */
static uint32_t FN(rev_decode_float)(orc_stream *s, TFLOAT *f, int size, const orc_config *c)
{
  if (orc_read_bits(s, 1) == 0) {
    for (int i = 0; i < size; i++) f[i] = 0;
    return orc_skip_to(s, 1, c->minbits);
  }
  uint32_t bits = 1;
  uint64_t reinterp = orc_read_bits(s, 1);
  bits += 1;
  TINT iblock[256];
  if (reinterp) {
    bits += FN(rev_decode_ints)(s, iblock, size, orc_sat_sub(c->maxbits, bits));
    FN(inv_xform)(iblock, size, FN(rev_inv_lift));
    for (int i = 0; i < size; i++) {
      TUINT b = (TUINT)iblock[i];
      if (iblock[i] < 0) b ^= (TUINT)(~((TUINT)1 << (W - 1)));
      memcpy(&f[i], &b, sizeof(b));
    }
  } else {
    int emax = (int)orc_read_bits(s, EBITS) - EBIAS;
    bits += EBITS;
    bits += FN(rev_decode_ints)(s, iblock, size, orc_sat_sub(c->maxbits, bits));
    FN(inv_xform)(iblock, size, FN(rev_inv_lift));
    FN(inv_cast)(iblock, f, size, emax);
  }
  return orc_skip_to(s, bits, c->minbits);
}
