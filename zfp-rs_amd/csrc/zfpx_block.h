// zfpx_block.h -- the ZFP block codec as per-thread code for gfx950.
//
// One GPU thread codes one 4^d block held in registers: block-floating-point
// cast, decorrelating lift, sequency permutation, negabinary, and the embedded
// bit-plane group-tested coder.  Every loop over coefficients or planes is
// unrolled so that block values, planes and permutations live in VGPRs with
// static indices; the plane loop keeps per-lane state (budget, significant
// count) and lanes that finish early simply idle.
//
// The same source compiles for the host (ZX_HD expands to `inline`), which is
// how the test suite checks this exact code against the oracle on the CPU.
//
// Bit-exactness follows zfp-rs (reference file:line):
//   bit-plane coder          src/codec/bitplane.rs:509-770
//   lifting transforms       src/codec/transform.rs:74-142, axis order :173-234
//   permutation tables       src/codec/encode/core.rs:12-39
//   negabinary, rounding     src/codec/encode/core.rs:53-109, decode/core.rs:25-82
//   exponent, cast           src/codec/encode/core.rs:231-423, decode/core.rs:162-175
//   float / int blocks       src/codec/encode/{float,integer}.rs, decode/core.rs:97-260
//   reversible blocks        src/codec/encode/reversible.rs, decode/reversible.rs
// The casts are done in exact integer arithmetic from the IEEE bits rather
// than by scaling in floating point; the two agree for every input (the
// scaled value is exact whenever it is at least one in magnitude).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define ZX_HD __host__ __device__ __forceinline__
#define ZX_DEVICE_BUILD 1
#else
#include <math.h>
#include <string.h>
#define ZX_HD inline
#define ZX_DEVICE_BUILD 0
#endif

#if !defined(__HIP_DEVICE_COMPILE__)
#define ZX_UNROLL _Pragma("GCC unroll 128")
#else
#define ZX_UNROLL _Pragma("unroll")
#endif

namespace zx {

enum { ROUND_NEVER = 0, ROUND_FIRST = 1, ROUND_LAST = 2 };

// Expert parameters of a stream plus the per-call rounding (config.rs:163-174).
struct Cfg {
  uint32_t minbits, maxbits, maxprec;
  int32_t minexp;
  int32_t rounding;
  int32_t tight;
};

// ---------------------------------------------------------------- bit helpers

ZX_HD uint32_t ctz64(uint64_t x) { return x ? (uint32_t)__builtin_ctzll(x) : 64u; }
ZX_HD uint32_t clz32(uint32_t x) { return x ? (uint32_t)__builtin_clz(x) : 32u; }
ZX_HD uint32_t clz64(uint64_t x) { return x ? (uint32_t)__builtin_clzll(x) : 64u; }
ZX_HD uint64_t lowmask(uint32_t n) { return n < 64 ? (1ull << n) - 1 : ~0ull; }
ZX_HD uint32_t ctz32(uint32_t x) { return x ? (uint32_t)__builtin_ctz(x) : 32u; }
ZX_HD uint32_t lowmask32(uint32_t n) { return n < 32 ? (1u << n) - 1 : ~0u; }
ZX_HD uint32_t popc32(uint32_t x) { return (uint32_t)__builtin_popcount(x); }
ZX_HD uint32_t popc64(uint64_t x) { return (uint32_t)__builtin_popcountll(x); }
ZX_HD uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }
ZX_HD uint32_t sat_sub(uint32_t a, uint32_t b) { return a > b ? a - b : 0u; }

// (maxprec + 1) * size > maxbits + 1 (encode/core.rs:183-185)
ZX_HD bool with_maxbits(uint32_t maxbits, uint32_t maxprec, uint32_t size)
{
  return ((uint64_t)maxprec + 1) * (uint64_t)size > (uint64_t)maxbits + 1;
}

ZX_HD uint32_t f2u(float f)
{
#if ZX_DEVICE_BUILD
  return __float_as_uint(f);
#else
  uint32_t u; memcpy(&u, &f, 4); return u;
#endif
}
ZX_HD float u2f(uint32_t u)
{
#if ZX_DEVICE_BUILD
  return __uint_as_float(u);
#else
  float f; memcpy(&f, &u, 4); return f;
#endif
}
ZX_HD uint64_t d2u(double d)
{
#if ZX_DEVICE_BUILD
  return (uint64_t)__double_as_longlong(d);
#else
  uint64_t u; memcpy(&u, &d, 8); return u;
#endif
}
ZX_HD double u2d(uint64_t u)
{
#if ZX_DEVICE_BUILD
  return __longlong_as_double((long long)u);
#else
  double d; memcpy(&d, &u, 8); return d;
#endif
}

// ---------------------------------------------------------------- traits

template <int P> struct Ints;
template <> struct Ints<32> {
  typedef int32_t I;
  typedef uint32_t U;
  static constexpr uint32_t NB = 0xaaaaaaaau;
};
template <> struct Ints<64> {
  typedef int64_t I;
  typedef uint64_t U;
  static constexpr uint64_t NB = 0xaaaaaaaaaaaaaaaaull;
};

template <class S> struct ST;
template <> struct ST<int32_t> {
  static constexpr int P = 32, PBITS = 5;
  static constexpr bool FLT = false;
};
template <> struct ST<int64_t> {
  static constexpr int P = 64, PBITS = 6;
  static constexpr bool FLT = false;
};
template <> struct ST<float> {
  static constexpr int P = 32, PBITS = 5, EBITS = 8, EBIAS = 127, MANT = 23, MINCAST = -97;
  static constexpr bool FLT = true;
};
template <> struct ST<double> {
  static constexpr int P = 64, PBITS = 6, EBITS = 11, EBIAS = 1023, MANT = 52, MINCAST = -961;
  static constexpr bool FLT = true;
};

template <int D> struct Blk { static constexpr int SIZE = 1 << (2 * D); };

// ---------------------------------------------------------------- permutation

// PERM_1..4 (encode/core.rs:12-39): coefficient order by sequency.
ZX_HD constexpr int perm(int d, int j)
{
  constexpr uint8_t p4[256] = {
      0,   1,   4,   16,  64,  5,   80,  17,  68,  65,  20,  2,   8,   32,  128, 84,  81,  69,  21,
      6,   18,  66,  24,  72,  9,   96,  33,  36,  129, 132, 144, 3,   12,  48,  192, 85,  82,  70,
      22,  73,  25,  88,  37,  100, 97,  148, 145, 133, 10,  160, 34,  136, 130, 40,  7,   19,  67,
      28,  76,  13,  112, 49,  52,  193, 196, 208, 86,  89,  101, 149, 161, 137, 41,  134, 38,  164,
      26,  152, 146, 104, 98,  74,  83,  71,  23,  77,  29,  92,  53,  116, 113, 212, 209, 197, 11,
      35,  131, 44,  140, 14,  176, 50,  56,  194, 200, 224, 90,  165, 102, 153, 150, 105, 168, 162,
      138, 42,  87,  93,  117, 213, 27,  75,  99,  39,  135, 147, 108, 45,  141, 156, 30,  78,  177,
      180, 54,  114, 120, 57,  198, 210, 216, 201, 225, 228, 15,  240, 51,  204, 195, 60,  169, 166,
      154, 106, 91,  103, 151, 109, 157, 94,  181, 118, 121, 214, 217, 229, 163, 139, 43,  142, 46,
      172, 58,  184, 178, 232, 226, 202, 241, 205, 61,  199, 55,  244, 31,  220, 211, 124, 115, 79,
      170, 167, 155, 107, 158, 110, 173, 122, 185, 182, 233, 230, 218, 95,  245, 119, 221, 215, 125,
      242, 206, 62,  203, 59,  248, 47,  236, 227, 188, 179, 143, 171, 174, 186, 234, 246, 222, 126,
      219, 123, 249, 111, 237, 231, 189, 183, 159, 252, 243, 207, 63,  175, 250, 187, 238, 235, 190,
      253, 247, 223, 127, 254, 251, 239, 191, 255};
  constexpr uint8_t p2[16] = {0, 1, 4, 5, 2, 8, 6, 9, 3, 12, 10, 7, 13, 11, 14, 15};
  constexpr uint8_t p3[64] = {
      0,  1,  4,  16, 20, 17, 5,  2,  8,  32, 21, 6,  18, 24, 9,  33, 36, 3,  12, 48, 22, 25,
      37, 40, 34, 10, 7,  19, 28, 13, 49, 52, 41, 38, 26, 23, 29, 53, 11, 35, 44, 14, 50, 56,
      42, 27, 39, 45, 30, 54, 57, 60, 51, 15, 43, 46, 58, 61, 55, 31, 62, 59, 47, 63};
  return d == 1 ? j : d == 2 ? p2[j] : d == 3 ? p3[j] : p4[j];
}

// ---------------------------------------------------------------- lifting

template <class I> ZX_HD I wadd(I a, I b)
{
  typedef typename Ints<sizeof(I) * 8>::U U;
  return (I)((U)a + (U)b);
}
template <class I> ZX_HD I wsub(I a, I b)
{
  typedef typename Ints<sizeof(I) * 8>::U U;
  return (I)((U)a - (U)b);
}

// forward lift of one line (transform.rs:74-86)
template <class I> ZX_HD void fwd_lift(I &x, I &y, I &z, I &w)
{
  x = wadd(x, w) >> 1;
  w = wsub(w, x);
  z = wadd(z, y) >> 1;
  y = wsub(y, z);
  x = wadd(x, z) >> 1;
  z = wsub(z, x);
  w = wadd(w, y) >> 1;
  y = wsub(y, w);
  w = wadd(w, (I)(y >> 1));
  y = wsub(y, (I)(w >> 1));
}

// inverse lift of one line (transform.rs:98-110).  Each of its last four
// pairs, b += a then a -= b - a, leaves a = a - b_old in wrapping arithmetic
// (a - ((b + a) - a)), which is written out so that each pair is two adds.
template <class I> ZX_HD void inv_lift(I &x, I &y, I &z, I &w)
{
  y = wadd(y, (I)(w >> 1));
  w = wsub(w, (I)(y >> 1));
  I t = y;
  y = wadd(y, w);
  w = wsub(w, t);
  t = z;
  z = wadd(z, x);
  x = wsub(x, t);
  t = y;
  y = wadd(y, z);
  z = wsub(z, t);
  t = w;
  w = wadd(w, x);
  x = wsub(x, t);
}

// reversible lifts (transform.rs:118-142)
template <class I> ZX_HD void rev_fwd_lift(I &x, I &y, I &z, I &w)
{
  w = wsub(w, z);
  z = wsub(z, y);
  y = wsub(y, x);
  w = wsub(w, z);
  z = wsub(z, y);
  w = wsub(w, z);
}
template <class I> ZX_HD void rev_inv_lift(I &x, I &y, I &z, I &w)
{
  w = wadd(w, z);
  z = wadd(z, y);
  w = wadd(w, z);
  y = wadd(y, x);
  z = wadd(z, y);
  w = wadd(w, z);
}

// Lift every line along the axis of stride S (transform.rs:151-170).
template <int SIZE, int S, class I, class F> ZX_HD void lift_axis(I (&p)[SIZE], F lift)
{
  ZX_UNROLL
  for (int g = 0; g < SIZE; g += 4 * S) {
    ZX_UNROLL
    for (int i = g; i < g + S; i++) lift(p[i], p[i + S], p[i + 2 * S], p[i + 3 * S]);
  }
}

struct FwdL { template <class I> ZX_HD void operator()(I &x, I &y, I &z, I &w) const { fwd_lift(x, y, z, w); } };
struct InvL { template <class I> ZX_HD void operator()(I &x, I &y, I &z, I &w) const { inv_lift(x, y, z, w); } };
struct RFwdL { template <class I> ZX_HD void operator()(I &x, I &y, I &z, I &w) const { rev_fwd_lift(x, y, z, w); } };
struct RInvL { template <class I> ZX_HD void operator()(I &x, I &y, I &z, I &w) const { rev_inv_lift(x, y, z, w); } };

// x, then y, then z, then w (forward); w, z, y, x (inverse)
template <int D, class I, class L> ZX_HD void fwd_xform(I (&p)[Blk<D>::SIZE], L lift)
{
  lift_axis<Blk<D>::SIZE, 1>(p, lift);
  if constexpr (D >= 2) lift_axis<Blk<D>::SIZE, 4>(p, lift);
  if constexpr (D >= 3) lift_axis<Blk<D>::SIZE, 16>(p, lift);
  if constexpr (D >= 4) lift_axis<Blk<D>::SIZE, 64>(p, lift);
}
template <int D, class I, class L> ZX_HD void inv_xform(I (&p)[Blk<D>::SIZE], L lift)
{
  if constexpr (D >= 4) lift_axis<Blk<D>::SIZE, 64>(p, lift);
  if constexpr (D >= 3) lift_axis<Blk<D>::SIZE, 16>(p, lift);
  if constexpr (D >= 2) lift_axis<Blk<D>::SIZE, 4>(p, lift);
  lift_axis<Blk<D>::SIZE, 1>(p, lift);
}

// ---------------------------------------------------------------- wave votes

// True if the predicate holds on any lane of the wave (on the host: this lane).
ZX_HD bool wave_any(bool p)
{
#if defined(__HIP_DEVICE_COMPILE__)
  return __ballot(p) != 0;
#else
  return p;
#endif
}

// True when p holds on every active lane of the wave.
ZX_HD bool wave_all(bool p) { return !wave_any(!p); }

// ---------------------------------------------------------------- bit planes

// Byte select of v_perm_b32: byte k of the result is byte sel_k of the
// eight bytes hi:lo (0-3 from lo, 4-7 from hi); selector values 0-7 only.
ZX_HD uint32_t perm_bytes(uint32_t hi, uint32_t lo, uint32_t sel)
{
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_perm(hi, lo, sel);
#else
  const uint64_t v = ((uint64_t)hi << 32) | lo;
  uint32_t r = 0;
  for (int k = 0; k < 4; k++) r |= (uint32_t)((v >> (8 * ((sel >> (8 * k)) & 7))) & 0xffu) << (8 * k);
  return r;
#endif
}

// Selector that keeps bytes 0-3 of lo except byte k, which takes byte b of hi:lo.
ZX_HD constexpr uint32_t byte_insert_sel(int k, int b)
{
  uint32_t s = 0;
  for (int i = 0; i < 4; i++) s |= (uint32_t)(i == k ? b : i) << (8 * i);
  return s;
}

// Stages of the 32x32 bit-matrix transpose (row r bit c <-> row c bit r).
// The stage of width j swaps index bit log2(j) of rows and columns; the five
// stages commute, so any subset of rows can be finished independently.
template <int J, int N>
ZX_HD void tr_stage(uint32_t (&a)[N])
{
  constexpr uint32_t M = J == 8 ? 0x00ff00ffu : J == 4 ? 0x0f0f0f0fu : J == 2 ? 0x33333333u : 0x55555555u;
  ZX_UNROLL
  for (int k = 0; k < N; k += 2 * J) {
    ZX_UNROLL
    for (int i = k; i < k + J; i++) {
      uint32_t lo = a[i], hi = a[i + J];
      if constexpr (J == 8) {
        // byte interleaves, one v_perm each: {lo.b0, hi.b0, lo.b2, hi.b2}
        // and {lo.b1, hi.b1, lo.b3, hi.b3}
        a[i] = perm_bytes(hi, lo, 0x06020400u);
        a[i + J] = perm_bytes(hi, lo, 0x07030501u);
      } else {
        a[i] = (lo & M) | ((hi << J) & ~M);
        a[i + J] = ((lo >> J) & M) | (hi & ~M);
      }
    }
  }
}

// Plane word: one bit per coefficient.
template <int SIZE> struct PWord { typedef uint32_t T; };
template <> struct PWord<64> { typedef uint64_t T; };

// Planes are handled CH at a time (a chunk).  The decoder keeps a chunk of
// decoded planes live beside the coefficients, and smaller chunks help it fit
// four waves per SIMD; the encoder transposes 16 planes at a time.
constexpr int kEncodeChunk = 16;
constexpr int kDecodeChunk = 8;

// Planes 16c .. 16c+15 of a block come from the transpose of the 32 x 32 bit
// matrix of each 32 coefficients (row r = coefficient r), of which only the 16
// rows needed are computed.  The first step packs, per group of 32
// coefficients, the 16 bits of chunk c of coefficients i and i + 16 into one
// word (i < 16): chunk_halves.  The encoder takes this step for every chunk
// before coding any, so that it holds these packed halves rather than the
// coefficients while it codes.
template <int SIZE> struct Halves {
  static constexpr int G = SIZE > 32 ? 2 : 1;
  uint32_t o[G][16];
};

template <int P, int SIZE>
ZX_HD void chunk_halves(const typename Ints<P>::U (&u)[SIZE], int c, Halves<SIZE> &hv)
{
  const int first = 16 * c;
  const int h = first >> 5;             // 32-bit half of the coefficients
  const bool hi16 = (first & 16) != 0;  // rows 16..31 of the half
  ZX_UNROLL
  for (int g = 0; g < Halves<SIZE>::G; g++) {
    uint32_t a[32];
    ZX_UNROLL
    for (int r = 0; r < 32; r++) {
      const int i = 32 * g + r;
      a[r] = i < SIZE ? (uint32_t)((uint64_t)u[i < SIZE ? i : 0] >> (32 * h)) : 0u;
    }
    ZX_UNROLL
    for (int i = 0; i < 16; i++) {
      // the 16-bit halves of two words in one byte permute: low halves
      // {a[i].b0, a[i].b1, a[i+16].b0, a[i+16].b1}, high halves likewise
      hv.o[g][i] = perm_bytes(a[i + 16], a[i], hi16 ? 0x07060302u : 0x05040100u);
#if defined(__HIP_DEVICE_COMPILE__)
      // computed here, not sunk into the chunk's branch (where the
      // coefficients would have to stay live until then)
      asm volatile("" : "+v"(hv.o[g][i]));
#endif
    }
  }
}

// The rest of the transpose: out[i] bit j = bit 16c+i of coefficient j.
template <int SIZE>
ZX_HD void chunk_planes(const Halves<SIZE> &hv, typename PWord<SIZE>::T (&out)[16])
{
  typedef typename PWord<SIZE>::T W;
  ZX_UNROLL
  for (int g = 0; g < Halves<SIZE>::G; g++) {
    uint32_t q[16];
    ZX_UNROLL
    for (int i = 0; i < 16; i++) q[i] = hv.o[g][i];
    tr_stage<8>(q);
    tr_stage<4>(q);
    tr_stage<2>(q);
    tr_stage<1>(q);
    ZX_UNROLL
    for (int i = 0; i < 16; i++) {
      if (g == 0) out[i] = (W)q[i];
      else out[i] |= (W)((uint64_t)q[i] << 32);
    }
  }
}

// ORs the decoded planes CH*c .. CH*c+CH-1 (pl[i] bit j = bit CH*c+i of
// coefficient j) into the coefficients: per 32 coefficients, the CH plane
// words are the top rows of an otherwise zero 32 x 32 bit matrix, and the
// stages below CH leave coefficient CH*m + i in field m (CH bits wide) of
// word i.
template <int P, int SIZE, int CH>
ZX_HD void chunk_to_coeffs(const typename PWord<SIZE>::T (&pl)[CH], int c,
                           typename Ints<P>::U (&u)[SIZE])
{
  static_assert(CH == 8 || CH == 16, "chunk of 8 or 16 planes");
  typedef typename Ints<P>::U U;
  constexpr int G = SIZE > 32 ? 2 : 1;
  constexpr uint32_t FM = CH == 16 ? 0xffffu : 0xffu;
  ZX_UNROLL
  for (int g = 0; g < G; g++) {
    uint32_t t[CH];
    ZX_UNROLL
    for (int i = 0; i < CH; i++) t[i] = (uint32_t)((uint64_t)pl[i] >> (32 * g));
    if constexpr (CH == 16) tr_stage<8>(t);
    tr_stage<4>(t);
    tr_stage<2>(t);
    tr_stage<1>(t);
    ZX_UNROLL
    for (int i = 0; i < CH; i++) {
      ZX_UNROLL
      for (int f = 0; f < 32 / CH; f++) {
        const int j = 32 * g + CH * f + i;
        if (j >= SIZE) continue;
        U &d = u[j < SIZE ? j : 0];
        if constexpr (CH == 8 && sizeof(U) == 4) {
          // byte f of t[i] into byte c of the coefficient (empty until now)
          d = (U)perm_bytes(t[i], (uint32_t)d, byte_insert_sel(c, 4 + f));
        } else {
          // (64-bit coefficients: an OR into the one dword the field lands
          // in; rebuilding the coefficient from its halves kept two copies
          // of every coefficient live across the chunks)
          d |= (U)((t[i] >> (CH * f)) & FM) << (CH * c);
        }
      }
    }
  }
}

// ---------------------------------------------------------------- bit sinks

// Append the low n bits (no bits set above them) of a 128-bit value (lo, hi),
// n <= 128, to a 64-bit buffer holding `bits` < 64 bits; returns the number
// of completed words (0..2) in w0, w1 and leaves the remainder in buf.
ZX_HD uint32_t append128(uint64_t &buf, uint32_t &bits, uint64_t lo, uint64_t hi, uint32_t n,
                         uint64_t &w0, uint64_t &w1)
{
  const uint32_t b = bits;
  w0 = buf | (lo << b);
  const uint64_t lo_sp = (lo >> 1) >> (63 - b);  // lo >> (64 - b), 0 when b == 0
  w1 = lo_sp | (hi << b);
  const uint64_t hi_sp = (hi >> 1) >> (63 - b);
  const uint32_t t = b + n;
  const uint32_t full = t >> 6;
  buf = full == 0 ? w0 : full == 1 ? w1 : hi_sp;
  bits = t & 63;
  return full;
}

// Writes a block's bits into a word array.  Words the block shares with its
// neighbours (the first, if the block starts mid-word, and the last partial
// one) are merged with an atomic OR into words zeroed beforehand; every other
// word belongs to this block alone and is stored whole.  Words at or past
// `cap` are dropped, as the reference drops writes past a stream's end.
struct WordSink {
  static constexpr bool kSpillPastBudget = false;
  uint64_t *words;
  uint64_t cap;
  uint64_t idx;
  uint64_t buf;
  uint32_t bits;
  bool shared_first;

  ZX_HD void init(uint64_t *w, uint64_t capacity, uint64_t bitpos)
  {
    words = w;
    cap = capacity;
    idx = bitpos >> 6;
    bits = (uint32_t)(bitpos & 63);
    buf = 0;
    shared_first = bits != 0;
  }
  ZX_HD static void merge(uint64_t *p, uint64_t v)
  {
#if defined(__HIP_DEVICE_COMPILE__)
    atomicOr((unsigned long long *)p, (unsigned long long)v);
#else
    *p |= v;
#endif
  }
  ZX_HD void emit(uint64_t w)
  {
    if (idx < cap) {
      if (shared_first) merge(words + idx, w);
      else words[idx] = w;
    }
    shared_first = false;
    idx++;
  }
  ZX_HD void put128(uint64_t lo, uint64_t hi, uint32_t n)
  {
    uint64_t w0, w1;
    const uint32_t full = append128(buf, bits, lo, hi, n, w0, w1);
    if (full >= 1) emit(w0);
    if (full >= 2) emit(w1);
  }
  ZX_HD void put(uint64_t v, uint32_t n) { put128(v, 0, n); }
  ZX_HD void zeros(uint32_t n)
  {
    while (n > 128) { put128(0, 0, 128); n -= 128; }
    put128(0, 0, n);
  }
  ZX_HD void finish()
  {
    if (bits && idx < cap) merge(words + idx, buf);
  }
};

// Writes a block's bits into a word array in LDS (or any memory one thread
// owns), branch free: every append stores its two candidate words and the
// index advances by the number completed.  Word i lives at
// base[i + (i >> skew)] (a pad word per 2^skew words keeps lanes on separate
// banks).  With SHARED, stores are atomic ORs into zeroed words, for blocks
// that share boundary words with their neighbours.
template <bool SHARED>
struct LdsSink {
  // Without SHARED the sink owns its block's whole words and one spare word
  // after them, and blocks end on a word boundary: bits of a code past the
  // budget may be appended unmasked, since they only reach the spare word.
  static constexpr bool kSpillPastBudget = !SHARED;
  uint64_t *base;
  uint32_t idx;
  uint32_t skew;
  uint64_t buf;
  uint32_t bits;

  ZX_HD void init(uint64_t *b, uint32_t sk, uint32_t bitpos)
  {
    base = b;
    skew = sk;
    idx = bitpos >> 6;
    bits = bitpos & 63;
    buf = 0;
  }
  ZX_HD uint64_t *at(uint32_t i) const
  {
    // an owned slice has no pad slots between a block's words
    if constexpr (SHARED) return base + i + (i >> skew);
    else return base + i;
  }
  ZX_HD void store(uint32_t i, uint64_t v)
  {
    if constexpr (SHARED) {
#if defined(__HIP_DEVICE_COMPILE__)
      atomicOr((unsigned long long *)at(i), (unsigned long long)v);
#else
      *at(i) |= v;
#endif
    } else {
      *at(i) = v;
    }
  }
  ZX_HD void put128(uint64_t lo, uint64_t hi, uint32_t n)
  {
    uint64_t w0, w1;
    const uint32_t full = append128(buf, bits, lo, hi, n, w0, w1);
    if constexpr (SHARED) {
      // OR only completed words; a partial word is carried in buf
      if (full >= 1) store(idx, w0);
      if (full >= 2) store(idx + 1, w1);
    } else {
      // candidates: a word stored before it is complete is rewritten later
      store(idx, w0);
      store(idx + 1, w1);
    }
    idx += full;
  }
  ZX_HD void put(uint64_t v, uint32_t n) { put128(v, 0, n); }
  ZX_HD void zeros(uint32_t n)
  {
    if (n == 0) return;  // no candidate stores once the block's words are done
    while (n > 128) { put128(0, 0, 128); n -= 128; }
    put128(0, 0, n);
  }
  ZX_HD void finish()
  {
    if (bits) store(idx, buf);
  }
};

// WordSink for the blocks of a group written straight to global memory (the
// one-pass compress, for a group too large to stage): the group's first and
// last words, which it shares with the neighbouring groups, go to two edge
// words (LDS on the device) by atomic OR instead of the stream; the caller
// has zeroed the group's other words that blocks share.
struct EdgeSink {
  static constexpr bool kSpillPastBudget = false;
  WordSink ws;
  uint64_t head, tail;  // stream word indices of the group's shared edge words, or ~0
  uint64_t *edge;       // edge[0]: head word bits, edge[1]: tail word bits

  ZX_HD void init(uint64_t *w, uint64_t capacity, uint64_t bitpos, uint64_t h, uint64_t t, uint64_t *e)
  {
    ws.init(w, capacity, bitpos);
    head = h;
    tail = t;
    edge = e;
  }
  ZX_HD void out(uint64_t i, uint64_t v, bool shared)
  {
    if (i == head || i == tail) {
      uint64_t *p = edge + (i == head ? 0 : 1);
#if defined(__HIP_DEVICE_COMPILE__)
      atomicOr((unsigned long long *)p, (unsigned long long)v);
#else
      *p |= v;
#endif
    } else if (i < ws.cap) {
      if (shared) WordSink::merge(ws.words + i, v);
      else ws.words[i] = v;
    }
  }
  ZX_HD void put128(uint64_t lo, uint64_t hi, uint32_t n)
  {
    uint64_t w0, w1;
    const uint32_t full = append128(ws.buf, ws.bits, lo, hi, n, w0, w1);
    if (full >= 1) {
      out(ws.idx, w0, ws.shared_first);
      ws.shared_first = false;
      ws.idx++;
    }
    if (full >= 2) {
      out(ws.idx, w1, false);
      ws.idx++;
    }
  }
  ZX_HD void put(uint64_t v, uint32_t n) { put128(v, 0, n); }
  ZX_HD void zeros(uint32_t n)
  {
    while (n > 128) { put128(0, 0, 128); n -= 128; }
    put128(0, 0, n);
  }
  ZX_HD void finish()
  {
    if (ws.bits) out(ws.idx, ws.buf, true);
  }
};

// Counts bits only (the length pass of variable-rate compression).
struct CountSink {
  static constexpr bool kSpillPastBudget = true;
  static constexpr bool kCountOnly = true;  // plane codes need only their lengths
  uint64_t n;
  ZX_HD void put(uint64_t, uint32_t k) { n += k; }
  ZX_HD void put128(uint64_t, uint64_t, uint32_t k) { n += k; }
  ZX_HD void zeros(uint32_t k) { n += k; }
  ZX_HD void finish() {}
};

// Word memories for sources: global words with a bound (past it reads zero,
// bitstream/core.rs:151-157), or a staged copy in LDS.
struct GlobalWords {
  const uint64_t *w;
  uint64_t n;
  ZX_HD uint64_t operator()(uint64_t i) const { return i < n ? w[i] : 0ull; }
};
// A staged copy in LDS.  Word i sits at slot i + (i >> skew), and the pad
// slot before each word whose index is a multiple of 2^skew holds a second
// copy of that word, so the two slots from word i's on hold words i and i + 1.
// FLAT: `w` is the slot of the reader's first word, and its words occupy the
// slots after it with no pad between them (a block of whole words that no
// multiple of 2^skew splits); the pad copy covers the word after the last.
template <bool FLAT>
struct StagedWords {
  const uint64_t *w;
  uint32_t skew;
  // words i and i + 1 for the word i holding bit p
  ZX_HD void pair_at(uint32_t p, uint64_t &lo, uint64_t &hi) const
  {
    const uint64_t *q = FLAT ? w + (p >> 6) : w + ((p >> 6) + (p >> (skew + 6)));
    lo = q[0];
    hi = q[1];
  }
  // words i, i + 1 and i + 2 for the word i holding bit p (the third may lie
  // past the words staged for the reader's block: a neighbour's word or a
  // spare slot, whose bits end no plane the reader parses).  Not FLAT: when
  // i + 1 is a multiple of 2^skew, the slot after word i holds word i + 1's
  // pad copy and word i + 2 is one slot further.
  ZX_HD void triple_at(uint32_t p, uint64_t &w0, uint64_t &w1, uint64_t &w2) const
  {
    const uint32_t i = p >> 6;
    const uint64_t *q = FLAT ? w + i : w + (i + (i >> skew));
    w0 = q[0];
    w1 = q[1];
    if constexpr (FLAT) w2 = q[2];
    else w2 = q[((i + 1) & ((1u << skew) - 1)) == 0 ? 3 : 2];
  }
};

// Reads a block's bits: a buffer of `bits` bits plus the next two words (the
// second is loaded a word ahead of use, so a load's latency overlaps the
// decoding of the word before).  `idx` is the index of the next word, which
// equals the reference cursor's word_pos.
template <class Mem>
struct Source {
  Mem mem;
  uint64_t idx;
  uint64_t buf;
  uint64_t nxt, nxt2;
  uint32_t bits;

  ZX_HD void init(const Mem &m, uint64_t bitpos)
  {
    mem = m;
    idx = bitpos >> 6;
    const uint32_t s = (uint32_t)(bitpos & 63);
    if (s) {
      buf = mem(idx) >> s;
      bits = 64 - s;
      idx++;
    } else {
      buf = 0;
      bits = 0;
    }
    nxt = mem(idx);
    nxt2 = mem(idx + 1);
  }
  ZX_HD uint64_t peek() const { return buf | (nxt << bits); }
  ZX_HD void consume(uint32_t n)
  {
    if (n <= bits) {
      buf >>= n;
      bits -= n;
    } else {
      bits = bits + 64 - n;
      buf = (nxt >> 1) >> (63 - bits);
      idx++;
      nxt = nxt2;
      nxt2 = mem(idx + 1);
    }
  }
  ZX_HD uint64_t read(uint32_t n)
  {
    uint64_t v = peek();
    consume(n);
    return v & lowmask(n);
  }
  // read(n), and the 64 bits that follow in `ahead`
  ZX_HD uint64_t read_peek(uint32_t n, uint64_t &ahead)
  {
    const uint64_t v = read(n);
    ahead = peek();
    return v;
  }
  ZX_HD uint64_t pos() const { return idx * 64 - bits; }
  ZX_HD uint64_t words_loaded() const { return idx; }
};

// Over a staged copy in LDS the reader keeps only its bit position: each peek
// loads the two words under it, so a step is one add and no refill branch.
// words_loaded() is the reference cursor's word_pos, ceil(pos / 64).
template <bool FLAT>
struct Source<StagedWords<FLAT>> {
  StagedWords<FLAT> mem;
  uint32_t p;

  ZX_HD void init(const StagedWords<FLAT> &m, uint64_t bitpos)
  {
    mem = m;
    p = (uint32_t)bitpos;
  }
  ZX_HD uint64_t peek() const
  {
    const uint32_t s = p & 63;
    uint64_t lo, hi;
    mem.pair_at(p, lo, hi);
    return (lo >> s) | ((hi << 1) << (63 - s));
  }
  ZX_HD void consume(uint32_t n) { p += n; }
  ZX_HD uint64_t read(uint32_t n)
  {
    const uint64_t v = peek();
    consume(n);
    return v & lowmask(n);
  }
  // read(n) (n <= 64), and the 64 bits that follow in `ahead`: both come
  // from one load of the three words under them, so a plane's verbatim bits
  // and its first test window cost one LDS round trip
  ZX_HD uint64_t read_peek(uint32_t n, uint64_t &ahead)
  {
    const uint32_t s = p & 63;
    uint64_t w0, w1, w2;
    mem.triple_at(p, w0, w1, w2);
    const uint64_t v = (w0 >> s) | ((w1 << 1) << (63 - s));
    const uint32_t t = s + n;  // 0 .. 127
    const uint32_t u = t & 63;
    const uint64_t a = t < 64 ? w0 : w1, b = t < 64 ? w1 : w2;
    ahead = (a >> u) | ((b << 1) << (63 - u));
    p += n;
    return v & lowmask(n);
  }
  ZX_HD uint64_t pos() const { return p; }
  ZX_HD uint64_t words_loaded() const { return (p + 63) >> 6; }
};

typedef Source<GlobalWords> WordSource;

// ---------------------------------------------------------------- plane coder

// expand8(b): the one-bits of the byte b, bit i (the r-th one) moved to i + r.
ZX_HD constexpr uint32_t expand8(uint32_t b)
{
  uint32_t e = 0, r = 0;
  for (uint32_t i = 0; i < 8; i++)
    if ((b >> i) & 1u) e |= 1u << (i + r++);
  return e;
}

struct Expand8Table {
  uint16_t v[256];
};

ZX_HD constexpr Expand8Table make_expand8_table()
{
  Expand8Table t{};
  for (uint32_t b = 0; b < 256; b++) t.v[b] = (uint16_t)expand8(b);
  return t;
}

static constexpr Expand8Table kExpand8Host = make_expand8_table();

#if ZX_DEVICE_BUILD
// The table in device memory, and the workgroup's copy in LDS (64-bit
// entries: a load gives the entry already widened for the 64-bit shift that
// places it, with no register to zero), which kernels that encode fill
// (expand8_fill) before any thread codes a block.
static __constant__ Expand8Table kExpand8Dev = make_expand8_table();
static __shared__ uint64_t zx_expand8_lds[256];
#endif

ZX_HD uint64_t expand8_of(uint32_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
  return zx_expand8_lds[b];
#else
  return kExpand8Host.v[b];
#endif
}

// compress8(b): the one-bits of the byte b, bit i (the r-th one) moved down to
// i - r; on bytes with no two adjacent one-bits it inverts expand8.
ZX_HD constexpr uint32_t compress8(uint32_t b)
{
  uint32_t v = 0, r = 0;
  for (uint32_t i = 0; i < 8; i++)
    if ((b >> i) & 1u) v |= 1u << (i - r++);
  return v;
}

struct alignas(8) Compress8Table {
  uint8_t v[256];
};

ZX_HD constexpr Compress8Table make_compress8_table()
{
  Compress8Table t{};
  for (uint32_t b = 0; b < 256; b++) t.v[b] = (uint8_t)compress8(b);
  return t;
}

static constexpr Compress8Table kCompress8Host = make_compress8_table();

#if ZX_DEVICE_BUILD
// The table in device memory, and the workgroup's copy in LDS, which kernels
// that decode fill (compress8_fill) before any thread decodes a block.
static __constant__ Compress8Table kCompress8Dev = make_compress8_table();
static __shared__ __align__(8) uint8_t zx_compress8_lds[256];
#endif

ZX_HD uint32_t compress8_of(uint32_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
  return zx_compress8_lds[b];
#else
  return kCompress8Host.v[b];
#endif
}

// The one-bits of v, none adjacent and all below bit `nbits`, the j-th from
// the bottom moved down by j: byte by byte through the table, byte m placed
// 8m - (one-bits below byte m) up.  The four low bytes' lookups are issued
// together; the high four only when a lane of the wave has bits there.
ZX_HD uint64_t compress_ones(uint64_t v, uint32_t nbits)
{
  const uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
  uint32_t t[8];
  ZX_UNROLL
  for (int m = 0; m < 4; m++) t[m] = compress8_of((lo >> (8 * m)) & 0xffu);
  uint64_t out = t[0];
  uint32_t cnt = popc32(lo & 0xffu);
  ZX_UNROLL
  for (int m = 1; m < 4; m++) {
    out |= (uint64_t)t[m] << (8 * m - cnt);
    cnt += popc32((lo >> (8 * m)) & 0xffu);
  }
  if (wave_any(nbits > 32)) {
    ZX_UNROLL
    for (int m = 4; m < 8; m++) t[m] = compress8_of((hi >> (8 * (m - 4))) & 0xffu);
    ZX_UNROLL
    for (int m = 4; m < 8; m++) {
      out |= (uint64_t)t[m] << (8 * m - cnt);
      cnt += popc32((hi >> (8 * (m - 4))) & 0xffu);
    }
  }
  return out;
}

// Dimensionalities whose decoders parse planes through the compress8 table
// (64-bit and 32-bit windows for 3-D blocks, 32-bit windows for 2-D blocks,
// 64-bit windows for 4-D blocks); a kernel that decodes them fills the
// workgroup's copy first.
template <int D> constexpr bool kDecodeTable = D >= 2;

// compress_ones of a value below 2^32: the four byte lookups only.
ZX_HD uint32_t compress_ones32(uint32_t v)
{
  uint32_t t[4];
  ZX_UNROLL
  for (int m = 0; m < 4; m++) t[m] = compress8_of((v >> (8 * m)) & 0xffu);
  uint32_t out = t[0];
  uint32_t cnt = popc32(v & 0xffu);
  ZX_UNROLL
  for (int m = 1; m < 4; m++) {
    out |= t[m] << (8 * m - cnt);
    cnt += popc32((v >> (8 * m)) & 0xffu);
  }
  return out;
}

// The one-bits of u (SIZE bits at most), the j-th from the bottom moved up
// by j, as a 128-bit value (lo, hi).  Bytes are expanded through the table
// and placed 8m + (one-bits below byte m) up; a byte range that is zero on
// every lane of the wave is skipped.
template <int SIZE>
ZX_HD void expand_ones(uint64_t u, uint64_t &lo, uint64_t &hi)
{
  constexpr int BYTES = SIZE >= 32 ? 4 : (SIZE + 7) / 8;  // bytes of the low 32 bits
  const uint32_t w0 = (uint32_t)u;
  uint64_t e = expand8_of(w0 & 0xffu);
  if constexpr (BYTES > 1) {
    if (wave_any((w0 >> 8) != 0)) {
      uint32_t r = popc32(w0 & 0xffu);
      uint32_t b[BYTES];
      uint64_t t[BYTES];
      ZX_UNROLL
      for (int m = 1; m < BYTES; m++) {
        b[m] = (w0 >> (8 * m)) & 0xffu;
        t[m] = expand8_of(b[m]);
      }
      ZX_UNROLL
      for (int m = 1; m < BYTES; m++) {
        e |= t[m] << (8 * m + r);  // entry below bit 15, shift at most 24 + 24
        r += popc32(b[m]);
      }
    }
  }
  lo = e;
  hi = 0;
  if constexpr (SIZE > 32) {
    const uint32_t w1 = (uint32_t)(u >> 32);
    if (wave_any(w1 != 0)) {
      uint32_t b[4];
      uint64_t t[4];
      ZX_UNROLL
      for (int m = 0; m < 4; m++) {
        b[m] = (w1 >> (8 * m)) & 0xffu;
        t[m] = expand8_of(b[m]);
      }
      uint64_t f = 0;
      uint32_t r = 0;
      ZX_UNROLL
      for (int m = 0; m < 4; m++) {
        f |= t[m] << (8 * m + r);
        r += popc32(b[m]);
      }
      const uint32_t sh = 32 + popc32(w0);  // where the upper half's expansion starts, 32..64
      lo |= sh < 64 ? f << (sh & 63) : 0ull;
      hi = f >> (64 - sh);
    }
  }
}

// Code of one bit plane x of a block whose first n coefficients are already
// significant (bitplane.rs:560-605): the n low bits verbatim, then the group
// tests over the rest.  Returns the code's length; the code is (lo, hi) and
// nn is the new n.
// plane_code's length and new significant count only (the bits of the code
// are not built): a counting pass need not expand the plane.
template <int SIZE>
ZX_HD uint32_t plane_len(uint64_t x, uint32_t n, uint32_t &nn)
{
  nn = n;
  if (n >= (uint32_t)SIZE) return SIZE;
  const uint64_t rest = x >> n;
  if (rest == 0) return n + 1;
  const uint32_t q = 63u - clz64(rest);
  const uint32_t tb = q + popc64(rest) - 1;
  const bool implied = n + q == (uint32_t)SIZE - 1;
  nn = n + q + 1;
  return n + (implied ? tb + 1 : tb + 3);
}

// Whether a sink only counts bits (CountSink::kCountOnly).
template <class Sink, class = void> struct sink_counts { static constexpr bool value = false; };
template <class Sink> struct sink_counts<Sink, decltype((void)Sink::kCountOnly)> {
  static constexpr bool value = Sink::kCountOnly;
};

template <int SIZE>
ZX_HD uint32_t plane_code(uint64_t x, uint32_t n, uint64_t &clo, uint64_t &chi, uint32_t &nn)
{
  nn = n;
  chi = 0;
  if (n >= (uint32_t)SIZE) {
    // every coefficient is significant: the plane verbatim
    clo = x;
    return SIZE;
  }
  const uint64_t rest = x >> n;
  clo = x & lowmask(n);
  if (rest == 0) return n + 1;  // one negative group test
  // The tests find the one-bits of rest in order; the j-th, at q_j, puts its
  // value bit at q_j + j + 1 of the tests' code and a positive test right
  // after it.  E holds bit q_j + j for every j; the code is 1, E << 1 (the
  // one-bits) and P << 2, P being E less its top bit (the positive tests).
  uint64_t elo, ehi;
  expand_ones<SIZE>(rest, elo, ehi);
  const uint32_t q = 63u - clz64(rest);      // the last one-bit
  const uint32_t tb = q + popc64(rest) - 1;  // E's top bit
  // a one-bit at the last position is implied, with no test after it
  const bool implied = n + q == (uint32_t)SIZE - 1;
  // E's one-bits are at least two apart, so E << 1 and E << 2 share no bit
  // and 1 | 6E is 1 | E << 1 | E << 2: every value bit with a positive test
  // after it.  The last one-bit's test is negative (bit tb + 2 cleared) or,
  // when implied, it has neither value nor test (bits tb + 1 and up cleared).
  const uint64_t slo = (elo << 2) | (elo << 1);
  const uint64_t shi = (ehi << 2) | (ehi << 1) | (uint64_t)(((uint32_t)(elo >> 32) >> 30) |
                                                            ((uint32_t)(elo >> 32) >> 31));
  const uint32_t keep = implied ? tb + 1 : tb + 2;  // bits of 1 | 6E kept, 1 .. 127
  // low (2 << (k - 1)) - 1 for k = min(keep, 64); the high word's when keep > 64
  const uint64_t glo = (slo | 1ull) & ((2ull << (umin(keep, 64u) - 1)) - 1);
  const uint64_t ghi = keep > 64 ? shi & ((2ull << (keep - 65)) - 1) : 0ull;
  const uint32_t glen = implied ? tb + 1 : tb + 3;
  clo |= glo << n;
  chi = (ghi << n) | ((glo >> 1) >> (63 - n));
  nn = n + q + 1;
  return n + glen;
}

// 4-D blocks: planes of 256 coefficients (bitplane.rs:161-218 keeps such a
// plane in four words), coded as C's encode_ints is, one group test at a
// time, and decoded a 64-bit window of tests at a time (plane_tests256).
// The planes come from the 64-coefficient chunk
// transposes applied to each quarter of the block, and decoded planes go back
// through the 64-coefficient chunk folds, so every coefficient access has a
// fixed index and the block stays in registers.
template <class T> using Quarter = T[64];
template <class T> ZX_HD Quarter<T> &quarter(T (&a)[256], int g)
{
  return *reinterpret_cast<Quarter<T> *>(a + 64 * g);
}
template <class T> ZX_HD const Quarter<T> &quarter(const T (&a)[256], int g)
{
  return *reinterpret_cast<const Quarter<T> *>(a + 64 * g);
}

// Position of the first one-bit at or after n in a 256-bit plane, or 256.
ZX_HD uint32_t next_one256(const uint64_t (&x)[4], uint32_t n)
{
  uint32_t q = 256;
  ZX_UNROLL
  for (int c = 3; c >= 0; c--) {
    const uint32_t lo = 64u * c;
    const uint64_t keep = n <= lo ? ~0ull : n >= lo + 64 ? 0ull : ~0ull << (n - lo);
    const uint64_t v = x[c] & keep;
    q = v ? lo + ctz64(v) : q;
  }
  return q;
}

// One 256-coefficient plane x, n coefficients already significant, at most
// `bits` written (encode_planes, bitplane.rs:561-610).
template <class Sink>
ZX_HD void code_plane256(Sink &w, uint32_t &bits, uint32_t &n, const uint64_t (&x)[4])
{
  // the n coefficients already significant, verbatim
  const uint32_t m = umin(n, bits);
  bits -= m;
  ZX_UNROLL
  for (int h = 0; h < 2; h++) {
    const uint32_t base = 128u * h;
    if (m > base) {
      const uint32_t t = umin(m - base, 128u);
      const uint64_t lo = t >= 64 ? x[2 * h] : x[2 * h] & lowmask(t);
      const uint64_t hi = t > 64 ? x[2 * h + 1] & lowmask(t - 64) : 0ull;
      w.put128(lo, hi, t);
    }
  }
  // group tests: 1, the zeros before the next one-bit, and its 1 (implied
  // for the last coefficient); 0 when no one-bit is left
  while (bits != 0 && n < 256) {
    const uint32_t q = next_one256(x, n);
    if (q == 256) {
      w.zeros(1);
      bits -= 1;
      break;
    }
    const uint32_t zeros = q - n;
    const bool last = q == 255;
    const uint32_t len = zeros + 2 - (last ? 1u : 0u);
    const uint32_t take = umin(len, bits);
    w.put128(1ull, 0ull, 1);
    w.zeros(umin(zeros, take - 1));
    if (take == len && !last) w.put128(1ull, 0ull, 1);
    bits -= take;
    n = q + 1;
  }
}

template <int P, class Sink>
ZX_HD uint32_t encode_ints256(Sink &w, uint32_t maxbits, uint32_t maxprec,
                              const typename Ints<P>::U (&u)[256])
{
  typedef typename Ints<P>::U U;
  const uint32_t budget = with_maxbits(maxbits, maxprec, 256) ? maxbits : 0xffffffffu;
  const uint32_t kmin = (uint32_t)P > maxprec ? (uint32_t)P - maxprec : 0u;
  U acc = 0;
  ZX_UNROLL
  for (int i = 0; i < 256; i++) acc |= u[i];
  const uint32_t top = P == 32 ? 32u - clz32((uint32_t)acc) : 64u - clz64((uint64_t)acc);
  uint32_t bits = budget;
  uint32_t empty = (uint32_t)P - (top > kmin ? top : kmin);
  if (empty > bits) empty = bits;
  w.zeros(empty);
  bits -= empty;
  // planes kstart - 1 down to kmin are coded while budget remains
  const uint32_t kstart = (uint32_t)P - empty;
  uint32_t n = 0;
  constexpr int CH = kEncodeChunk;
  ZX_UNROLL
  for (int c = P / CH - 1; c >= 0; c--) {
    const bool need = bits != 0 && kstart > (uint32_t)(CH * c) && kmin <= (uint32_t)(CH * c + CH - 1);
    if (!wave_any(need)) continue;
    uint64_t xs[CH][4];
    ZX_UNROLL
    for (int g = 0; g < 4; g++) {
      Halves<64> hv;
      chunk_halves<P, 64>(quarter(u, g), c, hv);
      uint64_t pl[CH];
      chunk_planes<64>(hv, pl);
      ZX_UNROLL
      for (int i = 0; i < CH; i++) xs[i][g] = pl[i];
    }
    ZX_UNROLL
    for (int i = CH - 1; i >= 0; i--) {
      const uint32_t k = (uint32_t)(CH * c + i);
      if (bits != 0 && k < kstart && k >= kmin) code_plane256(w, bits, n, xs[i]);
    }
  }
  return budget - bits;
}

// v (new one-bits from coefficient n on, fewer than 64) ORed into a
// 256-bit plane
ZX_HD void or_at256(uint64_t (&x)[4], uint64_t v, uint32_t n)
{
  const uint32_t w = n >> 6, sh = n & 63;
  const uint64_t lo = v << sh;
  const uint64_t hi = (v >> 1) >> (63 - sh);  // v >> (64 - sh), 0 when sh == 0
  ZX_UNROLL
  for (int q = 0; q < 4; q++) x[q] |= (w == (uint32_t)q ? lo : 0ull) | (w + 1 == (uint32_t)q ? hi : 0ull);
}

// The group tests of a 256-coefficient plane: plane_tests' 64-bit window
// parse (below), each lane on its own, with the plane held in four words.
// The reader sits at the plane's first test, n < 256 and bits > 0.
template <class Src>
ZX_HD void plane_tests256(Src &r, uint32_t &n, uint32_t &bits, uint64_t (&x)[4], bool &one)
{
  constexpr uint64_t ODD = 0xaaaaaaaaaaaaaaaaull;
  one = false;
  for (uint32_t s = 1;; s = 0) {
    const uint64_t w = r.peek();
    if (s && !(w & 1u)) {  // negative first test
      r.consume(1);
      bits -= 1;
      return;
    }
    const uint64_t y = w >> s;
    const uint32_t W = 64 - s;
    const uint64_t z = (y + (y & ~(y << 1) & ODD)) ^ ODD;
    const uint64_t val = y & z;
    const uint64_t term = ~y & (y << 1) & z & lowmask(W);
    const bool done = term != 0;
    const uint32_t e = ctz64(term);
    const uint32_t c = done ? e + 1 : W - (uint32_t)((val >> (W - 1)) & 1u);
    const uint32_t vr = done ? e : c;
    const uint64_t vals = val & lowmask(vr);
    const uint64_t xn = compress_ones(vals, vr);
    const uint32_t coef = done ? c - popc64(vals) : c - popc64(val & lowmask(c - 1));
    // a zero scan does not read coefficient 255: it is implied
    const uint32_t q = 255 - n;
    const bool size_hit = n + coef >= 256;  // then q < coef <= 64
    const uint32_t len = size_hit ? q + popc64(xn & lowmask(q)) : c;
    const uint32_t yb = bits - s;
    if (len > yb) {
      // the budget ends after yb bits: a one is implied unless the last bit
      // read was a value
      const uint32_t nt = yb ? yb - popc64(val & lowmask(yb - 1)) : 0u;
      const bool lastv = yb != 0 && ((val >> (yb - 1)) & 1u);
      or_at256(x, xn & lowmask(nt), n);
      n += nt;
      if (!lastv) or_at256(x, 1ull, n++);
      r.consume(s + yb);
      bits = 0;
      one = true;
      return;
    }
    if (size_hit) {
      or_at256(x, xn & lowmask(q), n);
      x[3] |= 1ull << 63;
      n = 256;
      r.consume(s + len);
      bits -= s + len;
      one = true;
      return;
    }
    or_at256(x, xn, n);
    n += coef;
    r.consume(s + c);
    bits -= s + c;
    if (done) return;
  }
}

template <int P, class Src>
ZX_HD uint32_t decode_ints256(Src &r, uint32_t maxbits, uint32_t maxprec, bool round_last,
                              typename Ints<P>::U (&u)[256], bool &skip)
{
  typedef typename Ints<P>::U U;
  const bool constrained = with_maxbits(maxbits, maxprec, 256);
  const uint32_t budget = constrained ? maxbits : 0xffffffffu;
  const uint32_t kmin = (uint32_t)P > maxprec ? (uint32_t)P - maxprec : 0u;
  ZX_UNROLL
  for (int i = 0; i < 256; i++) u[i] = 0;
  uint32_t bits = budget, n = 0, m = 0, k = P;
  bool act = true;
  constexpr int CH = kDecodeChunk;
  ZX_UNROLL
  for (int c = P / CH - 1; c >= 0; c--) {
    if (!wave_any(act)) break;
    uint64_t pl[4][CH];
    ZX_UNROLL
    for (int i = CH - 1; i >= 0; i--) {
      // as decode_ints: a plane is read while budget remains and k > kmin
      const bool live = act && bits != 0;
      k -= live ? 1u : 0u;  // == CH*c + i when live
      m = live ? 0u : m;
      act = live && k >= kmin;
      uint64_t x[4] = {0, 0, 0, 0};
      bool tested = false, one = false;
      if (act) {
        m = umin(n, bits);
        bits -= m;
        ZX_UNROLL
        for (int q = 0; q < 4; q++) {
          const uint32_t lo = 64u * q;
          if (m > lo) x[q] = r.read(umin(m - lo, 64u));
        }
        tested = bits != 0 && n < 256;
        if (tested) plane_tests256(r, n, bits, x, one);
      }
      if (tested) m = one ? n : 256u;
      ZX_UNROLL
      for (int q = 0; q < 4; q++) pl[q][i] = x[q];
    }
    ZX_UNROLL
    for (int g = 0; g < 4; g++) {
      uint64_t any = 0;
      ZX_UNROLL
      for (int i = 0; i < CH; i++) any |= pl[g][i];
      if (wave_any(any != 0)) chunk_to_coeffs<P, 64, CH>(pl[g], c, quarter(u, g));
    }
  }
  if (act && bits != 0) {  // ran through plane 0 with budget left (kmin == 0)
    m = 0;
    k -= 1;
  }
  if (round_last) {
    const uint32_t rm = constrained ? m : 0u;
    const uint32_t rp = constrained ? (uint32_t)P - k : (uint32_t)P - (kmin - 1u);
    if (rp < (uint32_t)P - 1) {
      const U hi = (U)((Ints<P>::NB >> 2) >> rp);
      const U lo = (U)((Ints<P>::NB >> 1) >> rp);
      ZX_UNROLL
      for (int i = 0; i < 256; i++) u[i] += (uint32_t)i < rm ? hi : lo;
    }
  }
  skip = n == 0 && !round_last;
  return budget - bits;
}

// The coefficients of a block of 4 to 64 values in the form the plane coder
// reads them: packed 16-bit halves of each 16-plane chunk, and the bit length
// of their OR (planes above it are empty).
// Returns the OR of the coefficients.
// DNB ("deferred negabinary"): u holds the coefficients before the
// negabinary XOR (q + NB), which encode_halves applies to the transposed
// planes instead (the XOR with NB inverts every odd plane), and top is P:
// coding the leading empty planes one by one gives the same bits as skipping
// them (each is a lone negative test while no coefficient is significant),
// and a float block's planes start at or near the top anyway.
template <int P, int SIZE, bool DNB = false>
ZX_HD typename Ints<P>::U prepare_ints(const typename Ints<P>::U (&u)[SIZE],
                                       Halves<SIZE> (&hv)[P / kEncodeChunk], uint32_t &top)
{
  typedef typename Ints<P>::U U;
  U acc = 0;
  if constexpr (DNB) {
    top = P;
  } else {
    ZX_UNROLL
    for (int i = 0; i < SIZE; i++) acc |= u[i];
    top = P == 32 ? 32u - clz32((uint32_t)acc) : 64u - clz64((uint64_t)acc);
  }
  ZX_UNROLL
  for (int c = 0; c < P / kEncodeChunk; c++) chunk_halves<P, SIZE>(u, c, hv[c]);
  return acc;
}

template <int P, int SIZE, bool DNB = false, class Sink>
ZX_HD uint32_t encode_halves(Sink &w, uint32_t maxbits, uint32_t maxprec,
                             const Halves<SIZE> (&hv)[P / kEncodeChunk], uint32_t top);

template <int P, int SIZE, bool DNB = false, class Sink>
ZX_HD uint32_t encode_ints(Sink &w, uint32_t maxbits, uint32_t maxprec,
                           const typename Ints<P>::U (&u)[SIZE])
{
  if constexpr (SIZE == 256) {
    static_assert(!DNB, "4-D blocks take the XOR before their planes");
    return encode_ints256<P>(w, maxbits, maxprec, u);
  } else {
    Halves<SIZE> hv[P / kEncodeChunk];
    uint32_t top;
    prepare_ints<P, SIZE, DNB>(u, hv, top);
    return encode_halves<P, SIZE, DNB>(w, maxbits, maxprec, hv, top);
  }
}

// encode_ints and encode_planes (bitplane.rs:539-610) from prepared chunks.
template <int P, int SIZE, bool DNB, class Sink>
ZX_HD uint32_t encode_halves(Sink &w, uint32_t maxbits, uint32_t maxprec,
                             const Halves<SIZE> (&hv)[P / kEncodeChunk], uint32_t top)
{
  {
  typedef typename PWord<SIZE>::T PW;
  const uint32_t budget = with_maxbits(maxbits, maxprec, SIZE) ? maxbits : 0xffffffffu;
  const uint32_t kmin = (uint32_t)P > maxprec ? (uint32_t)P - maxprec : 0u;

  uint32_t bits = budget;
  uint32_t empty = (uint32_t)P - (top > kmin ? top : kmin);
  if (empty > bits) empty = bits;
  w.zeros(empty);
  bits -= empty;
  const uint32_t kstart = (uint32_t)P - empty;  // planes below kstart remain
  uint32_t n = 0;
  constexpr int CH = kEncodeChunk;
  static_assert(CH == 16, "the encoder's chunks are 16 planes");
  // plane k's code, cut to the budget
  auto code_plane = [&](PW plane, uint32_t k) {
    if (bits != 0 && k < kstart && k >= kmin) {
      uint64_t clo = 0, chi = 0;
      uint32_t nn;
      uint32_t len;
      if constexpr (sink_counts<Sink>::value) len = plane_len<SIZE>((uint64_t)plane, n, nn);
      else len = plane_code<SIZE>((uint64_t)plane, n, clo, chi, nn);
      const uint32_t take = umin(len, bits);
      if constexpr (Sink::kSpillPastBudget)
        w.put128(clo, chi, take);
      else
        w.put128(clo & lowmask(take), take > 64 ? chi & lowmask(take - 64) : 0ull, take);
      bits -= take;
      n = nn;
    }
  };
  // the XOR with NB deferred to the planes (DNB) inverts every odd plane
  auto fix_planes = [&](PW (&ch)[CH]) {
    if constexpr (DNB) {
      static_assert(SIZE == 64, "planes of whole words");
      ZX_UNROLL
      for (int i = 1; i < CH; i += 2) ch[i] = ~ch[i];
    }
  };
  ZX_UNROLL
  for (int c = P / CH - 1; c >= 0; c--) {
    const bool need = bits != 0 && kstart > (uint32_t)(CH * c) && kmin <= (uint32_t)(CH * c + CH - 1);
    if (!wave_any(need)) continue;
    PW ch[CH];
    chunk_planes<SIZE>(hv[c], ch);
    fix_planes(ch);
    ZX_UNROLL
    for (int i = CH - 1; i >= 0; i--) code_plane(ch[i], (uint32_t)(CH * c + i));
  }
  return budget - bits;
  }
}

// The group tests of one plane (bitplane.rs:675-770), parsed a 64-bit window
// at a time instead of one test at a time, so that a lane's cost does not
// grow with the plane's one-bits.  The code after the plane's verbatim bits
// is: a test bit; after a positive test, zeros up to a value one-bit (each
// zero a coefficient); then the next test.  Past the first test every run of
// one-bits is therefore value, test, value, test, ..., and the plane ends
// right after the first run of odd length (a value followed by a negative
// test).  Run parities come from one carry chain (adding the lowest bit of a
// run clears the run and sets the bit after it); value bits
// are the ones at even offsets in their runs, and their coefficients are
// their positions less the tests below them (compress_ones).  A run that
// reaches the last coefficient or a budget that ends inside the window is
// cut with population counts; a plane longer than the window continues in
// the next, cut where a zero scan is under way.
//
// On lanes with `go` the reader sits at the plane's first test, n < SIZE and
// bits > 0.  x gains the plane's new one-bits; n and bits advance; `one` is
// whether the last test read was positive (the plane was cut short by the
// budget or by the last coefficient, round_state's m).
//
// Blocks of 4 or 16 values keep a loop of one test per step (plane_steps):
// their planes hold few one-bits, and a step costs less than a window.
template <int SIZE, class Src>
ZX_HD void plane_steps(Src &r, uint32_t &n, uint32_t &bits, uint64_t &x, bool &one, bool go)
{
  one = false;
  while (go) {
    const uint64_t ahead = r.peek();
    one = (ahead & 1u) != 0;
    // zeros before the next one-bit, up to the last coefficient (whose
    // one-bit is implied) or the budget left after the test bit
    const uint32_t limit = umin((uint32_t)SIZE - 1 - n, bits - 1);  // <= 63
    const uint32_t zeros = ctz64(ahead >> 1);
    const bool hit = zeros < limit;
    const uint32_t adv = hit ? zeros : limit;
    const uint32_t step = one ? adv + 1 + (hit ? 1u : 0u) : 1u;
    r.consume(step);
    bits -= step;
    x |= (ahead & 1u) << (n + adv);
    n += one ? adv + 1 : 0u;
    go = one && bits != 0 && n < (uint32_t)SIZE;
  }
}

// plane_tests for a wave whose lanes all have at most 16 coefficients left
// (n >= SIZE - 16; every plane of a 2-D block): each positive test is
// followed by at most two bits per remaining coefficient, the last one
// implied, so the tests after the first take at most 30 bits and one 32-bit
// window ends every lane's plane; the same parse in 32-bit arithmetic.
template <int SIZE, class Src>
ZX_HD void plane_tests_narrow(Src &r, uint32_t &n, uint32_t &bits, uint64_t &x, bool &one, bool go,
                              uint64_t ahead)
{
  constexpr uint32_t ODD = 0xaaaaaaaau;
  constexpr uint32_t W = 31;  // window bits after the first test
  one = false;
  if (!go) return;
  const uint32_t w = (uint32_t)ahead;
  if (!(w & 1u)) {  // negative first test
    r.consume(1);
    bits -= 1;
    return;
  }
  const uint32_t y = w >> 1;
  const uint32_t z = (y + (y & ~(y << 1) & ODD)) ^ ODD;
  const uint32_t val = y & z;
  const uint32_t term = ~y & (y << 1) & z & lowmask32(W);
  const bool done = term != 0;
  const uint32_t e = ctz32(term);
  const uint32_t c = done ? e + 1 : W - ((val >> (W - 1)) & 1u);
  const uint32_t vr = done ? e : c;
  const uint32_t vals = val & lowmask32(vr);
  const uint32_t xn = compress_ones32(vals);
  const uint32_t coef = done ? c - popc32(vals) : c - popc32(val & lowmask32(c - 1));
  const uint32_t q = (uint32_t)SIZE - 1 - n;
  const bool size_hit = n + coef >= (uint32_t)SIZE;
  const uint32_t len = size_hit ? q + popc32(xn & lowmask32(q)) : c;
  const uint32_t yb = bits - 1;
  if (len > yb) {
    const uint32_t nt = yb ? yb - popc32(val & lowmask32(yb - 1)) : 0u;
    const bool lastv = yb != 0 && ((val >> (yb - 1)) & 1u);
    x |= (uint64_t)(xn & lowmask32(nt)) << n;
    n += nt;
    if (!lastv) x |= 1ull << n++;
    r.consume(1 + yb);
    bits = 0;
    one = true;
  } else if (size_hit) {
    x |= ((uint64_t)(xn & lowmask32(q)) << n) | (1ull << (SIZE - 1));
    n = SIZE;
    r.consume(1 + len);
    bits -= 1 + len;
    one = true;
  } else {
    // the window holds the whole plane: the negative test ends it
    x |= (uint64_t)xn << n;
    n += coef;
    r.consume(1 + c);
    bits -= 1 + c;
  }
}

template <int SIZE, class Src>
ZX_HD void plane_tests(Src &r, uint32_t &n, uint32_t &bits, uint64_t &x, bool &one, bool go,
                       uint64_t ahead)
{
  if constexpr (SIZE < 64) {
    plane_steps<SIZE>(r, n, bits, x, one, go);
    return;
  }
  constexpr uint64_t ODD = 0xaaaaaaaaaaaaaaaaull;
  uint32_t s = 1;  // the first window (`ahead`) starts with the first test bit
  one = false;
  while (wave_any(go)) {
    if (go) {
      const uint64_t w = s ? ahead : r.peek();
      if (s && !(w & 1u)) {  // negative first test
        r.consume(1);
        bits -= 1;
        go = false;
      } else {
        // zero scan from y's bit 0; its top W bits are the stream's
        const uint64_t y = w >> s;
        const uint32_t W = 64 - s;
        // adding the lowest bit of each run that starts at an odd position
        // clears it; z then marks, in every run, the bits at even offsets
        // from its start (values), and after every run of odd length, the
        // bit that follows it
        const uint64_t z = (y + (y & ~(y << 1) & ODD)) ^ ODD;
        const uint64_t val = y & z;                                   // value bits
        const uint64_t term = ~y & (y << 1) & z & lowmask(W);        // negative tests
        const bool done = term != 0;
        const uint32_t e = ctz64(term);  // the negative test that ends the plane
        // y bits consumed; a cut leaves the next window in a zero scan
        const uint32_t c = done ? e + 1 : W - (uint32_t)((val >> (W - 1)) & 1u);
        const uint32_t vr = done ? e : c;
        const uint64_t vals = val & lowmask(vr);
        const uint64_t xn = compress_ones(vals, vr);
        // coefficients passed: every bit but the tests (one after each value
        // but the last before the end, or before the cut)
        const uint32_t coef = done ? c - popc64(vals) : c - popc64(val & lowmask(c - 1));
        // a zero scan does not read coefficient SIZE - 1: it is implied
        const uint32_t q = (uint32_t)SIZE - 1 - n;
        const bool size_hit = n + coef >= (uint32_t)SIZE;
        const uint32_t len = size_hit ? q + popc64(xn & lowmask(q)) : c;
        const uint32_t yb = bits - s;  // budget left for y
        if (len > yb) {
          // the budget ends after yb bits: a one is implied unless the last
          // bit read was a value
          const uint32_t nt = yb ? yb - popc64(val & lowmask(yb - 1)) : 0u;
          const bool lastv = yb != 0 && ((val >> (yb - 1)) & 1u);
          x |= (xn & lowmask(nt)) << n;
          n += nt;
          if (!lastv) x |= 1ull << n++;
          r.consume(s + yb);
          bits = 0;
          one = true;
          go = false;
        } else if (size_hit) {
          x |= ((xn & lowmask(q)) << n) | (1ull << (SIZE - 1));
          n = SIZE;
          r.consume(s + len);
          bits -= s + len;
          one = true;
          go = false;
        } else {
          x |= xn << n;
          n += coef;
          r.consume(s + c);
          bits -= s + c;
          go = !done;
        }
      }
    }
    s = 0;
  }
}

// decode_ints + decode_planes + round_state (bitplane.rs:626-770), one plane
// at a time as C's decode_ints is (its group tests by plane_tests); the
// reference's run-of-empty-planes and whole-plane fast paths leave the same
// cursor, planes and (m, prec).  Returns bits read; `skip` is set when the
// block is zero and unbiased.
template <int P, int SIZE, class Src>
ZX_HD uint32_t decode_ints(Src &r, uint32_t maxbits, uint32_t maxprec, bool round_last,
                           typename Ints<P>::U (&u)[SIZE], bool &skip)
{
  if constexpr (SIZE == 256) {
    return decode_ints256<P>(r, maxbits, maxprec, round_last, u, skip);
  } else {
  typedef typename Ints<P>::U U;
  typedef typename PWord<SIZE>::T PW;
  const bool constrained = with_maxbits(maxbits, maxprec, SIZE);
  const uint32_t budget = constrained ? maxbits : 0xffffffffu;
  const uint32_t kmin = (uint32_t)P > maxprec ? (uint32_t)P - maxprec : 0u;
  ZX_UNROLL
  for (int i = 0; i < SIZE; i++) u[i] = 0;
  uint32_t bits = budget, n = 0, m = 0, k = P;
  bool act = true;
  constexpr int CH = kDecodeChunk;
  ZX_UNROLL
  for (int c = P / CH - 1; c >= 0; c--) {
    if (!wave_any(act)) break;
    // planes CH*c .. CH*c+CH-1, folded into u once the chunk is decoded
    PW pl[CH];
    ZX_UNROLL
    for (int i = CH - 1; i >= 0; i--) {
      // a plane is coded while budget remains and k > kmin; stopping at
      // kmin still takes the plane off k (round_state's precision)
      const bool live = act && bits != 0;
      k -= live ? 1u : 0u;  // == CH*c + i when live
      m = live ? 0u : m;
      act = live && k >= kmin;
      uint64_t x = 0, ahead = 0;
      bool tested = false;
      if (act) {
        m = umin(n, bits);
        bits -= m;
        if constexpr (SIZE >= 16) x = r.read_peek(m, ahead);
        else x = r.read(m);
        tested = bits != 0 && n < (uint32_t)SIZE;
      }
      bool one;
      if constexpr (SIZE == 64) {
        if (wave_all(!tested || n >= (uint32_t)SIZE - 16))
          plane_tests_narrow<SIZE>(r, n, bits, x, one, tested, ahead);
        else
          plane_tests<SIZE>(r, n, bits, x, one, tested, ahead);
      } else if constexpr (SIZE == 16) {
        // every plane of a 16-coefficient block fits one 32-bit window
        static_assert(kDecodeTable<2>, "2-D decoders read the compress8 table");
        plane_tests_narrow<SIZE>(r, n, bits, x, one, tested, ahead);
      } else {
        plane_tests<SIZE>(r, n, bits, x, one, tested, 0);
      }
      // round_state's m: the count after a plane cut short, or all
      if (tested) m = one ? n : (uint32_t)SIZE;
      pl[i] = (PW)x;
    }
    PW any = 0;
    ZX_UNROLL
    for (int i = 0; i < CH; i++) any |= pl[i];
    if (wave_any(any != 0)) chunk_to_coeffs<P, SIZE, CH>(pl, c, u);
  }
  if (act && bits != 0) {  // ran through plane 0 with budget left (kmin == 0)
    m = 0;
    k -= 1;
  }
  const bool zero = n == 0;
  if (round_last) {
    // inv_round (decode/core.rs:62-82) with round_state (bitplane.rs:661-667)
    uint32_t rm = constrained ? m : 0u;
    uint32_t rp = constrained ? (uint32_t)P - k : (uint32_t)P - (kmin - 1u);
    if (rp < (uint32_t)P - 1) {
      const U hi = (U)((Ints<P>::NB >> 2) >> rp);
      const U lo = (U)((Ints<P>::NB >> 1) >> rp);
      ZX_UNROLL
      for (int i = 0; i < SIZE; i++) u[i] += (uint32_t)i < rm ? hi : lo;
    }
  }
  skip = zero && !round_last;
  return budget - bits;
  }
}

// ---------------------------------------------------------------- integer blocks

template <int P, int D, bool DNB = false, class Sink>
ZX_HD uint32_t encode_int_block(Sink &w, typename Ints<P>::I (&q)[Blk<D>::SIZE], uint32_t minbits,
                                uint32_t maxbits, uint32_t maxprec, int rounding)
{
  typedef typename Ints<P>::I I;
  typedef typename Ints<P>::U U;
  constexpr int SIZE = Blk<D>::SIZE;
  fwd_xform<D>(q, FwdL());
  if (rounding == ROUND_FIRST && maxprec < (uint32_t)P) {
    // fwd_round (encode/core.rs:53-74)
    const I bias = (I)((Ints<P>::NB >> 2) >> maxprec);
    ZX_UNROLL
    for (int i = 0; i < SIZE; i++) q[i] = (maxprec & 1u) ? wadd(q[i], bias) : wsub(q[i], bias);
  }
  U u[SIZE];
  ZX_UNROLL
  for (int j = 0; j < SIZE; j++)
    u[j] = DNB ? (U)q[perm(D, j)] + Ints<P>::NB : ((U)q[perm(D, j)] + Ints<P>::NB) ^ Ints<P>::NB;
  uint32_t bits = encode_ints<P, SIZE, DNB>(w, maxbits, maxprec, u);
  if (bits < minbits) {
    w.zeros(minbits - bits);
    bits = minbits;
  }
  return bits;
}

template <int P, int D, class Src>
ZX_HD uint32_t decode_int_block(Src &r, typename Ints<P>::I (&q)[Blk<D>::SIZE], uint32_t maxbits,
                                uint32_t maxprec, int rounding)
{
  typedef typename Ints<P>::I I;
  typedef typename Ints<P>::U U;
  constexpr int SIZE = Blk<D>::SIZE;
  U u[SIZE];
  bool skip;
  uint32_t bits = decode_ints<P, SIZE>(r, maxbits, maxprec, rounding == ROUND_LAST, u, skip);
  if (skip) {
    ZX_UNROLL
    for (int i = 0; i < SIZE; i++) q[i] = 0;
  } else {
    ZX_UNROLL
    for (int j = 0; j < SIZE; j++) q[perm(D, j)] = (I)((u[j] ^ Ints<P>::NB) - Ints<P>::NB);
    inv_xform<D>(q, InvL());
  }
  return bits;
}

// ---------------------------------------------------------------- float helpers

// frexp exponent of the largest non-NaN magnitude (encode/core.rs:280-307).
// `special` is set when the block holds an infinity or a NaN.
template <class S, int SIZE>
ZX_HD int exponent_block(const typename Ints<ST<S>::P>::U (&b)[SIZE], bool &special)
{
  typedef typename Ints<ST<S>::P>::U U;
  constexpr int P = ST<S>::P, EB = ST<S>::EBITS, MANT = ST<S>::MANT;
  const U absmask = ~((U)1 << (P - 1));
  const U inf = (U)(((U)1 << EB) - 1) << MANT;
  // The largest magnitude's bits without masking each value: as signed
  // integers the non-negative values order by magnitude and exceed every
  // negative one, and as unsigned integers the negative values order by
  // magnitude and exceed every non-negative one; so the signed maximum is
  // the largest non-negative value (when there is one) and the unsigned
  // maximum, less its sign bit, the largest negative magnitude (when there
  // is one, else the largest non-negative value again).
  typedef typename Ints<P>::I I;
  I smx = (I)b[0];
  U umx = b[0];
  ZX_UNROLL
  for (int i = 1; i < SIZE; i++) {
    smx = (I)b[i] > smx ? (I)b[i] : smx;
    umx = b[i] > umx ? b[i] : umx;
  }
  const U pmx = smx < 0 ? (U)0 : (U)smx;
  const U nmx = umx & absmask;
  U mx = pmx > nmx ? pmx : nmx;
  special = mx >= inf;
  if (special) {
    // NaNs take no part in the maximum
    mx = 0;
    ZX_UNROLL
    for (int i = 0; i < SIZE; i++) {
      U a = b[i] & absmask;
      a = a > inf ? (U)0 : a;
      mx = a > mx ? a : mx;
    }
  }
  if (mx == 0) return -ST<S>::EBIAS;
  const int biased = (int)(mx >> MANT);
  if (biased == (1 << EB) - 1) return 0;
  return biased - (ST<S>::EBIAS - 1);
}

template <class S, int SIZE>
ZX_HD int exponent_block(const typename Ints<ST<S>::P>::U (&b)[SIZE])
{
  bool special;
  return exponent_block<S, SIZE>(b, special);
}

// 2^k as an exact normal float / double
ZX_HD float pow2f(int k) { return u2f((uint32_t)(k + 127) << 23); }
ZX_HD double pow2d(int k) { return u2d((uint64_t)(k + 1023) << 52); }

// precision_f (encode/core.rs:231-244)
ZX_HD uint32_t precision_f(int emax, uint32_t maxprec, int32_t minexp, int dims, int tight)
{
  int64_t raw = (int64_t)emax - (int64_t)minexp + 2 * (int64_t)dims + (tight ? 1 : 2);
  if (raw < 0) raw = 0;
  if (raw > (int64_t)maxprec) raw = maxprec;
  return (uint32_t)raw;
}

// trunc(x * 2^(P-2-emax)) from the IEEE bits; out of range, Inf and NaN give
// the minimum integer (x86-64 truncation, encode/core.rs:391-423).
template <class S>
ZX_HD typename Ints<ST<S>::P>::I cast_one(typename Ints<ST<S>::P>::U b, int emax, bool &exact)
{
  typedef typename Ints<ST<S>::P>::I I;
  typedef typename Ints<ST<S>::P>::U U;
  constexpr int P = ST<S>::P, MANT = ST<S>::MANT, EB = ST<S>::EBITS;
  const I imin = (I)((U)1 << (P - 1));
  const bool neg = (b >> (P - 1)) != 0;
  int ex = (int)((b >> MANT) & (((U)1 << EB) - 1));
  U mant = b & (((U)1 << MANT) - 1);
  exact = true;
  if (ex == (1 << EB) - 1) { exact = false; return imin; }
  if (ex) mant |= (U)1 << MANT;
  else ex = 1;
  if (mant == 0) { exact = !neg; return 0; }
  const int sh = ex - ST<S>::EBIAS - MANT + (P - 2) - emax;
  U mag;
  if (sh >= 0) {
    if (sh >= P - 1 || (mant >> (P - 1 - sh)) != 0) return imin;
    mag = mant << sh;
  } else {
    const int r = -sh;
    if (r > MANT + 1) { exact = false; return 0; }
    exact = (mant & (((U)1 << r) - 1)) == 0;
    mag = mant >> r;
  }
  return neg ? (I)(0 - mag) : (I)mag;
}

// exact round-to-nearest-even int64 -> double
ZX_HD double i64_to_f64(int64_t i)
{
  const double hi = (double)(int32_t)(i >> 32) * 4294967296.0;
  const double lo = (double)(uint32_t)(i & 0xffffffff);
  return hi + lo;
}

// 2^e in the type, subnormal scales included, zero below them (ldexp(1, e)
// as inv_cast computes its scale).
ZX_HD float pow2f_scale(int e)
{
  return e >= -126 ? pow2f(e) : e >= -149 ? u2f(1u << (e + 149)) : 0.0f;
}
ZX_HD double pow2d_scale(int e)
{
  return e >= -1022 ? pow2d(e) : e >= -1074 ? u2d(1ull << (e + 1074)) : 0.0;
}

// inv_cast (decode/core.rs:162-175): x = fl(s * fl(i)) with s = 2^e rounded
// to the type, e = emax - (P - 2): one conversion (to nearest even) and one
// multiply, subnormal results rounded by the hardware (IEEE denormals).
template <class S, int SIZE>
ZX_HD void inv_cast_block(const typename Ints<ST<S>::P>::I (&q)[SIZE], int e,
                          typename Ints<ST<S>::P>::U (&b)[SIZE])
{
  if constexpr (ST<S>::P == 32) {
    const float sc = pow2f_scale(e);
    ZX_UNROLL
    for (int i = 0; i < SIZE; i++) b[i] = f2u(sc * (float)q[i]);
  } else {
    const double sc = pow2d_scale(e);
    ZX_UNROLL
    for (int i = 0; i < SIZE; i++) b[i] = d2u(sc * i64_to_f64(q[i]));
  }
}

template <class S> struct Bits;
template <> struct Bits<float> {
  ZX_HD static uint32_t get(float f) { return f2u(f); }
  ZX_HD static float make(uint32_t u) { return u2f(u); }
};
template <> struct Bits<double> {
  ZX_HD static uint64_t get(double f) { return d2u(f); }
  ZX_HD static double make(uint64_t u) { return u2d(u); }
};
template <> struct Bits<int32_t> {
  ZX_HD static uint32_t get(int32_t f) { return (uint32_t)f; }
  ZX_HD static int32_t make(uint32_t u) { return (int32_t)u; }
};
template <> struct Bits<int64_t> {
  ZX_HD static uint64_t get(int64_t f) { return (uint64_t)f; }
  ZX_HD static int64_t make(uint64_t u) { return (int64_t)u; }
};

// float -> int32 conversion with the device instruction's defined results:
// truncation, saturation at the int32 limits, 0 for NaN (v_cvt_i32_f32).
ZX_HD int32_t cvt_i32_sat(float v)
{
#if defined(__HIP_DEVICE_COMPILE__)
  int32_t r;
  asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(v));
  return r;
#else
  if (v != v) return 0;
  if (v >= 0x1p31f) return INT32_MAX;
  if (v < -0x1p31f) return INT32_MIN;
  return (int32_t)v;
#endif
}

// fwd_cast of a whole block (encode/core.rs:323-423): q = trunc(x * 2^(P-2-emax)),
// and the minimum integer wherever the scaled value is not within +-2^(P-1)
// (NaN, infinity, overflow), as x86's conversion gives.  The scaling is one
// ldexp (v_ldexp on the device), which is exact wherever the result is at
// least one in magnitude, whatever emax (subnormal inputs included); smaller
// results truncate to zero whatever their rounding, so ldexp and truncation
// give the exact result.  Every finite value is below 2^(emax+1) in
// magnitude, so only a block holding an infinity or a NaN (`special`, from
// exponent_block) can scale out of range.  32-bit values convert with
// saturation; in such a block NaNs are first made infinities, and the
// positive limit (which no value within range reaches: the largest float
// below 2^31 is 2^31 - 128) then becomes the minimum.
template <class S, int SIZE>
ZX_HD void cast_block(const typename Ints<ST<S>::P>::U (&b)[SIZE], int emax, bool special,
                      typename Ints<ST<S>::P>::I (&q)[SIZE])
{
  typedef typename Ints<ST<S>::P>::I I;
  typedef typename Ints<ST<S>::P>::U U;
  constexpr int P = ST<S>::P;
  const I imin = (I)((U)1 << (P - 1));
  const int k = P - 2 - emax;
  if constexpr (P == 32) {
    U x[SIZE];
    ZX_UNROLL
    for (int i = 0; i < SIZE; i++) x[i] = b[i];
    if (special) {
      ZX_UNROLL
      for (int i = 0; i < SIZE; i++) {
        const float v = ldexpf(u2f((uint32_t)x[i]), k);
        x[i] = fabsf(v) < 0x1p31f ? x[i] : 0xff800000u;
      }
    }
    ZX_UNROLL
    for (int i = 0; i < SIZE; i++) q[i] = (I)cvt_i32_sat(ldexpf(u2f((uint32_t)x[i]), k));
  } else {
    ZX_UNROLL
    for (int i = 0; i < SIZE; i++) {
      const double v = ldexp(u2d((uint64_t)b[i]), k);
      const bool in = fabs(v) < 0x1p63;
      q[i] = in ? (I)(in ? v : 0.0) : imin;
    }
  }
}

// ---------------------------------------------------------------- lossy blocks

// Encode one lossy block of raw value bits (encode/float.rs:22-113,
// encode/integer.rs:50-93).  Returns bits written, padding included.
template <class S, int D, class Sink>
ZX_HD uint32_t encode_lossy(Sink &w, const typename Ints<ST<S>::P>::U (&b)[Blk<D>::SIZE],
                            const Cfg &c)
{
  constexpr int P = ST<S>::P, SIZE = Blk<D>::SIZE;
  typedef typename Ints<P>::I I;
  I q[SIZE];
  if constexpr (!ST<S>::FLT) {
    ZX_UNROLL
    for (int i = 0; i < SIZE; i++) q[i] = (I)b[i];
    return encode_int_block<P, D>(w, q, c.minbits, c.maxbits, c.maxprec, c.rounding);
  } else {
    bool special;
    const int emax = exponent_block<S, SIZE>(b, special);
    const uint32_t prec = precision_f(emax, c.maxprec, c.minexp, D, c.tight);
    const uint32_t e = prec ? (uint32_t)(emax + ST<S>::EBIAS) : 0u;
    if (e) {
      const uint32_t hb = ST<S>::EBITS + 1;
      w.put(2ull * e + 1, hb);
      cast_block<S, SIZE>(b, emax, special, q);
      // 3-D float blocks take the negabinary XOR on their planes (DNB)
      return hb + encode_int_block<P, D, D == 3>(w, q, sat_sub(c.minbits, hb), sat_sub(c.maxbits, hb),
                                                 prec, c.rounding);
    }
    w.put(0, 1);
    if (c.minbits > 1) {
      w.zeros(c.minbits - 1);
      return c.minbits;
    }
    return 1;
  }
}

// Decode one lossy block into raw value bits (decode/core.rs:97-260).
template <class S, int D, class Src>
ZX_HD void decode_lossy(Src &r, typename Ints<ST<S>::P>::U (&b)[Blk<D>::SIZE], const Cfg &c)
{
  constexpr int P = ST<S>::P, SIZE = Blk<D>::SIZE;
  typedef typename Ints<P>::I I;
  typedef typename Ints<P>::U U;
  I q[SIZE];
  if constexpr (!ST<S>::FLT) {
    decode_int_block<P, D>(r, q, c.maxbits, c.maxprec, c.rounding);
    ZX_UNROLL
    for (int i = 0; i < SIZE; i++) b[i] = (U)q[i];
  } else {
    if (r.read(1)) {
      constexpr uint32_t hb = 1 + ST<S>::EBITS;
      const int emax = (int)r.read(ST<S>::EBITS) - ST<S>::EBIAS;
      const uint32_t prec = precision_f(emax, c.maxprec, c.minexp, D, c.tight);
      decode_int_block<P, D>(r, q, sat_sub(c.maxbits, hb), prec, c.rounding);
      const int e = emax - (P - 2);
      inv_cast_block<S, SIZE>(q, e, b);
    } else {
      ZX_UNROLL
      for (int i = 0; i < SIZE; i++) b[i] = 0;
    }
  }
}

// ---------------------------------------------------------------- reversible blocks

// rev_encode_int_block (encode/reversible.rs:88-120) on a transformed block
template <int P, int D, class Sink>
ZX_HD uint32_t rev_encode_ints(Sink &w, const typename Ints<P>::I (&q)[Blk<D>::SIZE],
                               uint32_t maxbits, uint32_t maxprec)
{
  typedef typename Ints<P>::U U;
  constexpr int SIZE = Blk<D>::SIZE;
  constexpr uint32_t PB = P == 32 ? 5u : 6u;
  U u[SIZE];
  U acc = 0;
  ZX_UNROLL
  for (int j = 0; j < SIZE; j++) {
    u[j] = ((U)q[perm(D, j)] + Ints<P>::NB) ^ Ints<P>::NB;
    acc |= u[j];
  }
  uint32_t prec = acc ? (uint32_t)P - ctz64((uint64_t)acc) : 0u;
  prec = umin(prec, maxprec);
  if (prec < 1) prec = 1;
  w.put((uint64_t)(prec - 1), PB);
  return PB + encode_ints<P, SIZE>(w, sat_sub(maxbits, PB), prec, u);
}

template <int P, int D, class Src>
ZX_HD uint32_t rev_decode_ints(Src &r, typename Ints<P>::I (&q)[Blk<D>::SIZE], uint32_t maxbits)
{
  typedef typename Ints<P>::I I;
  typedef typename Ints<P>::U U;
  constexpr int SIZE = Blk<D>::SIZE;
  constexpr uint32_t PB = P == 32 ? 5u : 6u;
  const uint32_t prec = (uint32_t)r.read(PB) + 1;
  U u[SIZE];
  bool skip;
  uint32_t bits = decode_ints<P, SIZE>(r, sat_sub(maxbits, PB), prec, false, u, skip);
  ZX_UNROLL
  for (int j = 0; j < SIZE; j++) q[perm(D, j)] = (I)((u[j] ^ Ints<P>::NB) - Ints<P>::NB);
  return PB + bits;
}

// Reversible encode of one block (encode/reversible.rs:172-339).
template <class S, int D, class Sink>
ZX_HD uint32_t encode_reversible(Sink &w, const typename Ints<ST<S>::P>::U (&b)[Blk<D>::SIZE],
                                 const Cfg &c)
{
  constexpr int P = ST<S>::P, SIZE = Blk<D>::SIZE;
  typedef typename Ints<P>::I I;
  typedef typename Ints<P>::U U;
  I q[SIZE];
  uint32_t bits;
  if constexpr (!ST<S>::FLT) {
    ZX_UNROLL
    for (int i = 0; i < SIZE; i++) q[i] = (I)b[i];
    fwd_xform<D>(q, RFwdL());
    bits = rev_encode_ints<P, D>(w, q, c.maxbits, c.maxprec);
  } else {
    const int emax = exponent_block<S, SIZE>(b);
    const uint32_t e = (uint32_t)(emax + ST<S>::EBIAS);
    const bool tiny = e != 0 && emax < ST<S>::MINCAST;
    bool rev = !tiny;
    ZX_UNROLL
    for (int i = 0; i < SIZE; i++) {
      bool ex;
      q[i] = e ? cast_one<S>(b[i], emax, ex) : (I)0;
      if (!e) ex = b[i] == 0;  // an all-zero block reconstructs +0 only
      rev = rev && ex;
    }
    if (rev) {
      if (e) {
        constexpr uint32_t hb = 2 + ST<S>::EBITS;
        w.put(1, 2);  // 0b01
        w.put(e, ST<S>::EBITS);
        fwd_xform<D>(q, RFwdL());
        bits = hb + rev_encode_ints<P, D>(w, q, sat_sub(c.maxbits, hb), c.maxprec);
      } else {
        w.put(0, 1);
        bits = 1;
      }
    } else {
      const U tcmask = ~((U)1 << (P - 1));
      ZX_UNROLL
      for (int i = 0; i < SIZE; i++) q[i] = (I)((b[i] >> (P - 1)) ? (b[i] ^ tcmask) : b[i]);
      w.put(3, 2);  // 0b11
      fwd_xform<D>(q, RFwdL());
      bits = 2 + rev_encode_ints<P, D>(w, q, sat_sub(c.maxbits, 2), c.maxprec);
    }
  }
  if (bits < c.minbits) {
    w.zeros(c.minbits - bits);
    bits = c.minbits;
  }
  return bits;
}

// ---------------------------------------------------------------- prepared blocks

// A block of 4 to 64 values ready for coding, so that a kernel can count its
// bits and later write them from the same registers (the one-pass
// variable-rate compress): the header bits, the coefficients' plane chunks
// and encode_ints' parameters.  Every block code is the header, then (when
// `coded`) encode_ints(ibits, iprec), then zeros up to min_bits; the fields
// restate encode_lossy and encode_reversible above case by case.
template <class S, int D>
struct PreparedBlock {
  static constexpr int P = ST<S>::P;
  Halves<Blk<D>::SIZE> hv[P / kEncodeChunk];
  uint32_t top;
  uint64_t hdr;
  uint32_t hbits;
  uint32_t ibits;
  uint32_t iprec;
  bool coded;
};

// Sequency order and negabinary (encode/core.rs:12-109), then prepare_ints;
// returns the OR of the coefficients.
template <int P, int D>
ZX_HD typename Ints<P>::U prepare_coeffs(const typename Ints<P>::I (&q)[Blk<D>::SIZE],
                                         Halves<Blk<D>::SIZE> (&hv)[P / kEncodeChunk], uint32_t &top)
{
  typedef typename Ints<P>::U U;
  U u[Blk<D>::SIZE];
  ZX_UNROLL
  for (int j = 0; j < Blk<D>::SIZE; j++) u[j] = ((U)q[perm(D, j)] + Ints<P>::NB) ^ Ints<P>::NB;
  return prepare_ints<P, Blk<D>::SIZE>(u, hv, top);
}

// encode_int_block's transform and rounding (before its planes)
template <int P, int D>
ZX_HD void prepare_int_block(typename Ints<P>::I (&q)[Blk<D>::SIZE], uint32_t maxprec, int rounding,
                             Halves<Blk<D>::SIZE> (&hv)[P / kEncodeChunk], uint32_t &top)
{
  typedef typename Ints<P>::I I;
  fwd_xform<D>(q, FwdL());
  if (rounding == ROUND_FIRST && maxprec < (uint32_t)P) {
    const I bias = (I)((Ints<P>::NB >> 2) >> maxprec);
    ZX_UNROLL
    for (int i = 0; i < Blk<D>::SIZE; i++) q[i] = (maxprec & 1u) ? wadd(q[i], bias) : wsub(q[i], bias);
  }
  prepare_coeffs<P, D>(q, hv, top);
}

template <class S, int D, bool REV>
ZX_HD void prepare_block(const typename Ints<ST<S>::P>::U (&b)[Blk<D>::SIZE], const Cfg &c,
                         PreparedBlock<S, D> &pb)
{
  constexpr int P = ST<S>::P, SIZE = Blk<D>::SIZE;
  typedef typename Ints<P>::I I;
  typedef typename Ints<P>::U U;
  I q[SIZE];
  pb.hdr = 0;
  pb.hbits = 0;
  pb.coded = true;
  if constexpr (!REV) {
    if constexpr (!ST<S>::FLT) {
      ZX_UNROLL
      for (int i = 0; i < SIZE; i++) q[i] = (I)b[i];
      pb.ibits = c.maxbits;
      pb.iprec = c.maxprec;
      prepare_int_block<P, D>(q, c.maxprec, c.rounding, pb.hv, pb.top);
    } else {
      bool special;
      const int emax = exponent_block<S, SIZE>(b, special);
      const uint32_t prec = precision_f(emax, c.maxprec, c.minexp, D, c.tight);
      const uint32_t e = prec ? (uint32_t)(emax + ST<S>::EBIAS) : 0u;
      if (e) {
        const uint32_t hb = ST<S>::EBITS + 1;
        pb.hdr = 2ull * e + 1;
        pb.hbits = hb;
        pb.ibits = sat_sub(c.maxbits, hb);
        pb.iprec = prec;
        cast_block<S, SIZE>(b, emax, special, q);
        prepare_int_block<P, D>(q, prec, c.rounding, pb.hv, pb.top);
      } else {
        pb.hbits = 1;
        pb.coded = false;
      }
    }
  } else {
    constexpr uint32_t PB = P == 32 ? 5u : 6u;
    uint32_t lead = 0, lbits = 0;  // the header before the precision bits
    uint32_t room = c.maxbits;
    if constexpr (!ST<S>::FLT) {
      ZX_UNROLL
      for (int i = 0; i < SIZE; i++) q[i] = (I)b[i];
    } else {
      const int emax = exponent_block<S, SIZE>(b);
      const uint32_t e = (uint32_t)(emax + ST<S>::EBIAS);
      const bool tiny = e != 0 && emax < ST<S>::MINCAST;
      bool rev = !tiny;
      ZX_UNROLL
      for (int i = 0; i < SIZE; i++) {
        bool ex;
        q[i] = e ? cast_one<S>(b[i], emax, ex) : (I)0;
        if (!e) ex = b[i] == 0;
        rev = rev && ex;
      }
      if (rev && !e) {
        pb.hbits = 1;
        pb.coded = false;
        return;
      }
      if (rev) {
        lead = 1u | (e << 2);  // 0b01, then the exponent
        lbits = 2 + ST<S>::EBITS;
      } else {
        const U tcmask = ~((U)1 << (P - 1));
        ZX_UNROLL
        for (int i = 0; i < SIZE; i++) q[i] = (I)((b[i] >> (P - 1)) ? (b[i] ^ tcmask) : b[i]);
        lead = 3;  // 0b11
        lbits = 2;
      }
      room = sat_sub(c.maxbits, lbits);
    }
    fwd_xform<D>(q, RFwdL());
    // rev_encode_ints: the precision from the coefficients' lowest set bit
    const U acc = prepare_coeffs<P, D>(q, pb.hv, pb.top);
    uint32_t prec = acc ? (uint32_t)P - ctz64((uint64_t)acc) : 0u;
    prec = umin(prec, c.maxprec);
    if (prec < 1) prec = 1;
    pb.hdr = (uint64_t)lead | ((uint64_t)(prec - 1) << lbits);
    pb.hbits = lbits + PB;
    pb.ibits = sat_sub(room, PB);
    pb.iprec = prec;
  }
}

// The code of a prepared block through `w`; returns its bits, padding included.
template <class S, int D, class Sink>
ZX_HD uint32_t code_block(Sink &w, const PreparedBlock<S, D> &pb, const Cfg &c)
{
  w.put(pb.hdr, pb.hbits);
  uint32_t bits = pb.hbits;
  if (pb.coded) bits += encode_halves<ST<S>::P, Blk<D>::SIZE>(w, pb.ibits, pb.iprec, pb.hv, pb.top);
  if (bits < c.minbits) {
    w.zeros(c.minbits - bits);
    bits = c.minbits;
  }
  return bits;
}

// Reversible decode of one block (decode/reversible.rs:34-274).
template <class S, int D, class Src>
ZX_HD void decode_reversible(Src &r, typename Ints<ST<S>::P>::U (&b)[Blk<D>::SIZE], const Cfg &c)
{
  constexpr int P = ST<S>::P, SIZE = Blk<D>::SIZE;
  typedef typename Ints<P>::I I;
  typedef typename Ints<P>::U U;
  I q[SIZE];
  if constexpr (!ST<S>::FLT) {
    rev_decode_ints<P, D>(r, q, c.maxbits);
    inv_xform<D>(q, RInvL());
    ZX_UNROLL
    for (int i = 0; i < SIZE; i++) b[i] = (U)q[i];
  } else {
  if (r.read(1) == 0) {
    ZX_UNROLL
    for (int i = 0; i < SIZE; i++) b[i] = 0;
    return;
  }
  if (r.read(1)) {
    rev_decode_ints<P, D>(r, q, sat_sub(c.maxbits, 2));
    inv_xform<D>(q, RInvL());
    const U tcmask = ~((U)1 << (P - 1));
    ZX_UNROLL
    for (int i = 0; i < SIZE; i++) b[i] = q[i] < 0 ? ((U)q[i] ^ tcmask) : (U)q[i];
  } else {
    const int emax = (int)r.read(ST<S>::EBITS) - ST<S>::EBIAS;
    rev_decode_ints<P, D>(r, q, sat_sub(c.maxbits, 2 + ST<S>::EBITS));
    inv_xform<D>(q, RInvL());
    const int e = emax - (P - 2);
    inv_cast_block<S, SIZE>(q, e, b);
  }
  }
}

// ---------------------------------------------------------------- length reader

// The bits decode_ints reads, without its values: the plane reader of an
// index-less variable-rate stream (decode/staged.rs:216-306 reads lengths the
// same way before reconstructing blocks).  Each plane skips its n verbatim
// bits, then reads group tests with zero runs counted 64 bits at a time.
template <int P, int SIZE, class Src>
ZX_HD uint32_t skip_ints(Src &r, uint32_t maxbits, uint32_t maxprec)
{
  const bool constrained = with_maxbits(maxbits, maxprec, SIZE);
  const uint32_t budget = constrained ? maxbits : 0xffffffffu;
  const uint32_t kmin = (uint32_t)P > maxprec ? (uint32_t)P - maxprec : 0u;
  uint32_t bits = budget, n = 0, k = P;
  while (bits != 0 && k-- > kmin) {
    const uint32_t m = umin(n, bits);
    bits -= m;
    for (uint32_t left = m; left != 0;) {
      const uint32_t t = umin(left, 64);
      r.consume(t);
      left -= t;
    }
    while (bits != 0 && n < (uint32_t)SIZE) {
      bits -= 1;
      if (!r.read(1)) break;
      // zeros up to the next one-bit, the last coefficient (implied one) or
      // the budget
      for (uint32_t left = umin((uint32_t)SIZE - 1 - n, bits); left != 0;) {
        const uint32_t span = umin(left, 64);
        const uint32_t z = ctz64(r.peek());
        if (z < span) {
          r.consume(z + 1);
          bits -= z + 1;
          n += z;
          break;
        }
        r.consume(span);
        bits -= span;
        n += span;
        left -= span;
      }
      n++;
    }
  }
  return budget - bits;
}

// Bits of one block's code as decode_lossy / decode_reversible read them
// (before the caller raises the count to min_bits).
template <class S, int D, class Src>
ZX_HD uint32_t skip_lossy(Src &r, const Cfg &c)
{
  constexpr int P = ST<S>::P, SIZE = Blk<D>::SIZE;
  if constexpr (!ST<S>::FLT) {
    return skip_ints<P, SIZE>(r, c.maxbits, c.maxprec);
  } else {
    if (!r.read(1)) return 1;
    constexpr uint32_t hb = 1 + ST<S>::EBITS;
    const int emax = (int)r.read(ST<S>::EBITS) - ST<S>::EBIAS;
    const uint32_t prec = precision_f(emax, c.maxprec, c.minexp, D, c.tight);
    return hb + skip_ints<P, SIZE>(r, sat_sub(c.maxbits, hb), prec);
  }
}

template <int P, int D, class Src>
ZX_HD uint32_t rev_skip_ints(Src &r, uint32_t maxbits)
{
  constexpr uint32_t PB = P == 32 ? 5u : 6u;
  const uint32_t prec = (uint32_t)r.read(PB) + 1;
  return PB + skip_ints<P, Blk<D>::SIZE>(r, sat_sub(maxbits, PB), prec);
}

template <class S, int D, class Src>
ZX_HD uint32_t skip_reversible(Src &r, const Cfg &c)
{
  constexpr int P = ST<S>::P;
  if constexpr (!ST<S>::FLT) {
    return rev_skip_ints<P, D>(r, c.maxbits);
  } else {
    if (!r.read(1)) return 1;
    if (r.read(1)) return 2 + rev_skip_ints<P, D>(r, sat_sub(c.maxbits, 2));
    r.consume(ST<S>::EBITS);
    return 2 + ST<S>::EBITS + rev_skip_ints<P, D>(r, sat_sub(c.maxbits, 2 + ST<S>::EBITS));
  }
}

}  // namespace zx
