// zfpx_field.h -- block gather/scatter over a strided 1-D..3-D field, and the
// per-block encode/decode drivers the kernels call once per thread.
//
// Block order is x fastest (field_plan.rs:128-172); a block's origin is
// 4*(ix*sx + iy*sy + iz*sz) elements from element [0,0,0] and its extent along
// each axis is min(4, n - 4c) (field_plan.rs:183-200).  Partial blocks are
// padded as C's pad_block does: position i of a run with n real values reads
// position i, 0 when i == 3, else n - 1 (strided.rs:716-751).  Scatters store
// only the in-field values (decode/dim{1,2,3}.rs).
#pragma once

#include "zfpx_block.h"

namespace zx {

// Division by an invariant 32-bit divisor with one high multiply and shifts
// (Granlund and Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994, fig. 4.1): exact for every 32-bit dividend and
// every divisor 1 <= d < 2^32.
struct Divisor {
  uint32_t d, m, s1, s2;
};

ZX_HD Divisor make_divisor(uint32_t d)
{
  uint32_t l = 0;  // ceil(log2 d)
  while (l < 32 && ((uint64_t)1 << l) < d) l++;
  Divisor v;
  v.d = d;
  v.m = (uint32_t)(((((uint64_t)1 << l) - d) << 32) / d + 1);
  v.s1 = l < 1 ? l : 1u;
  v.s2 = l > 0 ? l - 1 : 0u;
  return v;
}

ZX_HD uint32_t mulhi32(uint32_t a, uint32_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
  return __umulhi(a, b);
#else
  return (uint32_t)(((uint64_t)a * b) >> 32);
#endif
}

ZX_HD uint32_t divide(uint32_t n, const Divisor &v)
{
  const uint32_t t = mulhi32(v.m, n);
  return (t + ((n - t) >> v.s1)) >> v.s2;
}

struct FieldDesc {
  void *data;          // element [0,0,0]; negative strides reach below it
  int64_t n[4];        // extents (1 for unused axes)
  int64_t s[4];        // element strides
  uint32_t nb[4];      // blocks per axis
  uint64_t nblocks;
  Divisor bx, by, bz;  // nb[0], nb[1], nb[2] as divisors (set_block_grid)
};

// Sets the block grid of a field whose extents are set.
ZX_HD void set_block_grid(FieldDesc &f)
{
  f.nblocks = 1;
  for (int a = 0; a < 4; a++) {
    f.nb[a] = (uint32_t)((f.n[a] + 3) / 4);
    f.nblocks *= f.nb[a];
  }
  f.bx = make_divisor(f.nb[0]);
  f.by = make_divisor(f.nb[1]);
  f.bz = make_divisor(f.nb[2]);
}

ZX_HD uint32_t pad_src(uint32_t i, uint32_t n) { return i < n ? i : (i == 3 ? 0u : n - 1); }

// Between the rows of a scalar gather or scatter: keeps the compiler from
// computing all 64 element addresses up front (128 VGPRs of 64-bit pointers),
// which would set the register budget of every kernel that contains the path.
ZX_HD void row_fence()
{
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" ::: "memory");
#endif
}

struct BlockPos {
  int64_t origin;      // element offset of the block origin from data
  uint32_t len[4];     // values inside the field along each axis
  bool full;
};

template <int D>
ZX_HD BlockPos block_pos(const FieldDesc &f, uint64_t b)
{
  BlockPos p;
  uint32_t c[4] = {0, 0, 0, 0};
  const uint32_t bb = (uint32_t)b;
  if constexpr (D == 1) {
    c[0] = bb;
  } else if constexpr (D == 2) {
    c[1] = divide(bb, f.bx);
    c[0] = bb - c[1] * f.nb[0];
  } else if constexpr (D == 3) {
    const uint32_t t = divide(bb, f.bx);
    c[0] = bb - t * f.nb[0];
    c[2] = divide(t, f.by);
    c[1] = t - c[2] * f.nb[1];
  } else {
    const uint32_t t = divide(bb, f.bx);
    c[0] = bb - t * f.nb[0];
    const uint32_t u = divide(t, f.by);
    c[1] = t - u * f.nb[1];
    c[3] = divide(u, f.bz);
    c[2] = u - c[3] * f.nb[2];
  }
  p.origin = 0;
  p.full = true;
  ZX_UNROLL
  for (int a = 0; a < 4; a++) {
    if (a < D) {
      const int64_t o = 4 * (int64_t)c[a];
      const int64_t rem = f.n[a] - o;
      p.len[a] = rem < 4 ? (uint32_t)rem : 4u;
      p.full = p.full && p.len[a] == 4;
      p.origin += o * f.s[a];
    } else {
      p.len[a] = 1;
    }
  }
  return p;
}

// Load a block's raw value bits, padding partial blocks.
template <class S, int D>
ZX_HD void gather(const FieldDesc &f, const BlockPos &p,
                  typename Ints<ST<S>::P>::U (&v)[Blk<D>::SIZE])
{
  typedef typename Ints<ST<S>::P>::U U;
  const U *base = (const U *)f.data + p.origin;
  if (p.full) {
    ZX_UNROLL
    for (int i = 0; i < Blk<D>::SIZE; i++) {
      int64_t off = (int64_t)(i & 3) * f.s[0];
      if (D >= 2) off += (int64_t)((i >> 2) & 3) * f.s[1];
      if (D >= 3) off += (int64_t)((i >> 4) & 3) * f.s[2];
      if (D >= 4) off += (int64_t)((i >> 6) & 3) * f.s[3];
      v[i] = base[off];
    }
  } else {
    // per-axis offsets of the padded positions, then one row of four at a
    // time (see row_fence)
    int64_t o[4][4];
    ZX_UNROLL
    for (int a = 0; a < 4; a++) {
      ZX_UNROLL
      for (int k = 0; k < 4; k++) o[a][k] = a < D ? (int64_t)pad_src((uint32_t)k, p.len[a]) * f.s[a] : 0;
    }
    ZX_UNROLL
    for (int r = 0; r < Blk<D>::SIZE / 4; r++) {
      const U *row = base + o[1][r & 3] + o[2][(r >> 2) & 3] + (D >= 4 ? o[3][(r >> 4) & 3] : 0);
      ZX_UNROLL
      for (int x = 0; x < 4; x++) v[4 * r + x] = row[o[0][x]];
      row_fence();
    }
  }
}

// Store a block's raw value bits; only values inside the field.
template <class S, int D>
ZX_HD void scatter(const FieldDesc &f, const BlockPos &p,
                   const typename Ints<ST<S>::P>::U (&v)[Blk<D>::SIZE])
{
  typedef typename Ints<ST<S>::P>::U U;
  U *base = (U *)f.data + p.origin;
  ZX_UNROLL
  for (int r = 0; r < Blk<D>::SIZE / 4; r++) {
    const uint32_t y = r & 3, z = (r >> 2) & 3, w = (r >> 4) & 3;
    const bool row_in =
        p.full || ((D < 2 || y < p.len[1]) && (D < 3 || z < p.len[2]) && (D < 4 || w < p.len[3]));
    U *row = base + (D >= 2 ? (int64_t)y * f.s[1] : 0) + (D >= 3 ? (int64_t)z * f.s[2] : 0) +
             (D >= 4 ? (int64_t)w * f.s[3] : 0);
    ZX_UNROLL
    for (int x = 0; x < 4; x++)
      if (row_in && (p.full || (uint32_t)x < p.len[0])) row[(int64_t)x * f.s[0]] = v[4 * r + x];
    row_fence();
  }
}

template <class S, int D, bool REV, class Sink>
ZX_HD uint32_t encode_values(Sink &w, const typename Ints<ST<S>::P>::U (&v)[Blk<D>::SIZE],
                             const Cfg &c)
{
  if constexpr (REV) return encode_reversible<S, D>(w, v, c);
  else return encode_lossy<S, D>(w, v, c);
}

template <class S, int D, bool REV, class Src>
ZX_HD void decode_values(Src &r, typename Ints<ST<S>::P>::U (&v)[Blk<D>::SIZE], const Cfg &c)
{
  if constexpr (REV) decode_reversible<S, D>(r, v, c);
  else decode_lossy<S, D>(r, v, c);
}

// Bits block code occupies from the reader's position, min_bits padding
// included (what decode_values reads, raised as skip_to raises it).
template <class S, int D, bool REV, class Src>
ZX_HD uint32_t skip_values(Src &r, const Cfg &c)
{
  const uint32_t used = REV ? skip_reversible<S, D>(r, c) : skip_lossy<S, D>(r, c);
  return used < c.minbits ? c.minbits : used;
}

// The length reader over a run of blocks: offsets[i] = the start bit of block
// i of the run, the first at `pos`; returns the bit after the last.
template <class S, int D, bool REV>
ZX_HD uint64_t walk_offsets(const uint64_t *words, uint64_t nwords, uint64_t pos, uint64_t count,
                            const Cfg &c, uint64_t *offsets)
{
  for (uint64_t i = 0; i < count; i++) {
    offsets[i] = pos;
    Source<GlobalWords> r;
    r.init(GlobalWords{words, nwords}, pos);
    pos += skip_values<S, D, REV>(r, c);
  }
  return pos;
}

// Encode block b of the field through `w`; returns its bit count.
template <class S, int D, bool REV, class Sink>
ZX_HD uint32_t encode_field_block(const FieldDesc &f, uint64_t b, Sink &w, const Cfg &c)
{
  typename Ints<ST<S>::P>::U v[Blk<D>::SIZE];
  const BlockPos p = block_pos<D>(f, b);
  gather<S, D>(f, p, v);
  return encode_values<S, D, REV>(w, v, c);
}

// Decode block b of the field from `r`.
template <class S, int D, bool REV, class Src>
ZX_HD void decode_field_block(const FieldDesc &f, uint64_t b, Src &r, const Cfg &c)
{
  typename Ints<ST<S>::P>::U v[Blk<D>::SIZE];
  decode_values<S, D, REV>(r, v, c);
  const BlockPos p = block_pos<D>(f, b);
  scatter<S, D>(f, p, v);
}

}  // namespace zx
