// zfpx_alias.h -- the write order of a serial decode, for fields whose strides
// may alias (the kernels of zfpx_alias.hip; built for the CPU by the tests).
//
// The reference decodes such a field serially (src/decompress.rs:157-163,
// 268-271): blocks in order, each block's values in its own order (x
// fastest), so where two values land on one element the later one stays.
// Every value of the field gets a ticket, (block index) * 4^d + (index in its
// block) + 1, which grows with that order; the element keeps the value with
// the largest ticket aimed at it.
#pragma once

#include "zfpx_block.h"

namespace zx {

struct AliasArgs {
  const void *dense;  // count values, natural layout (x fastest)
  void *lowest;       // the field's lowest element (element [0,0,0,0] + imin)
  int64_t n[4];       // extents of the active axes
  int64_t s[4];       // element strides of the active axes
  int64_t imin;       // offset of the lowest element from element [0,0,0,0]
  uint64_t count;     // values
  uint64_t span;      // elements from the lowest to the highest address
  int dims;
  uint32_t esz;       // bytes per value: 4 or 8
};

struct AliasSlot {
  int64_t addr;     // element offset from the lowest address
  uint64_t ticket;  // serial write order + 1
};

// Where value i of the dense copy lands, and its ticket.
ZX_HD AliasSlot alias_slot(const AliasArgs &a, uint64_t i)
{
  uint64_t block = 0, local = 0, scale = 1, lscale = 1;
  int64_t addr = -a.imin;
  uint64_t r = i;
  for (int k = 0; k < a.dims; k++) {
    const uint64_t n = (uint64_t)a.n[k];
    const uint64_t c = r % n;
    r /= n;
    block += (c >> 2) * scale;
    scale *= (n + 3) >> 2;
    local += (c & 3) * lscale;
    lscale *= 4;
    addr += (int64_t)c * a.s[k];
  }
  return AliasSlot{addr, block * lscale + local + 1};
}

}  // namespace zx
