// One instantiation of the field codec's kernels (zfpx_kernels.inc.h), picked
// at compile time so that the 32 (scalar type, dimensionality, reversible)
// variants build in parallel: ZX_S is the scalar type, ZX_TAG its short name,
// ZX_D the dimensionality and ZX_REV 1 for reversible mode.
#include "zfpx_kernels.inc.h"

#define ZX_CAT5(a, b, c, d, e) a##_##b##_##c##_##d##_##e
#define ZX_NAME(op, tag, d, r) ZX_CAT5(zx, op, tag, d, r)
#define ZX_FN(op) ZX_NAME(op, ZX_TAG, ZX_D, ZX_REV)

hipError_t ZX_FN(enc)(const zx::EncodeArgs &a, hipStream_t s)
{
  return zx::launch_encode_t<ZX_S, ZX_D, ZX_REV != 0>(a, s);
}

hipError_t ZX_FN(dec)(const zx::DecodeArgs &a, hipStream_t s)
{
  return zx::launch_decode_t<ZX_S, ZX_D, ZX_REV != 0>(a, s);
}

hipError_t ZX_FN(benc)(const zx::BatchArgs &a, bool count, hipStream_t s)
{
  return zx::launch_batch_encode_t<ZX_S, ZX_D, ZX_REV != 0>(a, count, s);
}

hipError_t ZX_FN(bdec)(const zx::BatchArgs &a, bool walk, hipStream_t s)
{
  return zx::launch_batch_decode_t<ZX_S, ZX_D, ZX_REV != 0>(a, walk, s);
}

uint64_t ZX_FN(walk)(const uint64_t *words, uint64_t nwords, uint64_t pos, uint64_t count,
                     const zx::Cfg &c, uint64_t *offsets)
{
  return zx::walk_offsets<ZX_S, ZX_D, ZX_REV != 0>(words, nwords, pos, count, c, offsets);
}
