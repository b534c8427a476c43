/* zfp_block.h -- block-level entry points of zfp-rs-ffi (zfp-rs-ffi/src/block.rs:34-776),
 * which zfp-rs-ffi exports without declaring them in its zfp.h.  Each codes one block at the
 * stream's cursor with the stream's parameters and returns the block's bits (0 for a null
 * block, a length outside 1..4 in a partial block, or a failure).  Blocks run on the GPU as
 * one-block fields. */
#ifndef ZFP_BLOCK_H
#define ZFP_BLOCK_H

#include "zfp.h"

#ifdef __cplusplus
extern "C" {
#endif

/* int32 */
size_t zfp_encode_block_int32_1(zfp_stream *stream, const int32 *block);
size_t zfp_encode_block_int32_2(zfp_stream *stream, const int32 *block);
size_t zfp_encode_block_int32_3(zfp_stream *stream, const int32 *block);
size_t zfp_encode_block_int32_4(zfp_stream *stream, const int32 *block);
size_t zfp_decode_block_int32_1(zfp_stream *stream, int32 *block);
size_t zfp_decode_block_int32_2(zfp_stream *stream, int32 *block);
size_t zfp_decode_block_int32_3(zfp_stream *stream, int32 *block);
size_t zfp_decode_block_int32_4(zfp_stream *stream, int32 *block);
size_t zfp_encode_block_strided_int32_1(zfp_stream *stream, const int32 *block, ptrdiff_t sx);
size_t zfp_encode_block_strided_int32_2(zfp_stream *stream, const int32 *block, ptrdiff_t sx, ptrdiff_t sy);
size_t zfp_encode_block_strided_int32_3(zfp_stream *stream, const int32 *block, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz);
size_t zfp_encode_block_strided_int32_4(zfp_stream *stream, const int32 *block, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz, ptrdiff_t sw);
size_t zfp_decode_block_strided_int32_1(zfp_stream *stream, int32 *block, ptrdiff_t sx);
size_t zfp_decode_block_strided_int32_2(zfp_stream *stream, int32 *block, ptrdiff_t sx, ptrdiff_t sy);
size_t zfp_decode_block_strided_int32_3(zfp_stream *stream, int32 *block, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz);
size_t zfp_decode_block_strided_int32_4(zfp_stream *stream, int32 *block, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz, ptrdiff_t sw);
size_t zfp_encode_partial_block_strided_int32_1(zfp_stream *stream, const int32 *block, size_t nx, ptrdiff_t sx);
size_t zfp_encode_partial_block_strided_int32_2(zfp_stream *stream, const int32 *block, size_t nx, size_t ny, ptrdiff_t sx, ptrdiff_t sy);
size_t zfp_encode_partial_block_strided_int32_3(zfp_stream *stream, const int32 *block, size_t nx, size_t ny, size_t nz, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz);
size_t zfp_encode_partial_block_strided_int32_4(zfp_stream *stream, const int32 *block, size_t nx, size_t ny, size_t nz, size_t nw, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz, ptrdiff_t sw);
size_t zfp_decode_partial_block_strided_int32_1(zfp_stream *stream, int32 *block, size_t nx, ptrdiff_t sx);
size_t zfp_decode_partial_block_strided_int32_2(zfp_stream *stream, int32 *block, size_t nx, size_t ny, ptrdiff_t sx, ptrdiff_t sy);
size_t zfp_decode_partial_block_strided_int32_3(zfp_stream *stream, int32 *block, size_t nx, size_t ny, size_t nz, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz);
size_t zfp_decode_partial_block_strided_int32_4(zfp_stream *stream, int32 *block, size_t nx, size_t ny, size_t nz, size_t nw, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz, ptrdiff_t sw);

/* int64 */
size_t zfp_encode_block_int64_1(zfp_stream *stream, const int64 *block);
size_t zfp_encode_block_int64_2(zfp_stream *stream, const int64 *block);
size_t zfp_encode_block_int64_3(zfp_stream *stream, const int64 *block);
size_t zfp_encode_block_int64_4(zfp_stream *stream, const int64 *block);
size_t zfp_decode_block_int64_1(zfp_stream *stream, int64 *block);
size_t zfp_decode_block_int64_2(zfp_stream *stream, int64 *block);
size_t zfp_decode_block_int64_3(zfp_stream *stream, int64 *block);
size_t zfp_decode_block_int64_4(zfp_stream *stream, int64 *block);
size_t zfp_encode_block_strided_int64_1(zfp_stream *stream, const int64 *block, ptrdiff_t sx);
size_t zfp_encode_block_strided_int64_2(zfp_stream *stream, const int64 *block, ptrdiff_t sx, ptrdiff_t sy);
size_t zfp_encode_block_strided_int64_3(zfp_stream *stream, const int64 *block, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz);
size_t zfp_encode_block_strided_int64_4(zfp_stream *stream, const int64 *block, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz, ptrdiff_t sw);
size_t zfp_decode_block_strided_int64_1(zfp_stream *stream, int64 *block, ptrdiff_t sx);
size_t zfp_decode_block_strided_int64_2(zfp_stream *stream, int64 *block, ptrdiff_t sx, ptrdiff_t sy);
size_t zfp_decode_block_strided_int64_3(zfp_stream *stream, int64 *block, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz);
size_t zfp_decode_block_strided_int64_4(zfp_stream *stream, int64 *block, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz, ptrdiff_t sw);
size_t zfp_encode_partial_block_strided_int64_1(zfp_stream *stream, const int64 *block, size_t nx, ptrdiff_t sx);
size_t zfp_encode_partial_block_strided_int64_2(zfp_stream *stream, const int64 *block, size_t nx, size_t ny, ptrdiff_t sx, ptrdiff_t sy);
size_t zfp_encode_partial_block_strided_int64_3(zfp_stream *stream, const int64 *block, size_t nx, size_t ny, size_t nz, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz);
size_t zfp_encode_partial_block_strided_int64_4(zfp_stream *stream, const int64 *block, size_t nx, size_t ny, size_t nz, size_t nw, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz, ptrdiff_t sw);
size_t zfp_decode_partial_block_strided_int64_1(zfp_stream *stream, int64 *block, size_t nx, ptrdiff_t sx);
size_t zfp_decode_partial_block_strided_int64_2(zfp_stream *stream, int64 *block, size_t nx, size_t ny, ptrdiff_t sx, ptrdiff_t sy);
size_t zfp_decode_partial_block_strided_int64_3(zfp_stream *stream, int64 *block, size_t nx, size_t ny, size_t nz, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz);
size_t zfp_decode_partial_block_strided_int64_4(zfp_stream *stream, int64 *block, size_t nx, size_t ny, size_t nz, size_t nw, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz, ptrdiff_t sw);

/* float */
size_t zfp_encode_block_float_1(zfp_stream *stream, const float *block);
size_t zfp_encode_block_float_2(zfp_stream *stream, const float *block);
size_t zfp_encode_block_float_3(zfp_stream *stream, const float *block);
size_t zfp_encode_block_float_4(zfp_stream *stream, const float *block);
size_t zfp_decode_block_float_1(zfp_stream *stream, float *block);
size_t zfp_decode_block_float_2(zfp_stream *stream, float *block);
size_t zfp_decode_block_float_3(zfp_stream *stream, float *block);
size_t zfp_decode_block_float_4(zfp_stream *stream, float *block);
size_t zfp_encode_block_strided_float_1(zfp_stream *stream, const float *block, ptrdiff_t sx);
size_t zfp_encode_block_strided_float_2(zfp_stream *stream, const float *block, ptrdiff_t sx, ptrdiff_t sy);
size_t zfp_encode_block_strided_float_3(zfp_stream *stream, const float *block, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz);
size_t zfp_encode_block_strided_float_4(zfp_stream *stream, const float *block, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz, ptrdiff_t sw);
size_t zfp_decode_block_strided_float_1(zfp_stream *stream, float *block, ptrdiff_t sx);
size_t zfp_decode_block_strided_float_2(zfp_stream *stream, float *block, ptrdiff_t sx, ptrdiff_t sy);
size_t zfp_decode_block_strided_float_3(zfp_stream *stream, float *block, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz);
size_t zfp_decode_block_strided_float_4(zfp_stream *stream, float *block, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz, ptrdiff_t sw);
size_t zfp_encode_partial_block_strided_float_1(zfp_stream *stream, const float *block, size_t nx, ptrdiff_t sx);
size_t zfp_encode_partial_block_strided_float_2(zfp_stream *stream, const float *block, size_t nx, size_t ny, ptrdiff_t sx, ptrdiff_t sy);
size_t zfp_encode_partial_block_strided_float_3(zfp_stream *stream, const float *block, size_t nx, size_t ny, size_t nz, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz);
size_t zfp_encode_partial_block_strided_float_4(zfp_stream *stream, const float *block, size_t nx, size_t ny, size_t nz, size_t nw, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz, ptrdiff_t sw);
size_t zfp_decode_partial_block_strided_float_1(zfp_stream *stream, float *block, size_t nx, ptrdiff_t sx);
size_t zfp_decode_partial_block_strided_float_2(zfp_stream *stream, float *block, size_t nx, size_t ny, ptrdiff_t sx, ptrdiff_t sy);
size_t zfp_decode_partial_block_strided_float_3(zfp_stream *stream, float *block, size_t nx, size_t ny, size_t nz, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz);
size_t zfp_decode_partial_block_strided_float_4(zfp_stream *stream, float *block, size_t nx, size_t ny, size_t nz, size_t nw, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz, ptrdiff_t sw);

/* double */
size_t zfp_encode_block_double_1(zfp_stream *stream, const double *block);
size_t zfp_encode_block_double_2(zfp_stream *stream, const double *block);
size_t zfp_encode_block_double_3(zfp_stream *stream, const double *block);
size_t zfp_encode_block_double_4(zfp_stream *stream, const double *block);
size_t zfp_decode_block_double_1(zfp_stream *stream, double *block);
size_t zfp_decode_block_double_2(zfp_stream *stream, double *block);
size_t zfp_decode_block_double_3(zfp_stream *stream, double *block);
size_t zfp_decode_block_double_4(zfp_stream *stream, double *block);
size_t zfp_encode_block_strided_double_1(zfp_stream *stream, const double *block, ptrdiff_t sx);
size_t zfp_encode_block_strided_double_2(zfp_stream *stream, const double *block, ptrdiff_t sx, ptrdiff_t sy);
size_t zfp_encode_block_strided_double_3(zfp_stream *stream, const double *block, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz);
size_t zfp_encode_block_strided_double_4(zfp_stream *stream, const double *block, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz, ptrdiff_t sw);
size_t zfp_decode_block_strided_double_1(zfp_stream *stream, double *block, ptrdiff_t sx);
size_t zfp_decode_block_strided_double_2(zfp_stream *stream, double *block, ptrdiff_t sx, ptrdiff_t sy);
size_t zfp_decode_block_strided_double_3(zfp_stream *stream, double *block, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz);
size_t zfp_decode_block_strided_double_4(zfp_stream *stream, double *block, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz, ptrdiff_t sw);
size_t zfp_encode_partial_block_strided_double_1(zfp_stream *stream, const double *block, size_t nx, ptrdiff_t sx);
size_t zfp_encode_partial_block_strided_double_2(zfp_stream *stream, const double *block, size_t nx, size_t ny, ptrdiff_t sx, ptrdiff_t sy);
size_t zfp_encode_partial_block_strided_double_3(zfp_stream *stream, const double *block, size_t nx, size_t ny, size_t nz, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz);
size_t zfp_encode_partial_block_strided_double_4(zfp_stream *stream, const double *block, size_t nx, size_t ny, size_t nz, size_t nw, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz, ptrdiff_t sw);
size_t zfp_decode_partial_block_strided_double_1(zfp_stream *stream, double *block, size_t nx, ptrdiff_t sx);
size_t zfp_decode_partial_block_strided_double_2(zfp_stream *stream, double *block, size_t nx, size_t ny, ptrdiff_t sx, ptrdiff_t sy);
size_t zfp_decode_partial_block_strided_double_3(zfp_stream *stream, double *block, size_t nx, size_t ny, size_t nz, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz);
size_t zfp_decode_partial_block_strided_double_4(zfp_stream *stream, double *block, size_t nx, size_t ny, size_t nz, size_t nw, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz, ptrdiff_t sw);

#ifdef __cplusplus
}
#endif

#endif /* ZFP_BLOCK_H */
