/*
 * zfp_oracle.h -- CPU restatement of zfp-rs's codec, used ONLY as the test
 * oracle and as the timed CPU baseline.  Nothing in the shipped library links
 * or calls this code: tests/, the smoke check and bench.py's cpu_baseline leg
 * are its only users.
 *
 * It restates, in plain C, the algorithm of (reference file:line):
 *   bitstream cursor      src/bitstream/core.rs:151-542 (+ BitReader/BitWriter :549-705)
 *   bit-plane coder       src/codec/bitplane.rs:509-790
 *   lifting transforms    src/codec/transform.rs:74-142, axis order :151-234
 *   block encode          src/codec/encode/{core,float,integer,reversible}.rs
 *   block decode          src/codec/decode/{core,reversible}.rs
 *   gathers / scatters    src/codec/encode/dim{1,2,3,4}.rs, src/codec/decode/dim{1,2,3,4}.rs
 *   field drivers         src/compress.rs:22-115, src/decompress.rs:27-147
 *   parallel drivers      src/compress.rs:132-236, src/decompress.rs:205-330,
 *                         src/execution.rs:119-148
 *   config helpers        src/config.rs:85-102, 642-686
 *
 * Scope: 1-D to 3-D fields and 1-D to 4-D blocks of int32/int64/float/double,
 * every mode and rounding.
 */
#ifndef ZFP_ORACLE_H
#define ZFP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* scalar types use the zfp_type numbering */
enum { ORC_INT32 = 1, ORC_INT64 = 2, ORC_FLOAT = 3, ORC_DOUBLE = 4 };
/* rounding: ZfpRounding (src/config.rs:115-149) */
enum { ORC_ROUND_NEVER = 0, ORC_ROUND_FIRST = 1, ORC_ROUND_LAST = 2 };

typedef struct orc_stream {
  uint64_t *words;   /* backing words (capacity = nwords) */
  size_t nwords;
  size_t word_pos;   /* index of the next word to load/store */
  uint64_t buffer;   /* buffered bits */
  uint32_t bits;     /* number of buffered bits (< 64 between calls) */
  int overflowed;    /* a write fell past the end */
  int overread;      /* a load past the end happened before the last seek */
  size_t seek_word;  /* word the last seek moved to */
} orc_stream;

typedef struct orc_config {
  uint32_t minbits, maxbits, maxprec;
  int32_t minexp;
  int32_t rounding;  /* ORC_ROUND_* */
  int32_t tight;     /* tight_error */
} orc_config;

typedef struct orc_field {
  int32_t type;              /* ORC_* */
  size_t nx, ny, nz;         /* ny == 0 -> 1-D, nz == 0 -> 2-D */
  ptrdiff_t sx, sy, sz;      /* element strides, 0 = natural */
  void *data;                /* element [0,0,0] (C zfp convention) */
} orc_field;

/* ---- bit stream (src/bitstream/core.rs) ---- */
void orc_stream_init(orc_stream *s, uint64_t *words, size_t nwords);
uint64_t orc_read_bits(orc_stream *s, uint32_t n);
uint32_t orc_read_bit(orc_stream *s);
uint64_t orc_write_bits(orc_stream *s, uint64_t value, uint32_t n);
uint32_t orc_write_bit(orc_stream *s, uint32_t bit);
void orc_rseek(orc_stream *s, uint64_t offset);
void orc_wseek(orc_stream *s, uint64_t offset);
uint64_t orc_rtell(const orc_stream *s);
uint64_t orc_wtell(const orc_stream *s);
void orc_skip(orc_stream *s, uint64_t n);
void orc_pad(orc_stream *s, uint64_t n);
uint32_t orc_align(orc_stream *s);
uint32_t orc_flush(orc_stream *s);
void orc_rewind(orc_stream *s);
int orc_overread(const orc_stream *s);
void orc_reset_overread(orc_stream *s);

/* ---- config helpers (src/config.rs) ---- */
/* rate -> block bits (rate_bits, config.rs:85-102); returns 0 on failure */
uint32_t orc_rate_bits(double rate, int type, int dims, int align);
/* most bits a block can take (block_bits, config.rs:670-686) */
uint32_t orc_block_bits(const orc_config *c, int type, int dims);
/* maximum_size (config.rs:642-658) for a contiguous field; 0 on overflow */
size_t orc_maximum_size(const orc_config *c, int type, const size_t dims[3]);
int orc_is_reversible(const orc_config *c);

/* ---- block codec: one 4^d block (d = 1..4) through strides ----
 * `data` points at the block origin; lengths nx..nw (1..4) give partial
 * blocks, padded as C's pad_block does.  Entries past `dims` are not read.
 * Returns bits written/read. */
uint32_t orc_encode_block(orc_stream *s, const orc_config *c, int type, int dims,
                          const void *data, const ptrdiff_t strides[4],
                          const size_t lengths[4]);
uint32_t orc_decode_block(orc_stream *s, const orc_config *c, int type, int dims,
                          void *data, const ptrdiff_t strides[4],
                          const size_t lengths[4]);

/* ---- field drivers ----
 * compress: returns stream byte length after flush, or 0 with *err set
 * (1 = invalid field, 2 = buffer too small).  decompress: returns bytes
 * read after align, or 0 with *err set (1 = invalid field, 3 = truncated). */
size_t orc_compress(orc_stream *s, const orc_config *c, const orc_field *f, int *err);
size_t orc_decompress(orc_stream *s, const orc_config *c, const orc_field *f, int *err);

/* chunked parallel compress with C-OMP chunk ranges (identical bytes) */
size_t orc_compress_threads(orc_stream *s, const orc_config *c, const orc_field *f,
                            int threads, int *err);
/* fixed-rate parallel decompress by per-chunk seek; falls back to serial
 * where blocks have no fixed size */
size_t orc_decompress_threads(orc_stream *s, const orc_config *c, const orc_field *f,
                              int threads, int *err);

/* per-block bit offsets of a stream (the offset index): offsets[i] is the
 * bit where block i starts, offsets[nblocks] the end.  Decodes serially. */
int orc_block_offsets(orc_stream *s, const orc_config *c, const orc_field *f,
                      uint64_t *offsets);

size_t orc_num_blocks(const orc_field *f);

/* ---- test hooks: per-block primitives for the closed-form checks ---- */
int orc_exponent_block_f32(const float *f, int size);
int orc_exponent_block_f64(const double *f, int size);
void orc_fwd_cast_f32(int32_t *i, const float *f, int size, int emax);
void orc_fwd_cast_f64(int64_t *i, const double *f, int size, int emax);
void orc_inv_cast_f32(const int32_t *i, float *f, int size, int emax);
void orc_inv_cast_f64(const int64_t *i, double *f, int size, int emax);
int64_t orc_truncate_f64(double v);
uint32_t orc_precision_f(int emax, uint32_t maxprec, int32_t minexp, int dims, int tight);
void orc_xform_i32(int32_t *p, int size, int which);
void orc_xform_i64(int64_t *p, int size, int which);
void orc_fwd_round_i32(int32_t *b, int size, uint32_t maxprec);
void orc_fwd_round_i64(int64_t *b, int size, uint32_t maxprec);
void orc_inv_round_u32(uint32_t *d, int size, uint32_t m, uint32_t prec);
void orc_inv_round_u64(uint64_t *d, int size, uint32_t m, uint32_t prec);
void orc_to_planes(const uint64_t *u, int size, uint64_t planes[64]);

#ifdef __cplusplus
}
#endif
#endif
