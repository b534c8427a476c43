/*
 * A C program written against include/zfp.h alone (the zfp-rs-ffi header's
 * interface) and linked to libzfp_hip.so: what a C caller of zfp-rs-ffi would
 * do, unchanged.  Test infrastructure only.
 *
 *   dropin            every case (compress/decompress need the GPU)
 *   dropin --host     the calls that need no GPU: sizes, fields, headers and
 *                     bit I/O
 *
 * Exit status 0 when every check holds; each failure is printed.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "zfp.h"

/* the data-kind macros of the header, as a C caller may combine them */
_Static_assert((ZFP_DATA_UNUSED | ZFP_DATA_PADDING | ZFP_DATA_META | ZFP_DATA_MISC |
                ZFP_DATA_PAYLOAD | ZFP_DATA_INDEX | ZFP_DATA_CACHE | ZFP_DATA_HEADER) ==
                   ZFP_DATA_ALL,
               "ZFP_DATA_ALL is the union of the data kinds");

static int failures = 0;

#define CHECK(cond, ...)                                  \
  do {                                                    \
    if (!(cond)) {                                        \
      fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
      fprintf(stderr, __VA_ARGS__);                       \
      fputc('\n', stderr);                                \
      failures++;                                         \
    }                                                     \
  } while (0)

/* a smooth 3-D field, seeded */
static void fill_field(float *f, size_t nx, size_t ny, size_t nz)
{
  for (size_t z = 0; z < nz; z++)
    for (size_t y = 0; y < ny; y++)
      for (size_t x = 0; x < nx; x++)
        f[x + nx * (y + ny * z)] = (float)(sin(0.11 * (double)x) * cos(0.07 * (double)y) +
                                           0.3 * sin(0.05 * (double)(x + 2 * z)));
}

/* Host-only checks: configuration, sizes, headers and bit I/O. */
static void host_checks(void)
{
  float dummy[64];
  zfp_field *f = zfp_field_3d(dummy, zfp_type_float, 40, 33, 17);
  CHECK(f != NULL, "zfp_field_3d");
  CHECK(zfp_field_dimensionality(f) == 3, "dimensionality");
  CHECK(zfp_field_blocks(f) == 10u * 9u * 5u, "blocks %zu", zfp_field_blocks(f));
  CHECK(zfp_field_size(f, NULL) == 40u * 33u * 17u, "size");

  zfp_stream *z = zfp_stream_open(NULL);
  CHECK(z != NULL, "zfp_stream_open");
  double rate = zfp_stream_set_rate(z, 8.0, zfp_type_float, 3, zfp_false);
  CHECK(rate == 8.0, "rate %g", rate);
  /* 450 blocks of 512 bits plus the largest header, in whole words */
  CHECK(zfp_stream_maximum_size(z, f) == (450u * 512u + ZFP_HEADER_MAX_BITS + 63u) / 64u * 8u,
        "maximum_size %zu", zfp_stream_maximum_size(z, f));

  uint64_t words[32];
  memset(words, 0, sizeof words);
  bitstream *bs = stream_open(words, sizeof words);
  CHECK(bs != NULL, "stream_open");
  zfp_stream_set_bit_stream(z, bs);
  size_t hb = zfp_write_header(z, f, ZFP_HEADER_FULL);
  CHECK(hb == ZFP_MAGIC_BITS + ZFP_META_BITS + ZFP_MODE_SHORT_BITS, "header bits %zu", hb);
  zfp_stream_flush(z);

  /* read it back into a fresh field and stream */
  zfp_field *g = zfp_field_alloc();
  zfp_stream *y = zfp_stream_open(bs);
  zfp_stream_rewind(y);
  CHECK(zfp_read_header(y, g, ZFP_HEADER_FULL) == hb, "read header");
  CHECK(zfp_field_type(g) == zfp_type_float && g->nx == 40 && g->ny == 33 && g->nz == 17,
        "header field");
  CHECK(zfp_stream_rate(y, 3) == 8.0, "header rate");
  CHECK(zfp_stream_compression_mode(y) == zfp_mode_fixed_rate, "header mode");

  /* bit I/O round trip */
  stream_rewind(bs);
  stream_write_bits(bs, 0x2bu, 6);
  stream_write_bits(bs, 0x123456789abcdefull, 61);
  stream_flush(bs);
  stream_rewind(bs);
  CHECK(stream_read_bits(bs, 6) == 0x2bu, "read 6");
  CHECK(stream_read_bits(bs, 61) == 0x123456789abcdefull, "read 61");

  zfp_field_free(g);
  zfp_stream_close(y);
  zfp_field_free(f);
  zfp_stream_close(z);
  stream_close(bs);
}

/* Compress a field (with a full header) and decompress it into `out`
 * through a second stream that learns everything from the header. */
static size_t round_trip(zfp_field *in, zfp_stream *z, void *out, size_t *committed)
{
  size_t cap = zfp_stream_maximum_size(z, in);
  void *buf = calloc(cap, 1);
  bitstream *bs = stream_open(buf, cap);
  zfp_stream_set_bit_stream(z, bs);
  zfp_stream_rewind(z);
  size_t hb = zfp_write_header(z, in, ZFP_HEADER_FULL);
  size_t n = zfp_compress(z, in);
  CHECK(hb > 0 && n > 0, "compress (header %zu bits, %zu bytes)", hb, n);
  *committed = stream_size(bs);
  CHECK(*committed == n && n <= cap, "stream_size %zu vs %zu", *committed, n);

  zfp_stream *r = zfp_stream_open(bs);
  zfp_field *g = zfp_field_alloc();
  zfp_stream_rewind(r);
  CHECK(zfp_read_header(r, g, ZFP_HEADER_FULL) == hb, "read header");
  zfp_field_set_pointer(g, out);
  size_t m = zfp_decompress(r, g);
  CHECK(m == n, "decompress read %zu of %zu bytes", m, n);
  zfp_field_free(g);
  zfp_stream_close(r);
  stream_close(bs);
  free(buf);
  return n;
}

static void gpu_checks(void)
{
  /* 3-D float, fixed rate 8: the stream is blocks * 512 bits + header */
  {
    const size_t nx = 64, ny = 48, nz = 40;
    float *a = malloc(nx * ny * nz * sizeof(float)), *b = calloc(nx * ny * nz, sizeof(float));
    fill_field(a, nx, ny, nz);
    zfp_field *f = zfp_field_3d(a, zfp_type_float, nx, ny, nz);
    zfp_stream *z = zfp_stream_open(NULL);
    zfp_stream_set_rate(z, 8.0, zfp_type_float, 3, zfp_false);
    size_t committed;
    size_t n = round_trip(f, z, b, &committed);
    size_t blocks = (nx / 4) * (ny / 4) * (nz / 4);
    CHECK(n == (blocks * 512 + ZFP_MAGIC_BITS + ZFP_META_BITS + ZFP_MODE_SHORT_BITS + 63) / 64 * 8,
          "rate-8 size %zu", n);
    double err = 0;
    for (size_t i = 0; i < nx * ny * nz; i++) err = fmax(err, fabs((double)a[i] - (double)b[i]));
    CHECK(err < 0.05, "rate-8 error %g", err);
    printf("3-D float rate 8: %zu values -> %zu bytes, max error %.3g\n", nx * ny * nz, n, err);
    zfp_field_free(f);
    zfp_stream_close(z);
    free(a);
    free(b);
  }
  /* 1-D int32, reversible: lossless */
  {
    const size_t n = 1001;
    int32_t *a = malloc(n * 4), *b = calloc(n, 4);
    for (size_t i = 0; i < n; i++) a[i] = (int32_t)(17 * (int64_t)i * (int64_t)i - 23 * (int64_t)i + 5);
    zfp_field *f = zfp_field_1d(a, zfp_type_int32, n);
    zfp_stream *z = zfp_stream_open(NULL);
    zfp_stream_set_reversible(z);
    size_t committed;
    round_trip(f, z, b, &committed);
    CHECK(memcmp(a, b, n * 4) == 0, "reversible int32 not lossless");
    printf("1-D int32 reversible: %zu values -> %zu bytes, lossless\n", n, committed);
    zfp_field_free(f);
    zfp_stream_close(z);
    free(a);
    free(b);
  }
  /* 2-D double, fixed accuracy 1e-4, read through negative strides */
  {
    const size_t nx = 70, ny = 45;
    double *a = malloc(nx * ny * 8), *b = calloc(nx * ny, 8);
    for (size_t y = 0; y < ny; y++)
      for (size_t x = 0; x < nx; x++) a[x + nx * y] = 100.0 * sin(0.2 * (double)x + 0.1 * (double)y);
    /* the same values seen from the far corner: element [0,0] is the last */
    zfp_field *f = zfp_field_2d(a + nx * ny - 1, zfp_type_double, nx, ny);
    zfp_field_set_stride_2d(f, -1, -(ptrdiff_t)nx);
    zfp_stream *z = zfp_stream_open(NULL);
    const double tol = 1e-4;
    zfp_stream_set_accuracy(z, tol);
    size_t committed;
    round_trip(f, z, b, &committed);
    /* b holds the decoded field in its natural layout: b[x + nx*y] is the
       value at index [x, y] of the reversed view */
    double err = 0;
    for (size_t y = 0; y < ny; y++)
      for (size_t x = 0; x < nx; x++)
        err = fmax(err, fabs(b[x + nx * y] - a[(nx - 1 - x) + nx * (ny - 1 - y)]));
    CHECK(err <= tol, "accuracy error %g > %g", err, tol);
    printf("2-D double accuracy 1e-4 (negative strides): %zu values -> %zu bytes, max error %.3g\n",
           nx * ny, committed, err);
    zfp_field_free(f);
    zfp_stream_close(z);
    free(a);
    free(b);
  }
  /* a stream too small is reported as 0, as zfp-rs-ffi does */
  {
    float a[256];
    fill_field(a, 16, 16, 1);
    zfp_field *f = zfp_field_2d(a, zfp_type_float, 16, 16);
    zfp_stream *z = zfp_stream_open(NULL);
    zfp_stream_set_precision(z, 24);
    uint64_t small[2];
    bitstream *bs = stream_open(small, sizeof small);
    zfp_stream_set_bit_stream(z, bs);
    CHECK(zfp_compress(z, f) == 0, "undersized stream not refused");
    stream_close(bs);
    zfp_field_free(f);
    zfp_stream_close(z);
  }
}

int main(int argc, char **argv)
{
  const int host_only = argc > 1 && strcmp(argv[1], "--host") == 0;
  host_checks();
  if (!host_only) gpu_checks();
  if (failures) {
    fprintf(stderr, "%d check(s) failed\n", failures);
    return 1;
  }
  printf("dropin: all checks passed (%s)\n", host_only ? "host" : "host and GPU");
  return 0;
}
