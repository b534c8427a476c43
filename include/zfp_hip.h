/*
 * zfp_hip.h -- extensions of the MI355X ZFP codec beyond zfp-rs's C ABI.
 *
 * None of these has a counterpart in zfp-rs-ffi's C header; each replaces a
 * capability the reference offers through its Rust API:
 *  - rounding per stream: ZfpConfig::with_rounding (src/config.rs:115-174),
 *    which zfp-rs-ffi fixes at build time (zfp-rs-ffi/src/stream.rs:12-16);
 *  - the execution stream: ZfpExecution (src/execution.rs:12-93) picks a
 *    thread pool; here a HIP stream;
 *  - the offset index: the per-block bit lengths of a variable-rate stream,
 *    which the reference rediscovers serially on every decode
 *    (src/decompress/pipeline.rs:139-229) and this library records at
 *    compress time beside the unchanged stream.
 */
#ifndef ZFP_HIP_EXT_H
#define ZFP_HIP_EXT_H

#include "zfp.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Rounding of later compress/decompress calls through this stream:
 * ZFP_ROUND_NEVER (default), ZFP_ROUND_FIRST or ZFP_ROUND_LAST, with
 * ZFP_WITH_TIGHT_ERROR when tight_error is true.  Returns zfp_false for an
 * unknown mode.  Not recorded in the stream: decode with the same setting. */
zfp_bool zfp_stream_set_rounding(zfp_stream *stream, int rounding, zfp_bool tight_error);
int zfp_stream_rounding(const zfp_stream *stream, zfp_bool *tight_error);

/* HIP stream (hipStream_t) the codec's kernels run on; NULL = the default
 * stream.  zfp_compress/zfp_decompress still return only when done. */
zfp_bool zfp_stream_set_hip_stream(zfp_stream *stream, void *hip_stream);

/* Kernel timing: when on, the codec records HIP events on its stream around
 * its kernels; zfp_stream_kernel_ms returns the block-coding kernel's time of
 * the last compress/decompress in milliseconds (and, through call_ms, the
 * time from its first to its last kernel), or -1 when timing is off. */
zfp_bool zfp_stream_set_timing(zfp_stream *stream, zfp_bool on);
double zfp_stream_kernel_ms(const zfp_stream *stream, double *call_ms);

/* Offset index of the last variable-rate compression through a bit stream:
 * one uint32 bit length per block, in block order, for the blocks written
 * from bit `start_bit`.  zfp_decompress uses a matching index (same start,
 * block count and parameters) and otherwise walks the stream on the GPU.
 * export: returns the block count (copies at most `capacity` lengths when
 * `lengths` is non-null); 0 when the stream holds no index.
 * import: attach an index read from elsewhere; returns zfp_false on failure. */
size_t zfp_stream_index_export(const bitstream *stream, uint32_t *lengths, size_t capacity,
                               uint64 *start_bit);
zfp_bool zfp_stream_index_import(bitstream *stream, const uint32_t *lengths, size_t count,
                                 uint64 start_bit);

/* Batched codec, for callers that code many independent arrays at once
 * (serving, checkpoint writers).  Equivalent to zfp_compress(streams[i],
 * fields[i]) (or zfp_decompress) for i < count, in order: the same stream
 * bytes, cursors and return values (in sizes[i]; 0 on failure).  Arrays whose
 * data and stream words are in device memory are coded together with a few
 * kernel launches per phase; the others, and a stream's second and later
 * arrays, are coded by the single-array call after the batch.
 * ZFP_HIP_BATCH_DEVICE_FIELDS asserts that every field's data is device
 * memory (skips the pointer queries).  Returns the number of arrays coded. */
#define ZFP_HIP_BATCH_DEVICE_FIELDS 1u
size_t zfp_hip_compress_batch(zfp_stream *const *streams, const zfp_field *const *fields, size_t count,
                              size_t *sizes, uint flags);
size_t zfp_hip_decompress_batch(zfp_stream *const *streams, zfp_field *const *fields, size_t count,
                                size_t *sizes, uint flags);

/* Planned batches: a batch of the above planned once and replayed from a
 * captured HIP graph, for callers that code the same arrays again and again
 * (a checkpoint of the same fields every step, a fixed set of serving
 * buffers).  No counterpart in the reference; an extension beside
 * zfp_hip_compress_batch.
 * zfp_hip_batch_plan: plans the arrays (encode != 0: compress, else
 * decompress) from the streams' current cursors, tunes the launches' tiling
 * (blocks per workgroup, and the LDS slice of large fixed-rate arrays) by
 * timing the whole batch a few times -- those runs write what a run writes,
 * variable-rate offset indexes included, so a decompress planned next reads
 * through them -- unless ZFP_HIP_BATCH_NO_TUNE is set, and captures the
 * device work into a graph.  Returns NULL unless every array is one the batched launches take
 * (device data and stream words, one array per stream, at least one block).
 * zfp_hip_batch_run: rewinds each stream to its planned cursor, replays the
 * graph on the first stream's HIP stream, and leaves cursors and sizes as the
 * batch call would.  Returns the arrays coded; 0 (all sizes 0) when a stream,
 * its words, configuration or offset index is not the one planned.  The
 * fields' data pointers, shapes and strides are fixed at plan time.
 * zfp_hip_batch_tiling: the chosen blocks per workgroup and slice (0: the
 * default) and the tuned run's time in ms (-1 when not tuned). */
#define ZFP_HIP_BATCH_NO_TUNE 2u
typedef struct zfp_hip_batch zfp_hip_batch;
zfp_hip_batch *zfp_hip_batch_plan(int encode, zfp_stream *const *streams, zfp_field *const *fields,
                                  size_t count, uint flags);
size_t zfp_hip_batch_run(zfp_hip_batch *batch, size_t *sizes);
size_t zfp_hip_batch_count(const zfp_hip_batch *batch);
double zfp_hip_batch_tiling(const zfp_hip_batch *batch, uint *threads, uint *slice);
void zfp_hip_batch_free(zfp_hip_batch *batch);

/* Asynchronous batches, for serving loops that keep several batches in
 * flight.  zfp_hip_batch_submit: plans the arrays as zfp_hip_batch_plan does
 * (the same conditions; no tuning, no graph), enqueues their device work on
 * `hip_stream` (a hipStream_t; NULL: the first stream's) and returns without
 * waiting; NULL when it cannot take the arrays (cursors unchanged).  The
 * streams and arrays must not be used until the wait.  zfp_hip_batch_wait:
 * waits for that work, leaves cursors and sizes as the batch call would,
 * frees the submission and returns the arrays coded.  Submissions on
 * different HIP streams overlap on the device.  No counterpart in the
 * reference (an extension beside zfp_hip_compress_batch). */
zfp_hip_batch *zfp_hip_batch_submit(int encode, zfp_stream *const *streams, zfp_field *const *fields,
                                    size_t count, uint flags, void *hip_stream);
size_t zfp_hip_batch_wait(zfp_hip_batch *submission, size_t *sizes);
/* zfp_true when a submission's device work has completed (wait will not block). */
zfp_bool zfp_hip_batch_ready(const zfp_hip_batch *submission);

/* zfp_compress that allocates nothing.  No counterpart in the reference: for
 * callers that cannot allocate device or pinned memory (code under HIP graph
 * capture of its fixed-rate work, tight serving loops, harnesses that forbid
 * allocation).  The same stream bytes, cursor and return value as
 * zfp_compress, for a field whose data and stream words are both device
 * memory and that has at most ZFP_HIP_NOALLOC_MAX_BLOCKS blocks; anything
 * else returns 0 with the stream untouched.  Variable-rate streams take their
 * per-block lengths, offsets and scan tiles from one static device array per
 * GPU (about 12 MiB, held until the process ends), shared by every call on
 * that GPU: calls that overlap on one GPU race on it (this library serializes
 * its own calls on the host, but not work another process or an unsynchronized
 * caller leaves on the device).  The offset index is kept only where the
 * stream's index buffer already has room; otherwise the stream's old index is
 * marked stale, so a later zfp_decompress walks the stream. */
#define ZFP_HIP_NOALLOC_MAX_BLOCKS 1048576u
size_t zfp_hip_compress_noalloc(zfp_stream *stream, const zfp_field *field);
/* ZFP_HIP_NOALLOC_MAX_BLOCKS, as built into the library. */
uint64 zfp_hip_noalloc_max_blocks(void);

/* zfp_true when two index tuples of the field may address one element
 * (FieldPlan::strides_may_alias, src/field_plan.rs:92-118): the axes of two
 * or more values, sorted by |stride|, do not each step past the reach of the
 * smaller ones.  zfp_decompress into such a field leaves every element the
 * value the reference's serial decode writes last (src/decompress.rs:157-163):
 * the blocks are decoded into a dense device copy, then stored in that order.
 * zfp_false for a null or invalid field. */
zfp_bool zfp_hip_field_may_alias(const zfp_field *field);

/* Test hook, no counterpart in the C ABI (zfp-rs injects the same kind of
 * failure in its tests, src/decompress/pipeline.rs:76-94 FAIL_BUFFER_AFTER):
 * from this call on, the n-th reservation of workspace, staging or pinned
 * memory (0 = the next) fails once as an out-of-memory would; a call that
 * meets it returns 0 (or NULL) before it touches the caller's stream or
 * field.  A negative n disarms.  Returns the count that was left: -1 when
 * the armed failure has been used or none was armed. */
long zfp_hip_fail_after(long n);

/* Why the calling thread's last failed call (0 or NULL returned) failed inside
 * this library: the source line and the HIP error or reservation, as text;
 * empty before any failure.  No counterpart in the C ABI, whose calls return
 * 0 for every failure. */
const char *zfp_hip_last_error(void);

/* Device the library runs on (the current HIP device when first used). */
int zfp_hip_device(void);

#ifdef __cplusplus
}
#endif

#endif /* ZFP_HIP_EXT_H */
