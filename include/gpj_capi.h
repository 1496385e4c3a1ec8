/*
 * gpj_capi.h -- C ABI of the MI355X GpJSON engine (libgpj_hip.so).
 *
 * Plain pointers and sizes, no C++ types, no exceptions across the boundary.
 * Every entry point returns a gpj_status; on error gpj_last_error() returns a
 * thread-local message.  Device pointers are HIP device allocations; `stream`
 * is a hipStream_t (NULL = the null stream).  All work is stream-ordered.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference repository root).
 */
#ifndef GPJ_CAPI_H
#define GPJ_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t gpj_status;

enum {
  GPJ_OK = 0,
  GPJ_ERR_INVALID_ARGUMENT = -1, /* ImplementationError */
  GPJ_ERR_DEVICE = -2,           /* CudaError (src/gpjson/cuda/cuda.cpp:19-29) */
  GPJ_ERR_OUT_OF_MEMORY = -3,    /* CudaError from cudaMalloc */
  GPJ_ERR_QUERY_EXECUTION = -4,  /* QueryExecutionError (include/gpjson/query/error.hpp:17-20) */
  GPJ_ERR_MALFORMED_QUERY = -5,  /* MalformedQueryError (include/gpjson/query/error.hpp:7-10) */
  GPJ_ERR_UNSUPPORTED_QUERY = -6,/* UnsupportedQueryError (include/gpjson/query/error.hpp:12-15) */
  GPJ_ERR_CAPACITY = -7,         /* a caller-provided buffer is too small */
  GPJ_ERR_OUT_OF_RANGE = -8      /* std::out_of_range (an index literal overflows) */
};

/* Limits the query path shares with the reference
 * (include/gpjson/query/kernels/orig.hpp:9-10, leveled-bitmaps-index.cu:6). */
#define GPJ_MAX_QUERY_LEVELS 16
#define GPJ_MAX_QUERY_IR_BYTES 4096
#define GPJ_MAX_INDEX_LEVELS 22

/* Thread-local text of the last error on this thread ("" if none). */
const char *gpj_last_error(void);

/* ---- device runtime: replaces DeviceArray / check / device_available
 *      (include/gpjson/cuda/cuda.hpp:10-45, src/gpjson/cuda/cuda.cpp:33-125) */
int gpj_device_available(void);
gpj_status gpj_alloc(void **d_ptr, uint64_t bytes);
gpj_status gpj_free(void *d_ptr);
gpj_status gpj_h2d(void *d_dst, const void *h_src, uint64_t bytes, void *stream);
gpj_status gpj_d2h(void *h_dst, const void *d_src, uint64_t bytes, void *stream);
gpj_status gpj_stream_sync(void *stream);
gpj_status gpj_device_sync(void);

/* ---- asynchronous host runtime (no reference counterpart: the reference
 * copies synchronously from pageable memory, src/gpjson/file/file.cpp:30-36).
 * Pinned host memory makes gpj_h2d / gpj_d2h_async truly asynchronous, so
 * copies overlap kernels on other streams.  gpj_alloc_async / gpj_free_async
 * are stream-ordered (a device memory pool that keeps freed memory cached).
 * gpj_event_query returns 1 when the event has completed, 0 while pending,
 * or a negative status. */
gpj_status gpj_host_alloc(void **h_ptr, uint64_t bytes);
gpj_status gpj_host_free(void *h_ptr);
gpj_status gpj_alloc_async(void **d_ptr, uint64_t bytes, void *stream);
gpj_status gpj_free_async(void *d_ptr, void *stream);
gpj_status gpj_d2h_async(void *h_dst, const void *d_src, uint64_t bytes, void *stream);
/* d_dst[i] = *d_srcs[i] for i < count <= 16 (a null source gives 0): one
 * launch gathers a handful of device words so that one copy brings them to
 * the host.  d_srcs is a host array of device pointers.  Stream-ordered. */
gpj_status gpj_gather_words(uint64_t *d_dst, const uint64_t *const *d_srcs, int32_t count,
                            void *stream);
gpj_status gpj_event_create(void **event);
gpj_status gpj_event_record(void *event, void *stream);
gpj_status gpj_event_query(void *event);
gpj_status gpj_event_sync(void *event);
gpj_status gpj_event_elapsed_ms(void *start, void *end, float *ms);
gpj_status gpj_event_destroy(void *event);
/* Later work on `stream` waits for `event` (recorded on another stream). */
gpj_status gpj_stream_wait_event(void *stream, void *event);

/* ---- structural index: replaces IndexBuilder::build
 *      (include/gpjson/index/index_builder.hpp:55-63; kernels of
 *      src/gpjson/index/orig_index_builder.cu:224-395) */

typedef struct {
  /* Launch geometry of the reference options struct
   * (include/gpjson/index/index_builder.hpp:34-42).  Accepted and recorded;
   * results never depend on them. */
  int32_t grid_size;
  int32_t block_size;
  int32_t reduction_grid_size;
  int32_t reduction_block_size;
  /* Newline-index capacity hint in entries; 0 = default (n/64 + 2).  When the
   * input holds more lines, the build falls back to an exact second pass. */
  int64_t newline_capacity;
} gpj_index_opts;

/* Device-resident indices of one partition (include/gpjson/index/index.hpp:83-109). */
typedef struct {
  int64_t *d_newline;   /* int64[num_lines]: 0 then every '\n' offset */
  int64_t *d_num_lines; /* device copy of num_lines (int64) */
  int64_t num_lines;    /* 1 + count('\n') */
  uint64_t *d_string;   /* uint64[level_size] string mask */
  uint64_t *d_leveled;  /* uint64[num_levels][level_size], level-major */
  int64_t level_size;   /* ceil(n / 64) */
  int32_t num_levels;
  int32_t owns_memory;  /* set by gpj_index_build; gpj_index_free releases */
} gpj_index;

/* Allocate and build all three indices for n bytes at d_in.  Synchronizes
 * `stream` once to learn num_lines, as the reference does
 * (orig_index_builder.cu:325-330). */
gpj_status gpj_index_build(const uint8_t *d_in, uint64_t n, int32_t levels,
                           const gpj_index_opts *opts, gpj_index *out,
                           void *stream);
gpj_status gpj_index_free(gpj_index *index);

/* Bytes of device scratch gpj_index_build_into needs for n input bytes. */
uint64_t gpj_index_scratch_bytes(uint64_t n);

/* Stream-ordered build into caller-owned buffers; allocates nothing and never
 * synchronizes.  d_string: level_size words, d_leveled: levels*level_size
 * words, d_newline: newline_capacity entries, d_num_lines: 1 int64, d_scratch:
 * gpj_index_scratch_bytes(n) bytes.  If the input has more than
 * newline_capacity lines, *d_num_lines still receives the true count and only
 * the first newline_capacity entries are written; the caller checks.  If the
 * build failed (a tile's bounded wait for its predecessors' carry gave up; the
 * reference checks every launch instead, orig_index_builder.cu:107,115),
 * *d_num_lines receives -1 and gpj_index_status reports it. */
gpj_status gpj_index_build_into(const uint8_t *d_in, uint64_t n, int32_t levels,
                                uint64_t *d_string, uint64_t *d_leveled,
                                int64_t *d_newline, int64_t newline_capacity,
                                int64_t *d_num_lines, void *d_scratch,
                                void *stream);

/* ---- shards of one document (no counterpart in the reference, which
 *      indexes each partition from the start state,
 *      src/gpjson/index/orig_index_builder.cu:296-395).  A document cut at
 *      64-byte multiples into shards (one per GPU) is indexed shard by shard:
 *      (1) each shard is built alone (gpj_index_build_into) and
 *      gpj_index_shard_summary condenses it into its transfer table, 16 bytes;
 *      (2) the summaries, gathered in shard order, give every shard's carry-in
 *      (gpj_shard_carries, on the host); (3) gpj_index_build_shard rebuilds
 *      each shard from its carry-in.  Its string mask and leveled rows are
 *      then the document's words from the shard's first word on, and its
 *      newline index holds the shard's newlines as offsets from the shard's
 *      start (entry 0 = 0 as for any input; document entry
 *      carry.newlines + i is the shard's entry i, i >= 1, plus the shard's
 *      offset). */
typedef struct {
  int32_t escape;    /* the shard's first byte is escaped (0 or 1) */
  int32_t in_string; /* the shard starts inside a string (0 or 1) */
  int32_t depth;     /* nesting depth before it, int8 (the document starts at -1) */
  int32_t reserved;
  int64_t newlines;  /* newlines before the shard */
} gpj_carry;
/* The state change across a shard: byte s (s = escape | in_string << 1) of
 * bits 0..31 of `table` is the state after the shard from start state s
 * (escape | in_string << 1), byte s of bits 32..63 its depth change mod 256;
 * `newlines` counts the shard's newlines. */
typedef struct {
  uint64_t table;
  uint64_t newlines;
} gpj_shard_summary;
/* After a gpj_index_build_into or gpj_index_build_shard of n bytes with
 * d_scratch (n must be that build's size; check the build's status first,
 * gpj_index_status): writes that input's summary to device memory.
 * Stream-ordered. */
gpj_status gpj_index_shard_summary(const void *d_scratch, uint64_t n,
                                   gpj_shard_summary *d_summary, void *stream);
/* Host: carries[k] = the state before shard k (carries[0] = the document's
 * start, escape 0, in_string 0, depth -1, newlines 0; carries[count] = the
 * state after the last shard). */
gpj_status gpj_shard_carries(const gpj_shard_summary *summaries, int32_t count,
                             gpj_carry *carries);
/* gpj_index_build_into from the carry-in state instead of the document's
 * start (the carry's newline count is not used: the newline index stays
 * shard-relative). */
gpj_status gpj_index_build_shard(const uint8_t *d_in, uint64_t n, int32_t levels,
                                 uint64_t *d_string, uint64_t *d_leveled,
                                 int64_t *d_newline, int64_t newline_capacity,
                                 int64_t *d_num_lines, void *d_scratch,
                                 const gpj_carry *carry_in, void *stream);

/* gpj_index_build_into without a scratch argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops and
 * harnesses that forbid hipMalloc).  It has no counterpart in the reference.
 * The scratch is a static device array of the library,
 * GPJ_STATIC_SCRATCH_BYTES (16 MiB + 256 B) on every GPU it runs on, held
 * until the process ends; inputs above GPJ_STATIC_MAX_BYTES (16 GiB) return
 * GPJ_ERR_CAPACITY.  Calls nothing that allocates.  Every call on one GPU
 * shares that GPU's array and resets it first, so calls that overlap on one
 * GPU (on different streams or threads) race: serialize them.  Results and
 * failure reporting (-1 lines) are those of gpj_index_build_into. */
#define GPJ_STATIC_MAX_BYTES (16ull << 30)
#define GPJ_STATIC_SCRATCH_BYTES ((16ull << 20) + 256)
gpj_status gpj_index_build_static(const uint8_t *d_in, uint64_t n, int32_t levels,
                                  uint64_t *d_string, uint64_t *d_leveled,
                                  int64_t *d_newline, int64_t newline_capacity,
                                  int64_t *d_num_lines, void *stream);

/* Whether the last index build that used d_scratch (gpj_index_build_into
 * over n bytes, or gpj_batch_index over n total bytes) failed: *failed = 1
 * when a tile's bounded wait gave up, so its indices are not valid (the
 * stream-ordered line count reads -1 then).  Synchronizes `stream`.  Replaces
 * the error check after each reference launch
 * (src/gpjson/index/orig_index_builder.cu:107,115). */
gpj_status gpj_index_status(const void *d_scratch, uint64_t n, int32_t *failed, void *stream);

/* Debug (no counterpart in the reference): the number of polls a waiting
 * index tile makes before it gives up and flags the build as failed; negative
 * restores the default (2^24).  0 makes every tile that would wait give up at
 * once, so tests drive the failure path without any kernel hanging.
 * Process-wide; set it while no index build runs. */
gpj_status gpj_debug_set_wait_limit(int64_t polls);

/* Debug: escape and unescaped-quote bitmaps (reference kernels
 * src/gpjson/index/kernels/orig/uncombined/escape-index.cu and
 * src/gpjson/index/kernels/orig/quote-index.cu), ceil(n/64) words each. */
gpj_status gpj_debug_escape_quote(const uint8_t *d_in, uint64_t n,
                                  uint64_t *d_escape, uint64_t *d_quote,
                                  void *stream);

/* Calibration: a plain device-to-device copy kernel (no counterpart in the
 * reference).  Its HBM traffic counters give the scale factor that turns
 * profiler FETCH_SIZE / WRITE_SIZE readings into bytes.  16-byte aligned. */
gpj_status gpj_debug_copy(void *d_dst, const void *d_src, uint64_t bytes, void *stream);

/* Debug (no counterpart in the reference): the query kernels hold positions
 * as 32-bit offsets from each record's start and hand records that reach
 * `bytes` from it (or backward searches that reach `bytes` below it) to a
 * 64-bit kernel.  The default, 2^31 - 2^20, is restored with 0; tests set a
 * small value (>= 4096) to drive the 64-bit kernel on small inputs.  Results
 * never depend on it.  Process-wide; set it while no query runs. */
gpj_status gpj_debug_set_narrow_span(int64_t bytes);

/* Debug (no counterpart in the reference): a single-input index build of at
 * least `min_tiles` 16 KiB tiles runs two tiles per workgroup (one carry wait
 * for both); smaller inputs run one.  The default, 8192 tiles (128 MiB), is
 * restored with 0; tests set 1 (every input in pairs) or a huge count (none)
 * to drive either form on any input.  Results never depend on it.
 * Process-wide; set it while no index build runs. */
gpj_status gpj_debug_set_index_pairing(int64_t min_tiles);

/* Query walk (no counterpart in the reference, whose kernel runs one thread
 * per line, query.cu:144-185): GPJ_WALK_LANE runs one record per lane (64
 * records per wave), GPJ_WALK_WAVE one record per wave over the record staged
 * in LDS.  GPJ_WALK_AUTO (the default) picks the wave walk when the call's
 * record bound R (max_lines, or the record bound of the records calls) is at
 * most GPJ_WAVE_WALK_RECORDS, or at most GPJ_WAVE_WALK_LONG_RECORDS with at
 * least GPJ_WAVE_WALK_LONG_BYTES input bytes per record (n >= that * R), and
 * the lane walk otherwise (measured crossovers on MI355X).  Inputs whose
 * device address is not 16-byte aligned, and inputs above 2^31 - 2^20 bytes,
 * always take the lane walk.  Results
 * never depend on the walk.  Process-wide; set it while no query runs. */
#define GPJ_WALK_AUTO 0
#define GPJ_WALK_LANE 1
#define GPJ_WALK_WAVE 2
#define GPJ_WAVE_WALK_RECORDS 4096
#define GPJ_WAVE_WALK_LONG_RECORDS 24576
#define GPJ_WAVE_WALK_LONG_BYTES 2048
gpj_status gpj_query_set_walk(int32_t walk);

/* ---- query: replaces the query kernel launch of QueryExecutor::execute_batch
 *      (src/gpjson/query/kernels/orig/query_executor_host.cu:179-237,
 *      src/gpjson/query/kernels/orig/query.cu:144-454) */

/* Run one compiled query (IR already in device memory at d_ir, ir_len bytes)
 * over every line of the index.  d_spans receives int64
 * [max_lines][num_results][2] partition-local {start, end} pairs, -1 for no
 * match; lines beyond *index->d_num_lines are not touched.  max_lines bounds
 * the grid (use index->num_lines when known).  index->num_lines is the number
 * of entries d_newline holds (the exact count or the buffer's capacity): a
 * device count above it (an overflowed capacity) is clamped to it, so no
 * entry past the buffer is read, and a count of -1 (a failed index build)
 * runs no line.  Stream-ordered; allocates nothing. */
gpj_status gpj_query_run(const uint8_t *d_in, uint64_t n, const gpj_index *index,
                         const uint8_t *d_ir, uint32_t ir_len,
                         int32_t num_results, int64_t max_lines,
                         int64_t *d_spans, void *stream);

/* Extension mode (bytecode compiled with GPJ_COMPILE_EXTENSIONS; no
 * counterpart in the reference, whose compiler rejects `[*]` at
 * src/gpjson/query/query_compiler.cpp:287-290): a record may yield any number
 * of results, in execution order.  d_spans receives the first `capacity`
 * {start, end} pairs of each line ([max_lines][capacity][2]; entries past
 * the line's count are left untouched) and d_counts the full count per line;
 * a caller re-runs with a larger capacity when any count exceeds it.
 * Stream-ordered; allocates nothing. */
gpj_status gpj_query_run_ext(const uint8_t *d_in, uint64_t n, const gpj_index *index,
                             const uint8_t *d_ir, uint32_t ir_len, int32_t capacity,
                             int64_t max_lines, int64_t *d_spans, int64_t *d_counts,
                             void *stream);

/* A batch of compiled queries run over the same index in one launch: each
 * record is read once for all of them (the reference launches one kernel per
 * query, query_executor_host.cu:179-237; results are the same).  Up to
 * GPJ_MAX_BATCH_QUERIES queries with at most 8192 bytecode bytes in total.
 * variable = 0: reference fixed slots (capacity = slot count, -1 = no match);
 * variable = 1: extension mode as in gpj_query_run_ext (entries past a line's
 * count are left untouched).  `queries` is a host array.  Stream-ordered;
 * allocates nothing. */
#define GPJ_MAX_BATCH_QUERIES 8
typedef struct {
  const uint8_t *d_ir;
  uint32_t ir_len;
  int32_t capacity;
  int32_t variable;
  int32_t max_level; /* the compiler's max_depth (deepest level the bytecode
                        reaches); 0 = unknown.  Picks a smaller interpreter. */
  int64_t *d_spans;  /* [max_lines][capacity][2] */
  int64_t *d_counts; /* [max_lines]: variable mode, the full count per line;
                        fixed mode, optional (may be null): results stored */
} gpj_query_desc;
gpj_status gpj_query_run_batch(const uint8_t *d_in, uint64_t n, const gpj_index *index,
                               const gpj_query_desc *queries, int32_t num_queries,
                               int64_t max_lines, void *stream);

/* ---- extension: top-level array elements as records (no reference
 * counterpart; the reference treats a one-line array document as a single
 * record and, having no [*], cannot iterate it: query_compiler.cpp:287-290).
 * Record bounds are the structural characters of leveled row 0: for a
 * document whose top-level value is an array, its '[', the ',' between
 * elements and its ']'.  gpj_row_bounds writes their positions (the first
 * `capacity` of them) to d_bounds and their number M to *d_num_bounds (device
 * int64); there are M - 1 records and record r spans d_bounds[r] ..
 * d_bounds[r + 1].  Scratch: gpj_row_bounds_scratch_bytes(level_size) bytes.
 * Stream-ordered; allocates nothing. */
uint64_t gpj_row_bounds_scratch_bytes(int64_t level_size);
gpj_status gpj_row_bounds(const gpj_index *index, int64_t *d_bounds, int64_t capacity,
                          int64_t *d_num_bounds, void *d_scratch, void *stream);

/* Runs a query batch over those records (results [record][capacity][2]).  A
 * record is trimmed and interpreted exactly as a line, with interpreter level
 * k reading leveled row k + 1: the results equal line mode over the same
 * document with its top-level ',' replaced by '\n' and its '[' ']' by ' '.
 * Queries are compiled without the leading `$[*]` (e.g. `$[*].items[*].sku`
 * runs as `$.items[*].sku`); the index needs one more leveled row than the
 * query's max depth.  max_records must not exceed the bounds buffer's
 * capacity - 1 (record r reads d_bounds[r + 1]).  Stream-ordered; allocates
 * nothing. */
gpj_status gpj_query_run_records(const uint8_t *d_in, uint64_t n, const gpj_index *index,
                                 const int64_t *d_bounds, const int64_t *d_num_bounds,
                                 const gpj_query_desc *queries, int32_t num_queries,
                                 int64_t max_records, void *stream);

/* ---- extension: dense results, counted, scanned and emitted on the device
 * (no counterpart in the reference, which copies back a fixed
 * [lines][slots] array per query,
 * src/gpjson/query/kernels/orig/query_executor_host.cu:199-228).  Each
 * query's results land in one span array in record order plus CSR offsets:
 * record r's {start, end} pairs are d_spans[d_offsets[r] .. d_offsets[r + 1])
 * and d_offsets[max_records] is the total (entries past the record count
 * hold it too).  No per-record capacity: fixed-slot bytecode keeps its slot
 * limit (desc.capacity = the compiled slot count; its stored slots are always
 * the first ones, so its spans are those slots in order), extension bytecode
 * (desc.variable = 1) any number.  If a total exceeds its `capacity`, the
 * spans past it are not written (the offsets are complete) and the caller
 * reruns with capacity = the total.  desc.d_spans / d_counts are ignored.
 * d_scratch: gpj_query_dense_scratch_bytes(max_records, num_queries, sum of
 * the capacities) bytes.  Stream-ordered; allocates nothing. */
typedef struct {
  int64_t *d_offsets; /* [max_records + 1] */
  int64_t *d_spans;   /* [capacity][2] */
  int64_t capacity;
} gpj_dense_result;
uint64_t gpj_query_dense_scratch_bytes(int64_t max_records, int32_t num_queries,
                                       int64_t total_capacity);
/* Dense results over the lines of the index (as gpj_query_run_batch). */
gpj_status gpj_query_run_dense(const uint8_t *d_in, uint64_t n, const gpj_index *index,
                               const gpj_query_desc *queries, int32_t num_queries,
                               int64_t max_lines, const gpj_dense_result *results,
                               void *d_scratch, void *stream);
/* Dense results over top-level array elements (as gpj_query_run_records). */
gpj_status gpj_query_run_records_dense(const uint8_t *d_in, uint64_t n, const gpj_index *index,
                                       const int64_t *d_bounds, const int64_t *d_num_bounds,
                                       const gpj_query_desc *queries, int32_t num_queries,
                                       int64_t max_records, const gpj_dense_result *results,
                                       void *d_scratch, void *stream);

/* ---- extension: multi-file batches (no reference counterpart; the
 * reference engine loops over partitions one launch sequence at a time,
 * src/gpjson/engine.cpp:131-166).  Many independent files are indexed in one
 * launch and queried in one launch (and the pair can be captured in a HIP
 * graph with gpj_graph_*).  Layout: the files sit in one device buffer, file
 * f at files[f].offset (a multiple of GPJ_BATCH_ALIGN) with files[f].bytes
 * >= 1 bytes; bytes between files are ignored.  Per file, every index
 * structure and query result equals that of a single-file run over the file
 * alone (offsets are file-relative).
 *   string mask / leveled rows: in the batch's word space (W = total/64
 *     words; row r of the leveled rows starts at word r * W); file f's words
 *     start at files[f].offset / 64.
 *   newline index: file f's entries at [newline_offset, + newline_capacity)
 *     of d_newline; newline_capacity must be >= the file's newline count + 1.
 *   d_line_prefix [num_files + 1]: exclusive prefix of the files' line
 *     counts; record g of the batch is line g - prefix[f] of file f.
 *   query_mask: bit q = batch query q runs on this file, at most
 *     GPJ_BATCH_SLOTS bits; the file's k-th query writes result slot k. */
#define GPJ_BATCH_ALIGN 16384
#define GPJ_BATCH_SLOTS 2
typedef struct {
  uint64_t offset;
  uint64_t bytes;
  uint64_t newline_offset;
  uint64_t newline_capacity;
  uint32_t query_mask;
  uint32_t reserved;
} gpj_batch_file;

/* Scratch for gpj_batch_index: tile descriptors and counters. */
uint64_t gpj_batch_scratch_bytes(uint64_t total_bytes);
/* Host helper: tile_file[t] = the file holding tile t (tiles = total_bytes /
 * GPJ_BATCH_ALIGN rounded up); tiles no file covers get the preceding file. */
gpj_status gpj_batch_tile_map(const gpj_batch_file *files, int32_t num_files,
                              uint64_t total_bytes, uint32_t *tile_file);
/* Line counts of every file, counted on the device (no host pass over the
 * bytes): d_line_prefix[f] receives the exclusive prefix of the files' line
 * counts (1 + newlines inside the file) and d_line_prefix[num_files] the
 * total, from which a caller sizes the newline slices (newline_offset,
 * newline_capacity) before gpj_batch_index.  Those two fields of d_files are
 * not read here.  d_scratch: gpj_batch_scratch_bytes(total_bytes) bytes;
 * 64-byte aligned buffer.  Stream-ordered. */
gpj_status gpj_batch_line_counts(const uint8_t *d_in, uint64_t total_bytes,
                                 const gpj_batch_file *d_files, int32_t num_files,
                                 const uint32_t *d_tile_file, int64_t *d_line_prefix,
                                 void *d_scratch, void *stream);
/* One index launch over every file (d_files, d_tile_file: device arrays).
 * A failed build writes -1 into d_line_prefix[num_files] (the total line
 * count) and gpj_index_status(d_scratch, total_bytes, ...) reports it. */
gpj_status gpj_batch_index(const uint8_t *d_in, uint64_t total_bytes,
                           const gpj_batch_file *d_files, int32_t num_files,
                           const uint32_t *d_tile_file, int32_t levels, uint64_t *d_string,
                           uint64_t *d_leveled, int64_t *d_newline, int64_t *d_line_prefix,
                           void *d_scratch, void *stream);
/* One query launch over every record of every file.  queries: host array
 * (d_spans / d_counts of the descriptors are ignored); results: slot k
 * spans [total records][capacity][2] and counts [total records] (count 0
 * where the file has no k-th query).  Fixed-slot queries run with
 * `capacity` slots (>= their compiled slot count).  max_records sizes the
 * launch (records per pass of the grid); the launch covers every record. */
typedef struct {
  int64_t *d_spans[GPJ_BATCH_SLOTS];
  int64_t *d_counts[GPJ_BATCH_SLOTS];
  int32_t capacity;
} gpj_batch_results;
gpj_status gpj_batch_query(const uint8_t *d_in, uint64_t total_bytes,
                           const gpj_batch_file *d_files, int32_t num_files, int32_t levels,
                           const uint64_t *d_string, const uint64_t *d_leveled,
                           const int64_t *d_newline, const int64_t *d_line_prefix,
                           int64_t max_records, const gpj_query_desc *queries,
                           int32_t num_queries, const gpj_batch_results *results,
                           void *stream);
/* gpj_batch_query with the launch's workgroup shape chosen by the caller (no
 * counterpart in the reference; the batch autotuner measures it): each
 * one-wave workgroup takes records_per_wave (1..64) consecutive records at a
 * time.  gpj_batch_query takes 64.  Results never depend on it. */
gpj_status gpj_batch_query_shaped(const uint8_t *d_in, uint64_t total_bytes,
                                  const gpj_batch_file *d_files, int32_t num_files,
                                  int32_t levels, const uint64_t *d_string,
                                  const uint64_t *d_leveled, const int64_t *d_newline,
                                  const int64_t *d_line_prefix, int64_t max_records,
                                  int32_t records_per_wave, const gpj_query_desc *queries,
                                  int32_t num_queries, const gpj_batch_results *results,
                                  void *stream);
/* The same work as plain per-file launches (one index launch and one query
 * launch per file, like the reference's per-partition loop), writing the
 * same result layout; files: host array; line_prefix: host array of the
 * files' exact line-count prefix.  d_leveled_per_file holds levels * (sum of
 * the files' words) words; scratch: gpj_index_scratch_bytes(largest file). */
gpj_status gpj_batch_per_file(const uint8_t *d_in, const gpj_batch_file *files,
                              int32_t num_files, int32_t levels, uint64_t *d_string,
                              uint64_t *d_leveled_per_file, int64_t *d_newline,
                              int64_t *d_line_counts, const int64_t *line_prefix,
                              const gpj_query_desc *queries, int32_t num_queries,
                              const gpj_batch_results *results, void *d_scratch,
                              void *stream);

/* ---- HIP streams and graph capture of stream-ordered calls (for C callers
 * without HIP headers).  Streams have the default priority.  graph begin/end
 * bracket calls on `stream`; end instantiates. */
gpj_status gpj_stream_create(void **stream);
gpj_status gpj_stream_destroy(void *stream);
gpj_status gpj_graph_begin(void *stream);
gpj_status gpj_graph_end(void *stream, void **graph_exec);
gpj_status gpj_graph_launch(void *graph_exec, void *stream);
gpj_status gpj_graph_destroy(void *graph_exec);

/* ---- query compiler: replaces QueryCompiler::compile
 *      (src/gpjson/query/query_compiler.cpp:467-477).  Host only.
 * Writes the bytecode into ir (capacity ir_cap) and its length, the query's
 * max depth and its number of result slots.  flags: GPJ_COMPILE_EXTENSIONS
 * enables the `[*]` wildcard (which the reference rejects) and element
 * filters on arrays; such bytecode runs through gpj_query_run_ext and its
 * max depth is the number of leveled rows it reads. */
#define GPJ_COMPILE_EXTENSIONS 1u
gpj_status gpj_query_compile(const char *query, uint8_t *ir, uint32_t ir_cap,
                             uint32_t *ir_len, uint32_t *max_depth,
                             uint32_t *num_slots, uint32_t flags);

#ifdef __cplusplus
}
#endif

#endif /* GPJ_CAPI_H */
