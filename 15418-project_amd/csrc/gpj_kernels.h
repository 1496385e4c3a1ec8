// gpj_kernels.h -- kernel declarations shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gpj {

constexpr int kTileWords = 256;              // words (64 B each) per index tile
constexpr uint64_t kTileBytes = 64ull * kTileWords;
// An index workgroup takes NH (1 or kTileHalves) consecutive tiles, with one
// carry wait for all of them (in a multi-file batch they may belong to
// different files).  Single inputs of at least kPairMinTiles tiles take two
// per workgroup: more input bytes wait in each wave slot for their carry;
// smaller ones, whose tiles are all in flight at once anyway, take one.
constexpr int kTileHalves = 2;
constexpr uint64_t kPairMinTiles = 8192; // 128 MiB (measured crossover: 64-128 MiB)

__global__ void index_clear_kernel(uint64_t *desc, uint64_t count, uint32_t *counter);

// One file of a multi-file batch.  Files sit in one buffer, each starting on
// a tile boundary (offset % kTileBytes == 0) with zero bytes up to the next
// file; a file's newline index is the slice [nl_offset, nl_offset +
// nl_capacity) of the batch's newline array.  query_mask: bit q set = batch
// query q runs on this file (at most kBatchSlots bits).
struct BatchFile {
  uint64_t offset;
  uint64_t bytes;
  uint64_t nl_offset;
  uint64_t nl_capacity;
  uint32_t query_mask;
  uint32_t reserved;
};
constexpr int kBatchSlots = 2;

// Multi-file mode of the index kernel: carries restart at every file's first
// tile, newline positions are file-relative and go to the file's slice, and
// each file's last tile writes its line count.
struct BatchSeg {
  const uint32_t *tile_file; // [tiles] file of each tile
  const BatchFile *files;    // [num_files]
  int64_t *line_counts;      // [num_files] lines per file
};

// Every wait of the index kernel on another workgroup is bounded: after
// max_spins unsuccessful polls (kDefaultWaitLimit unless a test lowers it) the
// waiting tile sets the failure flag (tile_counter[1]), goes on with a dummy
// carry and exits; index_status_kernel then turns the flag into a line count
// of -1.  No wait is unbounded, so every launch ends.
constexpr uint32_t kDefaultWaitLimit = 1u << 24;

template <int LMAX, bool SEG, int NH>
__global__ void index_kernel(const uint8_t *in, uint64_t n, uint64_t words,
                             int levels, int aligned, uint64_t *string_out,
                             uint64_t *leveled_out, int64_t *newline_out,
                             int64_t newline_cap, int64_t *num_lines_out,
                             uint64_t *desc, uint32_t *tile_counter,
                             uint64_t num_tiles, BatchSeg seg, uint32_t max_spins,
                             uint64_t init);

// The state the index scan carries from byte to byte (escape parity,
// in-string parity, nesting depth as int8), encoded as the `init` argument of
// index_kernel: the state before the input's first byte.  A whole input
// starts at (0, 0, -1); a shard of a larger document starts at the state its
// predecessors leave (a multi-file batch always starts its files at
// (0, 0, -1)).
__host__ __device__ inline uint64_t index_start_state(uint32_t esc, uint32_t str, int32_t depth) {
  return (uint64_t)(esc & 1u) | ((uint64_t)(str & 1u) << 1) |
         ((uint64_t)((uint32_t)depth & 0xFFu) << 2) | (2ull << 62);
}

// After a single-input index launch: composes its tiles' transfer tables
// (still in the scratch's descriptors) into the input's own: out[0] the
// unpacked table (byte s of bits 0..31: the state after the input from
// start state s = esc | str << 1, as esc | str << 1; byte s of bits
// 32..63: the depth change from s, mod 256), out[1] the newline count.
__global__ void shard_summary_kernel(const uint64_t *desc, uint64_t num_tiles, uint64_t *out);

// After an index launch: when its failure flag is set, *lines_out = -1 (the
// single-input line count, or the total of a batch's line prefix), so every
// stream-ordered reader of the count sees the failure.
__global__ void index_status_kernel(const uint32_t *error_flag, int64_t *lines_out);

__global__ void newline_count_kernel(const uint8_t *in, uint64_t n, int aligned,
                                     uint64_t *tile_counts);
// Exclusive scan of per-tile counts in one workgroup; writes total + add.
__global__ void newline_scan_kernel(uint64_t *tile_counts, uint64_t tiles,
                                    int64_t *num_lines_out, int64_t add);
__global__ void newline_write_kernel(const uint8_t *in, uint64_t n, int aligned,
                                     const uint64_t *tile_offsets,
                                     int64_t *newline_out);
// Line counts of a multi-file batch: newlines per tile inside its file, then
// 1 + the sum over each file's tiles.
__global__ void batch_count_kernel(const uint8_t *in, const BatchFile *files,
                                   const uint32_t *tile_file, uint64_t *tile_counts);
__global__ void batch_file_lines_kernel(const BatchFile *files, int32_t num_files,
                                        const uint64_t *tile_counts, int64_t *lines);
__global__ void escape_quote_kernel(const uint8_t *in, uint64_t n, uint64_t words,
                                    uint64_t *escape_out, uint64_t *quote_out);

constexpr int kMaxBatchQueries = 8;
constexpr uint32_t kBatchIrBytes = 8192;

struct QuerySlot {
  const uint8_t *ir;
  uint32_t ir_len;
  int32_t cap;        // result slots (fixed mode) or inline capacity (variable)
  int32_t variable;   // 0: reference fixed slots, 1: extension variable-length
  int64_t *spans;     // [lines][cap][2]; dense: [out_cap][2] in record order
  int64_t *counts;    // [lines] or null (fixed mode); dense: [lines + 1], the
                      // counts, scanned in place into CSR offsets
  int32_t dense;      // 1: dense results (any mode): spans of record r at
                      // spans[offsets[r] .. offsets[r + 1])
  int64_t *tmp_off;   // dense: [lines] where the record's spans sit in the
                      // staging area, -1 = left to the dense fixup pass
  int64_t out_cap;    // dense: spans the output holds
};
// Staging area of dense results: 8 regions of `region` spans (int64 pairs);
// a wave reserves room in region blockIdx % 8 with the cursor at
// cursors[region * kCursorStride] (each in a 4 KiB block of its own, so the
// 8 cursors' atomics do not queue on one cache line).
constexpr int kDenseStage = 8; // spans per lane in the narrow kernels' LDS stage
constexpr int kDenseRegions = 8;
constexpr int kCursorStride = 512; // uint64 entries (4 KiB)
struct DenseStage {
  int64_t *tmp;
  uint64_t *cursors;
  int64_t region;
};
struct QueryBatch {
  QuerySlot q[kMaxBatchQueries];
  int32_t count;
  DenseStage ds;
};

// Dense results after the query passes: exclusive scan of each dense query's
// counts into CSR offsets ([max_lines + 1] entries; the record count is
// derived from the device line count as the query kernel does: chunk sums,
// their prefix, the chunks' scan), fused with the copy of staged spans into
// CSR order.
constexpr int kScanChunk = 4096; // counts per block of the scan (256 x 16)
struct DenseSet {
  int64_t *counts[kMaxBatchQueries];  // counts -> CSR offsets, in place
  int64_t *tmp_off[kMaxBatchQueries];
  int64_t *spans[kMaxBatchQueries];
  int64_t out_cap[kMaxBatchQueries];
  uint64_t *block_sums;               // [queries][blocks + 1]
  int64_t blocks;                     // scan blocks per query
  const int64_t *num_starts;          // device count of record starts
  int64_t starts_cap;
  int32_t drop_last;
  int64_t max_lines;
  const int64_t *tmp;                 // staging area (int64 pairs)
};
__global__ void dense_reduce_kernel(DenseSet d);
__global__ void dense_sums_kernel(DenseSet d);
__global__ void dense_scan_kernel(DenseSet d);
__global__ void fill_u64_kernel(uint64_t *p, int64_t count, uint64_t v);

// How records are cut: lines of a newline index (level_base 0, drop_last 0)
// or top-level array elements between row-0 bounds (level_base 1,
// drop_last 1); how many consecutive records one wave takes at a time
// (per_wave, 1..64; lanes past it idle); and how many entries the start
// array holds (starts_cap: a device count above it, e.g. a newline count
// beyond the buffer's capacity, is clamped, so no entry past it is read).
struct RecordMode {
  int32_t level_base;
  int32_t drop_last;
  int32_t per_wave;
  int64_t starts_cap;
};

// How the query kernels hold positions.  The narrow kernels (NARROW = true)
// keep 32-bit offsets from each record's 64-byte aligned start; a record
// that reaches `span` bytes from that start, or a backward search that would
// reach further than `span` bytes below it, is left to the wide kernel: the
// narrow kernel stores kWideMark in the first entry of the batch's first
// query (its count in variable mode, the start of its first span in fixed
// mode) and the wide kernel (64-bit file offsets) launched with fixup = 1
// reruns exactly the records so marked.  Neither can happen when the input
// (or every file of a batch) is at most `span` bytes long.
constexpr int64_t kNarrowSpan = (1ll << 31) - (1ll << 20);
constexpr int64_t kWideMark = -2;
struct PosMode {
  int64_t span;
  int32_t fixup;
};

// DENSE: the kernel can run dense queries (QuerySlot.dense); without it the
// dense paths are compiled out.
template <int MAXLV, bool NARROW, bool DENSE>
__global__ void query_batch_kernel(const uint8_t *file, int64_t n, const int64_t *newline,
                                   const int64_t *num_lines, int64_t max_lines,
                                   const uint64_t *string, const uint64_t *leveled,
                                   int64_t level_size, int32_t num_levels, RecordMode mode,
                                   PosMode pm, QueryBatch batch);

// The same queries with one wave per record (gpj_query_wave.hip): outputs as
// query_batch_kernel<MAXLV, false, true> with fixup 0 (64-bit positions, so
// no record is marked for a second pass); inputs 16-byte aligned.
template <int MAXLV>
__global__ void query_wave_kernel(const uint8_t *file, int64_t n, const int64_t *newline,
                                  const int64_t *num_lines, int64_t max_lines,
                                  const uint64_t *string, const uint64_t *leveled,
                                  int64_t level_size, int32_t num_levels, RecordMode mode,
                                  QueryBatch batch);

// Multi-file mode of the query kernel (see BatchFile).
struct FilesMode {
  const BatchFile *files;
  int32_t num_files;
  int32_t cap;                // entries per record in each slot's spans
  const int64_t *line_prefix; // [num_files + 1] exclusive prefix of lines
  int64_t *spans[kBatchSlots];
  int64_t *counts[kBatchSlots];
  int32_t per_wave;           // consecutive records per wave (1..64)
};

// Marks (PosMode) go to the record's count in result slot 0.
template <int MAXLV, bool NARROW>
__global__ void query_files_kernel(const uint8_t *batch_in, const int64_t *newline,
                                   const uint64_t *string, const uint64_t *leveled,
                                   int64_t level_stride, int32_t num_levels, FilesMode fm,
                                   PosMode pm, QueryBatch batch);

// Record bounds of a top-level array: positions of the set bits of one
// bitmap row (count per 256-word tile, scan, write).
__global__ void row_count_kernel(const uint64_t *row, uint64_t words, uint64_t *tile_counts);
__global__ void row_write_kernel(const uint64_t *row, uint64_t words,
                                 const uint64_t *tile_offsets, int64_t *out, int64_t capacity);

} // namespace gpj
