// gpj_query.hip -- JSONPath bytecode interpreter over the structural index.
//
// Semantics follow the reference query kernel
// (src/gpjson/query/kernels/orig/query.cu:144-454): each line between two
// newline-index entries is one record, trimmed to its first '{' and last '}';
// the bytecode walks the leveled rows and the string mask; STORE_RESULT
// writes {start, end} of the current value (spaces trimmed), -1 = no match.
// Two deliberate differences, both outside what the reference defines:
// bytes outside [0, n) read as 0 (the reference reads out of bounds there),
// and an unknown opcode or running off the bytecode ends the record.
// The bytecode comes in through a device pointer and is staged in LDS, so
// concurrent launches with different queries are independent.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gpj_kernels.h"

namespace gpj {

namespace {

enum : uint8_t {
  kEnd = 0, kStoreResult = 1, kMoveUp = 2, kMoveDown = 3, kMoveToKey = 4,
  kMoveToIndex = 5, kMoveToIndexReverse = 6, kMarkPosition = 7,
  kResetPosition = 8, kExpressionStringEquals = 9,
  kMoveToEach = 10, kJumpIfNotArray = 11 // extension opcodes
};
constexpr int kMaxLevels = 16;

// Per-lane interpreter context.  Two small register caches cut the number of
// scattered global loads per record: the last aligned 8 text bytes read, and
// the last two bitmap words read (the key search alternates between a leveled
// row and the string mask).  Both stay valid across the queries of a batch,
// which run over the same record.
struct Ctx {
  const uint8_t *file;
  int64_t n;
  const uint64_t *string;
  const uint64_t *leveled;
  int64_t level_size;
  int num_levels;
  const uint8_t *ir; // LDS copy of the current query
  uint32_t ir_len;
  int64_t text_at;   // aligned 8-byte block held in text_val (-1: none)
  uint64_t text_val;
  const uint64_t *w_ptr[2];
  uint64_t w_val[2];
  uint32_t w_next;   // slot to replace next
};

// Blocks are aligned 8-byte groups of device addresses; a block that is not
// wholly inside [file, file + n) is assembled byte by byte.
__device__ __forceinline__ uint8_t byte_at(Ctx &c, int64_t p) {
  if (p < 0 || p >= c.n) return 0;
  const uintptr_t base = reinterpret_cast<uintptr_t>(c.file);
  const int64_t blk = (int64_t)((base + (uintptr_t)p) >> 3);
  if (blk != c.text_at) {
    const uintptr_t lo = (uintptr_t)blk << 3;
    if (lo >= base && lo + 8 <= base + (uintptr_t)c.n) {
      c.text_val = *reinterpret_cast<const uint64_t *>(lo);
    } else {
      uint64_t v = 0;
      for (uintptr_t a = lo; a < lo + 8; ++a)
        if (a >= base && a < base + (uintptr_t)c.n)
          v |= (uint64_t)(*reinterpret_cast<const uint8_t *>(a)) << (8 * (a & 7));
      c.text_val = v;
    }
    c.text_at = blk;
  }
  return (uint8_t)(c.text_val >> (8 * ((base + (uintptr_t)p) & 7)));
}

__device__ __forceinline__ uint64_t word_at(Ctx &c, const uint64_t *p) {
  if (p == c.w_ptr[0]) return c.w_val[0];
  if (p == c.w_ptr[1]) return c.w_val[1];
  const uint64_t v = *p;
  const uint32_t s = c.w_next;
  c.w_ptr[s] = p;
  c.w_val[s] = v;
  c.w_next = s ^ 1u;
  return v;
}
__device__ __forceinline__ uint8_t ir_at(const Ctx &c, int64_t p) {
  return (p >= 0 && p < (int64_t)c.ir_len) ? c.ir[p] : (uint8_t)0;
}
__device__ __forceinline__ uint64_t through(int b) {
  return b >= 63 ? ~0ull : ((1ull << (b + 1)) - 1);
}

// Last set bit of `bm` in [lo, idx], or -1.
__device__ int64_t last_bit_in(Ctx &c, const uint64_t *bm, int64_t lo, int64_t idx) {
  if (idx < lo || idx < 0) return -1;
  int64_t wi = idx >> 6, wb = wi << 6;
  uint64_t word = word_at(c, bm + wi) & through((int)(idx - wb));
  while (wb + 63 >= lo) {
    uint64_t cand = word;
    const int64_t first = lo - wb;
    if (first > 0) cand &= ~through((int)(first - 1));
    if (cand) return wb + 63 - __builtin_clzll(cand);
    if (wi == 0) break;
    --wi;
    wb -= 64;
    if (wb + 63 < lo) break;
    word = word_at(c, bm + wi);
  }
  return -1;
}

// First structural bit of `level` in [idx, level_end], else level_end.
__device__ int64_t next_structural(Ctx &c, int64_t level_end, int64_t idx,
                                   int level) {
  if (idx >= level_end) return level_end;
  if (level < 0 || level >= c.num_levels) return level_end;
  if (idx < 0) idx = 0;
  const uint64_t *row = c.leveled + (uint64_t)c.level_size * (uint64_t)level;
  int64_t wi = idx >> 6, wb = wi << 6;
  uint64_t word = word_at(c, row + wi) & (~0ull << (idx - wb));
  while (wb <= level_end) {
    uint64_t cand = word;
    const int64_t last = level_end - wb;
    if (last < 63) cand &= through((int)last);
    if (cand) return wb + __builtin_ctzll(cand);
    ++wi;
    wb += 64;
    if (wb > level_end || wi >= c.level_size) break;
    word = word_at(c, row + wi);
  }
  return level_end;
}

__device__ int64_t prev_structural(Ctx &c, int64_t level_start,
                                   int64_t idx, int level) {
  if (idx <= level_start) return level_start;
  if (level < 0 || level >= c.num_levels) return level_start;
  const uint64_t *row = c.leveled + (uint64_t)c.level_size * (uint64_t)level;
  const int64_t s = last_bit_in(c, row, level_start, idx);
  return s == -1 ? level_start : s;
}

__device__ __forceinline__ int32_t read_varint(const Ctx &c, int64_t &qp) {
  uint32_t value = 0, byte = 0;
  int shift = 0;
  for (;;) {
    byte = ir_at(c, qp++);
    if (!(byte & 0x80u)) break;
    value |= (byte & 0x7Fu) << shift;
    shift += 7;
  }
  return (int32_t)(value | (byte << shift));
}

// One element iteration of MOVE_TO_EACH (extension): the array's level, the
// separator before the current element, where the path restarts, and the
// mark-stack depth to restore.
struct Frame {
  int lvl;
  int64_t sep;
  int64_t qp;
  int mi;
};
constexpr int kMaxFrames = 8;

// Interprets the bytecode over one record.  Fixed mode (reference): STORE i
// writes slot i; any failure ends the record.  Variable mode (extension):
// STORE appends (first `cap` stored, all counted) and a finished path
// resumes the innermost MOVE_TO_EACH at its next element.
// line_start/line_end are already trimmed to the record's first '{' and
// last '}' (line_start < line_end).
__device__ int64_t run_record(Ctx &c, int64_t line_start, int64_t line_end,
                              int32_t cap, bool variable, int64_t *res) {
  int64_t idx = line_start;
  int lvl = 0;
  int64_t qp = 0;
  int64_t level_end[kMaxLevels];
  int64_t curr_index[kMaxLevels];
  int64_t marked_pos[kMaxLevels];
  int marked_lvl[kMaxLevels];
  Frame fr[kMaxFrames];
  level_end[0] = line_end;
#pragma unroll
  for (int i = 1; i < kMaxLevels; ++i) level_end[i] = -1;
#pragma unroll
  for (int i = 0; i < kMaxLevels; ++i) curr_index[i] = -1;
  int mi = -1;
  int nres = 0;
  int nfr = 0;
  int64_t count = 0;

  for (;;) {
    if (qp >= (int64_t)c.ir_len) goto stop;
    {
      const uint8_t op = c.ir[qp++];
      switch (op) {
      case kEnd:
        goto stop;
      case kStoreResult: {
        if (!variable && nres >= cap) goto stop;
        int64_t end_str = level_end[lvl];
        while (byte_at(c, idx) == ' ' && idx < level_end[lvl]) ++idx;
        while (end_str > idx && byte_at(c, end_str - 1) == ' ') --end_str;
        const int64_t slot = variable ? count : nres;
        if (slot < cap) {
          res[slot * 2] = idx;
          res[slot * 2 + 1] = end_str;
        }
        if (variable) ++count;
        else ++nres;
        break;
      }
      case kMoveUp:
        if (lvl <= 0) goto stop;
        idx = level_end[lvl];
        level_end[lvl] = -1;
        --lvl;
        break;
      case kMoveDown:
        if (lvl + 1 >= kMaxLevels) goto stop;
        ++lvl;
        if (level_end[lvl] == -1) {
          level_end[lvl] = next_structural(c, level_end[lvl - 1], idx + 1, lvl - 1);
          while (byte_at(c, idx) == ' ') ++idx;
        }
        break;
      case kMoveToKey: {
        const int32_t key_len = read_varint(c, qp);
        const int64_t key_pos = qp;
        qp += key_len;
        if (byte_at(c, idx) != '{') goto stop;
        for (;;) {
          ++idx;
          idx = next_structural(c, level_end[lvl], idx, lvl);
          if (byte_at(c, idx) != ':') break;
          const int64_t string_end = last_bit_in(c, c.string, line_start + 1, idx - 1);
          const int64_t string_start = string_end - key_len;
          bool match = string_start >= line_start && byte_at(c, string_start) == '"';
          for (int32_t i = 0; match && i < key_len; ++i)
            match = ir_at(c, key_pos + i) == byte_at(c, string_start + i + 1);
          if (match) break;
          ++idx;
          idx = next_structural(c, level_end[lvl], idx, lvl);
          if (byte_at(c, idx) == '}') goto stop;
        }
        ++idx;
        break;
      }
      case kMoveToIndex: {
        const int32_t target = read_varint(c, qp);
        const uint8_t ch = byte_at(c, idx);
        if (ch != '[' && ch != ',' && ch != ']') goto stop;
        if (ch == '[') curr_index[lvl] = 0;
        else ++curr_index[lvl];
        for (;;) {
          if (curr_index[lvl] < target) {
            ++idx;
            idx = next_structural(c, level_end[lvl], idx, lvl);
            if (byte_at(c, idx) == ',') { ++curr_index[lvl]; continue; }
            goto stop;
          }
          if (curr_index[lvl] > target) {
            --idx;
            idx = prev_structural(c, 0, idx, lvl);
            --curr_index[lvl];
            continue;
          }
          break;
        }
        if (byte_at(c, idx) == ']') goto stop;
        ++idx;
        break;
      }
      case kMoveToIndexReverse: {
        const int32_t target = read_varint(c, qp);
        idx = level_end[lvl] - 1;
        while (idx >= 0 && byte_at(c, idx) == ' ') --idx;
        const uint8_t ch = byte_at(c, idx);
        if (ch != ']' && ch != ',' && ch != '[') goto stop;
        if (ch == ']') curr_index[lvl] = 0;
        else ++curr_index[lvl];
        for (;;) {
          if (curr_index[lvl] < target) {
            --idx;
            idx = prev_structural(c, 0, idx, lvl);
            if (byte_at(c, idx) == ',' || curr_index[lvl] + 1 == target) {
              ++curr_index[lvl];
              continue;
            }
            goto stop;
          }
          if (curr_index[lvl] > target) {
            ++idx;
            idx = next_structural(c, level_end[lvl], idx, lvl);
            --curr_index[lvl];
            continue;
          }
          break;
        }
        ++idx;
        break;
      }
      case kMarkPosition:
        if (mi + 1 >= kMaxLevels) goto stop;
        ++mi;
        marked_pos[mi] = idx;
        marked_lvl[mi] = lvl;
        break;
      case kResetPosition:
        if (mi < 0) goto stop;
        idx = marked_pos[mi];
        lvl = marked_lvl[mi];
        --mi;
        for (int i = lvl + 1; i < kMaxLevels; ++i) {
          if (level_end[i] == -1) break;
          level_end[i] = -1;
        }
        break;
      case kExpressionStringEquals: {
        const int32_t key_len = read_varint(c, qp);
        const int64_t key_pos = qp;
        qp += key_len;
        while (byte_at(c, idx) == ' ' && idx < level_end[lvl]) ++idx;
        int64_t string_end = level_end[lvl] - 1;
        while (idx < string_end && byte_at(c, string_end) == ' ') --string_end;
        if (string_end - idx + 1 != key_len) goto stop;
        bool eq = true;
        for (int32_t i = 0; eq && i < key_len; ++i)
          eq = ir_at(c, key_pos + i) == byte_at(c, idx + i);
        if (!eq) goto stop;
        break;
      }
      case kMoveToEach: {
        if (!variable || byte_at(c, idx) != '[' || nfr >= kMaxFrames) goto stop;
        const int64_t first = next_structural(c, level_end[lvl], idx + 1, lvl);
        int64_t p = idx + 1;
        while (p < first && byte_at(c, p) == ' ') ++p;
        if (p == first && byte_at(c, first) == ']') goto stop; // empty array
        fr[nfr].lvl = lvl;
        fr[nfr].sep = idx;
        fr[nfr].qp = qp;
        fr[nfr].mi = mi;
        ++nfr;
        ++idx;
        break;
      }
      case kJumpIfNotArray: {
        const int32_t off = read_varint(c, qp);
        if (byte_at(c, idx) != '[') qp += off;
        break;
      }
      default:
        goto stop;
      }
    }
    continue;
  stop:
    for (;;) {
      if (nfr == 0) return count;
      Frame &f = fr[nfr - 1];
      const int64_t sep = next_structural(c, level_end[f.lvl], f.sep + 1, f.lvl);
      if (byte_at(c, sep) == ',' && sep < level_end[f.lvl]) {
        f.sep = sep;
        lvl = f.lvl;
        for (int i = lvl + 1; i < kMaxLevels; ++i) level_end[i] = -1;
        mi = f.mi;
        idx = sep + 1;
        qp = f.qp;
        break;
      }
      --nfr;
    }
  }
}

} // namespace

// All queries of a batch over every line: each lane takes one record and runs
// the batch's bytecodes over it in turn, so the record's bytes and bitmap
// words are fetched once for the whole batch.
__global__ __launch_bounds__(256) void query_batch_kernel(
    const uint8_t *__restrict__ file, int64_t n,
    const int64_t *__restrict__ newline, const int64_t *__restrict__ num_lines,
    int64_t max_lines, const uint64_t *__restrict__ string,
    const uint64_t *__restrict__ leveled, int64_t level_size, int32_t num_levels,
    QueryBatch batch) {
  __shared__ uint8_t s_ir[kBatchIrBytes];
  uint32_t off[kMaxBatchQueries];
  uint32_t total = 0;
  for (int q = 0; q < batch.count; ++q) {
    off[q] = total;
    for (uint32_t i = threadIdx.x; i < batch.q[q].ir_len; i += blockDim.x)
      s_ir[total + i] = batch.q[q].ir[i];
    total += batch.q[q].ir_len;
  }
  __syncthreads();
  const int64_t all = *num_lines;
  const int64_t lines = all < max_lines ? all : max_lines;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t line = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; line < lines;
       line += stride) {
    Ctx c{file, n, string, leveled, level_size, num_levels, nullptr, 0, -1, 0,
          {nullptr, nullptr}, {0, 0}, 0};
    int64_t line_start = newline[line];
    int64_t line_end = (line + 1 < all) ? newline[line + 1] : n - 1;
    bool record = line_start >= 0 && line_start < n && line_end >= 0;
    if (record) {
      if (line_end >= n) line_end = n - 1;
      while (line_end > line_start && byte_at(c, line_end) != '}') --line_end;
      while (line_start < line_end && byte_at(c, line_start) != '{') ++line_start;
      record = line_start != line_end;
    }
    for (int q = 0; q < batch.count; ++q) {
      const QuerySlot &qs = batch.q[q];
      int64_t *res = qs.spans + line * 2 * (int64_t)qs.cap;
      if (!qs.variable)
        for (int i = 0; i < 2 * qs.cap; ++i) res[i] = -1;
      int64_t found = 0;
      if (record) {
        c.ir = s_ir + off[q];
        c.ir_len = qs.ir_len;
        found = run_record(c, line_start, line_end, qs.cap, qs.variable != 0, res);
      }
      if (qs.counts) qs.counts[line] = found;
    }
  }
}

} // namespace gpj
