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
constexpr int kMaxFrames = 8; // element iterations per record (as the oracle)

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
  const uint64_t *w_ptr0, *w_ptr1;
  uint64_t w_val0, w_val1;
  bool w_next1;      // slot to replace next
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
  if (p == c.w_ptr0) return c.w_val0;
  if (p == c.w_ptr1) return c.w_val1;
  const uint64_t v = *p;
  if (c.w_next1) {
    c.w_ptr1 = p;
    c.w_val1 = v;
  } else {
    c.w_ptr0 = p;
    c.w_val0 = v;
  }
  c.w_next1 = !c.w_next1;
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

// Fixed-size per-lane array indexed by a lane-varying value.  Every access is
// an unrolled compare/select over the N elements, so the array lives in
// registers (a plain indexed array would be placed in scratch memory).
template <typename T, int N> struct RegArray {
  T v[N];
  __device__ __forceinline__ T get(int i) const {
    T r = v[0];
#pragma unroll
    for (int k = 1; k < N; ++k) r = (k == i) ? v[k] : r;
    return r;
  }
  __device__ __forceinline__ void set(int i, T x) {
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = (k == i) ? x : v[k];
  }
  __device__ __forceinline__ void fill(T x) {
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = x;
  }
};

// One element iteration of MOVE_TO_EACH (extension): the array's level, the
// separator before the current element, where the path restarts, and the
// mark-stack depth to restore.
// (packed into registers: level, separator position, restart position,
// mark-stack depth)

// Interprets the bytecode over one record.  Fixed mode (reference): STORE i
// writes slot i; any failure ends the record.  Variable mode (extension):
// STORE appends (first `cap` stored, all counted) and a finished path
// resumes the innermost MOVE_TO_EACH at its next element.
// line_start/line_end are already trimmed to the record's first '{' and
// last '}' (line_start < line_end).
// MAXLV bounds the interpreter's level (the host derives it from the
// bytecode); levels, marks and element iterations each use at most MAXLV
// entries.
template <int MAXLV>
__device__ int64_t run_record(Ctx &c, int64_t line_start, int64_t line_end,
                              int32_t cap, bool variable, int64_t *res) {
  int64_t idx = line_start;
  int lvl = 0;
  int64_t qp = 0;
  RegArray<int64_t, MAXLV> level_end;
  RegArray<int32_t, MAXLV> curr_index;
  RegArray<int64_t, MAXLV> marked_pos;
  RegArray<int32_t, MAXLV> marked_lvl;
  constexpr int kF = MAXLV < kMaxFrames ? MAXLV : kMaxFrames;
  RegArray<int32_t, kF> fr_lvl, fr_qp, fr_mi;
  RegArray<int64_t, kF> fr_sep;
  level_end.fill(-1);
  level_end.set(0, line_end);
  curr_index.fill(-1);
  marked_pos.fill(0);
  marked_lvl.fill(0);
  fr_lvl.fill(0);
  fr_qp.fill(0);
  fr_mi.fill(0);
  fr_sep.fill(0);
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
        int64_t end_str = level_end.get(lvl);
        while (byte_at(c, idx) == ' ' && idx < level_end.get(lvl)) ++idx;
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
        idx = level_end.get(lvl);
        level_end.set(lvl, -1);
        --lvl;
        break;
      case kMoveDown:
        if (lvl + 1 >= MAXLV) goto stop;
        ++lvl;
        if (level_end.get(lvl) == -1) {
          level_end.set(lvl, next_structural(c, level_end.get(lvl - 1), idx + 1, lvl - 1));
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
          idx = next_structural(c, level_end.get(lvl), idx, lvl);
          if (byte_at(c, idx) != ':') break;
          const int64_t string_end = last_bit_in(c, c.string, line_start + 1, idx - 1);
          const int64_t string_start = string_end - key_len;
          bool match = string_start >= line_start && byte_at(c, string_start) == '"';
          for (int32_t i = 0; match && i < key_len; ++i)
            match = ir_at(c, key_pos + i) == byte_at(c, string_start + i + 1);
          if (match) break;
          ++idx;
          idx = next_structural(c, level_end.get(lvl), idx, lvl);
          if (byte_at(c, idx) == '}') goto stop;
        }
        ++idx;
        break;
      }
      case kMoveToIndex: {
        const int32_t target = read_varint(c, qp);
        const uint8_t ch = byte_at(c, idx);
        if (ch != '[' && ch != ',' && ch != ']') goto stop;
        const int64_t lend = level_end.get(lvl);
        int32_t ci = (ch == '[') ? 0 : curr_index.get(lvl) + 1;
        for (;;) {
          if (ci < target) {
            ++idx;
            idx = next_structural(c, lend, idx, lvl);
            if (byte_at(c, idx) == ',') { ++ci; continue; }
            curr_index.set(lvl, ci);
            goto stop;
          }
          if (ci > target) {
            --idx;
            idx = prev_structural(c, 0, idx, lvl);
            --ci;
            continue;
          }
          break;
        }
        curr_index.set(lvl, ci);
        if (byte_at(c, idx) == ']') goto stop;
        ++idx;
        break;
      }
      case kMoveToIndexReverse: {
        const int32_t target = read_varint(c, qp);
        const int64_t lend = level_end.get(lvl);
        idx = lend - 1;
        while (idx >= 0 && byte_at(c, idx) == ' ') --idx;
        const uint8_t ch = byte_at(c, idx);
        if (ch != ']' && ch != ',' && ch != '[') goto stop;
        int32_t ci = (ch == ']') ? 0 : curr_index.get(lvl) + 1;
        for (;;) {
          if (ci < target) {
            --idx;
            idx = prev_structural(c, 0, idx, lvl);
            if (byte_at(c, idx) == ',' || ci + 1 == target) {
              ++ci;
              continue;
            }
            curr_index.set(lvl, ci);
            goto stop;
          }
          if (ci > target) {
            ++idx;
            idx = next_structural(c, lend, idx, lvl);
            --ci;
            continue;
          }
          break;
        }
        curr_index.set(lvl, ci);
        ++idx;
        break;
      }
      case kMarkPosition:
        if (mi + 1 >= MAXLV) goto stop;
        ++mi;
        marked_pos.set(mi, idx);
        marked_lvl.set(mi, lvl);
        break;
      case kResetPosition: {
        if (mi < 0) goto stop;
        idx = marked_pos.get(mi);
        lvl = marked_lvl.get(mi);
        --mi;
        // clear the levels above, up to the first one already clear
        bool clearing = true;
#pragma unroll
        for (int i = 1; i < MAXLV; ++i) {
          if (i > lvl) {
            if (level_end.v[i] == -1) clearing = false;
            if (clearing) level_end.v[i] = -1;
          }
        }
        break;
      }
      case kExpressionStringEquals: {
        const int32_t key_len = read_varint(c, qp);
        const int64_t key_pos = qp;
        qp += key_len;
        while (byte_at(c, idx) == ' ' && idx < level_end.get(lvl)) ++idx;
        int64_t string_end = level_end.get(lvl) - 1;
        while (idx < string_end && byte_at(c, string_end) == ' ') --string_end;
        if (string_end - idx + 1 != key_len) goto stop;
        bool eq = true;
        for (int32_t i = 0; eq && i < key_len; ++i)
          eq = ir_at(c, key_pos + i) == byte_at(c, idx + i);
        if (!eq) goto stop;
        break;
      }
      case kMoveToEach: {
        if (!variable || byte_at(c, idx) != '[' || nfr >= kF) goto stop;
        const int64_t first = next_structural(c, level_end.get(lvl), idx + 1, lvl);
        int64_t p = idx + 1;
        while (p < first && byte_at(c, p) == ' ') ++p;
        if (p == first && byte_at(c, first) == ']') goto stop; // empty array
        fr_lvl.set(nfr, lvl);
        fr_sep.set(nfr, idx);
        fr_qp.set(nfr, (int32_t)qp);
        fr_mi.set(nfr, mi);
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
      const int f = nfr - 1;
      const int flvl = fr_lvl.get(f);
      const int64_t lend = level_end.get(flvl);
      const int64_t sep = next_structural(c, lend, fr_sep.get(f) + 1, flvl);
      if (byte_at(c, sep) == ',' && sep < lend) {
        fr_sep.set(f, sep);
        lvl = flvl;
#pragma unroll
        for (int i = 1; i < MAXLV; ++i)
          if (i > lvl) level_end.set(i, -1);
        mi = fr_mi.get(f);
        idx = sep + 1;
        qp = fr_qp.get(f);
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
template <int MAXLV>
__global__ __launch_bounds__(256) void query_batch_kernel(
    const uint8_t *__restrict__ file, int64_t n,
    const int64_t *__restrict__ newline, const int64_t *__restrict__ num_lines,
    int64_t max_lines, const uint64_t *__restrict__ string,
    const uint64_t *__restrict__ leveled, int64_t level_size, int32_t num_levels,
    QueryBatch batch) {
  __shared__ uint8_t s_ir[kBatchIrBytes];
  {
    uint32_t total = 0;
    for (int q = 0; q < batch.count; ++q) {
      for (uint32_t i = threadIdx.x; i < batch.q[q].ir_len; i += blockDim.x)
        s_ir[total + i] = batch.q[q].ir[i];
      total += batch.q[q].ir_len;
    }
  }
  __syncthreads();
  const int64_t all = *num_lines;
  const int64_t lines = all < max_lines ? all : max_lines;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t line = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; line < lines;
       line += stride) {
    Ctx c{file, n, string, leveled, level_size, num_levels, nullptr, 0, -1, 0,
          nullptr, nullptr, 0, 0, false};
    int64_t line_start = newline[line];
    int64_t line_end = (line + 1 < all) ? newline[line + 1] : n - 1;
    bool record = line_start >= 0 && line_start < n && line_end >= 0;
    if (record) {
      if (line_end >= n) line_end = n - 1;
      while (line_end > line_start && byte_at(c, line_end) != '}') --line_end;
      while (line_start < line_end && byte_at(c, line_start) != '{') ++line_start;
      record = line_start != line_end;
    }
    uint32_t ir_off = 0;
    for (int q = 0; q < batch.count; ++q) {
      const QuerySlot &qs = batch.q[q];
      int64_t *res = qs.spans + line * 2 * (int64_t)qs.cap;
      if (!qs.variable)
        for (int i = 0; i < 2 * qs.cap; ++i) res[i] = -1;
      int64_t found = 0;
      if (record) {
        c.ir = s_ir + ir_off;
        c.ir_len = qs.ir_len;
        found = run_record<MAXLV>(c, line_start, line_end, qs.cap, qs.variable != 0, res);
      }
      if (qs.counts) qs.counts[line] = found;
      ir_off += qs.ir_len;
    }
  }
}

template __global__ void query_batch_kernel<4>(const uint8_t *, int64_t, const int64_t *, const int64_t *, int64_t, const uint64_t *, const uint64_t *, int64_t, int32_t, QueryBatch);
template __global__ void query_batch_kernel<8>(const uint8_t *, int64_t, const int64_t *, const int64_t *, int64_t, const uint64_t *, const uint64_t *, int64_t, int32_t, QueryBatch);
template __global__ void query_batch_kernel<16>(const uint8_t *, int64_t, const int64_t *, const int64_t *, int64_t, const uint64_t *, const uint64_t *, int64_t, int32_t, QueryBatch);

} // namespace gpj
