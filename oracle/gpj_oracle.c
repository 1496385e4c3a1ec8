/*
 * gpj_oracle.c -- CPU restatement of the reference GpJSON hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls
 * this file: only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py use it, and only as the checker (or as the timed CPU baseline).
 *
 * Every function restates one piece of the reference algorithm and cites the
 * reference file:line it follows.  Paths are relative to the reference root.
 *
 * Pinning (what checks each function against something other than itself):
 *   gpjo_escape_index          -- tests/golden KAT vectors built from the
 *                                 reference's is_escaped_quote definition
 *                                 (tests/index/string_index_test.cpp:28-34)
 *                                 and brute force over all short strings.
 *   gpjo_escape_index_chunked  -- equals gpjo_escape_index whenever no
 *                                 backslash run covers a whole chunk; the
 *                                 quirk case is checked by a hand-derived KAT.
 *   gpjo_quote_index / gpjo_string_index
 *                              -- reference KAT "\"a\\\"b\"  \"x\"" ->
 *                                 0b01100011111 (string_index_test.cpp:86-93)
 *                                 and the closed form of string_index_test.cpp:36-59.
 *   gpjo_newline_index         -- KAT first\nsecond\nthird -> {0,5,12}
 *                                 (newline_index_test.cpp:64-69).
 *   gpjo_leveled_index         -- KAT {"a":[1,2]} L=3 (leveled_bitmap_index_test.cpp:123-131)
 *                                 and an independent stack-based recount.
 *   gpjo_query                 -- reference KATs (query_executor_test.cpp:155-245,
 *                                 engine_test.cpp:41-129) and an independent
 *                                 DOM evaluator on tiny compact records.
 *   gpjo_partition             -- file_reader_test.cpp:124-133 KATs.
 */
#include <pthread.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define GPJO_WORDS(n) (((n) + 63u) / 64u)

/* ------------------------------------------------------------------------ */
/* Structural index                                                          */
/* ------------------------------------------------------------------------ */

/* Escape bitmap, carry propagated from byte 0 (true JSON semantics).
 * Per-byte rule of src/gpjson/index/kernels/orig/uncombined/escape-index.cu:25-36:
 * the bit of byte i is the carry *before* byte i; a backslash flips the carry,
 * any other byte clears it. */
void gpjo_escape_index(const uint8_t *file, uint64_t n, uint64_t *out) {
  memset(out, 0, GPJO_WORDS(n) * 8u);
  int carry = 0;
  for (uint64_t i = 0; i < n; ++i) {
    if (carry) out[i / 64] |= 1ull << (i % 64);
    carry = (file[i] == '\\') ? (carry ^ 1) : 0;
  }
}

/* Escape bitmap exactly as the reference computes it for a launch of
 * `threads` threads: chunk = ceil(ceil(n/threads)/64)*64 bytes per thread
 * (escape-index.cu:11-16); the carry into chunk t is the trailing-run parity
 * of chunk t-1 counted from carry 0 (escape-carry-index.cu:17-27). */
void gpjo_escape_index_chunked(const uint8_t *file, uint64_t n,
                               uint64_t threads, uint64_t *out) {
  memset(out, 0, GPJO_WORDS(n) * 8u);
  if (threads == 0) return;
  uint64_t per = (n + threads - 1) / threads;
  uint64_t chunk = ((per + 63) / 64) * 64;
  if (chunk == 0) return;
  for (uint64_t t = 0; t < threads; ++t) {
    uint64_t start = t * chunk;
    if (start >= n) break;
    int carry = 0;
    if (t > 0) {
      for (uint64_t i = start - chunk; i < start && i < n; ++i)
        carry = (file[i] == '\\') ? (carry ^ 1) : 0;
    }
    for (uint64_t i = start; i < start + chunk && i < n; ++i) {
      if (carry) out[i / 64] |= 1ull << (i % 64);
      carry = (file[i] == '\\') ? (carry ^ 1) : 0;
    }
  }
}

/* Unescaped-quote bitmap (src/gpjson/index/kernels/orig/quote-index.cu:22-34). */
void gpjo_quote_index(const uint8_t *file, uint64_t n, const uint64_t *escape,
                      uint64_t *out) {
  memset(out, 0, GPJO_WORDS(n) * 8u);
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t bit = 1ull << (i % 64);
    if (file[i] == '"' && (escape[i / 64] & bit) == 0) out[i / 64] |= bit;
  }
}

/* String mask: inclusive prefix-XOR of the quote bitmap across the whole
 * partition (string-index.cu:12-29 plus the xor scan of
 * orig_index_builder.cu:129-162): opening quote and body set, closing quote
 * clear. */
void gpjo_string_index(const uint64_t *quote, uint64_t words, uint64_t *out) {
  uint64_t carry = 0;
  for (uint64_t w = 0; w < words; ++w) {
    uint64_t s = 0, in = carry;
    uint64_t q = quote[w];
    for (int b = 0; b < 64; ++b) {
      in ^= (q >> b) & 1u;
      s |= in << b;
    }
    out[w] = s;
    carry = in;
  }
}

/* Number of entries of the newline index: synthetic 0 plus one per '\n'
 * (orig_index_builder.cu:250-257). */
uint64_t gpjo_newline_count(const uint8_t *file, uint64_t n) {
  uint64_t c = 1;
  for (uint64_t i = 0; i < n; ++i) c += (file[i] == '\n');
  return c;
}

/* Newline index: out[0] = 0, then every '\n' offset in order
 * (newline-index.cu:16-20, orig_index_builder.cu:254-257). */
uint64_t gpjo_newline_index(const uint8_t *file, uint64_t n, int64_t *out) {
  uint64_t k = 0;
  out[k++] = 0;
  for (uint64_t i = 0; i < n; ++i)
    if (file[i] == '\n') out[k++] = (int64_t)i;
  return k;
}

/* Leveled bitmaps, level-major [levels][words]
 * (leveled-bitmaps-index.cu:28-71 with the depth carry of
 * leveled-bitmaps-carry-index.cu:21-39 and char-sum start -1 at
 * orig_index_builder.cu:376).  Depth is the reference's `signed char`, so it
 * wraps modulo 256; only bytes outside strings count. */
void gpjo_leveled_index(const uint8_t *file, uint64_t n, const uint64_t *string,
                        int levels, uint64_t *out) {
  uint64_t words = GPJO_WORDS(n);
  memset(out, 0, words * (uint64_t)levels * 8u);
  int8_t level = -1;
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t bit = 1ull << (i % 64);
    if (string[i / 64] & bit) continue;
    uint8_t c = file[i];
    if (c == '{' || c == '[') {
      level = (int8_t)(uint8_t)(level + 1);
      if (level >= 0 && level < levels) out[(uint64_t)level * words + i / 64] |= bit;
    } else if (c == '}' || c == ']') {
      if (level >= 0 && level < levels) out[(uint64_t)level * words + i / 64] |= bit;
      level = (int8_t)(uint8_t)(level - 1);
    } else if (c == ':' || c == ',') {
      if (level >= 0 && level < levels) out[(uint64_t)level * words + i / 64] |= bit;
    }
  }
}

/* All three outputs in one call (what IndexBuilder::build returns).
 * newline_out must hold gpjo_newline_count(file, n) entries. */
void gpjo_build_index(const uint8_t *file, uint64_t n, int levels,
                      uint64_t *string_out, uint64_t *leveled_out,
                      int64_t *newline_out) {
  uint64_t words = GPJO_WORDS(n);
  uint64_t *tmp = (uint64_t *)malloc((words ? words : 1) * 8u);
  gpjo_escape_index(file, n, tmp);
  gpjo_quote_index(file, n, tmp, string_out);
  memcpy(tmp, string_out, words * 8u);
  gpjo_string_index(tmp, words, string_out);
  free(tmp);
  gpjo_leveled_index(file, n, string_out, levels, leveled_out);
  gpjo_newline_index(file, n, newline_out);
}

/* ------------------------------------------------------------------------ */
/* Query interpreter                                                         */
/* Restates src/gpjson/query/kernels/orig/query.cu:34-453 (one record per   */
/* line, result[line][slot] = {start, end} or -1).  Differences, all on     */
/* inputs the reference reads out of bounds or on IR the compiler never     */
/* emits: bytes outside [0, n) read as 0, an unknown opcode or running off  */
/* the IR ends the path, and so does a negative key length or jump offset   */
/* (it would move the IR position backwards and never end).                 */
/* ------------------------------------------------------------------------ */

enum {
  OP_END = 0, OP_STORE_RESULT = 1, OP_MOVE_UP = 2, OP_MOVE_DOWN = 3,
  OP_MOVE_TO_KEY = 4, OP_MOVE_TO_INDEX = 5, OP_MOVE_TO_INDEX_REVERSE = 6,
  OP_MARK_POSITION = 7, OP_RESET_POSITION = 8, OP_EXPRESSION_STRING_EQUALS = 9
};
#define QMAX_LEVELS 16

typedef struct {
  const uint8_t *file;
  int64_t n;
  const uint64_t *string;
  const uint64_t *leveled;
  int64_t level_size;
  int num_levels;
  int level_base; /* leveled row of interpreter level 0 (0 for lines) */
} qctx;

static inline uint8_t qbyte(const qctx *c, int64_t p) {
  return (p >= 0 && p < c->n) ? c->file[p] : 0;
}

static inline uint64_t mask_through(int b) {
  return b >= 63 ? ~0ull : ((1ull << (b + 1)) - 1);
}

/* find_previous_bitmap_bit, query.cu:39-77 */
static int64_t prev_bit(const uint64_t *bm, int64_t lo, int64_t idx) {
  if (idx < lo || idx < 0) return -1;
  int64_t wi = idx / 64, wb = wi * 64;
  uint64_t word = bm[wi] & mask_through((int)(idx - wb));
  while (wb + 63 >= lo) {
    uint64_t cand = word;
    int64_t first = lo - wb;
    if (first > 0) cand &= ~mask_through((int)(first - 1));
    if (cand) return wb + 63 - __builtin_clzll(cand);
    if (wi == 0) break;
    --wi;
    wb -= 64;
    if (wb + 63 < lo) break;
    word = bm[wi];
  }
  return -1;
}

/* find_next_structural_char, query.cu:79-115 */
static int64_t next_struct(const qctx *c, int64_t level_end, int64_t idx, int level) {
  if (idx >= level_end) return level_end;
  if (level < 0 || level + c->level_base >= c->num_levels) return level_end; /* rows not built */
  if (idx < 0) idx = 0;
  const uint64_t *row =
      c->leveled + (uint64_t)c->level_size * (uint64_t)(level + c->level_base);
  int64_t wi = idx / 64, wb = wi * 64;
  uint64_t word = row[wi] & (~0ull << (idx - wb));
  while (wb <= level_end) {
    uint64_t cand = word;
    int64_t last = level_end - wb;
    if (last < 63) cand &= mask_through((int)last);
    if (cand) return wb + __builtin_ctzll(cand);
    ++wi;
    wb += 64;
    if (wb > level_end || wi >= c->level_size) break;
    word = row[wi];
  }
  return level_end;
}

/* find_previous_structural_char, query.cu:117-128 */
static int64_t prev_struct(const qctx *c, int64_t level_start, int64_t idx, int level) {
  if (idx <= level_start) return level_start;
  if (level < 0 || level + c->level_base >= c->num_levels) return level_start; /* rows not built */
  const uint64_t *row =
      c->leveled + (uint64_t)c->level_size * (uint64_t)(level + c->level_base);
  int64_t s = prev_bit(row, level_start, idx);
  return s == -1 ? level_start : s;
}

/* read_varint, query.cu:130-140 */
static int32_t read_varint(const uint8_t *ir, int64_t ir_len, int64_t *pos) {
  uint32_t value = 0;
  int shift = 0;
  uint32_t byte = 0;
  for (;;) {
    byte = (*pos < ir_len) ? ir[*pos] : 0;
    (*pos)++;
    if (!(byte & 0x80)) break;
    value |= (byte & 0x7Fu) << shift;
    shift += 7;
  }
  return (int32_t)(value | (byte << shift));
}

/* Extension opcodes (not in the reference; emitted only by the compiler's
 * extension mode):
 *   MOVE_TO_EACH (10): the current value must be a non-empty array; run the
 *     rest of the bytecode once per element, in order.  Whenever a path ends
 *     (END or any failure) the innermost iteration resumes at its next
 *     element; the record ends when no iteration is left.
 *   JUMP_IF_NOT_ARRAY (11) <varint off>: skip `off` bytecode bytes unless the
 *     current value starts with '['.
 * In variable mode STORE_RESULT appends to the record's result list (up to
 * `cap` spans stored, all counted). */
enum { OP_MOVE_TO_EACH = 10, OP_JUMP_IF_NOT_ARRAY = 11 };
#define QMAX_FRAMES 8

typedef struct {
  int lvl;
  int64_t sep;  /* '[' or ',' before the current element */
  int64_t qp;   /* bytecode position after MOVE_TO_EACH */
  int mi;
} qframe;

static void query_line(const qctx *c, const int64_t *nl, int64_t nl_size,
                       int64_t line, const uint8_t *ir, int64_t ir_len,
                       int num_results, int variable, int64_t *result,
                       int64_t *counts) {
  int64_t *res = result + line * 2 * (int64_t)num_results;
  for (int i = 0; i < 2 * num_results; ++i) res[i] = -1;
  int64_t count = 0;
  if (counts) counts[line] = 0;

  int64_t line_start = nl[line];
  int64_t line_end = (line + 1 < nl_size) ? nl[line + 1] : c->n - 1;
  if (line_start < 0 || line_start >= c->n || line_end < 0) return;
  if (line_end >= c->n) line_end = c->n - 1;
  while (line_end > line_start && qbyte(c, line_end) != '}') --line_end;
  while (line_start < line_end && qbyte(c, line_start) != '{') ++line_start;
  if (line_start == line_end) return;

  int64_t idx = line_start;
  int lvl = 0;
  int64_t qp = 0;
  int64_t level_end[QMAX_LEVELS];
  level_end[0] = line_end;
  for (int i = 1; i < QMAX_LEVELS; ++i) level_end[i] = -1;
  int64_t marked_pos[QMAX_LEVELS];
  int marked_lvl[QMAX_LEVELS];
  int mi = -1;
  int nres = 0;
  int64_t curr_index[QMAX_LEVELS];
  for (int i = 0; i < QMAX_LEVELS; ++i) curr_index[i] = -1;
  qframe fr[QMAX_FRAMES];
  int nfr = 0;

  for (;;) {
    if (qp >= ir_len) goto stop;
    {
    uint8_t op = ir[qp++];
    switch (op) {
    case OP_END:
      goto stop;
    case OP_STORE_RESULT: {
      if (!variable && nres >= num_results) goto stop;
      int64_t end_str = level_end[lvl];
      while (qbyte(c, idx) == ' ' && idx < level_end[lvl]) ++idx;
      while (end_str > idx && qbyte(c, end_str - 1) == ' ') --end_str;
      if (variable) {
        if (count < num_results) {
          res[count * 2] = idx;
          res[count * 2 + 1] = end_str;
        }
        ++count;
      } else {
        res[nres * 2] = idx;
        res[nres * 2 + 1] = end_str;
        ++nres;
      }
      break;
    }
    case OP_MOVE_UP:
      if (lvl <= 0) goto stop;
      idx = level_end[lvl];
      level_end[lvl] = -1;
      --lvl;
      break;
    case OP_MOVE_DOWN:
      if (lvl + 1 >= QMAX_LEVELS) goto stop;
      ++lvl;
      if (level_end[lvl] == -1) {
        level_end[lvl] = next_struct(c, level_end[lvl - 1], idx + 1, lvl - 1);
        while (qbyte(c, idx) == ' ') ++idx;
      }
      break;
    case OP_MOVE_TO_KEY: {
      int32_t key_len = read_varint(ir, ir_len, &qp);
      int64_t key_pos = qp;
      qp += key_len;
      if (key_len < 0) goto stop; /* malformed: would move qp backwards */
      if (qbyte(c, idx) != '{') goto stop;
      for (;;) { /* search_key */
        ++idx;
        idx = next_struct(c, level_end[lvl], idx, lvl);
        if (qbyte(c, idx) != ':') break;
        const int64_t cand = idx;
        int64_t string_end = prev_bit(c->string, line_start + 1, idx - 1);
        int64_t string_start = string_end - key_len;
        int match = (string_start >= line_start && qbyte(c, string_start) == '"');
        for (int32_t i = 0; match && i < key_len; ++i)
          if ((key_pos + i >= ir_len ? 0 : ir[key_pos + i]) != qbyte(c, string_start + i + 1))
            match = 0;
        if (match) break;
        ++idx;
        idx = next_struct(c, level_end[lvl], idx, lvl);
        if (qbyte(c, idx) == '}') goto stop;
        /* no progress: the candidate was the level end itself, a ':' that
         * only malformed input puts there; the reference would search it
         * again forever (query.cu:254-302); the path fails instead */
        if (idx == cand) goto stop;
      }
      ++idx;
      break;
    }
    case OP_MOVE_TO_INDEX: {
      int32_t target = read_varint(ir, ir_len, &qp);
      uint8_t ch = qbyte(c, idx);
      if (ch != '[' && ch != ',' && ch != ']') goto stop;
      if (ch == '[') curr_index[lvl] = 0;
      else ++curr_index[lvl];
      for (;;) { /* search_index */
        if (curr_index[lvl] < target) {
          ++idx;
          idx = next_struct(c, level_end[lvl], idx, lvl);
          if (qbyte(c, idx) == ',') { ++curr_index[lvl]; continue; }
          goto stop;
        }
        if (curr_index[lvl] > target) {
          --idx;
          idx = prev_struct(c, 0, idx, lvl);
          --curr_index[lvl];
          continue;
        }
        break;
      }
      if (qbyte(c, idx) == ']') goto stop;
      ++idx;
      break;
    }
    case OP_MOVE_TO_INDEX_REVERSE: {
      int32_t target = read_varint(ir, ir_len, &qp);
      idx = level_end[lvl] - 1;
      while (idx >= 0 && qbyte(c, idx) == ' ') --idx;
      uint8_t ch = qbyte(c, idx);
      if (ch != ']' && ch != ',' && ch != '[') goto stop;
      if (ch == ']') curr_index[lvl] = 0;
      else ++curr_index[lvl];
      for (;;) { /* search_index_reverse */
        if (curr_index[lvl] < target) {
          --idx;
          idx = prev_struct(c, 0, idx, lvl);
          if (qbyte(c, idx) == ',' || curr_index[lvl] + 1 == target) {
            ++curr_index[lvl];
            continue;
          }
          goto stop;
        }
        if (curr_index[lvl] > target) {
          ++idx;
          idx = next_struct(c, level_end[lvl], idx, lvl);
          --curr_index[lvl];
          continue;
        }
        break;
      }
      ++idx;
      break;
    }
    case OP_MARK_POSITION:
      if (mi + 1 >= QMAX_LEVELS) goto stop;
      ++mi;
      marked_pos[mi] = idx;
      marked_lvl[mi] = lvl;
      break;
    case OP_RESET_POSITION:
      if (mi < 0) goto stop;
      idx = marked_pos[mi];
      lvl = marked_lvl[mi];
      --mi;
      for (int i = lvl + 1; i < QMAX_LEVELS; ++i) {
        if (level_end[i] == -1) break;
        level_end[i] = -1;
      }
      break;
    case OP_EXPRESSION_STRING_EQUALS: {
      int32_t key_len = read_varint(ir, ir_len, &qp);
      int64_t key_pos = qp;
      qp += key_len;
      if (key_len < 0) goto stop; /* malformed: would move qp backwards */
      while (qbyte(c, idx) == ' ' && idx < level_end[lvl]) ++idx;
      int64_t string_end = level_end[lvl] - 1;
      while (idx < string_end && qbyte(c, string_end) == ' ') --string_end;
      if (string_end - idx + 1 != key_len) goto stop;
      for (int32_t i = 0; i < key_len; ++i)
        if ((key_pos + i >= ir_len ? 0 : ir[key_pos + i]) != qbyte(c, idx + i)) goto stop;
      break;
    }
    case OP_MOVE_TO_EACH: {
      if (!variable || qbyte(c, idx) != '[' || nfr >= QMAX_FRAMES) goto stop;
      int64_t first = next_struct(c, level_end[lvl], idx + 1, lvl);
      int64_t p = idx + 1;
      while (p < first && qbyte(c, p) == ' ') ++p;
      if (p == first && qbyte(c, first) == ']') goto stop; /* empty array */
      fr[nfr].lvl = lvl;
      fr[nfr].sep = idx;
      fr[nfr].qp = qp;
      fr[nfr].mi = mi;
      ++nfr;
      ++idx;
      break;
    }
    case OP_JUMP_IF_NOT_ARRAY: {
      int32_t off = read_varint(ir, ir_len, &qp);
      if (off < 0) goto stop; /* malformed: would move qp backwards */
      if (qbyte(c, idx) != '[') qp += off;
      break;
    }
    default:
      goto stop;
    }
    }
    continue;
  stop:
    /* path finished: resume the innermost element iteration, if any */
    for (;;) {
      if (nfr == 0) {
        if (counts) counts[line] = count;
        return;
      }
      qframe *f = &fr[nfr - 1];
      int64_t sep = next_struct(c, level_end[f->lvl], f->sep + 1, f->lvl);
      if (qbyte(c, sep) == ',' && sep < level_end[f->lvl]) {
        f->sep = sep;
        lvl = f->lvl;
        for (int i = lvl + 1; i < QMAX_LEVELS; ++i) level_end[i] = -1;
        mi = f->mi;
        idx = sep + 1;
        qp = f->qp;
        break;
      }
      --nfr;
    }
  }
}

/* One compiled query over one partition's indices (query.cu:144-454).
 * result holds num_lines * num_results * 2 int64 values. */
void gpjo_query(const uint8_t *file, int64_t n, const int64_t *newline,
                int64_t num_lines, const uint64_t *string,
                const uint64_t *leveled, int64_t level_size, int num_levels,
                const uint8_t *ir, int64_t ir_len, int num_results,
                int64_t *result) {
  qctx c = {file, n, string, leveled, level_size, num_levels, 0};
  for (int64_t line = 0; line < num_lines; ++line)
    query_line(&c, newline, num_lines, line, ir, ir_len, num_results, 0, result, NULL);
}

/* Extension mode: variable-length results.  result holds num_lines * cap * 2
 * int64 spans (first `cap` results of each line, in execution order), counts
 * the number of results of each line (which may exceed cap). */
void gpjo_query_ext(const uint8_t *file, int64_t n, const int64_t *newline,
                    int64_t num_lines, const uint64_t *string,
                    const uint64_t *leveled, int64_t level_size, int num_levels,
                    const uint8_t *ir, int64_t ir_len, int cap, int64_t *result,
                    int64_t *counts) {
  qctx c = {file, n, string, leveled, level_size, num_levels, 0};
  for (int64_t line = 0; line < num_lines; ++line)
    query_line(&c, newline, num_lines, line, ir, ir_len, cap, 1, result, counts);
}

/* Extension: top-level array elements as records.  `bounds` holds the
 * positions of every structural character of leveled row 0 (for a document
 * whose top-level value is an array: its '[', each ',' between elements and
 * its ']'); record r spans bounds[r] .. bounds[r + 1], so there are
 * num_bounds - 1 records.  Each record is trimmed and interpreted exactly as a
 * line, with interpreter level k reading leveled row k + 1.  Equivalently:
 * the result is that of line mode over the document with its top-level
 * ',' replaced by '\n' and its '[' ']' by ' ' (same byte offsets). */
int64_t gpjo_row_bounds(const uint64_t *row, int64_t words, int64_t *bounds, int64_t cap) {
  int64_t k = 0;
  for (int64_t w = 0; w < words; ++w) {
    uint64_t m = row[w];
    while (m) {
      const int b = __builtin_ctzll(m);
      m &= m - 1;
      if (k < cap) bounds[k] = w * 64 + b;
      ++k;
    }
  }
  return k;
}

void gpjo_query_records(const uint8_t *file, int64_t n, const int64_t *bounds,
                        int64_t num_bounds, const uint64_t *string,
                        const uint64_t *leveled, int64_t level_size, int num_levels,
                        const uint8_t *ir, int64_t ir_len, int cap, int variable,
                        int64_t *result, int64_t *counts) {
  qctx c = {file, n, string, leveled, level_size, num_levels, 1};
  for (int64_t r = 0; r + 1 < num_bounds; ++r)
    query_line(&c, bounds, num_bounds, r, ir, ir_len, cap, variable, result,
               variable ? counts : NULL);
}

/* ------------------------------------------------------------------------ */
/* Partitioning (src/gpjson/file/file_reader.cpp:18-77)                      */
/* Writes up to max_parts (start, end) pairs; returns the partition count,  */
/* or -1 when no newline precedes a cut (PartitionError).                    */
/* ------------------------------------------------------------------------ */
int64_t gpjo_partition(const uint8_t *file, uint64_t n, uint64_t part_size,
                       uint64_t *bounds, int64_t max_parts) {
  int64_t k = 0;
  if (part_size == 0 || n == 0 || n <= part_size) {
    if (max_parts > 0) { bounds[0] = 0; bounds[1] = n; }
    return 1;
  }
  uint64_t cur = 0;
  while (n - cur > part_size) {
    uint64_t off = cur + part_size, next = 0;
    int found = 0;
    for (;;) {
      if (file[off] == '\n') { next = off + 1; found = 1; break; }
      if (off == cur) break;
      --off;
    }
    if (!found) return -1;
    if (next >= n) break;
    if (k < max_parts) { bounds[2 * k] = cur; bounds[2 * k + 1] = next - 1; }
    ++k;
    cur = next;
  }
  if (k < max_parts) { bounds[2 * k] = cur; bounds[2 * k + 1] = n; }
  return k + 1;
}

/* ------------------------------------------------------------------------ */
/* Multi-threaded CPU baseline: the reference's chunked algorithm           */
/* (per-chunk escape carry, XOR scan of quote parity, sum scan of depth and */
/* of newline counts; orig_index_builder.cu:224-395) on std host threads.    */
/* Chunks are 64-byte aligned.  Output identical to gpjo_build_index.        */
/* ------------------------------------------------------------------------ */

typedef struct {
  const uint8_t *file;
  uint64_t n, start, end;
  int levels;
  uint64_t *string_out, *leveled_out;
  int64_t *newline_out;
  /* phase results */
  int all_bs, esc_out0, esc_in;
  int quote_parity, str_in;
  int depth_delta;
  int8_t depth_in;
  uint64_t nl_count, nl_base;
  int phase;
} mt_chunk;

static void *mt_worker(void *arg) {
  mt_chunk *ch = (mt_chunk *)arg;
  const uint8_t *f = ch->file;
  if (ch->phase == 0) {
    int carry = 0, all = 1;
    uint64_t nlc = 0;
    for (uint64_t i = ch->start; i < ch->end; ++i) {
      if (f[i] == '\\') carry ^= 1;
      else { carry = 0; all = 0; }
      nlc += (f[i] == '\n');
    }
    ch->all_bs = all;
    ch->esc_out0 = carry;
    ch->nl_count = nlc;
  } else if (ch->phase == 1) {
    /* quote bitmap + local prefix xor assuming str_in = 0 */
    int carry = ch->esc_in, in = 0;
    for (uint64_t w = ch->start / 64; w * 64 < ch->end; ++w) {
      uint64_t s = 0;
      uint64_t lim = (w + 1) * 64 < ch->end ? (w + 1) * 64 : ch->end;
      for (uint64_t i = w * 64; i < lim; ++i) {
        int esc = carry;
        carry = (f[i] == '\\') ? (carry ^ 1) : 0;
        if (f[i] == '"' && !esc) in ^= 1;
        s |= (uint64_t)in << (i % 64);
      }
      /* bits past the end of the input carry the final state, as the
         reference's in-word prefix XOR over zero quote bits does */
      if ((lim % 64) && in) s |= ~0ull << (lim % 64);
      ch->string_out[w] = s;
    }
    ch->quote_parity = in;
  } else if (ch->phase == 2) {
    /* apply string carry, depth delta */
    int delta = 0;
    for (uint64_t w = ch->start / 64; w * 64 < ch->end; ++w) {
      if (ch->str_in) ch->string_out[w] = ~ch->string_out[w];
      uint64_t lim = (w + 1) * 64 < ch->end ? (w + 1) * 64 : ch->end;
      uint64_t s = ch->string_out[w];
      for (uint64_t i = w * 64; i < lim; ++i) {
        if (s & (1ull << (i % 64))) continue;
        if (f[i] == '{' || f[i] == '[') ++delta;
        else if (f[i] == '}' || f[i] == ']') --delta;
      }
    }
    ch->depth_delta = delta;
  } else {
    uint64_t words = GPJO_WORDS(ch->n);
    int8_t level = ch->depth_in;
    uint64_t k = ch->nl_base;
    for (uint64_t w = ch->start / 64; w * 64 < ch->end; ++w) {
      for (int l = 0; l < ch->levels; ++l) ch->leveled_out[(uint64_t)l * words + w] = 0;
      uint64_t lim = (w + 1) * 64 < ch->end ? (w + 1) * 64 : ch->end;
      uint64_t s = ch->string_out[w];
      for (uint64_t i = w * 64; i < lim; ++i) {
        uint64_t bit = 1ull << (i % 64);
        if (f[i] == '\n') ch->newline_out[k++] = (int64_t)i;
        if (s & bit) continue;
        uint8_t c = f[i];
        if (c == '{' || c == '[') {
          level = (int8_t)(uint8_t)(level + 1);
          if (level >= 0 && level < ch->levels) ch->leveled_out[(uint64_t)level * words + w] |= bit;
        } else if (c == '}' || c == ']') {
          if (level >= 0 && level < ch->levels) ch->leveled_out[(uint64_t)level * words + w] |= bit;
          level = (int8_t)(uint8_t)(level - 1);
        } else if (c == ':' || c == ',') {
          if (level >= 0 && level < ch->levels) ch->leveled_out[(uint64_t)level * words + w] |= bit;
        }
      }
    }
  }
  return NULL;
}

static void mt_run(mt_chunk *chunks, int threads, int phase) {
  pthread_t *tid = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
  for (int t = 0; t < threads; ++t) {
    chunks[t].phase = phase;
    pthread_create(&tid[t], NULL, mt_worker, &chunks[t]);
  }
  for (int t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
  free(tid);
}

/* Returns the number of newline-index entries written (1 + count('\n')). */
uint64_t gpjo_build_index_mt(const uint8_t *file, uint64_t n, int levels,
                             int threads, uint64_t *string_out,
                             uint64_t *leveled_out, int64_t *newline_out) {
  if (threads < 1) threads = 1;
  uint64_t words = GPJO_WORDS(n);
  uint64_t per = (words + (uint64_t)threads - 1) / (uint64_t)threads;
  if (per == 0) per = 1;
  mt_chunk *ch = (mt_chunk *)calloc((size_t)threads, sizeof(mt_chunk));
  int used = 0;
  for (int t = 0; t < threads; ++t) {
    uint64_t s = (uint64_t)t * per * 64, e = s + per * 64;
    if (s >= n && !(n == 0 && t == 0)) break;
    if (e > n) e = n;
    memset(&ch[t], 0, sizeof ch[t]);
    ch[t].file = file;
    ch[t].n = n;
    ch[t].start = s;
    ch[t].end = e;
    ch[t].levels = levels;
    ch[t].string_out = string_out;
    ch[t].leveled_out = leveled_out;
    ch[t].newline_out = newline_out;
    ++used;
  }
  mt_run(ch, used, 0);
  int esc = 0;
  for (int t = 0; t < used; ++t) {
    ch[t].esc_in = esc;
    esc = ch[t].all_bs ? (esc ^ (int)((ch[t].end - ch[t].start) & 1)) : ch[t].esc_out0;
  }
  mt_run(ch, used, 1);
  int str = 0;
  for (int t = 0; t < used; ++t) {
    ch[t].str_in = str;
    str ^= ch[t].quote_parity;
  }
  mt_run(ch, used, 2);
  int8_t depth = -1;
  uint64_t nl = 1;
  newline_out[0] = 0;
  for (int t = 0; t < used; ++t) {
    ch[t].depth_in = depth;
    depth = (int8_t)(uint8_t)(depth + ch[t].depth_delta);
    ch[t].nl_base = nl;
    nl += ch[t].nl_count;
  }
  mt_run(ch, used, 3);
  free(ch);
  (void)words;
  return nl;
}

/* Query over all lines on `threads` host threads (lines are independent). */
typedef struct {
  qctx c;
  const int64_t *nl;
  int64_t nl_size, lo, hi;
  const uint8_t *ir;
  int64_t ir_len;
  int num_results;
  int variable;
  int64_t *result;
  int64_t *counts;
} mt_query;

static void *mt_query_worker(void *arg) {
  mt_query *q = (mt_query *)arg;
  for (int64_t line = q->lo; line < q->hi; ++line)
    query_line(&q->c, q->nl, q->nl_size, line, q->ir, q->ir_len, q->num_results, q->variable,
               q->result, q->counts);
  return NULL;
}

void gpjo_query_mt(const uint8_t *file, int64_t n, const int64_t *newline,
                   int64_t num_lines, const uint64_t *string,
                   const uint64_t *leveled, int64_t level_size, int num_levels,
                   const uint8_t *ir, int64_t ir_len, int num_results,
                   int64_t *result, int threads) {
  if (threads < 1) threads = 1;
  mt_query *q = (mt_query *)calloc((size_t)threads, sizeof(mt_query));
  pthread_t *tid = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
  int64_t per = (num_lines + threads - 1) / threads;
  for (int t = 0; t < threads; ++t) {
    int64_t lo = t * per, hi = lo + per;
    if (hi > num_lines) hi = num_lines;
    if (lo > hi) lo = hi;
    q[t] = (mt_query){{file, n, string, leveled, level_size, num_levels, 0}, newline, num_lines,
                      lo, hi, ir, ir_len, num_results, 0, result, NULL};
    pthread_create(&tid[t], NULL, mt_query_worker, &q[t]);
  }
  for (int t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
  free(tid);
  free(q);
}

/* Extension mode on `threads` host threads. */
void gpjo_query_ext_mt(const uint8_t *file, int64_t n, const int64_t *newline,
                       int64_t num_lines, const uint64_t *string,
                       const uint64_t *leveled, int64_t level_size, int num_levels,
                       const uint8_t *ir, int64_t ir_len, int cap, int64_t *result,
                       int64_t *counts, int threads) {
  if (threads < 1) threads = 1;
  mt_query *q = (mt_query *)calloc((size_t)threads, sizeof(mt_query));
  pthread_t *tid = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
  int64_t per = (num_lines + threads - 1) / threads;
  for (int t = 0; t < threads; ++t) {
    int64_t lo = t * per, hi = lo + per;
    if (hi > num_lines) hi = num_lines;
    if (lo > hi) lo = hi;
    q[t] = (mt_query){{file, n, string, leveled, level_size, num_levels, 0}, newline, num_lines,
                      lo, hi, ir, ir_len, cap, 1, result, counts};
    pthread_create(&tid[t], NULL, mt_query_worker, &q[t]);
  }
  for (int t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
  free(tid);
  free(q);
}
