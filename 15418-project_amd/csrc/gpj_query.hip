// gpj_query.hip -- JSONPath bytecode interpreter over the structural index.
//
// Semantics follow the reference query kernel
// (src/gpjson/query/kernels/orig/query.cu:144-454): each line between two
// newline-index entries is one record, trimmed to its first '{' and last '}';
// the bytecode walks the leveled rows and the string mask; STORE_RESULT
// writes {start, end} of the current value (spaces trimmed), -1 = no match.
// Deliberate differences, all outside what the reference defines: bytes
// outside [0, n) read as 0 (the reference reads out of bounds there); an
// unknown opcode or running off the bytecode ends the path; so does a
// negative key length or jump offset (which would move the bytecode position
// backwards and never end).
//
// Execution model.  Each lane interprets one record (a wave takes up to 64
// consecutive records), but the bytecode is decoded once per wave: every step
// picks the bytecode position all live lanes share, or else the furthest
// ahead, reads that instruction through the scalar unit, and runs it on
// exactly the lanes that sit there.  Lanes of one query normally move in
// lockstep, so an opcode costs one scalar decode instead of a divergent
// switch per lane; lanes that run ahead (past a filter) finish their step and
// loop back to where the others wait.  Some pairs always run in one dispatch
// (a step and the MOVE_DOWN after it, a store and END, a string test and
// RESET_POSITION, MARK_POSITION and what follows it).  Each lane still
// executes its own instruction sequence in order, so results are those of a
// per-record sequential interpreter.  The bytecode is read from
// its device buffer directly (no constant-memory copy), so concurrent
// launches with different queries are independent.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "gpj_bytecode.h"
#include "gpj_kernels.h"

namespace gpj {

namespace {

typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));
typedef int64_t i64x2a8 __attribute__((ext_vector_type(2), aligned(8)));

// Positions.  The interpreter holds byte positions as P: int32_t offsets from
// the record's 64-byte aligned start `base` in the narrow kernels, int64_t
// file offsets (base 0) in the wide ones.  The caches are tagged with the low
// 32 bits of file-absolute granule and word indices either way.
constexpr int32_t kUnset = INT32_MIN;            // level end not known yet
constexpr int32_t kNone = INT32_MIN + (1 << 19); // a backward search found nothing
// (narrow positions stay within span + 4 KiB of 0, span <= kNarrowSpan, so
// neither sentinel is ever a position, nor kNone minus a key length)

__device__ __forceinline__ u64x2 clear_past(u64x2 v, int64_t valid) {
  // keep the first `valid` bytes (0..16) of a granule
  const uint64_t m0 = valid >= 8 ? ~0ull : (valid <= 0 ? 0ull : (1ull << (8 * valid)) - 1);
  const int64_t v1 = valid - 8;
  const uint64_t m1 = v1 >= 8 ? ~0ull : (v1 <= 0 ? 0ull : (1ull << (8 * v1)) - 1);
  v.x &= m0;
  v.y &= m1;
  return v;
}

// Cache tags keep the low 32 bits of file-absolute granule and word indices.
// Word indices are >= -1 and below 2^32 - 2 (the host keeps the leveled rows
// below that), so the initial tag 2^32 - 2 never names a real word; granule
// indices of real bytes stay below 2^32 - 2^27 (inputs below 2^35 bytes), and
// a negative position that aliases a tag (the text pair's initial tags name
// granules -2 and -1) only ever meets zeros.
constexpr uint32_t kNoTag = 0xFFFFFFFEu;

// Per-lane view of one record's text and bitmaps, with caches for the last
// two consecutive text granules (filled together: a forward walk usually
// moves on into the next granule) and for the last granule read from even and
// from odd leveled rows (a walk alternates between a level and the one below
// it, so the two never evict each other).  The caches stay valid across the
// queries of a batch, which run over the same record.  String-mask words are
// read directly: a key lookup needs them only when the byte before its ':' is
// not an unescaped quote.
//
// Text granule G (file-absolute) holds file offsets [16G - head, 16G - head +
// 16), where head is the input address modulo 16: granules 0 < G < last hold
// input bytes only, granules below 0 or above last none (they read as 0), and
// the two edge granules are loaded (an aligned granule never crosses a page)
// with the bytes outside [0, n) cleared.  Position p lies in granule base / 16
// + (p + head) / 16 and in bitmap word base / 64 + p / 64.  Only base (held
// as gtag = base / 16, below 2^31 for inputs below 32 GiB) and the caches
// differ between lanes.
template <typename P> struct Ctx {
  const u64x2 *__restrict__ blocks; // granule 0 of the file
  const uint64_t *__restrict__ str;
  const uint64_t *__restrict__ lev; // leveled row k starts at lev + k * level_size
  int64_t n;
  int64_t level_size;
  int64_t span;     // positions reach at most span below 0 (see pos0)
  uint32_t lpar;    // address parity of leveled word 0 (words are read in 16-byte granules)
  uint32_t last;    // last granule holding an input byte (n > 0)
  uint32_t inner;   // last - 1, or 0 when last is 0: interior granules G - 1 < inner
  int32_t head;
  int32_t num_levels;
  int32_t lbase;    // leveled row of interpreter level 0
  uint32_t gtag;    // base / 16, base = file offset of position 0 (a multiple of 64)
  bool wide;        // a backward search reached pos0 while clamped: the record
                    // goes to the wide kernel
  uint32_t text_at; // granules text_at, text_at + 1 (absolute) in text_val, text_nxt
  u64x2 text_val;
  u64x2 text_nxt;
  uint32_t row_at0; // absolute word index (low 32 bits) of the granule in row_val0
  u64x2 row_val0;   // (a granule of an even row)
  uint32_t row_at1; // same for odd rows
  u64x2 row_val1;
};

template <typename P>
__device__ __forceinline__ Ctx<P> make_ctx(const uint8_t *file, int64_t n, const uint64_t *string,
                                           const uint64_t *leveled, int64_t level_size,
                                           int32_t num_levels, int32_t lbase, int64_t base,
                                           int64_t span) {
  Ctx<P> c;
  const int64_t head = (int64_t)(reinterpret_cast<uintptr_t>(file) & 15);
  const int64_t last = n > 0 ? (n - 1 + head) >> 4 : 0;
  c.blocks = reinterpret_cast<const u64x2 *>(file - head);
  c.str = string;
  c.lev = leveled;
  c.n = n;
  c.level_size = level_size;
  c.span = span;
  c.lpar = (uint32_t)(reinterpret_cast<uintptr_t>(leveled) >> 3) & 1u;
  c.last = (uint32_t)last;
  c.inner = last > 0 ? (uint32_t)(last - 1) : 0u;
  c.head = (int32_t)head;
  c.num_levels = num_levels;
  c.lbase = lbase;
  c.gtag = (uint32_t)(base >> 4);
  c.wide = false;
  c.text_at = kNoTag;
  c.text_val = u64x2{0, 0};
  c.text_nxt = u64x2{0, 0};
  c.row_at0 = kNoTag;
  c.row_val0 = u64x2{0, 0};
  c.row_at1 = kNoTag;
  c.row_val1 = u64x2{0, 0};
  return c;
}

template <typename P> __device__ __forceinline__ int64_t base_of(const Ctx<P> &c) {
  return (int64_t)c.gtag << 4;
}
// Bitmap word of relative word 0 (base / 64).
template <typename P> __device__ __forceinline__ int32_t bword(const Ctx<P> &c) {
  return (int32_t)(c.gtag >> 2);
}
// Lowest position a backward search reaches: file offset 0, or -span when
// that lies further down (clamped; narrow kernels only).
template <typename P> __device__ __forceinline__ bool clamped(const Ctx<P> &c) {
  return base_of(c) > c.span;
}
template <typename P> __device__ __forceinline__ P pos0(const Ctx<P> &c) {
  return (P)(clamped(c) ? -c.span : -base_of(c));
}
// Text granule g (relative); its absolute index gtag + g lies in (-2^31, 2^31).
template <typename P> __device__ __forceinline__ const u64x2 *granule_ptr(const Ctx<P> &c, P g) {
  return c.blocks + (int64_t)(int32_t)(c.gtag + (uint32_t)g);
}

// Granule g of the input (granules outside the input read as 0).
template <typename P> __device__ __forceinline__ u64x2 text_granule(const Ctx<P> &c, P g) {
  const uint32_t ga = c.gtag + (uint32_t)g;         // absolute granule, |G| < 2^31
  if (ga - 1u < c.inner) return *granule_ptr(c, g); // interior granule
  if (ga > c.last || c.n <= 0) return u64x2{0, 0}; // G < 0 or G > last
  // the first or the last granule: clear the bytes before offset 0 and
  // from offset n on
  const int64_t fo = 16 * (int64_t)(int32_t)ga - c.head; // file offset of byte 0
  u64x2 v = clear_past(*granule_ptr(c, g), c.n - fo);
  const int64_t lo = -fo; // bytes before offset 0
  if (lo > 0) {
    v.x &= lo >= 8 ? 0ull : (~0ull << (8 * lo));
    v.y &= lo >= 16 ? 0ull : (lo <= 8 ? ~0ull : (~0ull << (8 * (lo - 8))));
  }
  return v;
}

// Fills the lane's granule pair with granules g and g + 1 (one round trip).
template <typename P> __device__ __forceinline__ void text_fill(Ctx<P> &c, P g) {
  const uint32_t ga = c.gtag + (uint32_t)g;
  if (ga - 1u < c.inner && ga < c.inner) { // both interior
    const u64x2 *q = granule_ptr(c, g);
    c.text_val = q[0];
    c.text_nxt = q[1];
  } else {
    c.text_val = text_granule(c, g);
    c.text_nxt = text_granule(c, (P)(g + 1));
  }
  c.text_at = ga;
}
// Granule g from the lane's pair, filling the pair from g on a miss.
template <typename P> __device__ __forceinline__ u64x2 text_get(Ctx<P> &c, P g) {
  const uint32_t d = c.gtag + (uint32_t)g - c.text_at;
  if (d > 1u) {
    text_fill(c, g);
    return c.text_val;
  }
  return d ? c.text_nxt : c.text_val;
}

// Byte at position p; bytes outside [0, n) read as 0.
template <typename P> __device__ __forceinline__ uint8_t byte_at(Ctx<P> &c, P p) {
  const P a = p + c.head;
  const u64x2 v = text_get(c, (P)(a >> 4));
  const uint64_t w = (a & 8) ? v.y : v.x;
  return (uint8_t)(w >> (8 * (a & 7)));
}

// Little-endian 8 bytes from position p.
template <typename P> __device__ __forceinline__ uint64_t text8(Ctx<P> &c, P p) {
  const P a = p + c.head;
  const P g = a >> 4;
  const int32_t o = (int32_t)(a & 15);
  const u64x2 v0 = text_get(c, g);
  if (o == 0) return v0.x;
  if (o == 8) return v0.y;
  const uint32_t sh = 8 * (uint32_t)(o & 7);
  if (o < 8) return (v0.x >> sh) | (v0.y << (64 - sh));
  const u64x2 v1 = text_get(c, (P)(g + 1));
  return (v0.y >> sh) | (v1.x << (64 - sh));
}

// A leveled row as the searches read it: the index, counted from the start
// of the leveled rows, of the row's word that holds position 0, which row
// cache it uses (row parity), and valid = false for levels outside the built
// rows, which read as empty.
struct RowRef {
  int64_t w0;
  bool odd;
  bool valid;
};
template <typename P> __device__ __forceinline__ RowRef row_ref(const Ctx<P> &c, int level) {
  RowRef r;
  const int row = level + c.lbase;
  r.valid = level >= 0 && row < c.num_levels;
  r.w0 = c.level_size * (int64_t)row + bword(c);
  r.odd = (row & 1) != 0;
  return r;
}

// Word w of row r, read as part of its aligned 16-byte granule (the
// granule's other word may lie just outside the bitmap, inside the same
// aligned 16 bytes); the lane keeps one granule per row parity, tagged with
// the low 32 bits of the granule's word index.
template <typename P>
__device__ __forceinline__ uint64_t granule_word(const Ctx<P> &c, const RowRef &r, P w,
                                                 uint32_t &at, u64x2 &val) {
  const int64_t wi = r.w0 + (int64_t)w;
  const uint32_t hi = ((uint32_t)wi ^ c.lpar) & 1u; // wi is its granule's second word
  const int64_t g = wi - (int64_t)hi;
  if ((uint32_t)g != at) {
    val = *reinterpret_cast<const u64x2 *>(c.lev + g);
    at = (uint32_t)g;
  }
  return hi ? val.y : val.x;
}
template <typename P> __device__ __forceinline__ uint64_t row_word(Ctx<P> &c, const RowRef &r, P w) {
  return r.odd ? granule_word(c, r, w, c.row_at1, c.row_val1)
               : granule_word(c, r, w, c.row_at0, c.row_val0);
}
template <typename P> __device__ __forceinline__ uint64_t str_word(const Ctx<P> &c, P w) {
  return c.str[(int64_t)(bword(c) + (int32_t)w)];
}

__device__ __forceinline__ uint64_t through(int b) {
  return b >= 63 ? ~0ull : ((1ull << (b + 1)) - 1);
}

// Last set bit of the string mask in [lo, idx], or kNone.
// (lo >= pos0 in every call, so the search never leaves the file.)
template <typename P> __device__ __forceinline__ P last_string_bit(const Ctx<P> &c, P lo, P idx) {
  if (idx < lo) return (P)kNone;
  P wi = idx >> 6, wb = wi * 64;
  uint64_t word = str_word(c, wi) & through((int)(idx - wb));
  for (;;) {
    uint64_t cand = word;
    const P first = lo - wb;
    if (first > 0) cand &= ~through((int)(first - 1));
    if (cand) return wb + 63 - (P)__builtin_clzll(cand);
    if (bword(c) + (int32_t)wi <= 0 || wb - 1 < lo) return (P)kNone;
    --wi;
    wb -= 64;
    word = str_word(c, wi);
  }
}

// Last set bit of row r in [lo, idx], or kNone.
template <typename P>
__device__ __forceinline__ P last_row_bit(Ctx<P> &c, const RowRef &r, P lo, P idx) {
  if (idx < lo) return (P)kNone;
  P wi = idx >> 6, wb = wi * 64;
  uint64_t word = row_word(c, r, wi) & through((int)(idx - wb));
  for (;;) {
    uint64_t cand = word;
    const P first = lo - wb;
    if (first > 0) cand &= ~through((int)(first - 1));
    if (cand) return wb + 63 - (P)__builtin_clzll(cand);
    if (bword(c) + (int32_t)wi <= 0 || wb - 1 < lo) return (P)kNone;
    --wi;
    wb -= 64;
    word = row_word(c, r, wi);
  }
}

// First structural bit of row r in [idx, level_end], else level_end
// (find_next_structural_char, query.cu:79-115).  Every call passes a
// position at or above file offset 0 (one past a position >= -1), so the
// reference's clamp of idx to 0 never applies.
template <typename P>
__device__ __forceinline__ P next_structural(Ctx<P> &c, P level_end, P idx, const RowRef &r) {
  if (idx >= level_end) return level_end;
  if (!r.valid) return level_end;
  P wi = idx >> 6, wb = wi * 64;
  uint64_t word = row_word(c, r, wi) & (~0ull << (idx - wb));
  for (;;) {
    uint64_t cand = word;
    const P last = level_end - wb;
    if (last < 63) cand &= through((int)last);
    if (cand) return wb + (P)__builtin_ctzll(cand);
    ++wi;
    wb += 64;
    if (wb > level_end || (int64_t)(bword(c) + (int32_t)wi) >= c.level_size) return level_end;
    word = row_word(c, r, wi);
  }
}
template <typename P>
__device__ __forceinline__ P next_structural(Ctx<P> &c, P level_end, P idx, int level) {
  return next_structural(c, level_end, idx, row_ref(c, level));
}

// Last structural bit of the level's row in [file offset 0, idx], else file
// offset 0 (find_previous_structural_char, query.cu:117-128, always called
// with level start 0).  When file offset 0 lies below pos0 and the answer
// is pos0, the record goes to the wide kernel.
template <typename P> __device__ __forceinline__ P prev_structural(Ctx<P> &c, P idx, int level) {
  const P lo = pos0(c);
  P out = lo;
  if (idx > lo && level >= 0 && level + c.lbase < c.num_levels) {
    const P s = last_row_bit(c, row_ref(c, level), lo, idx);
    if (s != (P)kNone) out = s;
  }
  if (clamped(c) && out == lo) c.wide = true;
  return out;
}

// Per-lane array of N entries kept in LDS as [N][64] (one column per lane of
// the wave), so a lane-varying index is an ordinary LDS address and the
// interpreter's registers stay free.  Indices are in [0, N) by construction:
// the interpreter checks every level, mark and iteration depth against N
// before it moves there.
template <typename T, int N> struct LdsArray {
  T (*a)[64];
  int lane;
  __device__ __forceinline__ T get(int i) const { return a[i][lane]; }
  __device__ __forceinline__ void set(int i, T x) { a[i][lane] = x; }
  __device__ __forceinline__ void fill(T x) {
#pragma unroll
    for (int k = 0; k < N; ++k) a[k][lane] = x;
  }
  // entry k := x for every k > i (i >= 0)
  __device__ __forceinline__ void fill_above(int i, T x) {
#pragma unroll
    for (int k = 1; k < N; ++k)
      if (k > i) a[k][lane] = x;
  }
  // walking up from entry i + 1 (i >= 0): := x until the first entry that
  // already holds x (which stays, as do all after it)
  __device__ __forceinline__ void fill_above_until(int i, T x) {
    for (int k = i + 1; k < N; ++k) {
      if (a[k][lane] == x) break;
      a[k][lane] = x;
    }
  }
};

// Spans a lane keeps in the wave's LDS stage for dense results (the narrow
// kernels; the wide ones keep none): a record with more goes to the dense
// fixup pass.
template <typename P> struct StageSize {
  static constexpr int value = sizeof(P) == 4 ? kDenseStage : 1;
};

// LDS home of the per-lane arrays of one wave (with a dense-results stage of
// STAGE entries per lane in the kernels that run dense queries).
template <int MAXLV, typename P, int STAGE = 1> struct WaveArrays {
  static constexpr int kF = MAXLV < kMaxFrames ? MAXLV : kMaxFrames;
  P stage[STAGE][64][2]; // dense results: {start, end} per lane
  P level_end[MAXLV][64];
  P marked_pos[MAXLV][64];
  P fr_sep[kF][64];
  int32_t curr_index[MAXLV][64];
  uint32_t fr_pack[kF][64]; // qp | lvl << 16 | (mi + 1) << 21 | down << 31
  uint8_t marked_lvl[MAXLV][64];
};

__device__ __forceinline__ int32_t wave_max_i32(int32_t v) {
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) {
    const int32_t o = __shfl_xor(v, m, 64);
    v = o > v ? o : v;
  }
  return v;
}

// Per-lane interpreter state for one record and one query: the scalars in
// registers, the arrays in the wave's LDS block.
// MAXLV bounds the level (the host derives it from the bytecode); levels,
// marks and element iterations each use at most MAXLV entries.
template <int MAXLV, typename P> struct LaneState {
  static constexpr int kF = WaveArrays<MAXLV, P>::kF;
  P idx;
  int32_t lvl, qp, mi, nfr;
  P count; // results stored (fixed) or found (variable); at most one per byte
  LdsArray<P, MAXLV> level_end;
  LdsArray<int32_t, MAXLV> curr_index;
  LdsArray<P, MAXLV> marked_pos;
  LdsArray<uint8_t, MAXLV> marked_lvl;
  LdsArray<uint32_t, kF> fr_pack; // bytecode position (after MOVE_TO_EACH and
                                  // its MOVE_DOWN, so <= 4097), level, mark
                                  // depth + 1, and whether a MOVE_DOWN follows
                                  // MOVE_TO_EACH
  LdsArray<P, kF> fr_sep;
  template <int S>
  __device__ __forceinline__ LaneState(WaveArrays<MAXLV, P, S> &w, int lane)
      : level_end{w.level_end, lane}, curr_index{w.curr_index, lane},
        marked_pos{w.marked_pos, lane}, marked_lvl{w.marked_lvl, lane},
        fr_pack{w.fr_pack, lane}, fr_sep{w.fr_sep, lane} {}
};

// RESET_POSITION: back to the innermost mark, clearing the levels above it up
// to the first one already clear.  False: no mark (the path fails).
template <int MAXLV, typename P>
__device__ __forceinline__ bool reset_position(LaneState<MAXLV, P> &s) {
  if (s.mi < 0) return false;
  s.idx = s.marked_pos.get(s.mi);
  s.lvl = s.marked_lvl.get(s.mi);
  --s.mi;
  s.level_end.fill_above_until(s.lvl, (P)kUnset);
  return true;
}

// MOVE_DOWN: one level deeper; a level entered for the first time ends at
// the next structural character of the level above.  False: the path fails.
template <int MAXLV, typename P>
__device__ __forceinline__ bool move_down(Ctx<P> &c, LaneState<MAXLV, P> &s) {
  if (s.lvl + 1 >= MAXLV) return false;
  ++s.lvl;
  if (s.level_end.get(s.lvl) == (P)kUnset) {
    s.level_end.set(s.lvl, next_structural(c, s.level_end.get(s.lvl - 1), (P)(s.idx + 1),
                                           s.lvl - 1));
    while (byte_at(c, s.idx) == ' ') ++s.idx;
  }
  return true;
}

// A finished path (END or any failure): resume the innermost element
// iteration at its next element, or end the record.  Returns false when the
// record is done.  When the bytecode has a MOVE_DOWN right after MOVE_TO_EACH
// it runs here for the new element (its failure fails the element's path,
// which resumes the next element), so the element's next step is the
// instruction after it.
template <int MAXLV, typename P>
__device__ __forceinline__ bool resume(Ctx<P> &c, LaneState<MAXLV, P> &s) {
  for (;;) {
    if (s.nfr == 0) return false;
    const int f = s.nfr - 1;
    const uint32_t pack = s.fr_pack.get(f);
    const int flvl = (int)((pack >> 16) & 31u);
    const P lend = s.level_end.get(flvl);
    const P sep = next_structural(c, lend, (P)(s.fr_sep.get(f) + 1), flvl);
    if (byte_at(c, sep) == ',' && sep < lend) {
      s.fr_sep.set(f, sep);
      s.lvl = flvl;
      s.level_end.fill_above(s.lvl, (P)kUnset);
      s.mi = (int32_t)((pack >> 21) & 31u) - 1;
      s.idx = sep + 1;
      s.qp = (int32_t)(pack & 0xFFFFu);
      if ((pack >> 31) && !move_down(c, s)) continue;
      return true;
    }
    --s.nfr;
  }
}

// Where STORE_RESULT puts the spans of a lane's record: entries below `keep`
// of `res` (file offsets, int64 pairs), or, with a stage, the wave's LDS stage
// (positions, the first StageSize entries); the others are only counted.
template <typename P> struct Sink {
  int64_t *res;
  int64_t keep;
  P (*stage)[64][2];
};

// STORE_RESULT: the current value (spaces trimmed) goes to the next result
// entry.  False when fixed mode has no slot left (the path fails).
template <int MAXLV, typename P>
__device__ __forceinline__ bool store_result(Ctx<P> &c, LaneState<MAXLV, P> &s, int32_t cap,
                                             bool variable, const Sink<P> &out) {
  if (!variable && s.count >= cap) return false;
  const P lend = s.level_end.get(s.lvl);
  P end_str = lend;
  while (byte_at(c, s.idx) == ' ' && s.idx < lend) ++s.idx;
  while (end_str > s.idx && byte_at(c, (P)(end_str - 1)) == ' ') --end_str;
  if (out.stage) {
    if (s.count < (P)StageSize<P>::value) {
      const int lane = (int)(threadIdx.x & 63);
      out.stage[s.count][lane][0] = s.idx;
      out.stage[s.count][lane][1] = end_str;
    }
  } else if ((int64_t)s.count < out.keep) { // one 16-byte store (8-byte aligned array)
    *reinterpret_cast<i64x2a8 *>(out.res + (int64_t)s.count * 2) =
        i64x2a8{base_of(c) + (int64_t)s.idx, base_of(c) + (int64_t)end_str};
  }
  ++s.count;
  return true;
}

// ---- shared first step of a batch -----------------------------------------
// Every query whose bytecode starts with MOVE_TO_KEY searches the record's
// top-level object from its '{' (the trimmed record start) on level 0.  The
// candidates such a search visits, and where it stops when no key matches,
// depend on the record alone; only the match test depends on the key.  So
// when a batch holds two or more such queries, one scan of the object tests
// each candidate against all their distinct keys (the first kKeyMemo of them)
// and keeps, per lane and key, where that key's search alone would stop, or
// that it fails; each of those queries then starts after its first step,
// from that outcome, instead of searching.  Outcomes as MOVE_TO_KEY: the
// first candidate whose key matches (stop there), a candidate that is not a
// ':' (stop there, no failure), the object's closing '}' or no progress
// (failure).  When every query of a key follows it with MOVE_DOWN, the scan
// also takes that step: the value's end on level 0 is the next candidate
// search's first structural (the scan needs it anyway to go on), and the
// spaces after the stop are in the lane's text window.  (The kernels with a
// dense-results stage keep no level ends: three keys' positions keep their
// LDS within 10 KiB per wave, 16 waves per CU.)
constexpr int kKeyMemo = 3;
template <typename P, bool END> struct KeyMemo {
  P pos[kKeyMemo][64]; // where the query goes on (> 0), -1: the search fails
  P end[kKeyMemo][64]; // keys followed by MOVE_DOWN: the level end it sets
};
template <typename P> struct KeyMemo<P, false> {
  P pos[kKeyMemo][64];
};

// Which batch queries take their first step from the memo (wave-uniform):
// nibble q of `slot` is 1 + the memo key of query q (0: none), byte j of
// `query` the query whose bytecode holds key j; `keys` distinct keys; bit j
// of `down`: every query of key j has MOVE_DOWN right after it.  Off (slot 0)
// unless at least two queries qualify.
struct MemoPlan {
  uint32_t slot;
  uint32_t query;
  int32_t keys;
  uint32_t down; // (0 when the memo keeps no level ends)
};
__device__ __forceinline__ int32_t memo_slot(const MemoPlan &mp, int q) {
  return (int32_t)((mp.slot >> (4 * q)) & 15u) - 1;
}
__device__ __forceinline__ MemoPlan memo_plan(const QueryBatch &batch, bool ends) {
  MemoPlan mp{0, 0, 0, 0};
  uint32_t not_down = 0;
  int users = 0;
  for (int q = 0; q < batch.count; ++q) {
    const IrReader ir = IrReader::make(batch.q[q].ir, batch.q[q].ir_len);
    if (ir.len < 2 || ir.byte(0) != kMoveToKey) continue;
    uint32_t p = 1;
    const int32_t arg = ir.varint(p);
    if (arg < 0) continue;
    int32_t j = 0;
    for (; j < mp.keys; ++j) { // the same key as an earlier query's?
      const int qj = (int)((mp.query >> (8 * j)) & 0xFFu);
      const IrReader oj = IrReader::make(batch.q[qj].ir, batch.q[qj].ir_len);
      uint32_t pj = 1;
      if (oj.varint(pj) != arg) continue;
      bool same = true;
      for (int32_t k = 0; same && k < arg; k += 8) {
        const int32_t rest = arg - k;
        const uint64_t mk = rest >= 8 ? ~0ull : ((1ull << (8 * rest)) - 1);
        same = ((ir.u64(p + (uint32_t)k) ^ oj.u64(pj + (uint32_t)k)) & mk) == 0;
      }
      if (same) break;
    }
    if (j == mp.keys) {
      if (mp.keys == kKeyMemo) continue;
      mp.query |= (uint32_t)q << (8 * j);
      ++mp.keys;
    }
    mp.slot |= (uint32_t)(j + 1) << (4 * q);
    if (ir.byte(p + (uint32_t)arg) != kMoveDown) not_down |= 1u << j;
    ++users;
  }
  mp.down = ends ? ((1u << mp.keys) - 1u) & ~not_down : 0u;
  if (users < 2) mp = MemoPlan{0, 0, 0, 0};
  return mp;
}

// Where query q (with memo key j) starts: the bytecode position after its
// first MOVE_TO_KEY (and the MOVE_DOWN after it when the scan took that
// too), the record position and the level end of level 1 (kUnset: none); a
// failed search ends the record's walk (no element iteration is open yet).
template <typename P, bool END>
__device__ __forceinline__ int32_t memo_start(const KeyMemo<P, END> &m, const MemoPlan &mp,
                                              const IrReader &ir, int32_t j, bool &live,
                                              P &start_idx, P &start_end) {
  uint32_t p = 1;
  const int32_t arg = ir.varint(p); // >= 0 (memo_plan)
  const bool down = (mp.down >> j) & 1u;
  if (live) {
    const int lane = (int)(threadIdx.x & 63);
    const P pos = m.pos[j][lane];
    live = pos >= 0;
    start_idx = pos;
    if constexpr (END)
      if (down) start_end = m.end[j][lane];
  }
  return (int32_t)p + arg + (down ? 1 : 0);
}

// The scan (see KeyMemo): the loop of MOVE_TO_KEY (run_query) with one match
// test per pending key at each candidate.  lane_keys: the memo keys this
// lane's record needs (its scan ends once they are found).
template <typename P, bool END>
__device__ __forceinline__ void key_memo_scan(Ctx<P> &c, KeyMemo<P, END> &m, bool live,
                                              P line_start, P line_end, const QueryBatch &batch,
                                              const MemoPlan &mp, uint32_t lane_keys) {
  const int lane = (int)(threadIdx.x & 63);
  uint32_t pending = lane_keys;
  if (!live || !pending) return;
  const RowRef rr = row_ref(c, 0);
  // a search that stops at idx goes on at idx + 1; a MOVE_DOWN there sets
  // level 1's end to the first level-0 structural from idx + 2 (sep: the
  // first one from idx + 1) and skips the spaces from idx + 1
  auto stop = [&](uint32_t keys, P idx, P sep) {
    if (keys & mp.down) {
      const P end = sep == idx + 1 ? next_structural(c, line_end, (P)(idx + 2), rr) : sep;
      P at = idx + 1;
      while (byte_at(c, at) == ' ') ++at;
      for (int j = 0; j < mp.keys; ++j)
        if ((keys & mp.down) >> j & 1u) {
          m.pos[j][lane] = at;
          if constexpr (END) m.end[j][lane] = end;
        }
    }
    for (int j = 0; j < mp.keys; ++j)
      if ((keys & ~mp.down) >> j & 1u) m.pos[j][lane] = idx + 1;
  };
  P idx = next_structural(c, line_end, (P)(line_start + 1), rr);
  uint64_t w = text8(c, (P)(idx - 7)); // bytes idx - 7 .. idx
  for (;;) {
    if ((w >> 56) != ':') {
      stop(pending, idx, next_structural(c, line_end, (P)(idx + 1), rr));
      return;
    }
    const bool compact =
        idx - 2 >= line_start + 1 && ((w >> 48) & 0xFFu) == '"' && ((w >> 40) & 0xFFu) != '\\';
    P string_end = (P)kUnset; // found on the first key that needs it
    uint32_t matched = 0;
    for (int j = 0; j < mp.keys; ++j) {
      if (!((pending >> j) & 1u)) continue;
      const int q = (int)((mp.query >> (8 * j)) & 0xFFu);
      const IrReader ir = IrReader::make(batch.q[q].ir, batch.q[q].ir_len);
      uint32_t key_pos = 1;
      const int32_t arg = ir.varint(key_pos);
      const uint64_t key0 = ir.u64(key_pos);
      const uint64_t mask0 = arg >= 8 ? ~0ull : ((1ull << (8 * arg)) - 1);
      const bool short_key =
          arg <= 5 && (arg == 0 || ((key0 >> (8 * (arg - 1))) & 0xFFu) != '\\');
      bool match;
      if (compact && short_key) {
        const uint32_t jss = (uint32_t)(5 - arg);
        const uint64_t pattern = ((uint64_t)'"' << (8 * jss)) |
                                 ((key0 & mask0) << (8 * (jss + 1))) | (0x22ull << 48) |
                                 (0x3Aull << 56);
        const uint64_t pmask = ~0ull << (8 * jss);
        match = idx - 2 - arg >= line_start && ((w ^ pattern) & pmask) == 0;
      } else {
        if (string_end == (P)kUnset)
          string_end =
              compact ? (P)(idx - 2) : last_string_bit(c, (P)(line_start + 1), (P)(idx - 1));
        const P ss = string_end - arg;
        match = ss >= line_start && byte_at(c, ss) == '"';
        if (match && arg > 0) {
          match = ((text8(c, (P)(ss + 1)) ^ key0) & mask0) == 0;
          for (int32_t k = 8; match && k < arg; k += 8) {
            const int32_t rest = arg - k;
            const uint64_t mk = rest >= 8 ? ~0ull : ((1ull << (8 * rest)) - 1);
            match = ((text8(c, (P)(ss + 1 + k)) ^ ir.u64(key_pos + (uint32_t)k)) & mk) == 0;
          }
        }
      }
      if (match) matched |= 1u << j;
    }
    pending &= ~matched;
    const P sep = next_structural(c, line_end, (P)(idx + 1), rr);
    if (matched) stop(matched, idx, sep);
    if (!pending) return;
    const P nxt = next_structural(c, line_end, (P)(sep + 1), rr);
    const uint64_t wn = text8(c, (P)(nxt - 7));
    const P off = sep - (nxt - 7);
    const uint32_t sep_byte =
        (off >= 0 && off < 8) ? (uint32_t)(wn >> (8 * off)) & 0xFFu : (uint32_t)byte_at(c, sep);
    if (sep_byte == '}' || sep == idx) { // the searches of the pending keys fail
      for (int j = 0; j < mp.keys; ++j)
        if ((pending >> j) & 1u) m.pos[j][lane] = (P)-1;
      return;
    }
    idx = nxt;
    w = wn;
  }
}

// Runs one query over each lane's record.  `live` marks lanes holding a
// record (trimmed to line_start < line_end).  Fixed mode (reference): STORE i
// writes slot i; any failure ends the record.  Variable mode (extension):
// STORE appends (first `cap` stored, all counted) and a finished path
// resumes the innermost MOVE_TO_EACH at its next element.  A lane whose
// record needs the wide kernel (c.wide) stops.  Returns the lane's result
// count.  Stored spans are file offsets (base + position).  The walk starts
// at bytecode position start_qp (wave-uniform) and record position start_idx:
// 0 and line_start, or, for a query whose first steps key_memo_scan took,
// the position after them and where they left the record (on level 1 with
// that level ending at start_end when they include a MOVE_DOWN, else
// start_end is kUnset).
template <int MAXLV, typename P, int S>
__device__ __forceinline__ int64_t run_query(Ctx<P> &c, WaveArrays<MAXLV, P, S> &w, bool live,
                                             P line_start, P line_end, const IrReader ir,
                                             int32_t cap, bool variable, const Sink<P> &res,
                                             P start_idx, int32_t start_qp, P start_end) {
  static_assert(MAXLV >= 2, "a first MOVE_DOWN needs level 1");
  LaneState<MAXLV, P> s(w, (int)(threadIdx.x & 63));
  s.idx = start_idx;
  s.lvl = start_end == (P)kUnset ? 0 : 1;
  s.qp = start_qp;
  s.mi = -1;
  s.nfr = 0;
  s.count = 0;
  s.level_end.fill((P)kUnset);
  s.level_end.set(0, line_end);
  if (s.lvl) s.level_end.set(1, start_end);
  s.curr_index.fill(-1);

  for (;;) {
    const uint64_t act = __ballot(live);
    if (!act) break;
    // bytecode position to run: the one all live lanes share, else the
    // furthest ahead (lanes that run ahead, e.g. past a filter, then loop
    // back to where the others wait)
    const int32_t q0 = __builtin_amdgcn_readlane(s.qp, (int)__builtin_ctzll(act));
    int32_t qpu = q0;
    if (__ballot(live && s.qp == q0) != act) {
      const int32_t m =
          __builtin_amdgcn_readfirstlane(wave_max_i32(live ? s.qp : INT32_MIN));
      if (__ballot(live && s.qp == m)) qpu = m;
    }
    // decode the instruction at qpu (wave-uniform); a MARK_POSITION runs in
    // the same dispatch as the instruction after it
    uint32_t qn = (uint32_t)qpu;
    uint32_t op = qn >= ir.len ? (uint32_t)kOffEnd : ir.byte(qn++);
    const bool pre_mark = op == kMarkPosition;
    if (pre_mark) op = qn >= ir.len ? (uint32_t)kOffEnd : ir.byte(qn++);
    int32_t arg = 0;
    uint32_t key_pos = 0;
    if (op == kMoveToKey || op == kMoveToIndex || op == kMoveToIndexReverse ||
        op == kExpressionStringEquals || op == kJumpIfNotArray) {
      arg = ir.varint(qn);
      key_pos = qn;
      if (op == kMoveToKey || op == kExpressionStringEquals) qn += (uint32_t)arg;
    }
    // a key, index or element step that the bytecode follows with MOVE_DOWN,
    // a store followed by END and a string test followed by RESET_POSITION
    // run both in one dispatch (each lane still executes them in order)
    const uint32_t next_op = qn < ir.len ? ir.byte(qn) : (uint32_t)kOffEnd;
    const bool then_down = (op == kMoveToKey || op == kMoveToIndex ||
                            op == kMoveToIndexReverse || op == kMoveToEach) &&
                           next_op == kMoveDown;
    const bool then_end = op == kStoreResult && next_op == kEnd;
    const bool then_reset = op == kExpressionStringEquals && next_op == kResetPosition;
    if (then_down || then_end || then_reset) ++qn;
    // a key or index step with its MOVE_DOWN followed by STORE_RESULT and END
    // runs all four in one dispatch (not MOVE_TO_EACH: its frame keeps the
    // position after its MOVE_DOWN)
    const bool then_store = then_down && op != kMoveToEach && qn + 1 < ir.len &&
                            ir.byte(qn) == kStoreResult && ir.byte(qn + 1) == kEnd;
    if (then_store) qn += 2;
    if (!(live && s.qp == qpu)) continue;

    s.qp = (int32_t)qn;
    bool fail = false;
    if (pre_mark) {
      if (s.mi + 1 >= MAXLV) {
        fail = true;
      } else {
        ++s.mi;
        s.marked_pos.set(s.mi, s.idx);
        s.marked_lvl.set(s.mi, (uint8_t)s.lvl);
      }
    }
    if (!fail) switch (op) {
    case kStoreResult:
      fail = !store_result(c, s, cap, variable, res) || then_end; // END: the path is finished
      break;
    case kMoveUp:
      if (s.lvl <= 0) { fail = true; break; }
      s.idx = s.level_end.get(s.lvl);
      s.level_end.set(s.lvl, (P)kUnset);
      --s.lvl;
      break;
    case kMoveDown:
      fail = !move_down(c, s);
      break;
    case kMoveToKey: {
      if (arg < 0 || byte_at(c, s.idx) != '{') { fail = true; break; }
      const P lend = s.level_end.get(s.lvl);
      const uint64_t key0 = ir.u64(key_pos);
      const uint64_t mask0 = arg >= 8 ? ~0ull : ((1ull << (8 * arg)) - 1);
      // A key of at most 5 bytes that does not end in a backslash: the 8
      // bytes ending at a compact key's ':' hold its opening quote, the key,
      // its closing quote and the ':', so one text read decides the key
      // (wave-uniform pattern and mask over bytes 5 - arg .. 7).
      const bool short_key =
          arg <= 5 && (arg == 0 || ((key0 >> (8 * (arg - 1))) & 0xFFu) != '\\');
      const uint32_t jss = short_key ? (uint32_t)(5 - arg) : 0u;
      const uint64_t pattern = ((uint64_t)'"' << (8 * jss)) | ((key0 & mask0) << (8 * (jss + 1))) |
                               (0x22ull << 48) | (0x3Aull << 56);
      const uint64_t pmask = ~0ull << (8 * jss);
      const RowRef rr = row_ref(c, s.lvl);
      P idx = next_structural(c, lend, (P)(s.idx + 1), rr);
      uint64_t w = text8(c, (P)(idx - 7)); // bytes idx - 7 .. idx
      for (;;) {
        if ((w >> 56) != ':') break;
        // idx is outside strings, so an unescaped quote right before it closes
        // a string whose last in-string byte (body or opening quote) is
        // idx - 2 (compact): the string mask's last set bit in
        // [line_start + 1, idx - 1]
        const bool compact =
            idx - 2 >= line_start + 1 && ((w >> 48) & 0xFFu) == '"' && ((w >> 40) & 0xFFu) != '\\';
        bool match;
        if (compact && short_key) {
          match = idx - 2 - arg >= line_start && ((w ^ pattern) & pmask) == 0;
        } else {
          const P string_end =
              compact ? (P)(idx - 2) : last_string_bit(c, (P)(line_start + 1), (P)(idx - 1));
          const P ss = string_end - arg;
          match = ss >= line_start && byte_at(c, ss) == '"';
          if (match && arg > 0) {
            match = ((text8(c, (P)(ss + 1)) ^ key0) & mask0) == 0;
            for (int32_t k = 8; match && k < arg; k += 8) {
              const int32_t rest = arg - k;
              const uint64_t m = rest >= 8 ? ~0ull : ((1ull << (8 * rest)) - 1);
              match = ((text8(c, (P)(ss + 1 + k)) ^ ir.u64(key_pos + (uint32_t)k)) & m) == 0;
            }
          }
        }
        if (match) break;
        // the separator after the value ends the object if it is '}'; the
        // next key's 8-byte window usually holds it as well
        const P sep = next_structural(c, lend, (P)(idx + 1), rr);
        const P nxt = next_structural(c, lend, (P)(sep + 1), rr);
        const uint64_t wn = text8(c, (P)(nxt - 7));
        const P off = sep - (nxt - 7);
        const uint32_t sep_byte =
            (off >= 0 && off < 8) ? (uint32_t)(wn >> (8 * off)) & 0xFFu : (uint32_t)byte_at(c, sep);
        if (sep_byte == '}') {
          idx = sep;
          fail = true;
          break;
        }
        if (sep == idx) { // no progress: a ':' at the level end (malformed input;
          fail = true;    // the reference would search it forever)
          break;
        }
        idx = nxt;
        w = wn;
      }
      s.idx = idx + 1;
      if (!fail && then_down) fail = !move_down(c, s);
      break;
    }
    case kMoveToIndex: {
      const int32_t target = arg;
      const uint8_t ch = byte_at(c, s.idx);
      if (ch != '[' && ch != ',' && ch != ']') { fail = true; break; }
      const P lend = s.level_end.get(s.lvl);
      int32_t ci = (ch == '[') ? 0 : s.curr_index.get(s.lvl) + 1;
      P idx = s.idx;
      for (;;) {
        if (ci < target) {
          idx = next_structural(c, lend, (P)(idx + 1), s.lvl);
          if (byte_at(c, idx) == ',') { ++ci; continue; }
          fail = true;
          break;
        }
        if (ci > target) {
          idx = prev_structural(c, (P)(idx - 1), s.lvl);
          if (c.wide) break;
          --ci;
          continue;
        }
        break;
      }
      if (c.wide) break;
      s.curr_index.set(s.lvl, ci);
      if (!fail && byte_at(c, idx) == ']') fail = true;
      s.idx = idx + 1;
      if (!fail && then_down) fail = !move_down(c, s);
      break;
    }
    case kMoveToIndexReverse: {
      const int32_t target = arg;
      const P lend = s.level_end.get(s.lvl);
      P idx = lend - 1;
      const P lo = pos0(c);
      while (idx >= lo && byte_at(c, idx) == ' ') --idx;
      if (idx < lo && clamped(c)) { c.wide = true; break; }
      const uint8_t ch = byte_at(c, idx);
      if (ch != ']' && ch != ',' && ch != '[') { fail = true; break; }
      int32_t ci = (ch == ']') ? 0 : s.curr_index.get(s.lvl) + 1;
      for (;;) {
        if (ci < target) {
          idx = prev_structural(c, (P)(idx - 1), s.lvl);
          if (c.wide) break;
          if (byte_at(c, idx) == ',' || ci + 1 == target) { ++ci; continue; }
          fail = true;
          break;
        }
        if (ci > target) {
          idx = next_structural(c, lend, (P)(idx + 1), s.lvl);
          --ci;
          continue;
        }
        break;
      }
      if (c.wide) break;
      s.curr_index.set(s.lvl, ci);
      s.idx = idx + 1;
      if (!fail && then_down) fail = !move_down(c, s);
      break;
    }
    case kMarkPosition:
      if (s.mi + 1 >= MAXLV) { fail = true; break; }
      ++s.mi;
      s.marked_pos.set(s.mi, s.idx);
      s.marked_lvl.set(s.mi, (uint8_t)s.lvl);
      break;
    case kResetPosition:
      fail = !reset_position(s);
      break;
    case kExpressionStringEquals: {
      if (arg < 0) { fail = true; break; }
      const P lend = s.level_end.get(s.lvl);
      while (byte_at(c, s.idx) == ' ' && s.idx < lend) ++s.idx;
      P string_end = lend - 1;
      while (s.idx < string_end && byte_at(c, string_end) == ' ') --string_end;
      if (string_end - s.idx + 1 != (P)arg) { fail = true; break; }
      for (int32_t k = 0; k < arg; k += 8) {
        const int32_t rest = arg - k;
        const uint64_t m = rest >= 8 ? ~0ull : ((1ull << (8 * rest)) - 1);
        if (((text8(c, (P)(s.idx + k)) ^ ir.u64(key_pos + (uint32_t)k)) & m) != 0) {
          fail = true;
          break;
        }
      }
      if (!fail && then_reset) fail = !reset_position(s);
      break;
    }
    case kMoveToEach: {
      if (!variable || byte_at(c, s.idx) != '[' || s.nfr >= LaneState<MAXLV, P>::kF) {
        fail = true;
        break;
      }
      const P first = next_structural(c, s.level_end.get(s.lvl), (P)(s.idx + 1), s.lvl);
      P p = s.idx + 1;
      while (p < first && byte_at(c, p) == ' ') ++p;
      if (p == first && byte_at(c, first) == ']') { fail = true; break; } // empty array
      // the frame resumes after MOVE_TO_EACH and its fused MOVE_DOWN (which
      // resume() runs again for every element)
      s.fr_pack.set(s.nfr, (uint32_t)s.qp | ((uint32_t)s.lvl << 16) |
                               ((uint32_t)(s.mi + 1) << 21) | (then_down ? 1u << 31 : 0u));
      s.fr_sep.set(s.nfr, s.idx);
      ++s.nfr;
      ++s.idx;
      if (then_down) fail = !move_down(c, s);
      break;
    }
    case kJumpIfNotArray:
      if (arg < 0) { fail = true; break; }
      if (byte_at(c, s.idx) != '[') s.qp += arg;
      break;
    default: // END, unknown opcodes, running off the bytecode
      fail = true;
      break;
    }
    if (then_store && !fail && !c.wide) {
      store_result(c, s, cap, variable, res);
      fail = true; // END
    }
    if (c.wide) live = false; // the wide kernel redoes this record
    else if (fail) live = resume<MAXLV>(c, s);
  }
  return (int64_t)s.count;
}

// The record's bounds trimmed to its first '{' and last '}' (query.cu:160-172);
// false when that leaves nothing.
template <typename P> __device__ __forceinline__ bool trim_record(Ctx<P> &c, P &start, P &end) {
  while (end > start && byte_at(c, end) != '}') --end;
  while (start < end && byte_at(c, start) != '{') ++start;
  return start != end;
}

// Where a batch's marks live: the count of the first query in variable and
// dense mode, the start of its first span in fixed mode.
__device__ __forceinline__ int64_t *mark_slot(const QuerySlot &q0, int64_t line) {
  return (q0.variable || q0.dense) ? q0.counts + line : q0.spans + line * 2 * (int64_t)q0.cap;
}

// Inclusive add-scan over the wave.
__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t v) {
  const int lane = (int)(threadIdx.x & 63);
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  return v;
}

// Dense results, first pass: the lanes whose record's spans all sit in the
// wave's LDS stage move them to the staging area in one reservation (a wave
// takes room from one of its 8 regions with one atomic add) and record where
// (tmp_off); the others (more spans than the stage holds, a record left to the
// wide kernel, a full region) record -1 and go to the dense fixup pass.  Every
// valid lane writes its count (scanned into CSR offsets afterwards).
template <typename P>
__device__ __forceinline__ void flush_dense(const QuerySlot &qs, const DenseStage &ds,
                                            P (*stage)[64][2], bool valid, bool to_fixup,
                                            int64_t line, int64_t found, int64_t base) {
  const int lane = (int)(threadIdx.x & 63);
  const bool fits = valid && !to_fixup && found <= (int64_t)StageSize<P>::value;
  const uint32_t cnt = fits ? (uint32_t)found : 0u;
  const uint32_t incl = wave_incl_sum(cnt);
  const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
  int64_t at = 0;
  bool room = true;
  if (total) {
    const uint32_t region = blockIdx.x & 7u;
    uint64_t got = 0;
    if (lane == 0) got = atomicAdd(reinterpret_cast<unsigned long long *>(ds.cursors + region * kCursorStride),
                                   (unsigned long long)total);
    got = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(got >> 32)) << 32) |
          (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)got);
    room = got + total <= (uint64_t)ds.region;
    at = (int64_t)region * ds.region + (int64_t)got + (int64_t)(incl - cnt);
  }
  const bool staged = fits && (cnt == 0 || room);
  if (staged)
    for (uint32_t k = 0; k < cnt; ++k)
      *reinterpret_cast<i64x2a8 *>(ds.tmp + 2 * (at + (int64_t)k)) =
          i64x2a8{base + (int64_t)stage[k][lane][0], base + (int64_t)stage[k][lane][1]};
  if (valid) {
    qs.counts[line] = found;
    qs.tmp_off[line] = staged ? at : -1;
  }
}

} // namespace

// All queries of a batch over every record: each lane takes one record and runs
// the batch's bytecodes over it in turn, so the record's bytes and bitmap
// words are fetched once for the whole batch.  Every wave runs the same
// number of line iterations so the interpreter's wave-wide steps see all
// lanes.  One wave per block: the block's LDS holds that wave's arrays.
// Records are lines (record r spans newline[r] .. newline[r + 1], the last
// one runs to the end of the input) or, in top-level-array mode, the spans
// between consecutive leveled-row-0 characters, with interpreter level k on
// leveled row k + 1 (mode.level_base = 1, mode.drop_last = 1: the last bound
// opens no record).  Each wave takes mode.per_wave consecutive records at a
// time.  Positions as PosMode describes.  Dense queries (QuerySlot.dense) go
// through the wave's LDS stage and the staging area (flush_dense); the wide
// kernel's marked-record pass (fixup 1) counts their results and sends the
// records to the dense fixup pass (fixup 2, wide kernel), which reruns the
// records whose tmp_off is -1 and writes their spans straight to the CSR
// output (spans + 2 * offsets[line], scanned offsets).
template <int MAXLV, bool NARROW, bool DENSE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(MAXLV <= 6 ? 4 : 1))) void query_batch_kernel(
    const uint8_t *__restrict__ file, int64_t n,
    const int64_t *__restrict__ newline, const int64_t *__restrict__ num_lines,
    int64_t max_lines, const uint64_t *__restrict__ string,
    const uint64_t *__restrict__ leveled, int64_t level_size, int32_t num_levels,
    RecordMode mode, PosMode pm, QueryBatch batch) {
  typedef typename std::conditional<NARROW, int32_t, int64_t>::type P;
  const int64_t counted = *num_lines; // -1 after a failed index build: no record
  const int64_t all = counted < mode.starts_cap ? counted : mode.starts_cap;
  const int64_t avail = all - mode.drop_last;
  const int64_t lines = avail < max_lines ? avail : max_lines;
  // one wave per block, mode.per_wave records per wave
  const int64_t stride = (int64_t)gridDim.x * mode.per_wave;
  const int64_t wave_base = (int64_t)blockIdx.x * mode.per_wave;
  const int32_t lane = (int32_t)(threadIdx.x & 63);
  const int64_t span = NARROW ? pm.span : INT64_MAX;
  __shared__ WaveArrays<MAXLV, P, (DENSE && NARROW) ? kDenseStage : 1> arrays;
  // (the 64-bit passes rerun only marked records: they search on their own,
  // and their blocks, most of which have no record to run, skip the plan)
  constexpr bool kMemoEnds = !DENSE;
  __shared__ KeyMemo<P, kMemoEnds> memo;
  const MemoPlan mp = NARROW ? memo_plan(batch, kMemoEnds) : MemoPlan{0, 0, 0, 0};
  for (int64_t wb = wave_base; wb < lines; wb += stride) {
    const int64_t line = wb + lane;
    bool valid = lane < mode.per_wave && line < lines;
    uint32_t redo = 0; // dense fixup: the dense queries this record reruns
    if (!NARROW && pm.fixup == 1) {
      if (valid) valid = *mark_slot(batch.q[0], line) == kWideMark;
      if (!__ballot(valid)) continue; // (wave-uniform: lanes share wb)
    } else if (DENSE && !NARROW && pm.fixup == 2) {
      if (valid)
        for (int q = 0; q < batch.count; ++q)
          if (batch.q[q].dense && batch.q[q].tmp_off[line] < 0 &&
              batch.q[q].counts[line + 1] > batch.q[q].counts[line])
            redo |= 1u << q;
      valid = valid && redo != 0;
      if (!__ballot(valid)) continue;
    }
    int64_t ls = 0, le = 0;
    bool record = false, mark = false;
    if (valid) {
      ls = newline[line];
      le = (line + 1 < all) ? newline[line + 1] : n - 1;
      record = ls >= 0 && ls < n && le >= 0;
      if (record && le >= n) le = n - 1;
    }
    const int64_t base = (NARROW && record) ? (ls & ~63ll) : 0;
    if (NARROW && record && le - base >= span) {
      record = false;
      mark = true;
    }
    Ctx<P> c = make_ctx<P>(file, n, string, leveled, level_size, num_levels, mode.level_base,
                           base, span);
    P line_start = (P)(ls - base), line_end = (P)(le - base);
    if (record) record = trim_record(c, line_start, line_end);
    if (mp.slot)
      key_memo_scan(c, memo, record && !c.wide, line_start, line_end, batch, mp,
                       (1u << mp.keys) - 1u);
    for (int q = 0; q < batch.count; ++q) {
      const QuerySlot &qs = batch.q[q];
      const bool dense = DENSE && qs.dense; // (compile-time false without DENSE)
      const bool fix_dense = DENSE && !NARROW && pm.fixup == 2;
      if (fix_dense && !dense) continue;
      // where this query's spans go (one interpreter call site per kernel)
      bool live = record && !c.wide;
      Sink<P> out{nullptr, 0, nullptr};
      if (fix_dense) { // dense fixup: straight into the CSR output
        const bool mine = valid && ((redo >> q) & 1u);
        const int64_t off = mine ? qs.counts[line] : 0;
        const int64_t room = qs.out_cap - off;
        out.res = qs.spans + 2 * off;
        out.keep = mine ? (room < 0 ? 0 : room) : 0;
        live = record && mine;
      } else if (dense) { // narrow kernel: the wave's stage; wide: counted only
        out.stage = NARROW ? arrays.stage : nullptr;
      } else {
        out.res = qs.spans + line * 2 * (int64_t)qs.cap;
        out.keep = qs.cap;
        if (valid && !qs.variable)
          for (int i = 0; i < 2 * qs.cap; ++i) out.res[i] = -1;
      }
      const IrReader ir = IrReader::make(qs.ir, qs.ir_len);
      // a query whose first steps the memo holds starts after them (memo_start)
      P start_idx = line_start, start_end = (P)kUnset;
      const int32_t j = memo_slot(mp, q);
      const int32_t start_qp =
          j >= 0 ? memo_start(memo, mp, ir, j, live, start_idx, start_end) : 0;
      const int64_t found = run_query<MAXLV, P>(c, arrays, live, line_start, line_end, ir,
                                                qs.cap, qs.variable != 0, out, start_idx,
                                                start_qp, start_end);
      if (fix_dense) continue;
      if (dense) {
        if (NARROW)
          flush_dense<P>(qs, batch.ds, arrays.stage, valid, mark || c.wide, line, found,
                         base_of(c));
        else if (valid) {
          qs.counts[line] = found;
          qs.tmp_off[line] = -1;
        }
      } else if (valid && qs.counts) {
        qs.counts[line] = found;
      }
    }
    if (NARROW && (mark || c.wide)) *mark_slot(batch.q[0], line) = kWideMark;
  }
}

// A multi-file batch (see BatchFile): record g of the batch is line
// g - line_prefix[f] of the file f with line_prefix[f] <= g < line_prefix[f +
// 1]; each file runs the batch queries its query_mask selects, its k-th one
// writing result slot k ([total lines][cap][2] spans, [total lines] counts;
// a slot no query of the file uses gets count 0).  Per file, the results are
// those of a single-file launch over that file alone.  Positions are
// file-relative, as PosMode describes; marks go to the count of slot 0.
// (4 waves per SIMD up to 6 levels: 128 VGPRs with a 14-register spill ran
// the 10,000-file batch's queries in 5.95 ms against 6.73 ms at 3 waves.)
template <int MAXLV, bool NARROW>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(MAXLV <= 6 ? 4 : 1))) void query_files_kernel(
    const uint8_t *__restrict__ batch_in, const int64_t *__restrict__ newline,
    const uint64_t *__restrict__ string, const uint64_t *__restrict__ leveled,
    int64_t level_stride, int32_t num_levels, FilesMode fm, PosMode pm, QueryBatch batch) {
  typedef typename std::conditional<NARROW, int32_t, int64_t>::type P;
  const int64_t total = fm.line_prefix[fm.num_files];
  // one wave per block, fm.per_wave records per wave
  const int64_t stride = (int64_t)gridDim.x * fm.per_wave;
  const int64_t wave_base = (int64_t)blockIdx.x * fm.per_wave;
  const int32_t lane = (int32_t)(threadIdx.x & 63);
  const int64_t span = NARROW ? pm.span : INT64_MAX;
  __shared__ WaveArrays<MAXLV, P> arrays;
  __shared__ KeyMemo<P, true> memo; // (not in the 64-bit pass: see query_batch_kernel)
  const MemoPlan mp = NARROW ? memo_plan(batch, true) : MemoPlan{0, 0, 0, 0};
  for (int64_t wb = wave_base; wb < total; wb += stride) {
    const int64_t g = wb + lane;
    bool valid = lane < fm.per_wave && g < total;
    if (!NARROW && pm.fixup) {
      if (valid) valid = fm.counts[0][g] == kWideMark;
      if (!__ballot(valid)) continue; // (wave-uniform: lanes share wb)
    }
    // the file holding record g: last f with line_prefix[f] <= g
    int32_t lo = 0, hi = fm.num_files - 1;
    while (lo < hi) {
      const int32_t mid = (lo + hi + 1) >> 1;
      if (fm.line_prefix[mid] <= g) lo = mid;
      else hi = mid - 1;
    }
    const BatchFile bf = fm.files[lo];
    const int64_t n = (int64_t)bf.bytes;
    const int64_t line = g - fm.line_prefix[lo];
    int64_t all = fm.line_prefix[lo + 1] - fm.line_prefix[lo];
    if (all > (int64_t)bf.nl_capacity) all = (int64_t)bf.nl_capacity; // stay in the slice
    const int64_t *starts = newline + bf.nl_offset;
    int64_t ls = 0, le = 0;
    bool record = false, mark = false;
    if (valid && line < all) {
      ls = starts[line];
      le = (line + 1 < all) ? starts[line + 1] : n - 1;
      record = ls >= 0 && ls < n && le >= 0;
      if (record && le >= n) le = n - 1;
    }
    const int64_t base = (NARROW && record) ? (ls & ~63ll) : 0;
    if (NARROW && record && le - base >= span) {
      record = false;
      mark = true;
    }
    const uint64_t word_off = bf.offset >> 6;
    Ctx<P> c = make_ctx<P>(batch_in + bf.offset, n, string + word_off, leveled + word_off,
                           level_stride, num_levels, 0, base, span);
    P line_start = (P)(ls - base), line_end = (P)(le - base);
    if (record) record = trim_record(c, line_start, line_end);
    if (valid)
      for (int k = 0; k < kBatchSlots; ++k) fm.counts[k][g] = 0;
    if (mp.slot) {
      // the memo keys of the queries this record's file runs (in a result slot)
      uint32_t lane_keys = 0;
      for (int q = 0; q < batch.count; ++q) {
        const int32_t j = memo_slot(mp, q);
        if (j >= 0 && ((bf.query_mask >> q) & 1u) &&
            __popc(bf.query_mask & ((1u << q) - 1u)) < kBatchSlots)
          lane_keys |= 1u << j;
      }
      key_memo_scan(c, memo, valid && record && !c.wide, line_start, line_end, batch, mp,
                       lane_keys);
    }
    for (int q = 0; q < batch.count; ++q) {
      const QuerySlot &qs = batch.q[q];
      const bool uses = valid && ((bf.query_mask >> q) & 1u);
      const int slot = __popc(bf.query_mask & ((1u << q) - 1u));
      const bool mine = uses && slot < kBatchSlots;
      int64_t *res = (slot == 0 ? fm.spans[0] : fm.spans[1]) + g * 2 * (int64_t)fm.cap;
      if (mine && !qs.variable)
        for (int i = 0; i < 2 * fm.cap; ++i) res[i] = -1;
      const IrReader ir = IrReader::make(qs.ir, qs.ir_len);
      const int32_t cap = qs.variable ? fm.cap : qs.cap;
      const Sink<P> out{res, cap, nullptr};
      bool live = record && mine && !c.wide;
      P start_idx = line_start, start_end = (P)kUnset; // (memo: see query_batch_kernel)
      const int32_t j = memo_slot(mp, q);
      const int32_t start_qp =
          j >= 0 ? memo_start(memo, mp, ir, j, live, start_idx, start_end) : 0;
      const int64_t found = run_query<MAXLV, P>(c, arrays, live, line_start, line_end, ir, cap,
                                                qs.variable != 0, out, start_idx, start_qp,
                                                start_end);
      if (mine) (slot == 0 ? fm.counts[0] : fm.counts[1])[g] = found;
    }
    if (NARROW && (mark || c.wide)) fm.counts[0][g] = kWideMark;
  }
}

// ---- dense results: counts -> CSR offsets, staged spans -> CSR order ----

namespace {
// Records of the launch, as query_batch_kernel derives them.
__device__ __forceinline__ int64_t dense_records(const DenseSet &d) {
  const int64_t counted = *d.num_starts;
  const int64_t all = counted < d.starts_cap ? counted : d.starts_cap;
  const int64_t avail = all - d.drop_last;
  const int64_t lines = avail < d.max_lines ? avail : d.max_lines;
  return lines > 0 ? lines : 0;
}

__device__ __forceinline__ uint64_t wave_incl_sum64(uint64_t v) {
  const int lane = (int)(threadIdx.x & 63);
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  return v;
}

// Exclusive prefix over the block's 256 threads (4 waves); also the total.
__device__ __forceinline__ uint64_t block_excl_sum(uint64_t v, uint64_t &total) {
  __shared__ uint64_t s_wave[4];
  const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
  const uint64_t incl = wave_incl_sum64(v);
  if (lane == 63) s_wave[wv] = incl;
  __syncthreads();
  uint64_t before = 0;
  total = 0;
  for (uint32_t k = 0; k < 4; ++k) {
    if (k < wv) before += s_wave[k];
    total += s_wave[k];
  }
  __syncthreads();
  return before + incl - v;
}
} // namespace

// Per dense query (blockIdx.y) and chunk of kScanChunk entries: the sum of
// the record counts in the chunk (entries past the record count read 0).
// Coalesced: iteration k reads entries k * 256 + thread.
__global__ __launch_bounds__(256) void dense_reduce_kernel(DenseSet d) {
  const int q = (int)blockIdx.y;
  const int64_t lines = dense_records(d);
  const int64_t lo = (int64_t)blockIdx.x * kScanChunk + threadIdx.x;
  const int64_t *c = d.counts[q];
  uint64_t v = 0;
#pragma unroll
  for (int k = 0; k < kScanChunk / 256; ++k)
    if (lo + k * 256 < lines) v += (uint64_t)c[lo + k * 256];
  uint64_t total;
  (void)block_excl_sum(v, total);
  if (threadIdx.x == 0) d.block_sums[(int64_t)q * (d.blocks + 1) + blockIdx.x] = total;
}

// One workgroup per dense query (blockIdx.x): the chunk sums become their
// exclusive prefix, the total goes after them.
__global__ __launch_bounds__(1024) void dense_sums_kernel(DenseSet d) {
  __shared__ uint64_t s_part[1024];
  uint64_t *sums = d.block_sums + (int64_t)blockIdx.x * (d.blocks + 1);
  const uint32_t t = threadIdx.x;
  const int64_t per = (d.blocks + 1023) / 1024;
  const int64_t lo = (int64_t)t * per, hi = lo + per < d.blocks ? lo + per : d.blocks;
  uint64_t v = 0;
  for (int64_t i = lo; i < hi; ++i) v += sums[i];
  s_part[t] = v;
  __syncthreads();
  if (t == 0) {
    uint64_t run = 0;
    for (int i = 0; i < 1024; ++i) {
      const uint64_t x = s_part[i];
      s_part[i] = run;
      run += x;
    }
    sums[d.blocks] = run;
  }
  __syncthreads();
  uint64_t run = s_part[t];
  for (int64_t i = lo; i < hi; ++i) {
    const uint64_t x = sums[i];
    sums[i] = run;
    run += x;
  }
}

// Per dense query and chunk: entries [0, max_lines] of the counts become
// exclusive prefixes (the CSR offsets; entries from the record count on hold
// the total), and each record whose spans were staged has them copied to its
// CSR position (records left to the dense fixup pass are written there);
// nothing past the output's capacity.  The chunk goes through LDS so that
// global reads and writes stay coalesced while each thread scans 16
// consecutive entries.
__global__ __launch_bounds__(256) void dense_scan_kernel(DenseSet d) {
  __shared__ uint64_t s_val[kScanChunk];
  __shared__ uint64_t s_end;
  const int q = (int)blockIdx.y;
  const int64_t lines = dense_records(d);
  const int64_t base = (int64_t)blockIdx.x * kScanChunk;
  const uint32_t t = threadIdx.x;
  int64_t *c = d.counts[q];
  for (int k = 0; k < kScanChunk / 256; ++k) {
    const int64_t i = base + k * 256 + t;
    s_val[k * 256 + t] = i < lines ? (uint64_t)c[i] : 0;
  }
  __syncthreads();
  uint64_t sum = 0;
  for (int j = 0; j < 16; ++j) sum += s_val[16 * t + j];
  uint64_t total;
  uint64_t run = block_excl_sum(sum, total) + d.block_sums[(int64_t)q * (d.blocks + 1) + blockIdx.x];
  if (t == 255) s_end = run + sum; // the next chunk's first offset
  for (int j = 0; j < 16; ++j) {
    const uint64_t v = s_val[16 * t + j];
    s_val[16 * t + j] = run;
    run += v;
  }
  __syncthreads();
  const int64_t *toff = d.tmp_off[q];
  const int64_t cap = d.out_cap[q];
  int64_t *out = d.spans[q];
  for (int k = 0; k < kScanChunk / 256; ++k) {
    const int e = k * 256 + (int)t;
    const int64_t i = base + e;
    const int64_t o = (int64_t)s_val[e];
    if (i <= d.max_lines) c[i] = o;
    if (i >= lines) continue;
    const int64_t src = toff[i];
    if (src < 0) continue;
    const int64_t end = e + 1 < kScanChunk ? (int64_t)s_val[e + 1] : (int64_t)s_end;
    for (int64_t j = 0; j < end - o && o + j < cap; ++j)
      *reinterpret_cast<i64x2a8 *>(out + 2 * (o + j)) =
          *reinterpret_cast<const i64x2a8 *>(d.tmp + 2 * (src + j));
  }
}

__global__ void fill_u64_kernel(uint64_t *p, int64_t count, uint64_t v) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride)
    p[i] = v;
}

#define GPJ_QUERY_KERNELS(LV, NW)                                                              \
  template __global__ void query_files_kernel<LV, NW>(const uint8_t *, const int64_t *,        \
                                                      const uint64_t *, const uint64_t *,       \
                                                      int64_t, int32_t, FilesMode, PosMode,     \
                                                      QueryBatch);                              \
  template __global__ void query_batch_kernel<LV, NW, false>(                                  \
      const uint8_t *, int64_t, const int64_t *, const int64_t *, int64_t, const uint64_t *,    \
      const uint64_t *, int64_t, int32_t, RecordMode, PosMode, QueryBatch);                     \
  template __global__ void query_batch_kernel<LV, NW, true>(                                   \
      const uint8_t *, int64_t, const int64_t *, const int64_t *, int64_t, const uint64_t *,    \
      const uint64_t *, int64_t, int32_t, RecordMode, PosMode, QueryBatch);
GPJ_QUERY_KERNELS(4, true)
GPJ_QUERY_KERNELS(6, true)
GPJ_QUERY_KERNELS(8, true)
GPJ_QUERY_KERNELS(16, true)
GPJ_QUERY_KERNELS(4, false)
GPJ_QUERY_KERNELS(6, false)
GPJ_QUERY_KERNELS(8, false)
GPJ_QUERY_KERNELS(16, false)
#undef GPJ_QUERY_KERNELS

} // namespace gpj
