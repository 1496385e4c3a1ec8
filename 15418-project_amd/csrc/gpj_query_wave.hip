// gpj_query_wave.hip -- JSONPath bytecode interpreter, one wave per record.
//
// Same semantics and outputs as query_batch_kernel (gpj_query.hip), which
// follows the reference query kernel (src/gpjson/query/kernels/orig/
// query.cu:144-454) record by record; only the execution model differs.
//
// Execution model.  A wave takes one record at a time.  It stages a window of
// the input around the record in its LDS with coalesced loads: 4 KiB of text
// (plus 80 bytes past it), and the 64 words of every leveled row the
// queries can reach and of the string mask that cover those 4 KiB.  Every
// query of the batch then runs over the staged record.  The interpreter's
// state is wave-uniform (scalar registers; its per-level arrays are spread
// over the lanes of one vector register each), the bytecode is decoded
// through the scalar unit, and the bitmap searches are cooperative: each
// lane takes one of the window's 64 words, and a ballot picks the first (or,
// searching backwards, the last) word holding a set bit.  An access outside
// the window restages it around the position (records longer than the
// window, searches that leave it).  Positions are 32-bit file offsets: the
// host takes this walk for inputs of at most kNarrowSpan bytes only, so no
// record needs a second pass.
//
// One record costs the whole wave, so this kernel is meant for launches with
// few records (or long ones): the lane-per-record kernel keeps 64 records
// per wave and is faster once there are enough records to fill the GPU.
// Inputs must be 16-byte aligned (the host picks the other kernel otherwise).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gpj_bytecode.h"
#include "gpj_kernels.h"

namespace gpj {

namespace {

typedef uint64_t u64x2w __attribute__((ext_vector_type(2)));
typedef int32_t Pos; // a byte position (file offset), or a word index

constexpr Pos kWin = 4096;                     // staged text bytes (one word per lane)
constexpr Pos kWinPast = 80;                   // bytes staged past the window
constexpr int kTextWords = (int)((kWin + kWinPast) / 8) + 2;

template <int MAXLV> struct WaveLds {
  uint64_t text[kTextWords]; // bytes [base, base + kWin + kWinPast), then zeros
  uint64_t rows[MAXLV][64];  // interpreter level k: leveled row level_base + k
  uint64_t str[64];
};

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

__device__ __forceinline__ uint64_t through64(int b) {
  return b >= 63 ? ~0ull : ((1ull << (b + 1)) - 1);
}

__device__ __forceinline__ int64_t uni64(int64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ uint64_t lane64(uint64_t v, int l) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}

// A per-record array spread over the lanes: entry k lives in lane k.
struct LaneI32 {
  int32_t v;
  __device__ __forceinline__ int32_t get(int k) const { return __builtin_amdgcn_readlane(v, k); }
  __device__ __forceinline__ void set(int k, int32_t x) { v = lane_id() == k ? x : v; }
  // entry k := x for every k > i
  __device__ __forceinline__ void fill_above(int i, int32_t x) { v = lane_id() > i ? x : v; }
  // walking up from entry i + 1: := x until the first entry already x (at or
  // above `lim`, the array's end, nothing changes)
  __device__ __forceinline__ void fill_above_until(int i, int32_t x, int lim) {
    const uint64_t stop = __ballot(lane_id() > i && v == x);
    const int first = stop ? (int)__builtin_ctzll(stop) : lim;
    v = (lane_id() > i && lane_id() < first && lane_id() < lim) ? x : v;
  }
};

// The input and the staged window of one wave.
template <int MAXLV> struct Win {
  const uint8_t *file; // 16-byte aligned
  Pos n;
  Pos words;           // string-mask words
  const uint64_t *str;
  const uint64_t *lev;
  Pos level_size;
  int32_t num_levels;
  int32_t lbase;
  int32_t rows;        // staged rows (levels lbase .. lbase + rows - 1)
  Pos base;            // window start (a multiple of 64), -1 before the first stage
  WaveLds<MAXLV> *lds;

  // Granule at file offset off (a multiple of 16, >= 0); bytes from n on read 0.
  __device__ __forceinline__ u64x2w granule(Pos off) const {
    if (off >= n) return u64x2w{0, 0};
    u64x2w v = *reinterpret_cast<const u64x2w *>(file + off);
    const Pos valid = n - off;
    if (valid < 16) {
      if (valid <= 8) {
        v.y = 0;
        if (valid < 8) v.x &= (1ull << (8 * valid)) - 1;
      } else {
        v.y &= (1ull << (8 * (valid - 8))) - 1;
      }
    }
    return v;
  }

  __device__ __forceinline__ void stage(Pos b) {
    base = b;
    const int l = lane_id();
    const Pos off = b + 64 * l;
    u64x2w g[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) g[k] = granule(off + 16 * k);
    u64x2w e = u64x2w{0, 0};
    if (l < 5) e = granule(b + kWin + 16 * l);
    const Pos wi = (b >> 6) + l;
    uint64_t rv[MAXLV];
#pragma unroll
    for (int k = 0; k < MAXLV; ++k)
      rv[k] = (k < rows && wi < level_size) ? lev[(int64_t)(lbase + k) * level_size + wi] : 0ull;
    const uint64_t sv = wi < words ? str[wi] : 0ull;
    // (every lane's loads are done before it writes: the previous window's
    // readers in this wave have finished, since LDS operations of one wave
    // complete in order)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      lds->text[8 * l + 2 * k] = g[k].x;
      lds->text[8 * l + 2 * k + 1] = g[k].y;
    }
    if (l < 5) {
      lds->text[512 + 2 * l] = e.x;
      lds->text[512 + 2 * l + 1] = e.y;
    } else if (l < 7) {
      lds->text[512 + l + 5] = 0; // words 522, 523: past the staged bytes
    }
#pragma unroll
    for (int k = 0; k < MAXLV; ++k) lds->rows[k][l] = rv[k];
    lds->str[l] = sv;
  }

  // Text bytes [p, p + len) (len <= 16) inside the window; a forward walk
  // keeps 256 bytes behind p, a backward one 256 bytes past it.
  __device__ __forceinline__ void need_text(Pos p, Pos len, bool backward) {
    if (p >= base && p + len <= base + kWin + kWinPast - 8) return;
    Pos nb = backward ? ((p + len + 63) & ~63) + 256 - kWin : (p & ~63) - 256;
    if (nb < 0) nb = 0;
    stage(nb);
  }
  // Bitmap word wi inside the window.
  __device__ __forceinline__ void need_word(Pos wi, bool backward) {
    const Pos wb = base >> 6;
    if (base >= 0 && wi >= wb && wi < wb + 64) return;
    Pos nb = backward ? ((wi + 5) << 6) - kWin : (wi - 4) << 6;
    if (nb < 0) nb = 0;
    stage(nb);
  }

  // Byte at p; bytes outside [0, n) read as 0.
  __device__ __forceinline__ uint32_t byte_at(Pos p, bool backward = false) {
    if (p < 0 || p >= n) return 0;
    need_text(p, 1, backward);
    const uint8_t *t = reinterpret_cast<const uint8_t *>(lds->text);
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)t[p - base]);
  }
  // Little-endian 8 bytes from p (bytes outside [0, n) read as 0).
  __device__ __forceinline__ uint64_t text8(Pos p) {
    if (p >= n || p <= -8) return 0;
    const Pos q = p < 0 ? 0 : p;
    need_text(q, 8, false);
    const Pos a = q - base;
    const uint64_t lo = uni64((int64_t)lds->text[a >> 3]);
    const uint32_t sh = 8 * (uint32_t)(a & 7);
    uint64_t v = lo;
    if (sh) v = (lo >> sh) | ((uint64_t)uni64((int64_t)lds->text[(a >> 3) + 1]) << (64 - sh));
    if (p < 0) v <<= 8 * (uint32_t)(-p);
    return v;
  }

  // First set bit of staged row `which` (-1: the string mask) in [idx,
  // stop], or -1 (idx >= 0); the source has `limit` words.
  __device__ __forceinline__ Pos next_bit(int which, Pos idx, Pos stop, Pos limit) {
    for (;;) {
      need_word(idx >> 6, false);
      const Pos wb = base >> 6;
      const Pos rel = (idx >> 6) - wb, srel = (stop >> 6) - wb;
      const int l = lane_id();
      uint64_t m = which < 0 ? lds->str[l] : lds->rows[which][l];
      if (l < rel) m = 0;
      else if (l == rel) m &= ~0ull << (idx & 63);
      if (l > srel) m = 0;
      else if (l == srel) m &= through64((int)(stop & 63));
      const uint64_t b = __ballot(m != 0);
      if (b) {
        const int f = (int)__builtin_ctzll(b);
        return (wb + f) * 64 + (Pos)__builtin_ctzll(lane64(m, f));
      }
      if (srel <= 63 || wb + 64 >= limit) return -1;
      idx = (wb + 64) * 64;
    }
  }
  // Last set bit of `which` in [lo, idx], or -1 (lo >= 0).
  __device__ __forceinline__ Pos last_bit(int which, Pos lo, Pos idx) {
    if (idx < lo || idx < 0) return -1;
    for (;;) {
      need_word(idx >> 6, true);
      const Pos wb = base >> 6;
      const Pos rel = (idx >> 6) - wb, lrel = (lo >> 6) - wb;
      const int l = lane_id();
      uint64_t m = which < 0 ? lds->str[l] : lds->rows[which][l];
      if (l > rel) m = 0;
      else if (l == rel) m &= through64((int)(idx & 63));
      if (l < lrel) m = 0;
      else if (l == lrel) m &= ~0ull << (lo & 63);
      const uint64_t b = __ballot(m != 0);
      if (b) {
        const int f = 63 - (int)__builtin_clzll(b);
        return (wb + f) * 64 + 63 - (Pos)__builtin_clzll(lane64(m, f));
      }
      if (lrel >= 0 || wb == 0) return -1;
      idx = wb * 64 - 1;
    }
  }

  __device__ __forceinline__ bool row_built(int level) const {
    return level >= 0 && level + lbase < num_levels;
  }
  // First structural bit of the level's row in [idx, level_end], else
  // level_end (find_next_structural_char, query.cu:79-115; idx >= 0).
  __device__ __forceinline__ Pos next_structural(Pos level_end, Pos idx, int level) {
    if (idx >= level_end || !row_built(level)) return level_end;
    if (idx < 0) idx = 0;
    const Pos s = next_bit(level, idx, level_end, level_size);
    return s < 0 ? level_end : s;
  }
  // Last structural bit of the level's row in [0, idx], else 0
  // (find_previous_structural_char, query.cu:117-128, level start 0).
  __device__ __forceinline__ Pos prev_structural(Pos idx, int level) {
    if (idx <= 0 || !row_built(level)) return 0;
    const Pos s = last_bit(level, 0, idx);
    return s < 0 ? 0 : s;
  }
};

// A key as the 16 bytes ending at a compact key's ':' show it: opening
// quote, key, closing quote, ':' (keys of at most 13 bytes that do not end in
// a backslash; `fast` false otherwise).  lo holds the first 8 of the 16 bytes.
struct KeyPattern {
  uint64_t plo, phi, mlo, mhi;
  int32_t klen;
  bool fast;
  __device__ static __forceinline__ KeyPattern make(const IrReader &ir, uint32_t key_pos,
                                                    int32_t klen) {
    KeyPattern k;
    k.klen = klen;
    const uint64_t key0 = ir.u64(key_pos);
    const uint64_t key1 = klen > 8 ? ir.u64(key_pos + 8) : 0ull;
    const uint32_t last = klen == 0 ? 0u
                          : (uint32_t)(((klen > 8 ? key1 : key0) >> (8 * ((klen - 1) & 7))) & 0xFFu);
    k.fast = klen <= 13 && last != '\\';
    typedef unsigned __int128 u128;
    const int32_t kl = k.fast ? klen : 0;
    const u128 key = kl == 0 ? (u128)0
                     : (((u128)key1 << 64) | key0) & (((u128)1 << (8 * kl)) - 1);
    const u128 e = ((u128)'"' << (8 * (13 - kl))) | (key << (8 * (14 - kl))) |
                   ((u128)'"' << 112) | ((u128)':' << 120);
    const u128 m = ~(u128)0 << (8 * (13 - kl));
    k.plo = (uint64_t)e;
    k.phi = (uint64_t)(e >> 64);
    k.mlo = (uint64_t)m;
    k.mhi = (uint64_t)(m >> 64);
    return k;
  }
};

// Whether the key at the ':' at pos matches (the sequential test of the
// reference's search_key: the string mask's last set bit before pos ends the
// key's string, whose opening quote must lie at or after line_start).
template <int MAXLV>
__device__ __forceinline__ bool key_matches(Win<MAXLV> &w, const IrReader &ir, uint32_t key_pos,
                                            int32_t klen, const KeyPattern &kp, Pos pos,
                                            Pos line_start) {
  const uint64_t t1 = w.text8(pos - 7);
  const bool compact = pos - 2 >= line_start + 1 && ((t1 >> 48) & 0xFFu) == '"' &&
                       ((t1 >> 40) & 0xFFu) != '\\';
  if (compact && kp.fast)
    return pos - 2 - klen >= line_start && ((t1 ^ kp.phi) & kp.mhi) == 0 &&
           ((w.text8(pos - 15) ^ kp.plo) & kp.mlo) == 0;
  const Pos se = compact ? pos - 2 : w.last_bit(-1, line_start + 1, pos - 1);
  const Pos ss = se - klen;
  bool match = se >= 0 && ss >= line_start && w.byte_at(ss, true) == '"';
  for (int32_t k = 0; match && k < klen; k += 8) {
    const int32_t rest = klen - k;
    const uint64_t m = rest >= 8 ? ~0ull : ((1ull << (8 * rest)) - 1);
    match = ((w.text8(ss + 1 + k) ^ ir.u64(key_pos + (uint32_t)k)) & m) == 0;
  }
  return match;
}

__device__ __forceinline__ uint32_t wave_excl_sum(uint32_t v) {
  const int l = lane_id();
  uint32_t x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(x, d, 64);
    if (l >= d) x += o;
  }
  return x - v;
}

enum { kScanFound = 1, kScanFail = 2, kScanUnknown = 3, kScanExhausted = 4 };
struct KeyScan {
  int kind;     // kScan*
  Pos pos;      // the deciding character, or the last one scanned
  int phase;    // role of the character after pos (0: key position)
  Pos cand;     // the last key position scanned
};

// The key search over the row characters of level `lvl` after `pos` up to
// lend, whole windows at a time: each lane takes one word and runs its
// characters in order (rank parity from the wave's prefix count gives each
// its role); the first lane with a deciding character decides.  A key
// position decides when its byte is not ':' (found), when its compact key
// matches (found) or when the 16 bytes cannot decide it (unknown); a
// separator position decides when it is '}' (fail).
template <int MAXLV>
__device__ __forceinline__ KeyScan coop_key_scan(Win<MAXLV> &w, int lvl, Pos pos, int phase,
                                                 Pos cand, Pos lend, Pos line_start,
                                                 const KeyPattern &kp) {
  const int l = lane_id();
  for (;;) {
    const Pos start = pos + 1;
    if (start > lend) return KeyScan{kScanExhausted, pos, phase, cand};
    w.need_word(start >> 6, false);
    const Pos wb = w.base >> 6;
    const Pos rel = (start >> 6) - wb, srel = (lend >> 6) - wb;
    uint64_t m = w.lds->rows[lvl][l];
    if (l < rel) m = 0;
    else if (l == rel) m &= ~0ull << (start & 63);
    if (l > srel) m = 0;
    else if (l == srel) m &= through64((int)(lend & 63));
    const uint32_t cnt = (uint32_t)__popcll(m);
    const uint32_t pre = wave_excl_sum(cnt);
    const uint8_t *text = reinterpret_cast<const uint8_t *>(w.lds->text);
    int ev = 0;
    Pos evpos = 0;
    uint32_t r = (uint32_t)phase + pre;
    uint64_t mm = m;
    while (mm) {
      const int b = (int)__builtin_ctzll(mm);
      mm &= mm - 1;
      const Pos p = (wb + l) * 64 + b;
      const Pos a = p - w.base; // 0 .. 4095
      if ((r & 1u) == 0) {
        if (a < 15) { ev = kScanUnknown; evpos = p; break; }
        const uint64_t *tw = w.lds->text;
        const Pos a0 = a - 15, q = a0 >> 3;
        const uint32_t sh = 8 * (uint32_t)(a0 & 7);
        const uint64_t w0 = tw[q], w1 = tw[q + 1], w2 = tw[q + 2];
        const uint64_t t0 = sh ? (w0 >> sh) | (w1 << (64 - sh)) : w0;
        const uint64_t t1 = sh ? (w1 >> sh) | (w2 << (64 - sh)) : w1;
        if ((t1 >> 56) != ':') { ev = kScanFound; evpos = p; break; }
        const bool compact = p - 2 >= line_start + 1 && ((t1 >> 48) & 0xFFu) == '"' &&
                             ((t1 >> 40) & 0xFFu) != '\\';
        if (!(compact && kp.fast)) { ev = kScanUnknown; evpos = p; break; }
        if (p - 2 - kp.klen >= line_start && ((t0 ^ kp.plo) & kp.mlo) == 0 &&
            ((t1 ^ kp.phi) & kp.mhi) == 0) {
          ev = kScanFound;
          evpos = p;
          break;
        }
      } else if (text[a] == '}') {
        ev = kScanFail;
        evpos = p;
        break;
      }
      ++r;
    }
    const uint64_t evb = __ballot(ev != 0);
    if (evb) {
      const int f = (int)__builtin_ctzll(evb);
      const int kind = __builtin_amdgcn_readlane(ev, f);
      const Pos p = __builtin_amdgcn_readlane(evpos, f);
      // (a deciding key position leaves phase 0 and its own position as the
      // last key position; only kScanUnknown continues, from there)
      return KeyScan{kind, p, 0, p};
    }
    const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)(pre + cnt), 63);
    if (total) {
      const uint64_t nz = __ballot(m != 0);
      const int f = 63 - (int)__builtin_clzll(nz);
      const Pos last = (wb + f) * 64 + 63 - (Pos)__builtin_clzll(lane64(m, f));
      if ((((uint32_t)phase + total - 1) & 1u) == 0) cand = last; // last was a key position
      phase = (int)(((uint32_t)phase + total) & 1u);
    }
    // nothing decided: on from the window's end (or the tail past lend)
    if (srel <= 63) {
      return KeyScan{kScanExhausted, lend, phase, cand};
    }
    pos = (wb + 64) * 64 - 1;
  }
}

// Where a query's spans go: entries below `keep` of `res` (int64 pairs), or,
// in dense mode, the lanes of `stage_s` / `stage_e` (the first 64 spans).
struct WaveSink {
  int64_t *res;
  int32_t keep;
  bool dense;
};

// One query over one record (query_line of the oracle, the lane kernel's
// run_query): levels and marks below maxlv, element iterations below
// min(maxlv, kMaxFrames).  Returns the number of results (stored or counted).
template <int MAXLV>
__device__ __forceinline__ int32_t wave_query(Win<MAXLV> &w, Pos line_start, Pos line_end,
                              const IrReader ir, int32_t cap, bool variable, int32_t maxlv,
                              const WaveSink &out, LaneI32 &stage_s, LaneI32 &stage_e) {
  const int kf = maxlv < kMaxFrames ? maxlv : kMaxFrames;
  Pos idx = line_start;
  int32_t lvl = 0, mi = -1, nfr = 0;
  uint32_t qp = 0;
  int32_t count = 0;
  LaneI32 level_end{-1};
  level_end.set(0, line_end);
  LaneI32 marked_pos{0};
  LaneI32 marked_lvl{0};
  LaneI32 curr_index{-1};
  LaneI32 fr_lvl{0}, fr_mi{0};
  LaneI32 fr_qp{0};
  LaneI32 fr_sep{0};

  for (;;) {
    bool fail = false;
    const uint32_t op = qp >= ir.len ? (uint32_t)kOffEnd : ir.byte(qp++);
    switch (op) {
    case kStoreResult: {
      if (!variable && count >= cap) { fail = true; break; }
      const Pos lend = level_end.get(lvl);
      Pos end_str = lend;
      while (idx < lend && w.byte_at(idx) == ' ') ++idx;
      while (end_str > idx && w.byte_at(end_str - 1, true) == ' ') --end_str;
      if (out.dense) {
        if (count < 64) {
          stage_s.set(count, idx);
          stage_e.set(count, end_str);
        }
      } else if (count < out.keep && lane_id() == 0) {
        out.res[2 * count] = idx;
        out.res[2 * count + 1] = end_str;
      }
      ++count;
      break;
    }
    case kMoveUp:
      if (lvl <= 0) { fail = true; break; }
      idx = level_end.get(lvl);
      level_end.set(lvl, -1);
      --lvl;
      break;
    case kMoveDown:
      if (lvl + 1 >= maxlv) { fail = true; break; }
      ++lvl;
      if (level_end.get(lvl) == -1) {
        level_end.set(lvl, w.next_structural(level_end.get(lvl - 1), idx + 1, lvl - 1));
        while (w.byte_at(idx) == ' ') ++idx;
      }
      break;
    case kMoveToKey: {
      const int32_t klen = ir.varint(qp);
      const uint32_t key_pos = qp;
      qp += (uint32_t)klen;
      if (klen < 0 || w.byte_at(idx) != '{') { fail = true; break; }
      const Pos lend = level_end.get(lvl);
      const KeyPattern kp = KeyPattern::make(ir, key_pos, klen);
      // the reference's search_key loop (query.cu:254-302) visits the row's
      // structural characters after the '{' in order, alternating a key
      // position (':' expected) and a separator position, and after the last
      // one the level end forever.  The cooperative scan decides the real
      // characters window by window; a key it cannot decide from the 16 bytes
      // ending at its ':', and the level-end tail, go through the sequential
      // steps below.
      Pos pos = idx, cand = -1;
      int phase = 0; // 0: the next character is a key position, 1: a separator
      int state = 0; // 0: searching, 1: found at pos, 2: failed
      while (state == 0) {
        const KeyScan ks = w.row_built(lvl) ? coop_key_scan(w, lvl, pos, phase, cand, lend,
                                                             line_start, kp)
                                            : KeyScan{kScanExhausted, pos, phase, cand};
        pos = ks.pos;
        phase = ks.phase;
        cand = ks.cand;
        if (ks.kind == kScanFound) { state = 1; break; }
        if (ks.kind == kScanFail) { state = 2; break; }
        if (ks.kind == kScanUnknown) { // a key position the scan could not decide
          if ((w.text8(pos - 7) >> 56) != ':') { state = 1; break; }
          if (key_matches(w, ir, key_pos, klen, kp, pos, line_start)) { state = 1; break; }
          phase = 1;
          cand = pos;
          continue;
        }
        // past the row's last character: sequential steps (each reaches the
        // level end at once, so this ends within a few)
        for (;;) {
          pos = w.next_structural(lend, pos + 1, lvl);
          if (phase == 0) {
            if ((w.text8(pos - 7) >> 56) != ':') { state = 1; break; }
            cand = pos;
            if (key_matches(w, ir, key_pos, klen, kp, pos, line_start)) { state = 1; break; }
            phase = 1;
          } else {
            if (w.byte_at(pos) == '}' || pos == cand) { state = 2; break; }
            phase = 0;
          }
        }
      }
      if (state == 2) fail = true;
      idx = pos + 1;
      break;
    }
    case kMoveToIndex: {
      const int32_t target = ir.varint(qp);
      const uint32_t ch = w.byte_at(idx);
      if (ch != '[' && ch != ',' && ch != ']') { fail = true; break; }
      const Pos lend = level_end.get(lvl);
      int32_t ci = ch == '[' ? 0 : curr_index.get(lvl) + 1;
      for (;;) {
        if (ci < target) {
          idx = w.next_structural(lend, idx + 1, lvl);
          if (w.byte_at(idx) == ',') { ++ci; continue; }
          fail = true;
          break;
        }
        if (ci > target) {
          idx = w.prev_structural(idx - 1, lvl);
          --ci;
          continue;
        }
        break;
      }
      curr_index.set(lvl, ci);
      if (!fail && w.byte_at(idx) == ']') fail = true;
      ++idx;
      break;
    }
    case kMoveToIndexReverse: {
      const int32_t target = ir.varint(qp);
      const Pos lend = level_end.get(lvl);
      idx = lend - 1;
      while (idx >= 0 && w.byte_at(idx, true) == ' ') --idx;
      const uint32_t ch = w.byte_at(idx);
      if (ch != ']' && ch != ',' && ch != '[') { fail = true; break; }
      int32_t ci = ch == ']' ? 0 : curr_index.get(lvl) + 1;
      for (;;) {
        if (ci < target) {
          idx = w.prev_structural(idx - 1, lvl);
          if (w.byte_at(idx) == ',' || ci + 1 == target) { ++ci; continue; }
          fail = true;
          break;
        }
        if (ci > target) {
          idx = w.next_structural(lend, idx + 1, lvl);
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
      if (mi + 1 >= maxlv) { fail = true; break; }
      ++mi;
      marked_pos.set(mi, idx);
      marked_lvl.set(mi, lvl);
      break;
    case kResetPosition:
      if (mi < 0) { fail = true; break; }
      idx = marked_pos.get(mi);
      lvl = marked_lvl.get(mi);
      --mi;
      level_end.fill_above_until(lvl, -1, maxlv);
      break;
    case kExpressionStringEquals: {
      const int32_t klen = ir.varint(qp);
      const uint32_t key_pos = qp;
      qp += (uint32_t)klen;
      if (klen < 0) { fail = true; break; }
      const Pos lend = level_end.get(lvl);
      while (idx < lend && w.byte_at(idx) == ' ') ++idx;
      Pos string_end = lend - 1;
      while (idx < string_end && w.byte_at(string_end, true) == ' ') --string_end;
      if (string_end - idx + 1 != klen) { fail = true; break; }
      for (int32_t k = 0; k < klen; k += 8) {
        const int32_t rest = klen - k;
        const uint64_t m = rest >= 8 ? ~0ull : ((1ull << (8 * rest)) - 1);
        if (((w.text8(idx + k) ^ ir.u64(key_pos + (uint32_t)k)) & m) != 0) {
          fail = true;
          break;
        }
      }
      break;
    }
    case kMoveToEach: {
      if (!variable || nfr >= kf || w.byte_at(idx) != '[') { fail = true; break; }
      const Pos first = w.next_structural(level_end.get(lvl), idx + 1, lvl);
      Pos p = idx + 1;
      while (p < first && w.byte_at(p) == ' ') ++p;
      if (p == first && w.byte_at(first) == ']') { fail = true; break; } // empty array
      fr_lvl.set(nfr, lvl);
      fr_sep.set(nfr, idx);
      fr_qp.set(nfr, (int32_t)qp);
      fr_mi.set(nfr, mi);
      ++nfr;
      ++idx;
      break;
    }
    case kJumpIfNotArray: {
      const int32_t off = ir.varint(qp);
      if (off < 0) { fail = true; break; }
      if (w.byte_at(idx) != '[') qp += (uint32_t)off;
      break;
    }
    default: // END, unknown opcodes, running off the bytecode
      fail = true;
      break;
    }
    if (!fail) continue;
    // the path is finished: resume the innermost element iteration
    bool resumed = false;
    while (nfr > 0) {
      const int f = nfr - 1;
      const int flvl = fr_lvl.get(f);
      const Pos lend = level_end.get(flvl);
      const Pos sep = w.next_structural(lend, fr_sep.get(f) + 1, flvl);
      if (sep < lend && w.byte_at(sep) == ',') {
        fr_sep.set(f, sep);
        lvl = flvl;
        level_end.fill_above(lvl, -1);
        mi = fr_mi.get(f);
        idx = sep + 1;
        qp = (uint32_t)fr_qp.get(f);
        resumed = true;
        break;
      }
      --nfr;
    }
    if (!resumed) return count;
  }
}

// The record's bounds trimmed to its first '{' and last '}' (query.cu:160-172).
template <int MAXLV>
__device__ __forceinline__ bool wave_trim(Win<MAXLV> &w, Pos &start, Pos &end) {
  while (end > start && w.byte_at(end, true) != '}') --end;
  while (start < end && w.byte_at(start) != '{') ++start;
  return start != end;
}

} // namespace

// Records and outputs as query_batch_kernel with NARROW = false and fixup 0
// (dense queries: the record's spans, up to 64, are staged in the staging
// area with one reservation; a record with more goes to the dense fixup
// pass).  One wave per block; records blockIdx.x, blockIdx.x + gridDim.x, ...
template <int MAXLV>
__global__ __launch_bounds__(64) void query_wave_kernel(
    const uint8_t *__restrict__ file, int64_t n, const int64_t *__restrict__ newline,
    const int64_t *__restrict__ num_lines, int64_t max_lines,
    const uint64_t *__restrict__ string, const uint64_t *__restrict__ leveled,
    int64_t level_size, int32_t num_levels, RecordMode mode, QueryBatch batch) {
  __shared__ WaveLds<MAXLV> lds;
  const int64_t counted = *num_lines; // -1 after a failed index build: no record
  const int64_t all = counted < mode.starts_cap ? counted : mode.starts_cap;
  const int64_t avail = all - mode.drop_last;
  const int64_t lines = avail < max_lines ? avail : max_lines;
  const int lane = lane_id();
  Win<MAXLV> w;
  w.file = file;
  w.n = (Pos)n; // (at most kNarrowSpan: the host checks)
  w.words = (Pos)((n + 63) >> 6);
  w.str = string;
  w.lev = leveled;
  w.level_size = level_size < (int64_t)INT32_MAX ? (Pos)level_size : INT32_MAX;
  w.num_levels = num_levels;
  w.lbase = mode.level_base;
  {
    const int32_t r = num_levels - mode.level_base;
    w.rows = r < 0 ? 0 : (r > MAXLV ? MAXLV : r);
  }
  w.base = -1;
  w.lds = &lds;
  for (int64_t line = blockIdx.x; line < lines; line += gridDim.x) {
    const int64_t ls64 = newline[line];
    int64_t le64 = (line + 1 < all) ? newline[line + 1] : n - 1;
    bool record = ls64 >= 0 && ls64 < n && le64 >= 0;
    if (record && le64 >= n) le64 = n - 1;
    Pos ls = record ? (Pos)ls64 : 0, le = record ? (Pos)le64 : 0;
    if (record) {
      const Pos b = (ls & ~63) - 64;
      w.stage(b < 0 ? 0 : b);
      record = wave_trim(w, ls, le);
    }
    for (int q = 0; q < batch.count; ++q) {
      const QuerySlot &qs = batch.q[q];
      const bool dense = qs.dense != 0;
      WaveSink out{nullptr, 0, dense};
      if (!dense) {
        out.res = qs.spans + line * 2 * (int64_t)qs.cap;
        out.keep = qs.cap;
        // (by the lane that stores the spans, so the fill lands first)
        if (!qs.variable && lane == 0)
          for (int i = 0; i < 2 * qs.cap; ++i) out.res[i] = -1;
      }
      LaneI32 stage_s{0}, stage_e{0};
      int64_t found = 0;
      if (record)
        found = wave_query<MAXLV>(w, ls, le, IrReader::make(qs.ir, qs.ir_len), qs.cap,
                                  qs.variable != 0, MAXLV, out, stage_s, stage_e);
      if (dense) {
        int64_t at = 0;
        bool staged = found == 0;
        if (found > 0 && found <= 64) {
          const uint32_t region = blockIdx.x & 7u;
          uint64_t got = 0;
          if (lane == 0)
            got = atomicAdd(reinterpret_cast<unsigned long long *>(batch.ds.cursors +
                                                                   region * kCursorStride),
                            (unsigned long long)found);
          got = (uint64_t)uni64((int64_t)lane64(got, 0));
          if (got + (uint64_t)found <= (uint64_t)batch.ds.region) {
            at = (int64_t)region * batch.ds.region + (int64_t)got;
            staged = true;
            if (lane < found) {
              batch.ds.tmp[2 * (at + lane)] = (int64_t)stage_s.v;
              batch.ds.tmp[2 * (at + lane) + 1] = (int64_t)stage_e.v;
            }
          }
        }
        if (lane == 0) {
          qs.counts[line] = found;
          qs.tmp_off[line] = staged ? at : -1;
        }
      } else if (qs.counts && lane == 0) {
        qs.counts[line] = found;
      }
    }
  }
}

template __global__ void query_wave_kernel<4>(const uint8_t *, int64_t, const int64_t *,
                                              const int64_t *, int64_t, const uint64_t *,
                                              const uint64_t *, int64_t, int32_t, RecordMode,
                                              QueryBatch);
template __global__ void query_wave_kernel<6>(const uint8_t *, int64_t, const int64_t *,
                                              const int64_t *, int64_t, const uint64_t *,
                                              const uint64_t *, int64_t, int32_t, RecordMode,
                                              QueryBatch);
template __global__ void query_wave_kernel<8>(const uint8_t *, int64_t, const int64_t *,
                                              const int64_t *, int64_t, const uint64_t *,
                                              const uint64_t *, int64_t, int32_t, RecordMode,
                                              QueryBatch);
template __global__ void query_wave_kernel<16>(const uint8_t *, int64_t, const int64_t *,
                                               const int64_t *, int64_t, const uint64_t *,
                                               const uint64_t *, int64_t, int32_t, RecordMode,
                                               QueryBatch);

} // namespace gpj
