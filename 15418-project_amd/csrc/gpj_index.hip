// gpj_index.hip -- structural-index kernels for gfx950 (MI355X).
//
// One pass over the input builds the three indices the reference builds with
// thirteen launches (src/gpjson/index/orig_index_builder.cu:296-395):
//   * string mask   uint64[W]      (string-index.cu:3-31 semantics)
//   * leveled rows  uint64[L][W]   (leveled-bitmaps-index.cu:9-73 semantics)
//   * newline index int64[1+#nl]   (newline-index.cu:5-21 semantics)
// W = ceil(n/64).  Work unit: one 64-byte "word" per lane, 256 words
// (16 KiB) per workgroup tile.  Each lane classifies its 64 bytes with
// byte-parallel SWAR compares and packs the per-class flags into 64-bit masks
// with v_dot4_u32_u8.  The three carried states (escape parity, in-string
// parity, nesting depth) plus the newline count cross tiles as follows: each
// tile publishes, before it knows its carry-in, a 4-entry transfer table
// indexed by (escape-in, string-in).  One extra workgroup scans those tables
// in tile order, 1,024 tiles per round, and writes the inclusive state of
// every 8th tile; a tile waits for the one before its group of 8.  (One
// atomic tile counter, or a look-back per tile, each cost about 11 ns per
// tile in series on MI355X: the chain, not the work, bounded the kernel.)  In
// a multi-file batch each file's first tile publishes its inclusive state at
// once (files start from the initial state) and the scan restarts there.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gpj_kernels.h"

namespace gpj {

namespace {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr uint64_t kEven = 0x5555555555555555ull; // bit positions 0,2,4,...
constexpr uint64_t kOdd = 0xAAAAAAAAAAAAAAAAull;  // bit positions 1,3,5,...

__device__ __forceinline__ uint64_t low_mask(int bits) {
  return bits >= 64 ? ~0ull : ((1ull << bits) - 1ull);
}
// Bits 0..b inclusive (b may be 64 = all).
__device__ __forceinline__ uint64_t mask_upto(int b) {
  return b >= 63 ? ~0ull : ((2ull << b) - 1ull);
}

__device__ __forceinline__ uint64_t prefix_xor64(uint64_t x) {
  x ^= x << 1;
  x ^= x << 2;
  x ^= x << 4;
  x ^= x << 8;
  x ^= x << 16;
  x ^= x << 32;
  return x;
}

// ---- byte classification ---------------------------------------------------
// For a 7-bit byte b and ASCII c: ((b ^ (0x7F ^ c)) + 1) has bit 7 set iff
// b == c, with no carry out of the byte.  `nh` keeps only bit 7 of bytes
// whose original value is < 0x80.  v_xad_u32 computes (a ^ b) + c in one
// instruction (the compiler splits it into xor + add when an operand is a
// literal); the class constant sits in a scalar register.
__device__ __forceinline__ uint32_t xad(uint32_t a, uint32_t k, uint32_t c) {
  uint32_t r;
  asm("v_xad_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(k), "v"(c));
  return r;
}
__device__ __forceinline__ uint32_t eq_bytes(uint32_t lo7, uint32_t c) {
  return (lo7 ^ (0x7F7F7F7Fu ^ (c * 0x01010101u))) + 0x01010101u;
}

// Two dwords (8 bytes) of 0x80/0x00 flags -> mask byte, returned << 7.
__device__ __forceinline__ uint32_t pack2(uint32_t f0, uint32_t f1) {
  return __builtin_amdgcn_udot4(
      f1, 0x80402010u, __builtin_amdgcn_udot4(f0, 0x08040201u, 0u, false),
      false);
}
// Eight dwords (32 bytes) of flags -> 32-bit mask in byte order.
__device__ __forceinline__ uint32_t pack8(const uint32_t *f) {
  return (pack2(f[0], f[1]) >> 7) | (pack2(f[2], f[3]) << 1) |
         (pack2(f[4], f[5]) << 9) | (pack2(f[6], f[7]) << 17);
}
__device__ __forceinline__ uint64_t pack16(const uint32_t *f) {
  return (uint64_t)pack8(f) | ((uint64_t)pack8(f + 8) << 32);
}

struct WordClasses {
  uint64_t quote, backslash, newline, open, close, sep;
};

__device__ __forceinline__ void load_word(const uint8_t *in, uint64_t n,
                                          uint64_t pos, bool aligned,
                                          uint32_t (&x)[16]) {
  if (pos + 64 <= n && aligned) {
    const u32x4 *p = reinterpret_cast<const u32x4 *>(in + pos);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const u32x4 v = __builtin_nontemporal_load(p + i);
      x[4 * i + 0] = v.x;
      x[4 * i + 1] = v.y;
      x[4 * i + 2] = v.z;
      x[4 * i + 3] = v.w;
    }
  } else {
    // the input's tail, or an input not 16-byte aligned: 4-byte aligned
    // dwords (each holding at least one input byte, so none crosses past the
    // input's last page) shifted into place, bytes from n on cleared
    const uintptr_t a = reinterpret_cast<uintptr_t>(in + pos);
    const uint32_t k = (uint32_t)(a & 3);
    const uint32_t *d = reinterpret_cast<const uint32_t *>(a - k);
    const int64_t first = (int64_t)pos - (int64_t)k; // file offset of d[0]'s byte 0
    uint32_t w[17];
#pragma unroll
    for (int j = 0; j < 17; ++j) w[j] = first + 4 * j < (int64_t)n ? d[j] : 0u;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const uint32_t v = __builtin_amdgcn_alignbyte(w[i + 1], w[i], k);
      const int64_t valid = (int64_t)n - (int64_t)(pos + 4 * i);
      x[i] = valid >= 4 ? v : (valid <= 0 ? 0u : v & ((1u << (8 * valid)) - 1u));
    }
  }
}

// ASCII: every byte of the word is below 0x80, so the compares need neither
// the high-bit mask on their inputs nor the per-class clear of bytes >= 0x80
// (the classify() of a wave whose words are all ASCII; bytes past the input
// are 0).
template <bool ASCII>
__device__ __forceinline__ WordClasses classify(const uint32_t (&x)[16]) {
  // class constants 0x7F ^ c in every byte ('[' and ']' fold onto '{' and '}'
  // once bit 5 is set)
  const uint32_t kq = 0x7F7F7F7Fu ^ ('"' * 0x01010101u);
  const uint32_t kb = 0x7F7F7F7Fu ^ ('\\' * 0x01010101u);
  const uint32_t kn = 0x7F7F7F7Fu ^ ('\n' * 0x01010101u);
  const uint32_t ko = 0x7F7F7F7Fu ^ ('{' * 0x01010101u);
  const uint32_t kc = 0x7F7F7F7Fu ^ ('}' * 0x01010101u);
  const uint32_t kl = 0x7F7F7F7Fu ^ (':' * 0x01010101u);
  const uint32_t km = 0x7F7F7F7Fu ^ (',' * 0x01010101u);
  const uint32_t ones = 0x01010101u;
  // per class: bits of the low and high 32 bytes, packed two dwords at a time
  uint32_t acc[6][2] = {};
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    uint32_t f[6][2];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const uint32_t v = x[2 * j + e];
      const uint32_t nh = ASCII ? 0x80808080u : (~v & 0x80808080u);
      const uint32_t lo = ASCII ? v : (v & 0x7F7F7F7Fu);
      const uint32_t lo20 = lo | 0x20202020u;
      f[0][e] = xad(lo, kq, ones) & nh;
      f[1][e] = xad(lo, kb, ones) & nh;
      f[2][e] = xad(lo, kn, ones) & nh;
      f[3][e] = xad(lo20, ko, ones) & nh;
      f[4][e] = xad(lo20, kc, ones) & nh;
      f[5][e] = (xad(lo, kl, ones) | xad(lo, km, ones)) & nh;
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const uint32_t p = pack2(f[c][0], f[c][1]); // mask byte << 7
      const int sh = 8 * (j & 3) - 7;
      acc[c][j >> 2] |= sh < 0 ? (p >> 7) : (p << sh);
    }
  }
  WordClasses w;
  w.quote = (uint64_t)acc[0][0] | ((uint64_t)acc[0][1] << 32);
  w.backslash = (uint64_t)acc[1][0] | ((uint64_t)acc[1][1] << 32);
  w.newline = (uint64_t)acc[2][0] | ((uint64_t)acc[2][1] << 32);
  w.open = (uint64_t)acc[3][0] | ((uint64_t)acc[3][1] << 32);
  w.close = (uint64_t)acc[4][0] | ((uint64_t)acc[4][1] << 32);
  w.sep = (uint64_t)acc[5][0] | ((uint64_t)acc[5][1] << 32);
  return w;
}

// Escaped-byte mask of one word for escape carry-in 0: bit i set iff the
// backslash run ending right before byte i (inside the word) has odd length.
// Runs starting at even positions escape their odd positions and vice versa;
// adding the run's start bit to B sweeps the run and marks the byte after it.
__device__ __forceinline__ uint64_t escaped_from_zero(uint64_t b) {
  const uint64_t starts = b & ~(b << 1);
  const uint64_t from_even = (b + (starts & kEven)) ^ b;
  const uint64_t from_odd = (b + (starts & kOdd)) ^ b;
  return (from_even & kOdd) | (from_odd & kEven);
}

// ---- wave / block primitives ----------------------------------------------
__device__ __forceinline__ uint32_t count_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Inclusive add-scan over the wave with DPP row shifts and row broadcasts
// (six VALU instructions; rows of 16 lanes, then row 0 -> 1, row 2 -> 3,
// rows 0..1 -> 2..3).  Lanes outside the source range add 0.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ uint32_t dpp_src(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWMASK, 0xF, false);
}
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  v += dpp_src<0x111, 0xF>(v); // row_shr:1
  v += dpp_src<0x112, 0xF>(v); // row_shr:2
  v += dpp_src<0x114, 0xF>(v); // row_shr:4
  v += dpp_src<0x118, 0xF>(v); // row_shr:8
  v += dpp_src<0x142, 0xA>(v); // row_bcast:15
  v += dpp_src<0x143, 0xC>(v); // row_bcast:31
  return v;
}
// Sum over the wave (every lane gets it).
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readlane((int)wave_incl_scan(v), 63);
}

// ---- tile descriptors -------------------------------------------------------
// Aggregate (flag 1): a transfer table over the four input states
//   s = esc_in | str_in << 1, packed: bits 0..7 the output state of input s
//   at bits 2s (esc_out | str_out << 1), bits 8..39 its depth delta (mod
//   256) in byte s; bits 40..61: newline count.
// Inclusive (flag 2): esc bit 0, str bit 1, depth bits 2..9 (int8),
//   newline count bits 10..61.
constexpr uint64_t kFlagShift = 62;
constexpr uint64_t kFlagAggregate = 1;
constexpr uint64_t kFlagInclusive = 2;
// Tables in registers are unpacked: bits 0..31 the output state of input s
// in byte s, bits 32..63 the depth delta of input s in byte s, so that
// composing two is two byte permutes and a bytewise add.
// Identity: s -> s, depth delta 0.
constexpr uint64_t kIdentityTable = 0x03020100ull;
// The state before an input's first byte as an inclusive descriptor: no
// escape, outside strings, depth -1, no newlines.
constexpr uint64_t kInitialInclusive = (0xFFull << 2) | (kFlagInclusive << kFlagShift);

struct Carry {
  uint32_t esc, str;
  int32_t depth; // int8 semantics, kept in [-128,127]
  uint64_t nl;
};

// Unpacks the 40-bit table field of an aggregate descriptor.
__device__ __forceinline__ uint64_t tbl_unpack(uint64_t d) {
  const uint32_t a2 = (uint32_t)d & 0xFFu;
  const uint32_t a = (a2 | (a2 << 6) | (a2 << 12) | (a2 << 18)) & 0x03030303u;
  return (uint64_t)a | (((d >> 8) & 0xFFFFFFFFull) << 32);
}
// (f o g)(s) = f(g(s)): apply g first.  v_perm_b32 picks, for each byte s,
// byte g(s) of f's states and of f's deltas; the deltas add bytewise mod 256.
__device__ __forceinline__ uint64_t tbl_compose(uint64_t f, uint64_t g) {
  const uint32_t ga = (uint32_t)g, gb = (uint32_t)(g >> 32);
  const uint32_t a = __builtin_amdgcn_perm(0u, (uint32_t)f, ga);
  const uint32_t x = __builtin_amdgcn_perm(0u, (uint32_t)(f >> 32), ga);
  const uint32_t b = ((gb & 0x7F7F7F7Fu) + (x & 0x7F7F7F7Fu)) ^ ((gb ^ x) & 0x80808080u);
  return (uint64_t)a | ((uint64_t)b << 32);
}
__device__ __forceinline__ Carry tbl_apply(uint64_t tbl, uint64_t nl_add,
                                           Carry c) {
  const uint32_t sh = 8u * (c.esc | (c.str << 1));
  const uint32_t st = ((uint32_t)tbl >> sh) & 3u;
  const uint32_t dd = ((uint32_t)(tbl >> 32) >> sh) & 0xFFu;
  Carry o;
  o.esc = st & 1u;
  o.str = st >> 1;
  o.depth = (int32_t)(int8_t)(uint8_t)((uint32_t)c.depth + dd);
  o.nl = c.nl + nl_add;
  return o;
}
__device__ __forceinline__ uint64_t encode_inclusive(Carry c) {
  return (uint64_t)c.esc | ((uint64_t)c.str << 1) |
         ((uint64_t)((uint32_t)c.depth & 0xFFu) << 2) |
         ((c.nl & ((1ull << 52) - 1)) << 10) | (kFlagInclusive << kFlagShift);
}
__device__ __forceinline__ Carry decode_inclusive(uint64_t d) {
  Carry c;
  c.esc = (uint32_t)d & 1u;
  c.str = (uint32_t)(d >> 1) & 1u;
  c.depth = (int32_t)(int8_t)(uint8_t)(d >> 2);
  c.nl = (d >> 10) & ((1ull << 52) - 1);
  return c;
}

__device__ __forceinline__ uint64_t desc_load(const uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void desc_store(uint64_t *p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

} // namespace

// Block barrier for LDS traffic only: global loads and stores in flight stay
// in flight across it (__syncthreads() waits for them as well).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// The scanner publishes the inclusive state of the last tile of each group
// of kIncGroup tiles only (its coherent stores all leave one XCD); a tile
// composes the aggregates of its group's earlier tiles itself.
constexpr uint64_t kIncGroup = 8;
// Tiles a scan round resolves per lane: 4 waves x 64 lanes x 4 = 1,024.
constexpr uint32_t kScanPer = 4;
constexpr uint32_t kScanWaves = kTileWords / 64;

// Tables (unpacked: states in a, deltas in b) and newline counts of the
// lane CTRL moves in, identity where no lane moves in (DPP row / wave moves).
template <int CTRL, int ROWMASK>
__device__ __forceinline__ void dpp_tbl(uint32_t a, uint32_t b, uint32_t nl, uint32_t &oa,
                                        uint32_t &ob, uint32_t &onl) {
  oa = (uint32_t)__builtin_amdgcn_update_dpp((int)kIdentityTable, (int)a, CTRL, ROWMASK, 0xF, false);
  ob = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROWMASK, 0xF, false);
  onl = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)nl, CTRL, ROWMASK, 0xF, false);
}
// One step of the scanner's wave scan.  SEG: a composite may be absolute (a
// file's first tile and anything after it: `abs`), which no older composite
// changes.
template <int CTRL, int ROWMASK, bool SEG>
__device__ __forceinline__ void scan_step(uint64_t &tab, uint32_t &nl, uint32_t &abs) {
  uint32_t oa, ob, onl;
  dpp_tbl<CTRL, ROWMASK>((uint32_t)tab, (uint32_t)(tab >> 32), nl, oa, ob, onl);
  if (SEG) {
    const uint32_t oabs =
        (uint32_t)__builtin_amdgcn_update_dpp(0, (int)abs, CTRL, ROWMASK, 0xF, false);
    if (!abs) {
      tab = tbl_compose(tab, (uint64_t)oa | ((uint64_t)ob << 32));
      nl += onl;
      abs = oabs;
    }
  } else {
    tab = tbl_compose(tab, (uint64_t)oa | ((uint64_t)ob << 32));
    nl += onl;
  }
}

// A multi-file batch's scan meets inclusive descriptors (a file's first tile
// publishes its state, from the initial one, at once): an absolute
// composite, the table mapping every input state to that state (and depth:
// all four delta bytes hold it, so composing later tables onto it is the
// ordinary composition) with its newline count.
__device__ __forceinline__ uint64_t abs_table(uint64_t d) {
  const uint64_t st = d & 3u, depth = (d >> 2) & 0xFFu;
  return st * 0x01010101ull | ((depth * 0x01010101ull) << 32);
}
__device__ __forceinline__ Carry abs_apply(uint64_t tab, uint64_t nl) {
  Carry o;
  o.esc = (uint32_t)tab & 1u;
  o.str = ((uint32_t)tab >> 1) & 1u;
  o.depth = (int32_t)(int8_t)(uint8_t)(tab >> 32);
  o.nl = nl;
  return o;
}

// A half-tile's transfer table (packed: output states in bits 0..7, depth
// deltas from bit 8) from its escape transfer (tail: 0, 1 constant, 2
// identity, 3 flip), whether an escape carry-in un-counts a quote at its
// first non-backslash byte (g), its quote parity and its depth sums for
// string carry-in 0 (D0) and over all brackets (TOT).
__device__ __forceinline__ uint64_t half_table(uint32_t g, uint32_t tail, uint32_t par, int D0,
                                               int TOT) {
  const uint32_t D0b = (uint32_t)D0 & 0xFFu, TOTb = (uint32_t)TOT & 0xFFu;
  uint64_t tb = 0;
#pragma unroll
  for (uint32_t X = 0; X < 2; ++X) {
#pragma unroll
    for (uint32_t Y = 0; Y < 2; ++Y) {
      const uint32_t eff = Y ^ (X & g);
      const uint32_t esc_out = (tail & 2u) ? (X ^ (tail & 1u)) : (tail & 1u);
      const uint32_t str_out = Y ^ par ^ (X & g);
      const uint32_t dd = eff ? ((TOTb - D0b) & 0xFFu) : D0b;
      const uint32_t in = X | (Y << 1); // input state
      tb |= (uint64_t)(esc_out | (str_out << 1)) << (2 * in);
      tb |= (uint64_t)dd << (8 + 8 * in);
    }
  }
  return tb;
}

// The prefix scanner (one workgroup): resolves tiles in order.  Each round reads the aggregates of the next 1,024 tiles, takes
// the longest run of them that is published, composes each lane's share, then
// the lanes' composites with a DPP wave scan and the waves' composites through
// LDS, applies them to the running carry and writes the inclusive state of
// each resolved group's last tile.  A tile waits only for its group's
// carry-in, so the carry chain costs one descriptor round trip per round
// instead of a look-back per tile.  Every wave takes the same branches (counts come
// through LDS), so the barriers always match.  SEG: a multi-file batch, whose
// files' first tiles publish inclusive descriptors; the carry restarts there.
template <bool SEG>
__device__ void scan_tiles(const uint64_t *agg, uint64_t *inc, uint64_t num_tiles,
                           uint32_t *error_flag, uint32_t max_spins, uint32_t lane,
                           uint32_t wv, uint64_t init) {
  __shared__ uint32_t s_ready[kScanWaves];
  __shared__ uint64_t s_tab[kScanWaves];
  __shared__ uint32_t s_nl[kScanWaves];
  __shared__ uint32_t s_abs[kScanWaves];
  Carry carry = decode_inclusive(init);
  uint64_t next = 0;
  uint32_t spins = 0;
  const uint64_t lo = (uint64_t)(wv * 64 + lane) * kScanPer; // this lane's first slot
  while (next < num_tiles) {
    const uint64_t b = next + lo;
    uint64_t d[kScanPer];
#pragma unroll
    for (uint32_t k = 0; k < kScanPer; ++k)
      d[k] = (b + k < num_tiles) ? desc_load(&agg[b + k]) : 0ull;
    // published tiles at the front of this lane's share, then of the wave's
    uint32_t cnt = kScanPer;
#pragma unroll
    for (int k = kScanPer - 1; k >= 0; --k)
      if ((d[k] >> kFlagShift) == 0) cnt = (uint32_t)k;
    const uint64_t partial = __ballot(cnt < kScanPer);
    const uint32_t fl = partial ? (uint32_t)__builtin_ctzll(partial) : 64u;
    const uint32_t wave_ready =
        fl * kScanPer + (fl < 64 ? (uint32_t)__builtin_amdgcn_readlane((int)cnt, (int)fl) : 0u);
    if (lane == 0) s_ready[wv] = wave_ready;
    lds_barrier();
    uint64_t ready = 0;
    for (uint32_t w = 0; w < kScanWaves; ++w) {
      const uint32_t r = s_ready[w];
      ready += r;
      if (r < 64 * kScanPer) break;
    }
    if (ready > num_tiles - next) ready = num_tiles - next;
    if (ready == 0) {
      lds_barrier(); // s_ready is rewritten next round
      if (spins >= max_spins) {
        if (wv == 0 && lane == 0)
          __hip_atomic_store(error_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
      }
      ++spins;
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    // this lane's resolved tiles: [b, b + m)
    const uint32_t m = lo >= ready ? 0u : (uint32_t)min((uint64_t)kScanPer, ready - lo);
    uint64_t tab = kIdentityTable;
    uint32_t nl = 0, abs = 0;
#pragma unroll
    for (uint32_t k = 0; k < kScanPer; ++k)
      if (k < m) {
        if (SEG && (d[k] >> kFlagShift) == kFlagInclusive) { // a file's first tile
          tab = abs_table(d[k]);
          nl = (uint32_t)(d[k] >> 10) & ((1u << 22) - 1);
          abs = 1;
        } else {
          tab = tbl_compose(tbl_unpack(d[k]), tab);
          nl += (uint32_t)(d[k] >> 40) & ((1u << 22) - 1);
        }
      }
    // inclusive scan over the wave's lanes, older lanes applied first
    scan_step<0x111, 0xF, SEG>(tab, nl, abs); // row_shr:1
    scan_step<0x112, 0xF, SEG>(tab, nl, abs); // row_shr:2
    scan_step<0x114, 0xF, SEG>(tab, nl, abs); // row_shr:4
    scan_step<0x118, 0xF, SEG>(tab, nl, abs); // row_shr:8
    scan_step<0x142, 0xA, SEG>(tab, nl, abs); // row_bcast:15
    scan_step<0x143, 0xC, SEG>(tab, nl, abs); // row_bcast:31
    if (lane == 63) {
      s_tab[wv] = tab;
      s_nl[wv] = nl;
      s_abs[wv] = abs;
    }
    lds_barrier();
    // carry before this wave, and after the whole round
    Carry before = carry;
    for (uint32_t w = 0; w < kScanWaves; ++w) {
      if (w == wv) before = carry;
      carry = (SEG && s_abs[w]) ? abs_apply(s_tab[w], s_nl[w])
                                : tbl_apply(s_tab[w], s_nl[w], carry);
    }
    uint32_t ea, eb, enl; // this lane's exclusive prefix (wave_shr:1)
    dpp_tbl<0x138, 0xF>((uint32_t)tab, (uint32_t)(tab >> 32), nl, ea, eb, enl);
    const uint64_t etab = (uint64_t)ea | ((uint64_t)eb << 32);
    Carry c = tbl_apply(etab, enl, before);
    if (SEG && __builtin_amdgcn_update_dpp(0, (int)abs, 0x138, 0xF, 0xF, false))
      c = abs_apply(etab, enl);
#pragma unroll
    for (uint32_t k = 0; k < kScanPer; ++k)
      if (k < m) {
        c = (SEG && (d[k] >> kFlagShift) == kFlagInclusive)
                ? decode_inclusive(d[k])
                : tbl_apply(tbl_unpack(d[k]), (d[k] >> 40) & ((1ull << 22) - 1), c);
        if (((b + k) & (kIncGroup - 1)) == kIncGroup - 1 || b + k + 1 == num_tiles)
          desc_store(&inc[b + k], encode_inclusive(c));
      }
    next += ready;
    spins = 0;
    lds_barrier(); // s_ready, s_tab and s_nl are rewritten next round
  }
}

// Clears the tile descriptors, a spare word and the failure flag (kept out
// of hipMemsetAsync so a captured graph replays it correctly).
__global__ void index_clear_kernel(uint64_t *desc, uint64_t count,
                                   uint32_t *counter) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t k = i; k < count; k += stride) desc[k] = 0;
  if (i == 0) {
    counter[0] = 0; // spare
    counter[1] = 0; // failure flag of the bounded waits
  }
}

// The fused single-pass index kernel, plus the scanner workgroup (block 0).
// A workgroup of 256 lanes takes whole 16 KiB tiles of kTileWords words, one
// word per lane: NH (1 or kTileHalves) consecutive tiles per workgroup (lane
// t holds word t of each); their
// carry-free work (classification, escapes, in-tile string parity, depth and
// newline sums) runs one tile after the other, each tile publishing its
// aggregate as soon as that is done, then the workgroup waits once, for its
// first tile's carry-in, and the next tile starts from the previous one's
// carry-out.  The carry takes ~10 us to arrive (a tile's publish, the
// scanner's read and write and the tile's poll are device-scope memory round
// trips under load), so the bytes a wave slot holds across that wait, not
// the work per byte, set the throughput: one tile per workgroup at 8 waves
// per SIMD ran the 8 GiB input in 4.3 ms.  SEG: multi-file batch (see
// BatchSeg): files start on tiles, a file's first tile starts from the
// initial state, and a workgroup's two tiles may belong to different files.
template <int LMAX, bool SEG, int NH>
__global__ __launch_bounds__(kTileWords) __attribute__((amdgpu_waves_per_eu((LMAX <= 4) ? 8 : 4))) void index_kernel(
    const uint8_t *__restrict__ in, uint64_t n, uint64_t words, int levels,
    int aligned, uint64_t *__restrict__ string_out,
    uint64_t *__restrict__ leveled_out, int64_t *__restrict__ newline_out,
    int64_t newline_cap, int64_t *__restrict__ num_lines_out,
    uint64_t *__restrict__ desc, uint32_t *__restrict__ tile_counter,
    uint64_t num_tiles, BatchSeg seg, uint32_t max_spins, uint64_t init) {
  static_assert(NH == 1 || NH == 2, "one or two tiles per workgroup");
  // With up to 4 levels the kernel runs 8 waves per SIMD (64 VGPRs): with
  // two tiles per workgroup the first tile's kept values wait in LDS (in the
  // leveled-row slots, which only step 8 uses) while the second tile's steps
  // 1-4 run.
  constexpr bool STASH = NH == 2 && LMAX <= 4;
  constexpr int kKeptWords = 12; // dwords a tile keeps per lane (see below)
  constexpr int kRowWords = (LMAX + 1) * 2 * kTileWords;
  constexpr int kWorkWords =
      (STASH && kKeptWords * kTileWords > kRowWords) ? kKeptWords * kTileWords : kRowWords;
  uint32_t *const error_flag = tile_counter + 1;
  if (SEG) init = kInitialInclusive; // every file of a batch starts fresh
  __shared__ uint32_t s_tile;
  __shared__ uint8_t s_bsinfo[NH][kTileWords];  // bit0 trailing parity, bit1 all-backslash
  __shared__ uint8_t s_lead[NH][kTileWords];    // leading backslash run, bit7 = quote at its end
  __shared__ uint8_t s_cin[NH][kTileWords];     // escape carry-in per word (half carry-in 0)
  __shared__ uint32_t s_wany[NH][kTileWords / 64]; // per wave: a word of only backslashes
  __shared__ uint32_t s_wpar[NH][kTileWords / 64]; // per wave: quote parity
  __shared__ uint32_t s_wsum[NH][kTileWords / 64]; // per wave: sum of (d0a+64) | (tot+64) << 16
  __shared__ uint32_t s_wnl[NH][kTileWords / 64];  // per wave: newlines
  __shared__ uint32_t s_wscan[NH][kTileWords / 64];
  __shared__ uint32_t s_first[NH];              // k | lead_k << 16 | g << 24
  // per-lane leveled half-words (+ sink) of step 8: s_rows[LMAX + 1][2][256];
  // before that, the first tile's kept values ([kKeptWords][256])
  __shared__ uint32_t s_work[kWorkWords];
  uint32_t(*const s_rows)[2][kTileWords] = reinterpret_cast<uint32_t(*)[2][kTileWords]>(s_work);
  __shared__ uint32_t s_tail_esc[NH];           // half escape transfer: 0, 1 constant, 2 identity, 3 flip
  __shared__ Carry s_carry[NH];                 // carry-in of each half

  const uint32_t t = threadIdx.x;
  const uint32_t lane = t & 63u;
  const uint32_t wv = t >> 6;

  // 0. tile id from the block: workgroups are dispatched in block order, so
  //    the tile a wait depends on has been dispatched or is next in line.
  //    Block 0 is the prefix scanner, block b takes tiles NH (b - 1) ..  (An
  //    atomic tile counter costs ~11 ns per tile in series on MI355X.)
  if (blockIdx.x == 0) {
    scan_tiles<SEG>(desc, desc + num_tiles, num_tiles, error_flag, max_spins, lane, wv, init);
    return;
  }
  // The file of the workgroup's tile h (the whole input when !SEG;
  // wave-uniform): its first and last tiles, byte range (bytes past its end
  // read as 0; a tile past the last one reads none), newline slice and line
  // count.  SEG: read once into LDS.
  struct TileFile {
    uint64_t first, last, base, end;
    int64_t *nl_out;
    int64_t nl_cap;
    int64_t *lines_out;
  };
  __shared__ TileFile s_file[SEG ? NH : 1];
  if (t == 0) s_tile = NH * (blockIdx.x - 1);
  if (SEG && t < NH) {
    const uint64_t tt = NH * (uint64_t)(blockIdx.x - 1) + t;
    const uint32_t fi = seg.tile_file[tt < num_tiles ? tt : num_tiles - 1];
    const BatchFile bf = seg.files[fi];
    s_file[SEG ? t : 0] =
        TileFile{bf.offset / kTileBytes, (bf.offset + bf.bytes - 1) / kTileBytes, bf.offset,
                 tt < num_tiles ? bf.offset + bf.bytes : 0, newline_out + bf.nl_offset,
                 (int64_t)bf.nl_capacity, seg.line_counts + fi};
  }
  lds_barrier();
  // tile and word base are wave-uniform (scalar registers); per-lane offsets
  // are formed from them where needed rather than kept alive in VGPRs
  const uint64_t tile = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_tile);
  const uint64_t wbase = tile * (uint64_t)kTileWords;
  auto uni = [](uint64_t x) -> uint64_t { // (a wave-uniform LDS value into scalar registers)
    return ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(x >> 32)) << 32) |
           (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
  };
  auto file_of = [&](int h) -> TileFile {
    if (!SEG) return TileFile{0, num_tiles - 1, 0, n, newline_out, newline_cap, num_lines_out};
    const TileFile &f = s_file[SEG ? h : 0];
    return TileFile{uni(f.first), uni(f.last), uni(f.base), uni(f.end),
                    reinterpret_cast<int64_t *>(uni(reinterpret_cast<uint64_t>(f.nl_out))),
                    (int64_t)uni((uint64_t)f.nl_cap),
                    reinterpret_cast<int64_t *>(uni(reinterpret_cast<uint64_t>(f.lines_out)))};
  };

  // What each half keeps from its carry-free work until the carry arrives:
  // per lane the word's openers, closers, separators, newlines and string
  // mask for string carry-in 0 at the half's start, and its in-wave scans;
  // per wave (scalar registers) the half's transfer-table inputs and the
  // sums over the waves before this one.
  uint64_t kO[NH], kC[NH], kP[NH], kN[NH], kS0[NH];
  uint32_t kDex[NH], kNlEx[NH]; // exclusive in-wave depth and newline sums
  uint32_t uFirst[NH], uTail[NH], uNL[NH], uNlBefore[NH], uMyPar[NH];
  int32_t uDBefore[NH], uTotBefore[NH];
  uint64_t htbl[NH]; // each half's transfer table (wave 0)

#pragma unroll
  for (int h = 0; h < NH; ++h) {
    const uint64_t hb = wbase + (uint64_t)h * kTileWords; // the tile's first word
    const uint64_t pos = (hb + t) * 64;
    const uint64_t file_end = file_of(h).end;

    // 1. classify this lane's 64 bytes
    WordClasses m = {0, 0, 0, 0, 0, 0};
    if (pos < file_end) {
      uint32_t x[16];
      load_word(in, file_end, pos, aligned != 0, x);
      // a wave whose words hold no byte >= 0x80 (wave-uniform) classifies
      // without the high-bit masks
      uint32_t high = 0;
#pragma unroll
      for (int i = 0; i < 16; ++i) high |= x[i];
      if (__ballot((high & 0x80808080u) != 0) == 0)
        m = classify<true>(x);
      else
        m = classify<false>(x);
    }

    // 2. escapes, assuming the half starts with escape carry 0.  Each word's
    //    escape transfer: the parity of the backslash run ending at its last
    //    byte (a constant), or, for a word of only backslashes, its carry-in
    //    passed through (64 of them) or flipped (an odd number at the
    //    input's end).  Words past a single input's end pass the carry
    //    through, so the last tile's escape-out is the input's (a shard's
    //    successor needs it).
    const uint64_t B = m.backslash;
    const bool real = pos < file_end;
    const uint32_t valid = real ? (uint32_t)min((uint64_t)64, file_end - pos) : 0u;
    const uint64_t e0 = escaped_from_zero(B);
    const int lead = (!SEG && !real) ? 64 : ((B == ~0ull) ? 64 : __builtin_ctzll(~B));
    uint32_t mode; // 0, 1: escape-out constant; 2: carry passes; 3: carry flips
    if (!real) {
      mode = SEG ? 0u : 2u;
    } else {
      const uint64_t bv = valid == 64 ? B : (B << (64 - valid)); // last byte at bit 63
      const uint32_t run = (~bv == 0ull) ? 64u : (uint32_t)__builtin_clzll(~bv);
      mode = run >= valid ? (2u | (valid & 1u)) : (run & 1u);
    }
    const bool all_bs = mode >= 2u;
    s_bsinfo[h][t] = (uint8_t)mode;
    s_lead[h][t] = (uint8_t)(lead | ((lead < 64 && ((m.quote >> lead) & 1ull)) ? 0x80 : 0));
    const uint64_t any_ballot = __ballot(all_bs);
    if (lane == 0) s_wany[h][wv] = any_ballot != 0;
    lds_barrier();
    uint32_t any_all_bs = 0;
#pragma unroll
    for (uint32_t k = 0; k < kTileWords / 64; ++k) any_all_bs |= s_wany[h][k];
    uint32_t cin;
    if (any_all_bs) {
      // rare: a word of only backslashes (or past the input's end) passes
      // its carry through or flips it; one serial pass over the half's
      // words, from carry-in 0 and 1, gives every word's carry-in for 0 and
      // the half's transfer (constant, 2 = identity, 3 = flip)
      if (t == 0) {
        uint32_t c = 0, c1 = 1;
        for (uint32_t k = 0; k < kTileWords; ++k) {
          s_cin[h][k] = (uint8_t)c;
          const uint32_t md = s_bsinfo[h][k];
          c = md == 2u ? c : (md == 3u ? c ^ 1u : md);
          c1 = md == 2u ? c1 : (md == 3u ? c1 ^ 1u : md);
        }
        s_tail_esc[h] = c == c1 ? c : (2u | c);
        uint32_t k = 0;
        while (k < kTileWords && (s_lead[h][k] & 0x7F) >= 64) ++k;
        if (k == kTileWords) {
          s_first[h] = (uint32_t)kTileWords;
        } else {
          s_first[h] = k | ((uint32_t)(s_lead[h][k] & 0x7F) << 16) |
                       ((uint32_t)(s_lead[h][k] >> 7) << 24);
        }
      }
      lds_barrier();
      cin = s_cin[h][t];
      uTail[h] = __builtin_amdgcn_readfirstlane(s_tail_esc[h]);
      uFirst[h] = __builtin_amdgcn_readfirstlane(s_first[h]);
    } else {
      cin = t ? (s_bsinfo[h][t - 1] & 1u) : 0u;
      uTail[h] = __builtin_amdgcn_readfirstlane(s_bsinfo[h][kTileWords - 1] & 1u);
      const uint32_t l0 = s_lead[h][0];
      uFirst[h] = __builtin_amdgcn_readfirstlane(((l0 & 0x7Fu) << 16) | ((l0 >> 7) << 24));
    }
    const uint64_t escaped = cin ? (e0 ^ mask_upto(lead)) : e0;

    // 3. string mask for string carry-in 0 at this wave's first word (the
    //    parity the half's earlier waves leave is applied below)
    const uint64_t q = m.quote & ~escaped;
    const uint64_t px = prefix_xor64(q);
    const uint64_t par_ballot = __ballot((uint32_t)(px >> 63));
    const uint64_t s0a = px ^ ((count_below(par_ballot) & 1u) ? ~0ull : 0ull);

    // 4. depth delta of this word under s0a and its total, newline count;
    //    the in-wave inclusive scans serve both the half's aggregate and,
    //    once the carry-in is known, this word's depth and newline offsets
    const int d0a = __popcll(m.open & ~s0a) - __popcll(m.close & ~s0a);
    const int tot = __popcll(m.open) - __popcll(m.close);
    const uint32_t nlc = (uint32_t)__popcll(m.newline);
    // 16-bit fields: in-wave sums stay below 64 * 129
    const uint32_t dv = (uint32_t)(d0a + 64) | ((uint32_t)(tot + 64) << 16);
    const uint32_t dincl = wave_incl_scan(dv);
    const uint32_t nincl = wave_incl_scan(nlc);
    if (lane == 63) {
      s_wsum[h][wv] = dincl;
      s_wnl[h][wv] = nincl;
    }
    if (lane == 0) s_wpar[h][wv] = (uint32_t)__popcll(par_ballot) & 1u;
    lds_barrier();

    // half sums for string carry-in 0 (wave k's words see the parity of the
    // waves before it), and the same sums over the waves before this one
    // (wave-uniform: scalar registers)
    uint32_t run_par = 0, my_par = 0, NL = 0, nl_before = 0;
    int D0 = 0, TOT = 0, d_before = 0, tot_before = 0;
#pragma unroll
    for (uint32_t k = 0; k < kTileWords / 64; ++k) {
      const uint32_t ws = __builtin_amdgcn_readfirstlane(s_wsum[h][k]);
      const uint32_t wn = __builtin_amdgcn_readfirstlane(s_wnl[h][k]);
      const int d0k = (int)(ws & 0xFFFFu) - 64 * 64;
      const int totk = (int)(ws >> 16) - 64 * 64;
      const int dk = run_par ? totk - d0k : d0k;
      if (k < wv) {
        d_before += dk;
        tot_before += totk;
        nl_before += wn;
      }
      if (k == wv) my_par = run_par;
      D0 += dk;
      TOT += totk;
      NL += wn;
      run_par ^= __builtin_amdgcn_readfirstlane(s_wpar[h][k]);
    }
    kO[h] = m.open;
    kC[h] = m.close;
    kP[h] = m.sep;
    kN[h] = m.newline;
    kS0[h] = s0a ^ (my_par ? ~0ull : 0ull);
    kDex[h] = dincl - dv;
    kNlEx[h] = nincl - nlc;
    if (STASH && h == 0) {
      uint32_t *const st = &s_work[t];
      st[0 * kTileWords] = (uint32_t)kO[0];
      st[1 * kTileWords] = (uint32_t)(kO[0] >> 32);
      st[2 * kTileWords] = (uint32_t)kC[0];
      st[3 * kTileWords] = (uint32_t)(kC[0] >> 32);
      st[4 * kTileWords] = (uint32_t)kP[0];
      st[5 * kTileWords] = (uint32_t)(kP[0] >> 32);
      st[6 * kTileWords] = (uint32_t)kN[0];
      st[7 * kTileWords] = (uint32_t)(kN[0] >> 32);
      st[8 * kTileWords] = (uint32_t)kS0[0];
      st[9 * kTileWords] = (uint32_t)(kS0[0] >> 32);
      st[10 * kTileWords] = kDex[0];
      st[11 * kTileWords] = kNlEx[0];
    }
    uNL[h] = NL;
    uNlBefore[h] = nl_before;
    uMyPar[h] = my_par;
    uDBefore[h] = d_before;
    uTotBefore[h] = tot_before;
    if (wv == 0) {
      htbl[h] = half_table((uFirst[h] >> 24) & 1u, uTail[h], run_par, D0, TOT);
      // each tile publishes as soon as its aggregate is known (the second
      // tile's work overlaps the wait for the first's carry); a file's
      // first tile of a batch knows its inclusive state already
      if (lane == 0 && tile + h < num_tiles) {
        if (SEG && tile + h == file_of(h).first)
          desc_store(&desc[tile + h],
                     encode_inclusive(tbl_apply(tbl_unpack(htbl[h]), NL, decode_inclusive(init))));
        else
          desc_store(&desc[tile + h],
                     htbl[h] | ((uint64_t)NL << 40) | (kFlagAggregate << kFlagShift));
      }
    }
  }

  // 5. resolve the exclusive prefix of the workgroup's first tile (wave 0)
  if (wv == 0) {
    const uint64_t first_tile = file_of(0).first;
    Carry excl = decode_inclusive(init); // the first tile's carry-in
    if (tile != first_tile) {
      // wait for the inclusive state before this tile's group (the initial
      // state for group 0) and for the descriptors of the group's earlier
      // tiles, all in one load per poll (lanes 0 .. need-1 and 63); stop when
      // a failure is flagged.  SEG: when the tile's file starts inside the
      // group, its first tile's inclusive descriptor is the start instead.
      const uint64_t *inc = desc + num_tiles;
      const uint64_t g0 = tile & ~(kIncGroup - 1);
      const uint32_t need = (uint32_t)(tile - g0);
      uint32_t spins = 0;
      uint64_t start = 0;
      uint32_t from = 0; // the group's first tile composed onto start
      uint64_t v = 0;
      for (;;) {
        if (spins >= max_spins) { // bounded wait: give up, flag the failure
          if (lane == 0)
            __hip_atomic_store(error_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          start = kInitialInclusive;
          from = need;
          break;
        }
        v = 0;
        if (lane < need) v = desc_load(&desc[g0 + lane]);
        else if (lane == 63) v = g0 > 0 ? desc_load(&inc[g0 - 1]) : init;
        const uint64_t missing = __ballot(lane < need && (v >> kFlagShift) == 0);
        const uint64_t restart = SEG ? __ballot(lane < need && (v >> kFlagShift) == kFlagInclusive)
                                     : 0ull;
        if (missing == 0 && restart) {
          const int k = 63 - __builtin_clzll(restart);
          start = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), k) << 32) |
                  (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, k);
          from = (uint32_t)k + 1;
          break;
        }
        const uint64_t d = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), 63) << 32) |
                           (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, 63);
        if (missing == 0 && (d >> kFlagShift) == kFlagInclusive) {
          start = d;
          from = 0;
          break;
        }
        // another tile's failure ends this wait as well
        if ((++spins & 1023u) == 0 &&
            __hip_atomic_load(error_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
          spins = max_spins;
        __builtin_amdgcn_s_sleep(1);
      }
      uint64_t acc = kIdentityTable, acc_nl = 0;
      for (uint32_t i = from; i < need; ++i) {
        const uint64_t a = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), (int)i) << 32) |
                           (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, (int)i);
        acc = tbl_compose(tbl_unpack(a), acc);
        acc_nl += (a >> 40) & ((1ull << 22) - 1);
      }
      excl = tbl_apply(acc, acc_nl, decode_inclusive(start));
    }
    if (lane == 0) {
      Carry ch = excl;
#pragma unroll
      for (int h = 0; h < NH; ++h) {
        const TileFile f = file_of(h);
        if (SEG && h > 0 && tile + h == f.first) ch = decode_inclusive(init); // a new file
        s_carry[h] = ch;
        ch = tbl_apply(tbl_unpack(htbl[h]), uNL[h], ch);
        if (tile + h == f.last && tile + h < num_tiles) *f.lines_out = (int64_t)(1 + ch.nl);
      }
    }
  }
  lds_barrier();

#pragma unroll
  for (int h = 0; h < NH; ++h) {
    const uint64_t hb = wbase + (uint64_t)h * kTileWords; // the tile's first word
    const Carry c = s_carry[h];
    const TileFile f = file_of(h);
    if (STASH && h == 0) { // (each lane reads back only its own slots)
      const uint32_t *const st = &s_work[t];
      kO[0] = (uint64_t)st[0 * kTileWords] | ((uint64_t)st[1 * kTileWords] << 32);
      kC[0] = (uint64_t)st[2 * kTileWords] | ((uint64_t)st[3 * kTileWords] << 32);
      kP[0] = (uint64_t)st[4 * kTileWords] | ((uint64_t)st[5 * kTileWords] << 32);
      kN[0] = (uint64_t)st[6 * kTileWords] | ((uint64_t)st[7 * kTileWords] << 32);
      kS0[0] = (uint64_t)st[8 * kTileWords] | ((uint64_t)st[9 * kTileWords] << 32);
      kDex[0] = st[10 * kTileWords];
      kNlEx[0] = st[11 * kTileWords];
    }

    // 6. final string mask: string carry-in, plus the flip of everything from
    //    the half's first non-backslash byte on when an escape carry-in of 1
    //    un-counts a quote there
    const uint32_t first = uFirst[h];
    const uint32_t k_first = first & 0xFFFFu;
    const uint32_t lead_k = (first >> 16) & 0xFFu;
    const uint32_t g = (first >> 24) & 1u;
    const bool esc_flip = (c.esc & g) != 0; // uniform
    uint64_t flip = c.str ? ~0ull : 0ull;
    if (esc_flip) {
      if (t > k_first) flip ^= ~0ull;
      else if (t == k_first) flip ^= ~0ull << lead_k;
    }
    const uint64_t S = kS0[h] ^ flip;
    const uint64_t O = kO[h] & ~S;
    const uint64_t C = kC[h] & ~S;
    const uint64_t P = kP[h] & ~S;

    // 7. depth before this word and its newline offset
    const uint64_t nl_base = 1 + c.nl + uNlBefore[h] + kNlEx[h];
    int depth_sum;
    if (!esc_flip) {
      // every word keeps its polarity relative to s0a: the words of this
      // wave flip iff my_par ^ c.str, the earlier waves' sums flip iff c.str
      const uint32_t dex = kDex[h];
      const int d0_ex = (int)(dex & 0xFFFFu) - 64 * (int)lane;
      const int tot_ex = (int)(dex >> 16) - 64 * (int)lane;
      depth_sum = (c.str ? uTotBefore[h] - uDBefore[h] : uDBefore[h]) +
                  ((uMyPar[h] ^ c.str) ? tot_ex - d0_ex : d0_ex);
    } else {
      // rare: scan the final per-word deltas
      const int d = __popcll(O) - __popcll(C);
      const uint32_t v = (uint32_t)(d + 64);
      const uint32_t incl_v = wave_incl_scan(v);
      if (lane == 63) s_wscan[h][wv] = incl_v;
      lds_barrier();
      uint32_t excl_v = incl_v - v;
      for (uint32_t k = 0; k < wv; ++k) excl_v += s_wscan[h][k];
      depth_sum = (int)excl_v - 64 * (int)t;
    }
    const int depth_in = (int)(int8_t)(uint8_t)((uint32_t)c.depth + (uint32_t)depth_sum);

    // 8. leveled rows: walk this word's depth changes in order, one 32-bit
    //    half at a time; every structural bit between two changes belongs to
    //    the current depth (an opener's own bit to the depth it opens, a
    //    closer's to the depth it closes).  Each segment is ORed into this
    //    lane's LDS slot of its row; row LMAX is a sink for depths outside
    //    [0, LMAX).  Depth needs no 8-bit wrap inside the word: from an entry
    //    depth in [-128, 127], 64 changes cannot wrap into [0, LMAX).  (A
    //    lane's slots are its own, so the halves reuse them in turn.)
    const bool live = hb + t < words;
    if (live) {
#pragma unroll
      for (int k = 0; k < LMAX; ++k) {
        s_rows[k][0][t] = 0;
        s_rows[k][1][t] = 0;
      }
      const uint64_t E = O | C | P;
      int depth = depth_in;
#pragma unroll
      for (uint32_t hw = 0; hw < 2; ++hw) {
        const uint32_t Eh = (uint32_t)(E >> (32 * hw));
        const uint32_t Oh = (uint32_t)(O >> (32 * hw));
        uint32_t x = Oh | (uint32_t)(C >> (32 * hw));
        uint32_t done = 0;
        uint32_t *const slot = &s_rows[0][hw][t];
        while (x) {
          const uint32_t low = x & (0u - x);
          x ^= low;
          const uint32_t upto = (low - 1u) | (low & ~Oh);
          const uint32_t seg = upto & ~done & Eh;
          done = upto;
          const uint32_t row = min((uint32_t)depth, (uint32_t)LMAX);
          atomicOr(slot + row * (2 * kTileWords), seg);
          depth += (low & Oh) ? 1 : -1;
        }
        const uint32_t row = min((uint32_t)depth, (uint32_t)LMAX);
        atomicOr(slot + row * (2 * kTileWords), Eh & ~done);
      }
      (string_out + hb)[t] = S;
#pragma unroll
      for (int k = 0; k < LMAX; ++k)
        if (k < levels)
          (leveled_out + (uint64_t)k * words + hb)[t] =
              (uint64_t)s_rows[k][0][t] | ((uint64_t)s_rows[k][1][t] << 32);

      // 9. newline positions
      if (tile + h == f.first && t == 0 && f.nl_cap > 0) f.nl_out[0] = 0;
      uint64_t nlm = kN[h];
      int64_t o = (int64_t)nl_base;
      while (nlm) {
        const int b = __builtin_ctzll(nlm);
        nlm &= nlm - 1;
        if (o < f.nl_cap) f.nl_out[o] = (int64_t)(hb * 64 - f.base + t * 64 + (uint64_t)b);
        ++o;
      }
    }
  }
}

template __global__ void index_kernel<2, false, 1>(const uint8_t *, uint64_t, uint64_t, int, int, uint64_t *, uint64_t *, int64_t *, int64_t, int64_t *, uint64_t *, uint32_t *, uint64_t, BatchSeg, uint32_t, uint64_t);
template __global__ void index_kernel<2, false, 2>(const uint8_t *, uint64_t, uint64_t, int, int, uint64_t *, uint64_t *, int64_t *, int64_t, int64_t *, uint64_t *, uint32_t *, uint64_t, BatchSeg, uint32_t, uint64_t);
template __global__ void index_kernel<2, true, 2>(const uint8_t *, uint64_t, uint64_t, int, int, uint64_t *, uint64_t *, int64_t *, int64_t, int64_t *, uint64_t *, uint32_t *, uint64_t, BatchSeg, uint32_t, uint64_t);
template __global__ void index_kernel<4, false, 1>(const uint8_t *, uint64_t, uint64_t, int, int, uint64_t *, uint64_t *, int64_t *, int64_t, int64_t *, uint64_t *, uint32_t *, uint64_t, BatchSeg, uint32_t, uint64_t);
template __global__ void index_kernel<4, false, 2>(const uint8_t *, uint64_t, uint64_t, int, int, uint64_t *, uint64_t *, int64_t *, int64_t, int64_t *, uint64_t *, uint32_t *, uint64_t, BatchSeg, uint32_t, uint64_t);
template __global__ void index_kernel<4, true, 2>(const uint8_t *, uint64_t, uint64_t, int, int, uint64_t *, uint64_t *, int64_t *, int64_t, int64_t *, uint64_t *, uint32_t *, uint64_t, BatchSeg, uint32_t, uint64_t);
template __global__ void index_kernel<8, false, 1>(const uint8_t *, uint64_t, uint64_t, int, int, uint64_t *, uint64_t *, int64_t *, int64_t, int64_t *, uint64_t *, uint32_t *, uint64_t, BatchSeg, uint32_t, uint64_t);
template __global__ void index_kernel<8, false, 2>(const uint8_t *, uint64_t, uint64_t, int, int, uint64_t *, uint64_t *, int64_t *, int64_t, int64_t *, uint64_t *, uint32_t *, uint64_t, BatchSeg, uint32_t, uint64_t);
template __global__ void index_kernel<8, true, 2>(const uint8_t *, uint64_t, uint64_t, int, int, uint64_t *, uint64_t *, int64_t *, int64_t, int64_t *, uint64_t *, uint32_t *, uint64_t, BatchSeg, uint32_t, uint64_t);
template __global__ void index_kernel<16, false, 1>(const uint8_t *, uint64_t, uint64_t, int, int, uint64_t *, uint64_t *, int64_t *, int64_t, int64_t *, uint64_t *, uint32_t *, uint64_t, BatchSeg, uint32_t, uint64_t);
template __global__ void index_kernel<16, false, 2>(const uint8_t *, uint64_t, uint64_t, int, int, uint64_t *, uint64_t *, int64_t *, int64_t, int64_t *, uint64_t *, uint32_t *, uint64_t, BatchSeg, uint32_t, uint64_t);
template __global__ void index_kernel<16, true, 2>(const uint8_t *, uint64_t, uint64_t, int, int, uint64_t *, uint64_t *, int64_t *, int64_t, int64_t *, uint64_t *, uint32_t *, uint64_t, BatchSeg, uint32_t, uint64_t);
template __global__ void index_kernel<22, false, 1>(const uint8_t *, uint64_t, uint64_t, int, int, uint64_t *, uint64_t *, int64_t *, int64_t, int64_t *, uint64_t *, uint32_t *, uint64_t, BatchSeg, uint32_t, uint64_t);
template __global__ void index_kernel<22, false, 2>(const uint8_t *, uint64_t, uint64_t, int, int, uint64_t *, uint64_t *, int64_t *, int64_t, int64_t *, uint64_t *, uint32_t *, uint64_t, BatchSeg, uint32_t, uint64_t);
template __global__ void index_kernel<22, true, 2>(const uint8_t *, uint64_t, uint64_t, int, int, uint64_t *, uint64_t *, int64_t *, int64_t, int64_t *, uint64_t *, uint32_t *, uint64_t, BatchSeg, uint32_t, uint64_t);

// One workgroup: each thread composes a contiguous run of tile tables, then
// the runs are composed in order.
__global__ __launch_bounds__(256) void shard_summary_kernel(const uint64_t *desc,
                                                            uint64_t num_tiles,
                                                            uint64_t *out) {
  __shared__ uint64_t s_tab[256];
  __shared__ uint64_t s_nl[256];
  const uint32_t t = threadIdx.x;
  const uint64_t per = (num_tiles + 255) / 256;
  const uint64_t lo = (uint64_t)t * per;
  const uint64_t hi = lo + per < num_tiles ? lo + per : num_tiles;
  uint64_t tab = kIdentityTable, nl = 0;
  for (uint64_t k = lo; k < hi; ++k) {
    const uint64_t d = desc[k];
    tab = tbl_compose(tbl_unpack(d), tab);
    nl += (d >> 40) & ((1ull << 22) - 1);
  }
  s_tab[t] = tab;
  s_nl[t] = nl;
  __syncthreads();
  if (t == 0) {
    uint64_t all = kIdentityTable, total = 0;
    for (uint32_t k = 0; k < 256; ++k) {
      all = tbl_compose(s_tab[k], all);
      total += s_nl[k];
    }
    out[0] = all;
    out[1] = total;
  }
}

__global__ void index_status_kernel(const uint32_t *error_flag, int64_t *lines_out) {
  if (threadIdx.x == 0 && __hip_atomic_load(error_flag, __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT) != 0)
    *lines_out = -1;
}

// ---- exact newline index (fallback when the capacity hint was too small) ---

// Per-tile newline counts (256 words per block).
__global__ __launch_bounds__(kTileWords) void newline_count_kernel(
    const uint8_t *__restrict__ in, uint64_t n, int aligned,
    uint64_t *__restrict__ tile_counts) {
  __shared__ uint64_t s_sum[kTileWords / 64];
  const uint64_t pos = ((uint64_t)blockIdx.x * kTileWords + threadIdx.x) * 64;
  uint64_t c = 0;
  if (pos < n) {
    uint32_t x[16];
    load_word(in, n, pos, aligned != 0, x);
#pragma unroll
    for (int i = 0; i < 16; ++i)
      c += __popc(eq_bytes(x[i] & 0x7F7F7F7Fu, '\n') & ~x[i] & 0x80808080u);
  }
  c = wave_sum((uint32_t)c); // at most 64 * 64 per wave
  if ((threadIdx.x & 63u) == 0) s_sum[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t s = 0;
    for (int k = 0; k < kTileWords / 64; ++k) s += s_sum[k];
    tile_counts[blockIdx.x] = s;
  }
}

// Exclusive scan of tile counts in one workgroup; writes total + add (the
// line count for the newline index: add = 1 for the synthetic first entry).
__global__ __launch_bounds__(1024) void newline_scan_kernel(
    uint64_t *__restrict__ tile_counts, uint64_t tiles,
    int64_t *__restrict__ num_lines_out, int64_t add) {
  __shared__ uint64_t s_part[1024];
  const uint32_t t = threadIdx.x;
  const uint64_t per = (tiles + 1023) / 1024;
  const uint64_t lo = t * per, hi = (lo + per < tiles) ? lo + per : tiles;
  uint64_t s = 0;
  for (uint64_t i = lo; i < hi; ++i) s += tile_counts[i];
  s_part[t] = s;
  __syncthreads();
  if (t == 0) {
    uint64_t run = 0;
    for (int i = 0; i < 1024; ++i) {
      const uint64_t v = s_part[i];
      s_part[i] = run;
      run += v;
    }
    *num_lines_out = (int64_t)run + add;
  }
  __syncthreads();
  uint64_t run = s_part[t];
  for (uint64_t i = lo; i < hi; ++i) {
    const uint64_t v = tile_counts[i];
    tile_counts[i] = run;
    run += v;
  }
}

__global__ __launch_bounds__(kTileWords) void newline_write_kernel(
    const uint8_t *__restrict__ in, uint64_t n, int aligned,
    const uint64_t *__restrict__ tile_offsets, int64_t *__restrict__ newline_out) {
  __shared__ uint32_t s_wscan[kTileWords / 64];
  const uint32_t t = threadIdx.x, lane = t & 63u, wv = t >> 6;
  const uint64_t pos = ((uint64_t)blockIdx.x * kTileWords + t) * 64;
  uint64_t nlm = 0;
  if (pos < n) {
    uint32_t x[16], f[16];
    load_word(in, n, pos, aligned != 0, x);
#pragma unroll
    for (int i = 0; i < 16; ++i)
      f[i] = eq_bytes(x[i] & 0x7F7F7F7Fu, '\n') & ~x[i] & 0x80808080u;
    nlm = pack16(f);
  }
  const uint32_t c = (uint32_t)__popcll(nlm);
  const uint32_t incl = wave_incl_scan(c);
  if (lane == 63) s_wscan[wv] = incl;
  __syncthreads();
  uint64_t o = incl - c;
  for (uint32_t k = 0; k < wv; ++k) o += s_wscan[k];
  o += 1 + tile_offsets[blockIdx.x];
  if (blockIdx.x == 0 && t == 0) newline_out[0] = 0;
  while (nlm) {
    const int b = __builtin_ctzll(nlm);
    nlm &= nlm - 1;
    newline_out[o++] = (int64_t)(pos + (uint64_t)b);
  }
}

// ---- line counts of a multi-file batch -------------------------------------

// Per tile: the newlines among its bytes that lie inside its file (tiles past
// a file's end belong to that file and count nothing beyond it).
__global__ __launch_bounds__(kTileWords) void batch_count_kernel(
    const uint8_t *__restrict__ in, const BatchFile *__restrict__ files,
    const uint32_t *__restrict__ tile_file, uint64_t *__restrict__ tile_counts) {
  __shared__ uint64_t s_sum[kTileWords / 64];
  const BatchFile bf = files[tile_file[blockIdx.x]];
  const uint64_t pos = ((uint64_t)blockIdx.x * kTileWords + threadIdx.x) * 64;
  const uint64_t end = bf.offset + bf.bytes;
  uint64_t c = 0;
  if (pos >= bf.offset && pos < end) {
    uint32_t x[16];
    load_word(in, end, pos, true, x); // bytes from `end` on read as 0
#pragma unroll
    for (int i = 0; i < 16; ++i)
      c += __popc(eq_bytes(x[i] & 0x7F7F7F7Fu, '\n') & ~x[i] & 0x80808080u);
  }
  c = wave_sum((uint32_t)c); // at most 64 * 64 per wave
  if ((threadIdx.x & 63u) == 0) s_sum[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t sum = 0;
    for (int k = 0; k < kTileWords / 64; ++k) sum += s_sum[k];
    tile_counts[blockIdx.x] = sum;
  }
}

// Per file: 1 + the newlines of its tiles (its line count).
__global__ void batch_file_lines_kernel(const BatchFile *__restrict__ files, int32_t num_files,
                                        const uint64_t *__restrict__ tile_counts,
                                        int64_t *__restrict__ lines) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= num_files) return;
  const BatchFile bf = files[f];
  uint64_t sum = 1;
  for (uint64_t t = bf.offset / kTileBytes; t <= (bf.offset + bf.bytes - 1) / kTileBytes; ++t)
    sum += tile_counts[t];
  lines[f] = (int64_t)sum;
}

// ---- set-bit positions of one bitmap row ----------------------------------
// Record bounds of a top-level array document are the set bits of leveled
// row 0 (its '[', the ',' between elements and its ']').

__global__ __launch_bounds__(kTileWords) void row_count_kernel(
    const uint64_t *__restrict__ row, uint64_t words, uint64_t *__restrict__ tile_counts) {
  __shared__ uint64_t s_sum[kTileWords / 64];
  const uint64_t w = (uint64_t)blockIdx.x * kTileWords + threadIdx.x;
  uint64_t c = w < words ? (uint64_t)__popcll(row[w]) : 0;
  c = wave_sum((uint32_t)c); // at most 64 * 64 per wave
  if ((threadIdx.x & 63u) == 0) s_sum[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t s = 0;
    for (int k = 0; k < kTileWords / 64; ++k) s += s_sum[k];
    tile_counts[blockIdx.x] = s;
  }
}

__global__ __launch_bounds__(kTileWords) void row_write_kernel(
    const uint64_t *__restrict__ row, uint64_t words,
    const uint64_t *__restrict__ tile_offsets, int64_t *__restrict__ out, int64_t capacity) {
  __shared__ uint32_t s_wscan[kTileWords / 64];
  const uint32_t t = threadIdx.x, lane = t & 63u, wv = t >> 6;
  const uint64_t w = (uint64_t)blockIdx.x * kTileWords + t;
  uint64_t m = w < words ? row[w] : 0;
  const uint32_t c = (uint32_t)__popcll(m);
  const uint32_t incl = wave_incl_scan(c);
  if (lane == 63) s_wscan[wv] = incl;
  __syncthreads();
  int64_t o = (int64_t)(incl - c);
  for (uint32_t k = 0; k < wv; ++k) o += s_wscan[k];
  o += (int64_t)tile_offsets[blockIdx.x];
  while (m) {
    const int b = __builtin_ctzll(m);
    m &= m - 1;
    if (o < capacity) out[o] = (int64_t)(w * 64 + (uint64_t)b);
    ++o;
  }
}

// ---- debug: escape and quote bitmaps (true carry semantics) ----------------
__global__ __launch_bounds__(256) void escape_quote_kernel(
    const uint8_t *__restrict__ in, uint64_t n, uint64_t words,
    uint64_t *__restrict__ escape_out, uint64_t *__restrict__ quote_out) {
  const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= words) return;
  const uint64_t pos = w * 64;
  uint32_t x[16];
  load_word(in, n, pos, false, x);
  const WordClasses m = classify<false>(x);
  // carry-in: parity of the backslash run ending right before this word
  uint32_t cin = 0;
  for (uint64_t p = pos; p > 0 && in[p - 1] == '\\'; --p) cin ^= 1u;
  const uint64_t B = m.backslash;
  const int lead = (B == ~0ull) ? 64 : __builtin_ctzll(~B);
  uint64_t e = escaped_from_zero(B);
  if (cin) e ^= mask_upto(lead);
  const uint64_t valid = (pos + 64 <= n) ? ~0ull : low_mask((int)(n - pos));
  escape_out[w] = e & valid;
  quote_out[w] = m.quote & ~e & valid;
}

} // namespace gpj
