// gpj_bytecode.h -- query bytecode opcodes and the wave-uniform bytecode
// reader shared by the query kernels (device code only).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gpj {
namespace {

enum : uint32_t {
  kEnd = 0, kStoreResult = 1, kMoveUp = 2, kMoveDown = 3, kMoveToKey = 4,
  kMoveToIndex = 5, kMoveToIndexReverse = 6, kMarkPosition = 7,
  kResetPosition = 8, kExpressionStringEquals = 9,
  kMoveToEach = 10, kJumpIfNotArray = 11, // extension opcodes
  kOffEnd = 256                           // past the bytecode: ends the path
};
constexpr int kMaxFrames = 8; // element iterations per record (as the oracle)

typedef const __attribute__((address_space(4))) uint32_t *ConstU32;

// Wave-uniform reads of one query's bytecode through the scalar data cache,
// in 32-bit arithmetic (bytecode positions stay below 2^32).  Only aligned
// dwords that hold at least one byte of the buffer are loaded; bytes at or
// past `len` read as 0.
struct IrReader {
  ConstU32 words; // aligned dwords covering the bytecode
  uint32_t head;  // bytecode address modulo 4
  uint32_t len;
  __device__ static __forceinline__ IrReader make(const uint8_t *ir, uint32_t len) {
    const uintptr_t a = reinterpret_cast<uintptr_t>(ir);
    return IrReader{reinterpret_cast<ConstU32>(a & ~(uintptr_t)3), (uint32_t)(a & 3), len};
  }
  __device__ __forceinline__ uint32_t dword(uint32_t w) const {
    return w < ((head + len + 3) >> 2) ? words[w] : 0u;
  }
  __device__ __forceinline__ uint32_t byte(uint32_t p) const {
    if (p >= len) return 0;
    const uint32_t o = p + head;
    return (words[o >> 2] >> (8 * (o & 3))) & 0xFFu;
  }
  // Little-endian 8 bytes from p (bytes at or past len read as 0).
  __device__ __forceinline__ uint64_t u64(uint32_t p) const {
    if (p >= len) return 0;
    const uint32_t o = p + head, w = o >> 2, sh = 8 * (o & 3);
    const uint64_t lo = (uint64_t)dword(w) | ((uint64_t)dword(w + 1) << 32);
    const uint64_t hi = dword(w + 2);
    uint64_t v = sh ? ((lo >> sh) | (hi << (64 - sh))) : lo;
    const uint32_t valid = len - p;
    if (valid < 8) v &= (1ull << (8 * valid)) - 1;
    return v;
  }
  __device__ __forceinline__ int32_t varint(uint32_t &qp) const {
    uint32_t value = 0, b = 0;
    int shift = 0;
    for (;;) {
      b = byte(qp++);
      if (!(b & 0x80u)) break;
      value |= (b & 0x7Fu) << shift;
      shift += 7;
    }
    return (int32_t)(value | (b << shift));
  }
};

} // namespace
} // namespace gpj
