// hash.hpp -- hash functions of the star-join path, shared by host and device.
//
//  * pht_slot   : HASH(X,Y,Z) ((X-Z) % Y)   crystal/join.cuh:27
//  * bloom MOD  : HASH1(X,Y,Z) ((X-Z) % (Y)) crystal/bloom.cuh:7, Y = words*32
//  * bloom MULT : multiplicative hashing (Knuth, TAOCP vol. 3, sec. 6.4):
//                 h_i(x) = (x * A_i mod 2^32) >> (32 - log2 m); not in the
//                 reference, defined here.
//
// The reference computes X-Z in 32-bit int and takes the C remainder, so a key
// below keys_min yields a negative slot and an out-of-bounds access.  Here the
// difference is exact (64-bit) and a negative result is reported as "no slot"
// (-1): a probe misses, a build skips the key and counts it.
//
// Remainders by a runtime divisor use the direct-remainder method of Lemire,
// Kaser and Kurz, "Faster Remainder by Direct Computation" (2019): with
// M = floor((2^64-1)/d) + 1, a % d = mulhi64(M*a mod 2^64, d) for all 32-bit a, d.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define LJ_HD __host__ __device__ __forceinline__
#else
#define LJ_HD inline
#endif

namespace lj {

struct FastMod {
  uint64_t m;  // floor((2^64-1)/d) + 1
  uint32_t d;  // divisor (>= 1)
  uint32_t pow2_mask;  // d-1 when d is a power of two, else 0
};

inline FastMod make_fastmod(uint32_t d) {
  FastMod f;
  f.d = d ? d : 1;
  f.m = UINT64_MAX / f.d + 1;
  f.pow2_mask = ((f.d & (f.d - 1)) == 0) ? f.d - 1 : 0u;
  return f;
}

LJ_HD uint64_t mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul64hi(a, b);
#else
  return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

LJ_HD uint32_t fastmod_u32(uint32_t a, const FastMod& f) {
  return (uint32_t)mulhi64(f.m * (uint64_t)a, f.d);
}

// HASH(key, len, keys_min) of crystal/join.cuh:27; -1 when key < keys_min.
LJ_HD int64_t pht_slot(int32_t key, const FastMod& len, int32_t keys_min) {
  int64_t x = (int64_t)key - (int64_t)keys_min;
  if (x < 0) return -1;
  return (int64_t)fastmod_u32((uint32_t)x, len);
}

// Hash-table kinds of a dimension: the reference's perfect hashes (PHT_2
// pairs, PHT_1 keys; crystal/join.cuh) and a textbook open-addressed table.
enum HtKind : int32_t { HT_PHT2 = 0, HT_PHT1 = 1, HT_TEXTBOOK = 2 };

// Textbook table: home slot of a key among mask + 1 (a power of two) slots,
// Knuth's multiplicative hash (TAOCP vol. 3, sec. 6.4); keys are stored as
// key ^ 0x80000000 so that an all-zero table is empty.
LJ_HD uint32_t tb_home(int32_t key, uint32_t mask) {
#if defined(__HIP_DEVICE_COMPILE__)
  const uint32_t lg = (uint32_t)__popc(mask);
#else
  const uint32_t lg = (uint32_t)__builtin_popcount(mask);
#endif
  return lg == 0 ? 0u : ((uint32_t)key * 0x9E3779B1u) >> (32 - lg);
}
LJ_HD uint32_t tb_encode(int32_t key) { return (uint32_t)key ^ 0x80000000u; }
constexpr int32_t kTbExcludedKey = (int32_t)0x80000000;  // INT32_MIN encodes to the empty slot

// Multipliers of the multiplicative Bloom family: Knuth's golden-ratio
// constant and the three odd 32-bit xxHash primes.
#define LJ_MULT_A0 0x9E3779B1u
#define LJ_MULT_A1 0x85EBCA77u
#define LJ_MULT_A2 0xC2B2AE3Du
#define LJ_MULT_A3 0x27D4EB2Fu

LJ_HD uint32_t mult_const(int i) {
  return i == 0 ? LJ_MULT_A0 : i == 1 ? LJ_MULT_A1 : i == 2 ? LJ_MULT_A2 : LJ_MULT_A3;
}

enum BloomKind : int32_t {
  BLOOM_NONE = 0,
  BLOOM_MOD_POW2 = 1,   // HASH1 with words*32 a power of two: h = x & (m-1)
  BLOOM_MOD = 2,        // HASH1, general m: fast remainder
  BLOOM_MULT_STD = 3,   // multiplicative, k positions over the array
  BLOOM_MULT_BLOCKED = 4// multiplicative, k positions inside one 32-bit word
};

// Device-side description of one Bloom filter's hash.
struct BloomHash {
  int32_t kind;      // BloomKind
  int32_t k;         // hashes per key (1..3)
  int32_t keys_min;
  uint32_t shift;    // MULT_STD: 32 - log2(m); MULT_BLOCKED: 32 - log2(m/32)
  FastMod mod;       // MOD: m = words*32
};

// Bit positions of `key`; returns the number of positions (0: no position,
// i.e. a key below keys_min under HASH1).
LJ_HD int bloom_bits(const BloomHash& b, int32_t key, uint32_t (&bits)[3]) {
  if (b.kind == BLOOM_MOD_POW2 || b.kind == BLOOM_MOD) {
    int64_t x = (int64_t)key - (int64_t)b.keys_min;
    if (x < 0) return 0;
    bits[0] = b.kind == BLOOM_MOD_POW2 ? ((uint32_t)x & b.mod.pow2_mask) : fastmod_u32((uint32_t)x, b.mod);
    return 1;
  }
  uint32_t x = (uint32_t)key - (uint32_t)b.keys_min;
  if (b.kind == BLOOM_MULT_STD) {
    for (int i = 0; i < 3; i++)
      if (i < b.k) bits[i] = (x * mult_const(i)) >> b.shift;
    return b.k;
  }
  if (b.kind == BLOOM_MULT_BLOCKED) {
    uint32_t w = (x * mult_const(0)) >> b.shift;
    for (int i = 0; i < 3; i++)
      if (i < b.k) bits[i] = (w << 5) | ((x * mult_const(i + 1)) >> 27);
    return b.k;
  }
  return 0;
}

// Word index and in-word mask of a BLOCKED key (one word holds all k bits).
LJ_HD bool bloom_blocked_word(const BloomHash& b, int32_t key, uint32_t& word, uint32_t& mask) {
  uint32_t x = (uint32_t)key - (uint32_t)b.keys_min;
  word = (x * mult_const(0)) >> b.shift;
  mask = 0;
  for (int i = 0; i < 3; i++)
    if (i < b.k) mask |= 1u << ((x * mult_const(i + 1)) >> 27);
  return true;
}

}  // namespace lj
