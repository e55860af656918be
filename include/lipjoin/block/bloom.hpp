// bloom.hpp -- Bloom filter build/probe (crystal/bloom.cuh:7-174), wave64.
//
// BlockBuildBloomFilter / BlockProbeBloomFilter keep the reference's
// signatures: `bloom_filter_size` is in 32-bit words and the bit of a key is
// HASH1(key, words*32, keys_min) = (key-keys_min) % (words*32); build sets it
// with atomicOr, probe ANDs "bit is set" into the flag of selected items.
// The BloomHash overloads take any hash of hash.hpp (MOD, or the
// multiplicative family with k = 1..3, standard or one-word-blocked) and a
// filter that may live in LDS or in global memory.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hash.hpp"

namespace lj {

// True iff every bit of `key` is set in `bf` (global or LDS pointer).
// `Ptr` is a pointer to the filter's words: a plain pointer, or one qualified
// with the LDS / global address space so that the reads compile to ds_read /
// global_load rather than flat loads (a flat load waits on every outstanding
// memory operation of the wave).
template <typename Ptr>
__device__ __forceinline__ bool bloom_test(const BloomHash& b, Ptr bf, int32_t key) {
  if (b.kind == BLOOM_MULT_BLOCKED) {
    uint32_t w, m;
    bloom_blocked_word(b, key, w, m);
    return (bf[w] & m) == m;
  }
  uint32_t bits[3];
  int k = bloom_bits(b, key, bits);
  if (k == 0) return false;
  bool ok = true;
  for (int i = 0; i < 3; i++)
    if (i < k) ok = ok && ((bf[bits[i] >> 5] >> (bits[i] & 31)) & 1u);
  return ok;
}

// Sets the bits of `key`; returns false when the key has no bit (HASH1 below keys_min).
__device__ __forceinline__ bool bloom_insert(const BloomHash& b, uint32_t* bf, int32_t key) {
  if (b.kind == BLOOM_MULT_BLOCKED) {
    uint32_t w, m;
    bloom_blocked_word(b, key, w, m);
    atomicOr(&bf[w], m);
    return true;
  }
  uint32_t bits[3];
  int k = bloom_bits(b, key, bits);
  for (int i = 0; i < 3; i++)
    if (i < k) atomicOr(&bf[bits[i] >> 5], 1u << (bits[i] & 31));
  return k > 0;
}

__device__ __forceinline__ BloomHash bloom_mod_hash(int bloom_filter_size, int32_t keys_min) {
  BloomHash b;
  b.kind = BLOOM_MOD;
  b.k = 1;
  b.keys_min = keys_min;
  b.shift = 0;
  uint32_t m = (uint32_t)bloom_filter_size * 32u;
  b.mod.d = m ? m : 1u;
  b.mod.m = 0;
  b.mod.pow2_mask = ((m & (m - 1)) == 0 && m) ? m - 1 : 0u;
  return b;
}

__device__ __forceinline__ int64_t bloom_mod_bit(int32_t key, int bloom_filter_size, int32_t keys_min) {
  int64_t x = (int64_t)key - (int64_t)keys_min;
  int64_t m = (int64_t)bloom_filter_size * 32;
  if (x < 0 || m <= 0) return -1;
  return x % m;
}

}  // namespace lj

template <typename K, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockBuildBloomFilter(K (&keys)[ITEMS_PER_THREAD],
                                                      int (&selection_flags)[ITEMS_PER_THREAD],
                                                      uint32_t* bloom_filter, int bloom_filter_size, K keys_min,
                                                      int num_items) {
#pragma unroll
  for (int ITEM = 0; ITEM < ITEMS_PER_THREAD; ITEM++) {
    if ((int)(threadIdx.x + ITEM * BLOCK_THREADS) < num_items && selection_flags[ITEM]) {
      int64_t h = lj::bloom_mod_bit(keys[ITEM], bloom_filter_size, keys_min);
      if (h >= 0) atomicOr(&bloom_filter[h >> 5], 1u << (h & 31));
    }
  }
}

template <typename K, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockBuildBloomFilter(K (&keys)[ITEMS_PER_THREAD],
                                                      int (&selection_flags)[ITEMS_PER_THREAD],
                                                      uint32_t* bloom_filter, int bloom_filter_size, int num_items) {
  BlockBuildBloomFilter<K, BLOCK_THREADS, ITEMS_PER_THREAD>(keys, selection_flags, bloom_filter, bloom_filter_size,
                                                            0, num_items);
}

template <typename K, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockProbeBloomFilter(K (&keys)[ITEMS_PER_THREAD],
                                                      int (&selection_flags)[ITEMS_PER_THREAD],
                                                      const uint32_t* bloom_filter, int bloom_filter_size,
                                                      K keys_min, int num_items) {
#pragma unroll
  for (int ITEM = 0; ITEM < ITEMS_PER_THREAD; ITEM++) {
    if ((int)(threadIdx.x + ITEM * BLOCK_THREADS) < num_items && selection_flags[ITEM]) {
      int64_t h = lj::bloom_mod_bit(keys[ITEM], bloom_filter_size, keys_min);
      selection_flags[ITEM] = (h >= 0 && ((bloom_filter[h >> 5] >> (h & 31)) & 1u)) ? 1 : 0;
    }
  }
}

template <typename K, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockProbeBloomFilter(K (&keys)[ITEMS_PER_THREAD],
                                                      int (&selection_flags)[ITEMS_PER_THREAD],
                                                      const uint32_t* bloom_filter, int bloom_filter_size,
                                                      int num_items) {
  BlockProbeBloomFilter<K, BLOCK_THREADS, ITEMS_PER_THREAD>(keys, selection_flags, bloom_filter, bloom_filter_size,
                                                            0, num_items);
}

// Any-hash forms (MOD or multiplicative; global or LDS filter).
template <typename K, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockBuildBloomFilter(K (&keys)[ITEMS_PER_THREAD],
                                                      int (&selection_flags)[ITEMS_PER_THREAD],
                                                      uint32_t* bloom_filter, const lj::BloomHash& hash,
                                                      int num_items) {
#pragma unroll
  for (int ITEM = 0; ITEM < ITEMS_PER_THREAD; ITEM++)
    if ((int)(threadIdx.x + ITEM * BLOCK_THREADS) < num_items && selection_flags[ITEM])
      lj::bloom_insert(hash, bloom_filter, keys[ITEM]);
}

template <typename K, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockProbeBloomFilter(K (&keys)[ITEMS_PER_THREAD],
                                                      int (&selection_flags)[ITEMS_PER_THREAD],
                                                      const uint32_t* bloom_filter, const lj::BloomHash& hash,
                                                      int num_items) {
#pragma unroll
  for (int ITEM = 0; ITEM < ITEMS_PER_THREAD; ITEM++)
    if ((int)(threadIdx.x + ITEM * BLOCK_THREADS) < num_items && selection_flags[ITEM])
      selection_flags[ITEM] = lj::bloom_test(hash, bloom_filter, keys[ITEM]) ? 1 : 0;
}
