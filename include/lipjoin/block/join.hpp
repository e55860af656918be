// join.hpp -- perfect-hash join build/probe (crystal/join.cuh:27-335), wave64.
//
// PHT_1: int32 ht[len], slot = (key-keys_min) % len, 0 = empty.
//   build: atomicCAS(&ht[slot], 0, key)          (join.cuh:191-258)
//   probe: flag = (ht[slot] != 0), no key compare (join.cuh:29-106)
// PHT_2: int32 ht[2*len] of (key, val) pairs.
//   build: atomicCAS(&ht[2*slot], 0, key); ht[2*slot+1] = val, unconditionally
//          (join.cuh:260-335; last writer wins on a shared slot)
//   probe: one 64-bit load of the pair; != 0 -> payload = val, else flag = 0
//          (join.cuh:108-189)
// Only items with a set flag are touched; items at or beyond num_items are
// ignored.  Difference from the reference: a key below keys_min (negative C
// remainder, an out-of-bounds access there) misses on probe and is skipped on
// build.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hash.hpp"

namespace lj {
__device__ __forceinline__ int64_t pht_slot_div(int32_t key, int ht_len, int32_t keys_min) {
  int64_t x = (int64_t)key - (int64_t)keys_min;
  if (x < 0 || ht_len <= 0) return -1;
  return (int64_t)((uint32_t)x % (uint32_t)ht_len);
}

// 64-bit PHT_2 probe of one key: returns true and the payload on a hit.
__device__ __forceinline__ bool pht2_probe(const int32_t* ht, const FastMod& len, int32_t keys_min,
                                           int32_t key, int32_t& payload) {
  int64_t s = pht_slot(key, len, keys_min);
  if (s < 0) return false;
  uint64_t slot = *reinterpret_cast<const uint64_t*>(ht + 2 * s);
  if (slot == 0) return false;
  payload = (int32_t)(uint32_t)(slot >> 32);
  return true;
}
}  // namespace lj

template <typename K, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockProbeAndPHT_1(K (&items)[ITEMS_PER_THREAD],
                                                   int (&selection_flags)[ITEMS_PER_THREAD], const K* ht,
                                                   int ht_len, K keys_min, int num_items) {
#pragma unroll
  for (int ITEM = 0; ITEM < ITEMS_PER_THREAD; ITEM++) {
    if ((int)(threadIdx.x + ITEM * BLOCK_THREADS) < num_items && selection_flags[ITEM]) {
      int64_t s = lj::pht_slot_div(items[ITEM], ht_len, keys_min);
      selection_flags[ITEM] = (s >= 0 && ht[s] != 0) ? 1 : 0;
    }
  }
}

template <typename K, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockProbeAndPHT_1(K (&items)[ITEMS_PER_THREAD],
                                                   int (&selection_flags)[ITEMS_PER_THREAD], const K* ht,
                                                   int ht_len, int num_items) {
  BlockProbeAndPHT_1<K, BLOCK_THREADS, ITEMS_PER_THREAD>(items, selection_flags, ht, ht_len, 0, num_items);
}

template <typename K, typename V, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockProbeAndPHT_2(K (&keys)[ITEMS_PER_THREAD], V (&res)[ITEMS_PER_THREAD],
                                                   int (&selection_flags)[ITEMS_PER_THREAD], const K* ht,
                                                   int ht_len, K keys_min, int num_items) {
#pragma unroll
  for (int ITEM = 0; ITEM < ITEMS_PER_THREAD; ITEM++) {
    if ((int)(threadIdx.x + ITEM * BLOCK_THREADS) < num_items && selection_flags[ITEM]) {
      int64_t s = lj::pht_slot_div(keys[ITEM], ht_len, keys_min);
      uint64_t slot = s >= 0 ? *reinterpret_cast<const uint64_t*>(ht + 2 * s) : 0ull;
      if (slot != 0) res[ITEM] = (V)(int32_t)(uint32_t)(slot >> 32);
      else selection_flags[ITEM] = 0;
    }
  }
}

template <typename K, typename V, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockProbeAndPHT_2(K (&keys)[ITEMS_PER_THREAD], V (&res)[ITEMS_PER_THREAD],
                                                   int (&selection_flags)[ITEMS_PER_THREAD], const K* ht,
                                                   int ht_len, int num_items) {
  BlockProbeAndPHT_2<K, V, BLOCK_THREADS, ITEMS_PER_THREAD>(keys, res, selection_flags, ht, ht_len, 0, num_items);
}

template <typename K, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockBuildSelectivePHT_1(K (&keys)[ITEMS_PER_THREAD],
                                                         int (&selection_flags)[ITEMS_PER_THREAD], K* ht,
                                                         int ht_len, K keys_min, int num_items) {
#pragma unroll
  for (int ITEM = 0; ITEM < ITEMS_PER_THREAD; ITEM++) {
    if ((int)(threadIdx.x + ITEM * BLOCK_THREADS) < num_items && selection_flags[ITEM]) {
      int64_t s = lj::pht_slot_div(keys[ITEM], ht_len, keys_min);
      if (s >= 0) atomicCAS(&ht[s], (K)0, keys[ITEM]);
    }
  }
}

template <typename K, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockBuildSelectivePHT_1(K (&keys)[ITEMS_PER_THREAD],
                                                         int (&selection_flags)[ITEMS_PER_THREAD], K* ht,
                                                         int ht_len, int num_items) {
  BlockBuildSelectivePHT_1<K, BLOCK_THREADS, ITEMS_PER_THREAD>(keys, selection_flags, ht, ht_len, 0, num_items);
}

template <typename K, typename V, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockBuildSelectivePHT_2(K (&keys)[ITEMS_PER_THREAD], V (&res)[ITEMS_PER_THREAD],
                                                         int (&selection_flags)[ITEMS_PER_THREAD], K* ht,
                                                         int ht_len, K keys_min, int num_items) {
#pragma unroll
  for (int ITEM = 0; ITEM < ITEMS_PER_THREAD; ITEM++) {
    if ((int)(threadIdx.x + ITEM * BLOCK_THREADS) < num_items && selection_flags[ITEM]) {
      int64_t s = lj::pht_slot_div(keys[ITEM], ht_len, keys_min);
      if (s >= 0) {
        atomicCAS(&ht[2 * s], (K)0, keys[ITEM]);
        ht[2 * s + 1] = (K)res[ITEM];
      }
    }
  }
}

template <typename K, typename V, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockBuildSelectivePHT_2(K (&keys)[ITEMS_PER_THREAD], V (&res)[ITEMS_PER_THREAD],
                                                         int (&selection_flags)[ITEMS_PER_THREAD], K* ht,
                                                         int ht_len, int num_items) {
  BlockBuildSelectivePHT_2<K, V, BLOCK_THREADS, ITEMS_PER_THREAD>(keys, res, selection_flags, ht, ht_len, 0,
                                                                  num_items);
}
