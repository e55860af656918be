// term.hpp -- early termination test (crystal/term.cuh:25-33).
//
// IsTerm is per thread, as in the reference: true when none of this thread's
// items is selected.  The reference's callers `return` on it, which is only
// safe when no barrier follows (bloom_join.cu:91 returns before BlockSum's
// __syncthreads).  IsTermWave is the barrier-safe form: one wave-wide vote, so
// a whole wave leaves together, and callers that still reach a block barrier
// must use BlockAnyAlive instead.
#pragma once

#include <hip/hip_runtime.h>

template <typename T, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ bool IsTerm(int (&selection_flags)[ITEMS_PER_THREAD]) {
  int count = 0;
#pragma unroll
  for (int ITEM = 0; ITEM < ITEMS_PER_THREAD; ITEM++) count += selection_flags[ITEM];
  return count == 0;
}

template <typename T, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ bool IsTermWave(int (&selection_flags)[ITEMS_PER_THREAD]) {
  int any = 0;
#pragma unroll
  for (int ITEM = 0; ITEM < ITEMS_PER_THREAD; ITEM++) any |= selection_flags[ITEM];
  return __ballot(any != 0) == 0;
}

// True in every thread iff some thread of the block still has a selected item.
template <int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ bool BlockAnyAlive(int (&selection_flags)[ITEMS_PER_THREAD]) {
  int any = 0;
#pragma unroll
  for (int ITEM = 0; ITEM < ITEMS_PER_THREAD; ITEM++) any |= selection_flags[ITEM];
  return __syncthreads_or(any) != 0;
}
