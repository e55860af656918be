// store.hpp -- striped block-wide tile store (crystal/store.cuh:25-69):
// thread t writes items[i] to tile[t + i*BLOCK_THREADS], guarded by num_items
// on a partial tile.
#pragma once

#include <hip/hip_runtime.h>

template <typename T, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockStoreDirect(int tid, T* block_itr, T (&items)[ITEMS_PER_THREAD]) {
  T* thread_itr = block_itr + tid;
#pragma unroll
  for (int ITEM = 0; ITEM < ITEMS_PER_THREAD; ITEM++) thread_itr[ITEM * BLOCK_THREADS] = items[ITEM];
}

template <typename T, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockStoreDirect(int tid, T* block_itr, T (&items)[ITEMS_PER_THREAD],
                                                 int num_items) {
  T* thread_itr = block_itr + tid;
#pragma unroll
  for (int ITEM = 0; ITEM < ITEMS_PER_THREAD; ITEM++)
    if (tid + ITEM * BLOCK_THREADS < num_items) thread_itr[ITEM * BLOCK_THREADS] = items[ITEM];
}

template <typename T, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockStore(T* out, T (&items)[ITEMS_PER_THREAD], int num_items) {
  if (BLOCK_THREADS * ITEMS_PER_THREAD == num_items)
    BlockStoreDirect<T, BLOCK_THREADS, ITEMS_PER_THREAD>(threadIdx.x, out, items);
  else
    BlockStoreDirect<T, BLOCK_THREADS, ITEMS_PER_THREAD>(threadIdx.x, out, items, num_items);
}
