// reduce.hpp -- block-wide sum (crystal/reduce.cuh:25-75), re-derived for
// 64-lane wavefronts: a 6-step shuffle tree inside each wave, one partial per
// wave in `shared` (at least BLOCK_THREADS/64 entries), then a tree over the
// partials in wave 0.  The result is valid in thread 0 only, as in the
// reference.  Every thread of the block must call it (two barriers).
#pragma once

#include <hip/hip_runtime.h>

namespace lj {
template <typename T>
__device__ __forceinline__ T WaveSum(T val) {
#pragma unroll
  for (int offset = 32; offset > 0; offset >>= 1) val += __shfl_down(val, offset, 64);
  return val;  // valid in lane 0
}
}  // namespace lj

template <typename T, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ T BlockSum(T item, T* shared) {
  static_assert(BLOCK_THREADS % 64 == 0, "BLOCK_THREADS must be a multiple of the 64-lane wavefront");
  __syncthreads();
  T val = lj::WaveSum(item);
  const int lane = threadIdx.x % 64;
  const int wid = threadIdx.x / 64;
  if (lane == 0) shared[wid] = val;
  __syncthreads();
  val = (threadIdx.x < BLOCK_THREADS / 64) ? shared[lane] : T(0);
  if (wid == 0) val = lj::WaveSum(val);
  return val;
}

template <typename T, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ T BlockSum(T (&items)[ITEMS_PER_THREAD], T* shared) {
  T thread_sum = 0;
#pragma unroll
  for (int ITEM = 0; ITEM < ITEMS_PER_THREAD; ITEM++) thread_sum += items[ITEM];
  return BlockSum<T, BLOCK_THREADS, ITEMS_PER_THREAD>(thread_sum, shared);
}
