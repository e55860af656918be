// load.hpp -- block-wide tile loads (crystal/load.cuh:25-116), wave64.
//
// Same names, template parameters and striped arrangement as the reference:
// thread t holds items[i] = tile[t + i*BLOCK_THREADS]; a partial last tile is
// guarded by num_items; BlockPredLoad skips items whose flag is 0 and leaves
// their registers untouched.  With BLOCK_THREADS a multiple of 64 every
// item-instruction of a wave covers 64 consecutive int32 (256 B, two 128-B
// lines), so the striped form is fully coalesced on CDNA.
//
// Native additions: WaveLoadVec4 / WavePredLoadVec4 give each lane 4
// consecutive rows through one 16-byte load (a wave covers 1 KiB per
// instruction), optionally non-temporal so a streamed fact column does not
// evict the hash tables and filters from L2/MALL.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lj {
typedef int32_t v4i32 __attribute__((ext_vector_type(4)));
}

template <typename T, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockPredLoadDirect(const unsigned int tid, const T* block_itr,
                                                    T (&items)[ITEMS_PER_THREAD],
                                                    int (&selection_flags)[ITEMS_PER_THREAD]) {
  const T* thread_itr = block_itr + tid;
#pragma unroll
  for (int ITEM = 0; ITEM < ITEMS_PER_THREAD; ITEM++)
    if (selection_flags[ITEM]) items[ITEM] = thread_itr[ITEM * BLOCK_THREADS];
}

template <typename T, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockPredLoadDirect(const unsigned int tid, const T* block_itr,
                                                    T (&items)[ITEMS_PER_THREAD], int num_items,
                                                    int (&selection_flags)[ITEMS_PER_THREAD]) {
  const T* thread_itr = block_itr + tid;
#pragma unroll
  for (int ITEM = 0; ITEM < ITEMS_PER_THREAD; ITEM++)
    if (selection_flags[ITEM] && (int)(tid + ITEM * BLOCK_THREADS) < num_items)
      items[ITEM] = thread_itr[ITEM * BLOCK_THREADS];
}

template <typename T, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockPredLoad(const T* inp, T (&items)[ITEMS_PER_THREAD], int num_items,
                                              int (&selection_flags)[ITEMS_PER_THREAD]) {
  if (BLOCK_THREADS * ITEMS_PER_THREAD == num_items)
    BlockPredLoadDirect<T, BLOCK_THREADS, ITEMS_PER_THREAD>(threadIdx.x, inp, items, selection_flags);
  else
    BlockPredLoadDirect<T, BLOCK_THREADS, ITEMS_PER_THREAD>(threadIdx.x, inp, items, num_items,
                                                            selection_flags);
}

template <typename T, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockLoadDirect(const unsigned int tid, const T* block_itr,
                                                T (&items)[ITEMS_PER_THREAD]) {
  const T* thread_itr = block_itr + tid;
#pragma unroll
  for (int ITEM = 0; ITEM < ITEMS_PER_THREAD; ITEM++) items[ITEM] = thread_itr[ITEM * BLOCK_THREADS];
}

template <typename T, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockLoadDirect(const unsigned int tid, const T* block_itr,
                                                T (&items)[ITEMS_PER_THREAD], int num_items) {
  const T* thread_itr = block_itr + tid;
#pragma unroll
  for (int ITEM = 0; ITEM < ITEMS_PER_THREAD; ITEM++)
    if ((int)(tid + ITEM * BLOCK_THREADS) < num_items) items[ITEM] = thread_itr[ITEM * BLOCK_THREADS];
}

template <typename T, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockLoad(const T* inp, T (&items)[ITEMS_PER_THREAD], int num_items) {
  if (BLOCK_THREADS * ITEMS_PER_THREAD == num_items)
    BlockLoadDirect<T, BLOCK_THREADS, ITEMS_PER_THREAD>(threadIdx.x, inp, items);
  else
    BlockLoadDirect<T, BLOCK_THREADS, ITEMS_PER_THREAD>(threadIdx.x, inp, items, num_items);
}

namespace lj {

// The 16-byte vector pointer of a column pointer, in the same address space.
template <typename ColPtr>
struct Vec4PtrOf;
template <>
struct Vec4PtrOf<const int32_t*> {
  typedef const v4i32* type;
};
template <>
struct Vec4PtrOf<int32_t*> {
  typedef const v4i32* type;
};
template <>
struct Vec4PtrOf<const __attribute__((address_space(1))) int32_t*> {
  typedef const __attribute__((address_space(1))) v4i32* type;
};

// Rows [row0, row0+4) of `col` into v (row0 % 4 == 0, col 16-byte aligned).
// Rows at or beyond n are not read and come back as 0.  `col` is a plain
// column pointer or one typed with the global address space (the loads are
// then global loads rather than flat ones).
template <bool NONTEMPORAL, typename ColPtr>
__device__ __forceinline__ v4i32 WaveLoadVec4(ColPtr col, int64_t row0, int64_t n) {
  v4i32 v;
  if (row0 + 4 <= n) {
    const typename Vec4PtrOf<ColPtr>::type p = reinterpret_cast<typename Vec4PtrOf<ColPtr>::type>(col + row0);
    if (NONTEMPORAL) v = __builtin_nontemporal_load(p);
    else v = *p;
  } else {
    v = v4i32{0, 0, 0, 0};
    if (row0 + 0 < n) v.x = col[row0 + 0];
    if (row0 + 1 < n) v.y = col[row0 + 1];
    if (row0 + 2 < n) v.z = col[row0 + 2];
  }
  return v;
}

// As WaveLoadVec4, but only when any of the 4 rows is alive (bit j of mask4).
template <typename ColPtr>
__device__ __forceinline__ v4i32 WavePredLoadVec4(ColPtr col, int64_t row0, int64_t n, uint32_t mask4) {
  v4i32 v = v4i32{0, 0, 0, 0};
  if (mask4) v = WaveLoadVec4<false>(col, row0, n);
  return v;
}

__device__ __forceinline__ int32_t vget(const v4i32& v, int j) {
  return j == 0 ? v.x : j == 1 ? v.y : j == 2 ? v.z : v.w;
}

}  // namespace lj
