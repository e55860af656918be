// pred.hpp -- selection flags and predicates (crystal/pred.cuh:25-356).
//
// InitFlags sets every flag; BlockPred* overwrites the flags with the functor
// result, BlockPredAnd* ANDs it in, BlockPredOr* ORs it in; items at or beyond
// num_items keep their flags.  Functors and the BlockPredLT ... BlockPredOrEQ
// wrappers have the reference's names and argument order.  PredProgram is the
// runtime form used by the C-ABI descriptors (a list of such calls).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

template <int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void InitFlags(int (&selection_flags)[ITEMS_PER_THREAD]) {
#pragma unroll
  for (int ITEM = 0; ITEM < ITEMS_PER_THREAD; ITEM++) selection_flags[ITEM] = 1;
}

template <typename T, typename SelectOp, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockPredDirect(int tid, T (&items)[ITEMS_PER_THREAD], SelectOp select_op,
                                                int (&selection_flags)[ITEMS_PER_THREAD], int num_items) {
#pragma unroll
  for (int ITEM = 0; ITEM < ITEMS_PER_THREAD; ITEM++)
    if (tid + ITEM * BLOCK_THREADS < num_items) selection_flags[ITEM] = select_op(items[ITEM]);
}

template <typename T, typename SelectOp, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockPred(T (&items)[ITEMS_PER_THREAD], SelectOp select_op,
                                          int (&selection_flags)[ITEMS_PER_THREAD], int num_items) {
  BlockPredDirect<T, SelectOp, BLOCK_THREADS, ITEMS_PER_THREAD>(threadIdx.x, items, select_op, selection_flags,
                                                               num_items);
}

template <typename T, typename SelectOp, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockPredAndDirect(int tid, T (&items)[ITEMS_PER_THREAD], SelectOp select_op,
                                                   int (&selection_flags)[ITEMS_PER_THREAD], int num_items) {
#pragma unroll
  for (int ITEM = 0; ITEM < ITEMS_PER_THREAD; ITEM++)
    if (tid + ITEM * BLOCK_THREADS < num_items)
      selection_flags[ITEM] = selection_flags[ITEM] && select_op(items[ITEM]);
}

template <typename T, typename SelectOp, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockPredAnd(T (&items)[ITEMS_PER_THREAD], SelectOp select_op,
                                             int (&selection_flags)[ITEMS_PER_THREAD], int num_items) {
  BlockPredAndDirect<T, SelectOp, BLOCK_THREADS, ITEMS_PER_THREAD>(threadIdx.x, items, select_op,
                                                                  selection_flags, num_items);
}

template <typename T, typename SelectOp, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockPredOrDirect(int tid, T (&items)[ITEMS_PER_THREAD], SelectOp select_op,
                                                  int (&selection_flags)[ITEMS_PER_THREAD], int num_items) {
#pragma unroll
  for (int ITEM = 0; ITEM < ITEMS_PER_THREAD; ITEM++)
    if (tid + ITEM * BLOCK_THREADS < num_items)
      selection_flags[ITEM] = selection_flags[ITEM] || select_op(items[ITEM]);
}

template <typename T, typename SelectOp, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockPredOr(T (&items)[ITEMS_PER_THREAD], SelectOp select_op,
                                            int (&selection_flags)[ITEMS_PER_THREAD], int num_items) {
  BlockPredOrDirect<T, SelectOp, BLOCK_THREADS, ITEMS_PER_THREAD>(threadIdx.x, items, select_op,
                                                                 selection_flags, num_items);
}

template <typename T> struct LessThan {
  T compare;
  __device__ __forceinline__ LessThan(T c) : compare(c) {}
  __device__ __forceinline__ bool operator()(const T& a) const { return a < compare; }
};
template <typename T> struct GreaterThan {
  T compare;
  __device__ __forceinline__ GreaterThan(T c) : compare(c) {}
  __device__ __forceinline__ bool operator()(const T& a) const { return a > compare; }
};
template <typename T> struct LessThanEq {
  T compare;
  __device__ __forceinline__ LessThanEq(T c) : compare(c) {}
  __device__ __forceinline__ bool operator()(const T& a) const { return a <= compare; }
};
template <typename T> struct GreaterThanEq {
  T compare;
  __device__ __forceinline__ GreaterThanEq(T c) : compare(c) {}
  __device__ __forceinline__ bool operator()(const T& a) const { return a >= compare; }
};
template <typename T> struct Eq {
  T compare;
  __device__ __forceinline__ Eq(T c) : compare(c) {}
  __device__ __forceinline__ bool operator()(const T& a) const { return a == compare; }
};

#define LJ_PRED_WRAPPER(NAME, KIND, OP)                                                               \
  template <typename T, int BLOCK_THREADS, int ITEMS_PER_THREAD>                                     \
  __device__ __forceinline__ void NAME(T(&items)[ITEMS_PER_THREAD], T compare,                        \
                                       int(&selection_flags)[ITEMS_PER_THREAD], int num_items) {      \
    OP<T> select_op(compare);                                                                         \
    KIND<T, OP<T>, BLOCK_THREADS, ITEMS_PER_THREAD>(items, select_op, selection_flags, num_items);   \
  }

LJ_PRED_WRAPPER(BlockPredLT, BlockPred, LessThan)
LJ_PRED_WRAPPER(BlockPredAndLT, BlockPredAnd, LessThan)
LJ_PRED_WRAPPER(BlockPredGT, BlockPred, GreaterThan)
LJ_PRED_WRAPPER(BlockPredAndGT, BlockPredAnd, GreaterThan)
LJ_PRED_WRAPPER(BlockPredLTE, BlockPred, LessThanEq)
LJ_PRED_WRAPPER(BlockPredAndLTE, BlockPredAnd, LessThanEq)
LJ_PRED_WRAPPER(BlockPredGTE, BlockPred, GreaterThanEq)
LJ_PRED_WRAPPER(BlockPredAndGTE, BlockPredAnd, GreaterThanEq)
LJ_PRED_WRAPPER(BlockPredEQ, BlockPred, Eq)
LJ_PRED_WRAPPER(BlockPredAndEQ, BlockPredAnd, Eq)
LJ_PRED_WRAPPER(BlockPredOrEQ, BlockPredOr, Eq)
#undef LJ_PRED_WRAPPER

namespace lj {

// Runtime predicate: up to 4 BlockPred / And / Or calls applied in order.
struct PredProgram {
  int32_t n_steps;
  int32_t comb[4];   // 0 SET, 1 AND, 2 OR
  int32_t cmp[4];    // 0 LT, 1 GT, 2 LTE, 3 GTE, 4 EQ
  int32_t value[4];
};

__device__ __forceinline__ bool pred_cmp(int32_t cmp, int32_t a, int32_t c) {
  return cmp == 0 ? a < c : cmp == 1 ? a > c : cmp == 2 ? a <= c : cmp == 3 ? a >= c : a == c;
}

__device__ __forceinline__ bool eval_pred(const PredProgram& p, int32_t attr) {
  bool f = true;
#pragma unroll
  for (int s = 0; s < 4; s++) {
    if (s < p.n_steps) {
      bool r = pred_cmp(p.cmp[s], attr, p.value[s]);
      f = p.comb[s] == 0 ? r : p.comb[s] == 1 ? (f && r) : (f || r);
    }
  }
  return f;
}

}  // namespace lj
