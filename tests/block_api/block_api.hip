// block_api.hip -- test-only kernels written against the wave64 block
// operator API (lipjoin/block/crystal.hpp), the drop-in for crystal/*.cuh.
//
// Two kernels per shape <BLOCK_THREADS, ITEMS_PER_THREAD> (128,4 and 256,2,
// one tile of BLOCK_THREADS * ITEMS_PER_THREAD rows per block, ragged last
// tile), each a plain composition of the API's calls:
//   ba_build : BlockLoad keys/vals -> InitFlags -> BlockPredGTE, BlockPredAndLT,
//              BlockPredOrEQ -> BlockBuildBloomFilter, BlockBuildSelectivePHT_1,
//              BlockBuildSelectivePHT_2
//   ba_probe : InitFlags -> BlockLoad fk -> BlockProbeBloomFilter -> IsTerm
//              (stored per thread) -> BlockProbeAndPHT_1 -> BlockProbeAndPHT_2
//              -> BlockPredLoad measure -> BlockPredAndLTE, BlockPredAndGT ->
//              BlockStore flags and payloads -> BlockSum of measure * payload
// The host entry points are extern "C" so the tests call them through ctypes
// with device addresses; they return hipError_t of the launch.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lipjoin/block/crystal.hpp"

template <int BLOCK_THREADS, int ITEMS_PER_THREAD>
__global__ __launch_bounds__(BLOCK_THREADS) void build_k(const int* keys, const int* vals, int n, int lo, int hi,
                                                          int eq, int keys_min, uint32_t* bf, int bf_words,
                                                          int* ht1, int* ht2, int ht_len) {
  const int tile_offset = blockIdx.x * TILE_SIZE;
  const int num_tile_items = n - tile_offset < TILE_SIZE ? n - tile_offset : TILE_SIZE;
  int items[ITEMS_PER_THREAD], vitems[ITEMS_PER_THREAD], flags[ITEMS_PER_THREAD];
  BlockLoad<int, BLOCK_THREADS, ITEMS_PER_THREAD>(keys + tile_offset, items, num_tile_items);
  BlockLoad<int, BLOCK_THREADS, ITEMS_PER_THREAD>(vals + tile_offset, vitems, num_tile_items);
  InitFlags<BLOCK_THREADS, ITEMS_PER_THREAD>(flags);
  BlockPredGTE<int, BLOCK_THREADS, ITEMS_PER_THREAD>(vitems, lo, flags, num_tile_items);
  BlockPredAndLT<int, BLOCK_THREADS, ITEMS_PER_THREAD>(vitems, hi, flags, num_tile_items);
  BlockPredOrEQ<int, BLOCK_THREADS, ITEMS_PER_THREAD>(vitems, eq, flags, num_tile_items);
  BlockBuildBloomFilter<int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, flags, bf, bf_words, keys_min, num_tile_items);
  BlockBuildSelectivePHT_1<int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, flags, ht1, ht_len, keys_min,
                                                                  num_tile_items);
  BlockBuildSelectivePHT_2<int, int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, vitems, flags, ht2, ht_len, keys_min,
                                                                       num_tile_items);
}

template <int BLOCK_THREADS, int ITEMS_PER_THREAD>
__global__ __launch_bounds__(BLOCK_THREADS) void probe_k(const int* fk, const int* meas, int n, const uint32_t* bf,
                                                          int bf_words, int keys_min, const int* ht1, const int* ht2,
                                                          int ht_len, int mhi, int mlo, int* flags_out,
                                                          int* pay_out, int* term_out,
                                                          unsigned long long* sum_out) {
  __shared__ long long shared[BLOCK_THREADS / 64];
  const int tile_offset = blockIdx.x * TILE_SIZE;
  const int num_tile_items = n - tile_offset < TILE_SIZE ? n - tile_offset : TILE_SIZE;
  int items[ITEMS_PER_THREAD], flags[ITEMS_PER_THREAD], pay[ITEMS_PER_THREAD], m[ITEMS_PER_THREAD];
  for (int i = 0; i < ITEMS_PER_THREAD; i++) {
    pay[i] = 0;
    m[i] = 0;
  }
  InitFlags<BLOCK_THREADS, ITEMS_PER_THREAD>(flags);
  BlockLoad<int, BLOCK_THREADS, ITEMS_PER_THREAD>(fk + tile_offset, items, num_tile_items);
  BlockProbeBloomFilter<int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, flags, bf, bf_words, keys_min, num_tile_items);
  const bool term = IsTerm<int, BLOCK_THREADS, ITEMS_PER_THREAD>(flags);
  term_out[blockIdx.x * BLOCK_THREADS + threadIdx.x] = term ? 1 : 0;
  // the reference's callers return here; a thread that would return skips
  // the rest instead, because BlockSum below has barriers
  if (!term) {
    BlockProbeAndPHT_1<int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, flags, ht1, ht_len, keys_min, num_tile_items);
    BlockProbeAndPHT_2<int, int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, pay, flags, ht2, ht_len, keys_min,
                                                                   num_tile_items);
    BlockPredLoad<int, BLOCK_THREADS, ITEMS_PER_THREAD>(meas + tile_offset, m, num_tile_items, flags);
    BlockPredAndLTE<int, BLOCK_THREADS, ITEMS_PER_THREAD>(m, mhi, flags, num_tile_items);
    BlockPredAndGT<int, BLOCK_THREADS, ITEMS_PER_THREAD>(m, mlo, flags, num_tile_items);
  }
  BlockStore<int, BLOCK_THREADS, ITEMS_PER_THREAD>(flags_out + tile_offset, flags, num_tile_items);
  BlockStore<int, BLOCK_THREADS, ITEMS_PER_THREAD>(pay_out + tile_offset, pay, num_tile_items);
  long long prod[ITEMS_PER_THREAD];
  for (int i = 0; i < ITEMS_PER_THREAD; i++) {
    const bool in = (int)(threadIdx.x + i * BLOCK_THREADS) < num_tile_items;
    prod[i] = in && flags[i] ? (long long)m[i] * pay[i] : 0;
  }
  const long long s = BlockSum<long long, BLOCK_THREADS, ITEMS_PER_THREAD>(prod, shared);
  if (threadIdx.x == 0) atomicAdd(sum_out, (unsigned long long)s);
}

template <int BT, int IPT>
static int launch_build(const int* keys, const int* vals, int n, int lo, int hi, int eq, int kmin, uint32_t* bf,
                        int bfw, int* ht1, int* ht2, int len, hipStream_t s) {
  const int tiles = (n + BT * IPT - 1) / (BT * IPT);
  if (tiles > 0) hipLaunchKernelGGL((build_k<BT, IPT>), dim3(tiles), dim3(BT), 0, s, keys, vals, n, lo, hi, eq, kmin,
                                    bf, bfw, ht1, ht2, len);
  return (int)hipGetLastError();
}

template <int BT, int IPT>
static int launch_probe(const int* fk, const int* meas, int n, const uint32_t* bf, int bfw, int kmin, const int* ht1,
                        const int* ht2, int len, int mhi, int mlo, int* fo, int* po, int* to,
                        unsigned long long* so, hipStream_t s) {
  const int tiles = (n + BT * IPT - 1) / (BT * IPT);
  if (tiles > 0) hipLaunchKernelGGL((probe_k<BT, IPT>), dim3(tiles), dim3(BT), 0, s, fk, meas, n, bf, bfw, kmin, ht1,
                                    ht2, len, mhi, mlo, fo, po, to, so);
  return (int)hipGetLastError();
}

// shape 0: <128, 4>; shape 1: <256, 2>
extern "C" int ba_build(int shape, const int* keys, const int* vals, int n, int lo, int hi, int eq, int keys_min,
                        uint32_t* bf, int bf_words, int* ht1, int* ht2, int ht_len, void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (shape == 0) return launch_build<128, 4>(keys, vals, n, lo, hi, eq, keys_min, bf, bf_words, ht1, ht2, ht_len, s);
  if (shape == 1) return launch_build<256, 2>(keys, vals, n, lo, hi, eq, keys_min, bf, bf_words, ht1, ht2, ht_len, s);
  return (int)hipErrorInvalidValue;
}

extern "C" int ba_probe(int shape, const int* fk, const int* meas, int n, const uint32_t* bf, int bf_words,
                        int keys_min, const int* ht1, const int* ht2, int ht_len, int mhi, int mlo, int* flags_out,
                        int* pay_out, int* term_out, unsigned long long* sum_out, void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (shape == 0)
    return launch_probe<128, 4>(fk, meas, n, bf, bf_words, keys_min, ht1, ht2, ht_len, mhi, mlo, flags_out, pay_out,
                                term_out, sum_out, s);
  if (shape == 1)
    return launch_probe<256, 2>(fk, meas, n, bf, bf_words, keys_min, ht1, ht2, ht_len, mhi, mlo, flags_out, pay_out,
                                term_out, sum_out, s);
  return (int)hipErrorInvalidValue;
}
