// lj_kernels_misc.hip -- table fills, dimension build, result finalize,
// synthetic column generation.
//
// The dimension build follows build_kernel (src/ops/bloom_join.cu:43-65) and
// build_hashtable_{s,c,d} (src/ssb-jo/q33/0-lip.cu:114-203): load a tile of
// (key, attr), evaluate the predicate, insert selected keys into the Bloom
// filter and into PHT_2 (key -> payload).  All dimensions of one query are
// built by a single launch whose workgroups are partitioned by dimension.
#include <hip/hip_runtime.h>

#include "lipjoin/block/crystal.hpp"
#include "lj_gen.hpp"
#include "lj_internal.hpp"

namespace lj {

// ---------------------------------------------------------------- fill ----
// Plain vector stores (a fill kernel rather than hipMemsetAsync so the whole
// query can be captured in a HIP graph).
__global__ __launch_bounds__(256) void fill_kernel(FillArgs a) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int r = 0; r < a.n; r++) {
    uint32_t* p = a.r[r].ptr;
    const int64_t words = a.r[r].words;
    const uint32_t v = a.r[r].value;
    // head until 16-byte aligned, body as uint4, tail
    int64_t head = (int64_t)((16 - ((uintptr_t)p & 15)) & 15) / 4;
    if (head > words) head = words;
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid < head) p[tid] = v;
    uint4* q = reinterpret_cast<uint4*>(p + head);
    const int64_t nq = (words - head) / 4;
    const uint4 vv = make_uint4(v, v, v, v);
    for (int64_t i = tid; i < nq; i += stride) q[i] = vv;
    const int64_t tail0 = head + nq * 4;
    if (tid < words - tail0) p[tail0 + tid] = v;
  }
}

hipError_t launch_fill(const FillArgs& a, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  int64_t maxw = 0;
  for (int r = 0; r < a.n; r++) maxw = a.r[r].words > maxw ? a.r[r].words : maxw;
  int64_t blocks = (maxw / 4 + 255) / 256;
  if (blocks < 1) blocks = 1;
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}

// --------------------------------------------------------------- build ----
constexpr int kBuildThreads = 256;
constexpr int kBuildItems = 4;

__global__ __launch_bounds__(kBuildThreads) void build_kernel(BuildArgs a) {
  constexpr int BLOCK_THREADS = kBuildThreads;
  constexpr int ITEMS_PER_THREAD = kBuildItems;
  __shared__ unsigned long long red[BLOCK_THREADS / 64];

  int d = 0;
  for (int i = 1; i < a.n_dims; i++)
    if ((int64_t)blockIdx.x >= a.dim[i].block_begin) d = i;
  const BuildDim& D = a.dim[d];

  const int64_t tile_offset = ((int64_t)blockIdx.x - D.block_begin) * TILE_SIZE;
  int num_tile_items = TILE_SIZE;
  if (tile_offset + TILE_SIZE > D.n) num_tile_items = (int)(D.n - tile_offset);

  int keys[ITEMS_PER_THREAD];
  int attr[ITEMS_PER_THREAD];
  int vals[ITEMS_PER_THREAD];
  int selection_flags[ITEMS_PER_THREAD];

  BlockLoad<int, BLOCK_THREADS, ITEMS_PER_THREAD>(D.keys + tile_offset, keys, num_tile_items);
  if (D.pred.n_steps > 0) BlockLoad<int, BLOCK_THREADS, ITEMS_PER_THREAD>(D.attr + tile_offset, attr, num_tile_items);
  BlockLoad<int, BLOCK_THREADS, ITEMS_PER_THREAD>(D.payload + tile_offset, vals, num_tile_items);

  unsigned long long n_sel = 0, n_oob = 0;
#pragma unroll
  for (int i = 0; i < ITEMS_PER_THREAD; i++) {
    const bool in_tile = (int)(threadIdx.x + i * BLOCK_THREADS) < num_tile_items;
    selection_flags[i] = in_tile && (D.pred.n_steps == 0 || eval_pred(D.pred, attr[i]));
    n_sel += selection_flags[i];
  }
  if (D.bloom) BlockBuildBloomFilter<int, BLOCK_THREADS, ITEMS_PER_THREAD>(keys, selection_flags, D.bloom, D.bh,
                                                                          num_tile_items);
  // PHT_2 insert (BlockBuildSelectivePHT_2 with the fast remainder)
#pragma unroll
  for (int i = 0; i < ITEMS_PER_THREAD; i++) {
    if (selection_flags[i]) {
      const int64_t s = pht_slot(keys[i], D.ht_len, D.keys_min);
      if (s < 0) {
        n_oob++;
        continue;
      }
      atomicCAS(&D.ht[2 * s], 0, keys[i]);
      D.ht[2 * s + 1] = vals[i];
    }
  }
  const unsigned long long sel_total = BlockSum<unsigned long long, BLOCK_THREADS, ITEMS_PER_THREAD>(n_sel, red);
  const unsigned long long oob_total = BlockSum<unsigned long long, BLOCK_THREADS, ITEMS_PER_THREAD>(n_oob, red);
  if (threadIdx.x == 0) {
    if (sel_total) atomicAdd(&D.counters[0], sel_total);
    if (oob_total) atomicAdd(&D.counters[1], oob_total);
  }
}

hipError_t launch_build(const BuildArgs& a, int64_t n_blocks, hipStream_t s) {
  if (n_blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(build_kernel, dim3((unsigned)n_blocks), dim3(kBuildThreads), 0, s, a);
  return hipGetLastError();
}

int64_t build_tile_rows() { return (int64_t)kBuildThreads * kBuildItems; }

// ------------------------------------------------------------ finalize ----
// One workgroup: copy the touched groups out in touch order, reset their
// dense accumulators to zero for the next query, reduce the per-workgroup
// statistics of the probe.
__global__ __launch_bounds__(1024) void finalize_kernel(FinalizeArgs a) {
  __shared__ uint32_t n_sh;
  __shared__ long long part[1024 / kStats][kStats];
  if (threadIdx.x == 0) n_sh = *a.n_touched;
  // statistics: thread (r, k) sums word k over blocks r, r + 64, ...; then a
  // column sum over the 64 partial rows
  {
    const int k = threadIdx.x % kStats, r = threadIdx.x / kStats;
    long long acc = 0;
    if (k != ST_ORDER)
      for (int b = r; b < a.n_blocks; b += 1024 / kStats) acc += a.stats[(int64_t)b * kStats + k];
    part[r][k] = acc;
  }
  __syncthreads();
  const uint32_t n = n_sh;
  const uint32_t listed = n < a.touched_cap ? n : a.touched_cap;
  for (uint32_t i = threadIdx.x; i < listed; i += blockDim.x) {
    const uint32_t g = a.touched[i];
    if ((int64_t)i < a.out_cap) {
      GroupOut o;
      o.g = g;
      o.sum = (long long)a.gsum[g];
      o.count = a.gcount[g];
      o.prod = a.gprod ? a.gprod[g] : 0ull;
      for (int d = 0; d < kMaxDims; d++) o.pay[d] = a.gpay[(int64_t)g * kMaxDims + d];
      a.out[i] = o;
    }
    a.gsum[g] = 0;
    a.gcount[g] = 0;
    if (a.gprod) a.gprod[g] = 0;
  }
  for (int k = threadIdx.x; k < kStats; k += blockDim.x) {
    long long acc = 0;
    if (k == ST_ORDER) acc = a.n_blocks > 0 ? a.stats[k] : 0;
    else
      for (int r = 0; r < 1024 / kStats; r++) acc += part[r][k];
    a.out_hdr[1 + k] = acc;
  }
  if (threadIdx.x == 0) a.out_hdr[0] = n;
  if (threadIdx.x < 2 * kMaxDims) {
    const int d = threadIdx.x / 2;
    a.out_hdr[1 + kStats + threadIdx.x] = a.dim_counters[d] ? (long long)a.dim_counters[d][threadIdx.x % 2] : 0;
  }
  __syncthreads();
  if (threadIdx.x == 0) *a.n_touched = 0;
}

hipError_t launch_finalize(const FinalizeArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(1024), 0, s, a);
  return hipGetLastError();
}

// ------------------------------------------------------------ generate ----
__global__ __launch_bounds__(256) void gen_kernel(int32_t kind, uint64_t key, int64_t lo, int64_t hi,
                                                  int64_t first, int64_t n, int32_t* out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = gen_value(kind, key, lo, hi, first + i);
}

hipError_t launch_gen(int32_t kind, uint64_t key, int64_t lo, int64_t hi, int64_t first, int64_t n,
                      int32_t* out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(gen_kernel, dim3((unsigned)blocks), dim3(256), 0, s, kind, key, lo, hi, first, n, out);
  return hipGetLastError();
}

// ------------------------------------------------------------ popcount ----
__global__ __launch_bounds__(256) void popcount_kernel(const uint32_t* w, int64_t n, unsigned long long* out) {
  __shared__ unsigned long long red[4];
  unsigned long long c = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) c += __popc(w[i]);
  c = BlockSum<unsigned long long, 256, 1>(c, red);
  if (threadIdx.x == 0 && c) atomicAdd(out, c);
}

hipError_t launch_popcount(const uint32_t* words, int64_t n, unsigned long long* out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(popcount_kernel, dim3((unsigned)blocks), dim3(256), 0, s, words, n, out);
  return hipGetLastError();
}

}  // namespace lj
