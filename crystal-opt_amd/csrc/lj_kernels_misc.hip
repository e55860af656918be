// lj_kernels_misc.hip -- table fills, dimension build, synthetic column
// generation, Bloom popcount.
//
// The dimension build follows build_kernel (src/ops/bloom_join.cu:43-65) and
// build_hashtable_{s,c,d} (src/ssb-jo/q33/0-lip.cu:114-203): evaluate the
// predicate on (key, attr), insert selected keys into the Bloom filter and
// into PHT_2 (key -> payload).  All dimensions of one query are built by a
// single launch.
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

// ---------------------------------------------------------------- copy ----
// A kernel rather than hipMemcpyAsync for the serving loop's launch
// arguments: the host-to-device copy engine path could block the calling
// thread for milliseconds behind the stream's earlier work, and the sources
// here are small (pinned host memory read through its device address).
__global__ __launch_bounds__(256) void copy_kernel(CopyArgs a) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int r = 0; r < a.n; r++)
    for (int64_t i = tid; i < a.r[r].words; i += stride) a.r[r].dst[i] = a.r[r].src[i];
}

hipError_t launch_copy(const CopyArgs& a, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  int64_t maxw = 0;
  for (int r = 0; r < a.n; r++) maxw = a.r[r].words > maxw ? a.r[r].words : maxw;
  int64_t blocks = (maxw + 255) / 256;
  if (blocks < 1) blocks = 1;
  if (blocks > 64) blocks = 64;
  hipLaunchKernelGGL(copy_kernel, dim3((unsigned)blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}

// --------------------------------------------------------------- build ----
// Predicate, Bloom insert, PHT_2 insert (CAS of the key, then the
// unconditional value store; PHT_1: the CAS alone, join.cuh:191-258; or the
// textbook table's insert) -- build_kernel of src/ops/bloom_join.cu:43-65 and
// build_hashtable_{s,c,d} of src/ssb-jo/q33/0-lip.cu:114-203.  Every dimension
// of the query in one launch: each dimension owns a range of workgroups (all
// dimensions build at once), a thread takes 4 consecutive rows (1 in a
// dimension of fewer than kBuildSmall rows, so that a small dimension's
// inserts spread over more workgroups), and the selected / out-of-domain
// counts reduce per workgroup to one atomic each.
constexpr int kBuildThreads = 256;
constexpr int kBuildRows = 4;
constexpr int64_t kBuildSmall = 1 << 16;

__host__ __device__ inline int build_rows_per_thread(int64_t rows) { return rows < kBuildSmall ? 1 : kBuildRows; }

// Textbook insert: claim the first slot from the key's home slot on that is
// empty or already holds the key (compare-and-swap of the stored key), then
// store the value (a repeated key keeps its slot; the last value stays).
// False when the key is the excluded one or the table is full.
__device__ __forceinline__ bool tb_insert(int32_t* ht, uint32_t mask, int32_t key, int32_t val) {
  if (key == kTbExcludedKey) return false;
  const int32_t enc = (int32_t)tb_encode(key);
  uint32_t s = tb_home(key, mask);
  for (uint32_t i = 0; i <= mask; i++) {
    const int32_t old = atomicCAS(&ht[2 * (int64_t)s], 0, enc);
    if (old == 0 || old == enc) {
      ht[2 * (int64_t)s + 1] = val;
      return true;
    }
    s = (s + 1) & mask;
  }
  return false;
}

// Workgroup `blk` of dimension B: rows [blk * 256 * rpt, (blk + 1) * 256 * rpt).
__device__ __forceinline__ void build_block(const BuildDim& B, int64_t blk) {
  __shared__ unsigned int cnt[4];
  if (threadIdx.x < 4) cnt[threadIdx.x] = 0;
  __syncthreads();
  const int rpt = build_rows_per_thread(B.n);
  const int64_t r0 = (blk * kBuildThreads + threadIdx.x) * rpt;
  uint32_t nsel = 0, noob = 0, npast = 0, nout = 0;
#pragma unroll
  for (int j = 0; j < kBuildRows; j++) {
    const int64_t r = r0 + j;
    if (j < rpt && r < B.n) {
      const int32_t key = B.keys[r];
      const bool sel = B.pred.n_steps == 0 || eval_pred(B.pred, B.attr[r]);
      if (sel) {
        nsel++;
        if (B.bloom) {
          bloom_insert(B.bh, B.bloom, key);
          if (B.past_bit) {
            uint32_t bits[3];
            if (bloom_bits(B.bh, key, bits) && (int64_t)bits[0] >= B.past_bit) npast++;
          }
        }
        if (B.ht_kind == HT_TEXTBOOK) {
          if (!tb_insert(B.ht, B.ht_len.pow2_mask, key, B.payload[r])) noob++;
        } else {
          const int64_t s = pht_slot(key, B.ht_len, B.keys_min);
          if (s < 0) {
            noob++;
          } else {
            // a key from keys_min + ht_len on wraps into the table (the
            // reference's HASH); counted, inserted like the others
            if ((int64_t)key - (int64_t)B.keys_min >= (int64_t)B.ht_len.d) nout++;
            if (B.ht_kind == HT_PHT1) {
              atomicCAS(&B.ht[s], 0, key);
            } else {
              atomicCAS(&B.ht[2 * s], 0, key);
              B.ht[2 * s + 1] = B.payload[r];
            }
          }
        }
      }
    }
  }
  nsel = WaveSum(nsel);
  noob = WaveSum(noob);
  npast = WaveSum(npast);
  nout = WaveSum(nout);
  if (__lane_id() == 0) {
    if (nsel) atomicAdd(&cnt[0], nsel);
    if (noob) atomicAdd(&cnt[1], noob);
    if (npast) atomicAdd(&cnt[2], npast);
    if (nout) atomicAdd(&cnt[3], nout);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long* c = B.counters + 2 * (blockIdx.x % kCounterLanes);
    if (cnt[0]) atomicAdd(&c[0], (unsigned long long)cnt[0]);
    if (cnt[1]) atomicAdd(&c[1], (unsigned long long)cnt[1]);
    if (cnt[2]) atomicAdd(&B.counters[kCounterPast], (unsigned long long)cnt[2]);
    if (cnt[3]) atomicAdd(&B.counters[kCounterOutside], (unsigned long long)cnt[3]);
  }
}

__global__ __launch_bounds__(kBuildThreads) void build_kernel(BuildArgs args) {
#if defined(__HIP_DEVICE_COMPILE__)
  (void)args;
  const BuildArgs& a = *(const BuildArgs*)(__builtin_amdgcn_kernarg_segment_ptr());
#else
  const BuildArgs& a = args;
#endif
  if (a.clock && blockIdx.x == 0 && threadIdx.x == 0) *a.clock = (long long)wall_clock64();
  int d = 0;
  for (int k = 1; k < a.n_dims; k++)
    if ((int64_t)blockIdx.x >= a.dim[k].block_begin) d = k;
  build_block(a.dim[d], (int64_t)blockIdx.x - a.dim[d].block_begin);
}

// The dimensions of a batch (descriptors in device memory).
__global__ __launch_bounds__(kBuildThreads) void build_multi_kernel(const BuildDim* __restrict__ dims, int32_t n) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {  // last dimension with block_begin <= blockIdx.x (uniform)
    const int mid = (lo + hi + 1) >> 1;
    if (dims[mid].block_begin <= (int64_t)blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  build_block(dims[lo], (int64_t)blockIdx.x - dims[lo].block_begin);
}

int64_t build_blocks(int64_t rows) {
  const int64_t rows_per_block = (int64_t)kBuildThreads * build_rows_per_thread(rows);
  return (rows + rows_per_block - 1) / rows_per_block;
}

hipError_t launch_build_multi(const BuildDim* dims, int n, int64_t total_blocks, hipStream_t s) {
  if (n <= 0 || total_blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(build_multi_kernel, dim3((unsigned)total_blocks), dim3(kBuildThreads), 0, s, dims, (int32_t)n);
  return hipGetLastError();
}

// Assigns the workgroup ranges (block_begin) and launches.
hipError_t launch_build(BuildArgs& a, hipStream_t s) {
  int64_t blocks = 0;
  for (int d = 0; d < a.n_dims; d++) {
    a.dim[d].block_begin = blocks;
    blocks += build_blocks(a.dim[d].n);
  }
  if (blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(build_kernel, dim3((unsigned)blocks), dim3(kBuildThreads), 0, s, a);
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
