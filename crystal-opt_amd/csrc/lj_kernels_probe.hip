// lj_kernels_probe.hip -- the fused fact-table probe of the star join.
//
// Semantics (what is computed) follow the reference's probe kernels:
//   NAIVE  q33/{i}-naive.cu:28-91 -- per dimension in join order: (Pred)Load the
//          FK column, PHT_2 probe (payload), drop misses; then the measure and
//          the group add.
//   LIP    q33/{i}-lip.cu:28-112  -- per dimension in Bloom order: (Pred)Load
//          the FK column, Bloom probe; then PHT_2 probes in join order on the
//          FKs already loaded; then the measure and the group add.
//   single join (join.cu:65-108, bloom_join.cu:67-112) is the 1-dimension case
//          with one group and the extra (int32)(measure*payload) sum.
// Results are integer sums, so none of the structure below changes them.
//
// Structure (how it is computed on CDNA4):
//   * Persistent grid; each wave walks 256*VEC-row tiles with a grid stride.
//     Lane l of a wave owns rows 4l..4l+3 of every 256-row segment and loads
//     them with one 16-byte non-temporal load, prefetching the next tile's
//     first FK column while the current tile is processed.
//   * Stage 0 (the first FK column, read in full) probes its filter for every
//     row.  Bloom filters that the host placed in LDS are probed there.
//   * Survivors of stage 0 are compacted into a per-wave 64-entry LDS queue
//     (row, key, payload) by ballot/mbcnt; a full queue is processed with all
//     64 lanes busy: gather the next FK per survivor, probe, ... then PHT_2
//     probes, the measure gather and the group add.  A segment with more than
//     kDenseMin survivors is processed in place instead (predicated 16-byte
//     loads of the later columns), which is cheaper when most rows survive.
//   * LIP_ADAPTIVE: every wave first probes ALL Bloom filters on one 256-row
//     sample segment; the workgroup sums the pass counts in LDS and all its
//     waves order the filters by ascending pass count.
//   * Group-by: a 64-slot open-addressing table in LDS per workgroup holds
//     int64 sums/counts; it is flushed to the dense global arrays with 64-bit
//     atomics at the end; a slot miss goes straight to the global atomics.
//     A group's first global touch appends it to the `touched` list.
//   * Exact per-stage statistics: rows alive and 128-byte lines holding an
//     alive row when each column is read (the algorithmic byte count).
#include <hip/hip_runtime.h>

#include "lipjoin/block/crystal.hpp"
#include "lj_internal.hpp"

namespace lj {

constexpr int kDenseMin = 32;  // stage-0 survivors per segment above which the segment is processed in place
constexpr int32_t kEmptyGroup = -1;

// Wave-uniform counters.  Indices into v[] are compile-time constants after
// unrolling; the measure stage (index n_dims, a runtime value) has its own
// fields and is placed at flush time.
struct WaveStats {
  long long v[kStats];
  long long meas_rows;
  long long meas_lines;
};

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ int popc_ballot4(uint32_t m) {
  return __popcll(__ballot(m & 1u)) + __popcll(__ballot(m & 2u)) + __popcll(__ballot(m & 4u)) +
         __popcll(__ballot(m & 8u));
}

__device__ __forceinline__ uint32_t mbcnt(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// number of nonzero bytes of a 64-bit ballot over "lane holds an alive row"
// (8 lanes x 4 rows = one 32-row, 128-byte line)
__device__ __forceinline__ int lines_of_ballot(unsigned long long b) {
  b |= b >> 4;
  b |= b >> 2;
  b |= b >> 1;
  return __popcll(b & 0x0101010101010101ull);
}

// ---- filter probes ----------------------------------------------------------
template <int KIND, bool IN_LDS>
__device__ __forceinline__ bool bloom_probe_k(const ProbeDim& D, const uint32_t* lds, int32_t key) {
  const uint32_t* bf = IN_LDS ? lds + D.bloom_lds : D.bloom;
  if constexpr (KIND == BLOOM_NONE) {
    return true;
  } else if constexpr (KIND == BLOOM_MOD_POW2) {
    int64_t x = (int64_t)key - (int64_t)D.bh.keys_min;
    uint32_t h = (uint32_t)x & D.bh.mod.pow2_mask;
    return x >= 0 && ((bf[h >> 5] >> (h & 31)) & 1u);
  } else if constexpr (KIND == BLOOM_MOD) {
    int64_t x = (int64_t)key - (int64_t)D.bh.keys_min;
    uint32_t h = fastmod_u32((uint32_t)x, D.bh.mod);
    return x >= 0 && ((bf[h >> 5] >> (h & 31)) & 1u);
  } else {
    return bloom_test(D.bh, bf, key);
  }
}

__device__ __forceinline__ bool bloom_probe_rt(const ProbeDim& D, const uint32_t* lds, int32_t key) {
  if (D.bh.kind == BLOOM_NONE) return true;
  const uint32_t* bf = D.bloom_lds >= 0 ? lds + D.bloom_lds : D.bloom;
  return bloom_test(D.bh, bf, key);
}

__device__ __forceinline__ bool pht_probe(const ProbeDim& D, int32_t key, int32_t& pay) {
  return pht2_probe(D.ht, D.ht_len, D.keys_min, key, pay);
}

// ---- per-dimension register sets (uniform dimension index, no dynamic VGPR indexing)
struct Quad {
  int32_t v[kMaxDims];
  __device__ __forceinline__ void set(int d, int32_t x) {
    if (d == 0) v[0] = x;
    else if (d == 1) v[1] = x;
    else if (d == 2) v[2] = x;
    else v[3] = x;
  }
  __device__ __forceinline__ int32_t get(int d) const {
    return d == 0 ? v[0] : d == 1 ? v[1] : d == 2 ? v[2] : v[3];
  }
};

// ---- group table in LDS --------------------------------------------------
struct GroupLds {
  int32_t* key;                 // [kGroupSlots]
  unsigned long long* cnt;      // [kGroupSlots]
  unsigned long long* sum;      // [kGroupSlots]
  unsigned long long* prod;     // [kGroupSlots]
  int32_t* pay;                 // [kGroupSlots * 4]
};

__device__ __forceinline__ void global_group_add(const ProbeArgs& a, int64_t g, long long s, unsigned long long c,
                                                 unsigned long long p, const Quad& pay, int nd) {
  atomicAdd(&a.gsum[g], (unsigned long long)s);
  if (a.want_product) atomicAdd(&a.gprod[g], p);
  unsigned long long old = atomicAdd(&a.gcount[g], c);
  if (old == 0) {
    for (int d = 0; d < kMaxDims; d++)
      if (d < nd) a.gpay[g * kMaxDims + d] = pay.v[d];
    uint32_t idx = atomicAdd(a.n_touched, 1u);
    if (idx < a.touched_cap) a.touched[idx] = (uint32_t)g;
  }
}

// Add one row (alive lanes) to the workgroup's group table.  Called by every
// lane of the wave.  Returns the number of rows (wave-wide) whose group index
// fell outside [0, n_groups).
__device__ __forceinline__ int group_add(const ProbeArgs& a, GroupLds& gt, bool alive, const Quad& pay,
                                         int32_t m) {
  const int nd = a.n_dims;
  if (a.n_groups == 1) {
    // one group (single join): wave-reduce, one LDS atomic per wave
    long long s = alive ? (long long)m : 0;
    unsigned long long p = 0;
    if (a.want_product && alive) p = (unsigned long long)(long long)(int32_t)((uint32_t)m * (uint32_t)pay.v[0]);
    unsigned long long b = __ballot(alive);
    if (b == 0) return 0;
    s = WaveSum(s);
    if (a.want_product) p = WaveSum(p);
    if (lane_id() == 0) {
      atomicAdd(&gt.sum[0], (unsigned long long)s);
      atomicAdd(&gt.cnt[0], (unsigned long long)__popcll(b));
      if (a.want_product) atomicAdd(&gt.prod[0], p);
      if (gt.key[0] == kEmptyGroup) {
        gt.key[0] = 0;
        for (int d = 0; d < kMaxDims; d++) gt.pay[d] = pay.v[d];
      }
    }
    return 0;
  }
  int64_t g = 0;
  if (alive) {
    int64_t acc = a.goffset;
#pragma unroll
    for (int d = 0; d < kMaxDims; d++)
      if (d < nd) acc += (int64_t)pay.v[d] * a.dim[d].gmult;
    g = acc % a.n_groups;
  }
  const int n_bad = __popcll(__ballot(alive && g < 0));
  if (!alive || g < 0) return n_bad;
  const unsigned long long p =
      a.want_product ? (unsigned long long)(long long)(int32_t)((uint32_t)m * (uint32_t)pay.v[0]) : 0ull;
  uint32_t slot = ((uint32_t)g * 0x9E3779B1u) >> 26;  // 64 slots
  for (int probe = 0; probe < kGroupSlots; probe++) {
    int32_t k = gt.key[slot];
    if (k == kEmptyGroup) {
      int32_t old = atomicCAS(&gt.key[slot], kEmptyGroup, (int32_t)g);
      if (old == kEmptyGroup) {
        for (int d = 0; d < kMaxDims; d++) gt.pay[slot * kMaxDims + d] = pay.v[d];
        k = (int32_t)g;
      } else {
        k = old;
      }
    }
    if (k == (int32_t)g) {
      atomicAdd(&gt.sum[slot], (unsigned long long)(long long)m);
      atomicAdd(&gt.cnt[slot], 1ull);
      if (a.want_product) atomicAdd(&gt.prod[slot], p);
      return n_bad;
    }
    slot = (slot + 1) & (kGroupSlots - 1);
  }
  global_group_add(a, g, (long long)m, 1ull, p, pay, nd);
  return n_bad;
}

// ---- line statistics for a compacted batch (rows ascending with lane) -------
__device__ __forceinline__ int batch_lines(bool alive, int32_t row, int32_t& last_line) {
  // last_line: line of the last alive row of this stage in earlier batches
  int32_t line = alive ? (row >> 5) : -1;
  // inclusive prefix max over lanes
  int32_t pm = line;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    int32_t o = __shfl_up(pm, off, 64);
    if ((int)lane_id() >= off) pm = pm > o ? pm : o;
  }
  int32_t prev = __shfl_up(pm, 1, 64);
  if (lane_id() == 0) prev = -1;
  if (prev < 0) prev = last_line;
  const bool fresh = alive && line != prev;
  const int cnt = __popcll(__ballot(fresh));
  const int32_t top = __shfl(pm, 63, 64);
  if (top >= 0) last_line = top;
  return cnt;
}

// ---- per-wave context --------------------------------------------------------
// Named scalars (not arrays): every access below is by a compile-time stage
// number, so nothing has to live in scratch memory.
struct WaveCtx {
  int nd;
  int ld0, ld1, ld2, ld3;   // column load order (stage -> dim)
  int32_t* q_row;
  int32_t* q_key;
  int32_t* q_pay;
  int qcount;
  int32_t last1, last2, last3, last_meas;  // batch line bookkeeping per stage
  __device__ __forceinline__ int ld(int s) const { return s == 0 ? ld0 : s == 1 ? ld1 : s == 2 ? ld2 : ld3; }
  __device__ __forceinline__ void set_ld(int s, int d) {
    if (s == 0) ld0 = d;
    else if (s == 1) ld1 = d;
    else if (s == 2) ld2 = d;
    else ld3 = d;
  }
  __device__ __forceinline__ int32_t last(int s) const { return s == 1 ? last1 : s == 2 ? last2 : last3; }
  __device__ __forceinline__ void set_last(int s, int32_t x) {
    if (s == 1) last1 = x;
    else if (s == 2) last2 = x;
    else last3 = x;
  }
};

// Process `cnt` queued stage-0 survivors (lane i takes entry i).
__device__ __forceinline__ void process_batch(const ProbeArgs& a, const uint32_t* lds, GroupLds& gt, WaveCtx& w, int cnt,
                              WaveStats& st) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const int i = (int)lane_id();
  bool alive = i < cnt;
  int32_t row = 0;
  Quad key, pay;
  key.v[0] = key.v[1] = key.v[2] = key.v[3] = 0;
  pay = key;
  if (alive) {
    row = w.q_row[i];
    key.set(w.ld0, w.q_key[i]);
    pay.set(w.ld0, w.q_pay[i]);
  }
  const int nd = w.nd;
#pragma unroll
  for (int s = 1; s < kMaxDims; s++) {
    if (s >= nd) continue;
    const int d = w.ld(s);
    const ProbeDim& D = a.dim[d];
    st.v[ST_ROWS_IN + s] += __popcll(__ballot(alive));
    int32_t ls = w.last(s);
    st.v[ST_LINES_IN + s] += batch_lines(alive, row, ls);
    w.set_last(s, ls);
    int32_t k = 0;
    if (alive) k = D.fk[row];
    st.v[ST_OOB] += __popcll(__ballot(alive && k < D.keys_min));
    if (alive) {
      key.set(d, k);
      if (a.method == 0) {
        int32_t p = 0;
        alive = pht_probe(D, k, p);
        pay.set(d, p);
      } else {
        alive = bloom_probe_rt(D, lds, k);
      }
    }
  }
  if (a.method != 0) {
    st.v[ST_BLOOM_PASS] += __popcll(__ballot(alive));
#pragma unroll
    for (int t = 0; t < kMaxDims; t++) {
      if (t >= nd) continue;
      const int e = a.order[t];
      if (alive) {
        int32_t p = 0;
        alive = pht_probe(a.dim[e], key.get(e), p);
        pay.set(e, p);
      }
    }
  }
  st.meas_rows += __popcll(__ballot(alive));
  st.meas_lines += batch_lines(alive, row, w.last_meas);
  int32_t m = 0;
  if (alive) m = a.measure[row];
  st.v[ST_OOB] += group_add(a, gt, alive, pay, m);
}

// In-place processing of one segment whose stage-0 survivors are dense.
// key0: the lane's 4 stage-0 keys, pay0: their payloads (NAIVE), mask: alive bits.
__device__ __forceinline__ void process_dense(const ProbeArgs& a, const uint32_t* lds, GroupLds& gt, WaveCtx& w, int64_t row0,
                              const v4i32& key0, const int32_t (&pay0)[4], uint32_t mask, WaveStats& st) {
  const int nd = w.nd;
  Quad key[4], pay[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    key[j].v[0] = key[j].v[1] = key[j].v[2] = key[j].v[3] = 0;
    pay[j] = key[j];
    key[j].set(w.ld0, vget(key0, j));
    pay[j].set(w.ld0, pay0[j]);
  }
#pragma unroll
  for (int s = 1; s < kMaxDims; s++) {
    if (s >= nd) continue;
    const int d = w.ld(s);
    const ProbeDim& D = a.dim[d];
    st.v[ST_LINES_IN + s] += lines_of_ballot(__ballot(mask != 0));
    st.v[ST_ROWS_IN + s] += popc_ballot4(mask);
    const v4i32 kv = WavePredLoadVec4(D.fk, row0, a.n, mask);
    uint32_t nm = 0, bad = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) bad |= (uint32_t)(((mask >> j) & 1u) && vget(kv, j) < D.keys_min) << j;
    st.v[ST_OOB] += popc_ballot4(bad);
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if ((mask >> j) & 1u) {
        const int32_t k = vget(kv, j);
        key[j].set(d, k);
        bool ok;
        if (a.method == 0) {
          int32_t p = 0;
          ok = pht_probe(D, k, p);
          pay[j].set(d, p);
        } else {
          ok = bloom_probe_rt(D, lds, k);
        }
        nm |= (uint32_t)ok << j;
      }
    }
    mask = nm;
  }
  if (a.method != 0) {
    st.v[ST_BLOOM_PASS] += popc_ballot4(mask);
#pragma unroll
    for (int t = 0; t < kMaxDims; t++) {
      if (t >= nd) continue;
      const int e = a.order[t];
      uint32_t nm = 0;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        if ((mask >> j) & 1u) {
          int32_t p = 0;
          bool ok = pht_probe(a.dim[e], key[j].get(e), p);
          pay[j].set(e, p);
          nm |= (uint32_t)ok << j;
        }
      }
      mask = nm;
    }
  }
  st.meas_lines += lines_of_ballot(__ballot(mask != 0));
  st.meas_rows += popc_ballot4(mask);
  const v4i32 mv = WavePredLoadVec4(a.measure, row0, a.n, mask);
#pragma unroll
  for (int j = 0; j < 4; j++) st.v[ST_OOB] += group_add(a, gt, (mask >> j) & 1u, pay[j], vget(mv, j));
}

// Stage 0 over one segment: returns the alive mask of the lane's 4 rows.
template <int KIND, bool IN_LDS, bool IS_PHT>
__device__ __forceinline__ uint32_t stage0_probe(const ProbeDim& D, const uint32_t* lds, const v4i32& k,
                                                 int64_t row0, int64_t n, int32_t (&pay)[4], int& oob) {
  uint32_t mask = 0;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int32_t key = vget(k, j);
    const bool valid = row0 + j < n;
    oob += (valid && key < D.keys_min) ? 1 : 0;
    bool ok;
    if constexpr (IS_PHT) {
      int32_t p = 0;
      ok = valid && pht_probe(D, key, p);
      pay[j] = p;
    } else {
      ok = valid && bloom_probe_k<KIND, IN_LDS>(D, lds, key);
      pay[j] = 0;
    }
    mask |= (uint32_t)ok << j;
  }
  return mask;
}

template <int KIND, bool IN_LDS, bool IS_PHT, int VEC>
__device__ __forceinline__ void stage0_tile(const ProbeDim& D, const uint32_t* lds, const v4i32 (&k)[VEC],
                                            int64_t base, int64_t n, int32_t (&pay)[VEC][4], uint32_t (&mask)[VEC],
                                            int& oob) {
#pragma unroll
  for (int v = 0; v < VEC; v++) {
    const int64_t row0 = base + v * kSegRows + (int64_t)lane_id() * 4;
    mask[v] = stage0_probe<KIND, IN_LDS, IS_PHT>(D, lds, k[v], row0, n, pay[v], oob);
  }
}

template <int VEC>
__device__ __forceinline__ void stage0_dispatch(const ProbeArgs& a, const ProbeDim& D, const uint32_t* lds,
                                                const v4i32 (&k)[VEC], int64_t base, int32_t (&pay)[VEC][4],
                                                uint32_t (&mask)[VEC], int& oob) {
  const int64_t n = a.n;
  if (a.method == 0) {
    stage0_tile<0, false, true, VEC>(D, lds, k, base, n, pay, mask, oob);
    return;
  }
  const bool in_lds = D.bloom_lds >= 0;
  switch (D.bh.kind) {
    case BLOOM_NONE: stage0_tile<BLOOM_NONE, false, false, VEC>(D, lds, k, base, n, pay, mask, oob); break;
    case BLOOM_MOD_POW2:
      if (in_lds) stage0_tile<BLOOM_MOD_POW2, true, false, VEC>(D, lds, k, base, n, pay, mask, oob);
      else stage0_tile<BLOOM_MOD_POW2, false, false, VEC>(D, lds, k, base, n, pay, mask, oob);
      break;
    case BLOOM_MOD:
      if (in_lds) stage0_tile<BLOOM_MOD, true, false, VEC>(D, lds, k, base, n, pay, mask, oob);
      else stage0_tile<BLOOM_MOD, false, false, VEC>(D, lds, k, base, n, pay, mask, oob);
      break;
    default:
      if (in_lds) stage0_tile<BLOOM_MULT_STD, true, false, VEC>(D, lds, k, base, n, pay, mask, oob);
      else stage0_tile<BLOOM_MULT_STD, false, false, VEC>(D, lds, k, base, n, pay, mask, oob);
      break;
  }
}

// Append the alive rows of one segment to the wave queue (row order kept).
__device__ __forceinline__ void enqueue(WaveCtx& w, int64_t row0, const v4i32& k, const int32_t (&pay)[4],
                                        uint32_t mask, int cnt_total) {
  const unsigned long long b0 = __ballot(mask & 1u), b1 = __ballot(mask & 2u), b2 = __ballot(mask & 4u),
                           b3 = __ballot(mask & 8u);
  uint32_t pos = (uint32_t)w.qcount + mbcnt(b0) + mbcnt(b1) + mbcnt(b2) + mbcnt(b3);
#pragma unroll
  for (int j = 0; j < 4; j++) {
    if ((mask >> j) & 1u) {
      w.q_row[pos] = (int32_t)(row0 + j);
      w.q_key[pos] = vget(k, j);
      w.q_pay[pos] = pay[j];
      pos++;
    }
  }
  w.qcount += cnt_total;
}

template <int WG, int VEC>
__global__ __launch_bounds__(WG) void probe_kernel(ProbeArgs args) {
  // Read the arguments in place from the kernarg segment: indexing the
  // by-value copy with a (uniform) dimension index would copy it to scratch.
#if defined(__HIP_DEVICE_COMPILE__)
  (void)args;
  const ProbeArgs& a = *(const ProbeArgs*)(__builtin_amdgcn_kernarg_segment_ptr());
#else
  const ProbeArgs& a = args;
#endif
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  constexpr int NW = WG / 64;
  const int wave = threadIdx.x >> 6;
  const int lane = (int)lane_id();

  uint32_t* lds_filter = reinterpret_cast<uint32_t*>(lds_raw);
  size_t off = ((size_t)a.lds_filter_words * 4 + 15) & ~(size_t)15;
  int32_t* q_base = reinterpret_cast<int32_t*>(lds_raw + off) + wave * (3 * kQueue);
  off += (size_t)NW * 3 * kQueue * 4;
  GroupLds gt;
  gt.cnt = reinterpret_cast<unsigned long long*>(lds_raw + off);
  gt.sum = gt.cnt + kGroupSlots;
  gt.prod = gt.sum + kGroupSlots;
  off += 3 * kGroupSlots * 8;
  gt.key = reinterpret_cast<int32_t*>(lds_raw + off);
  off += kGroupSlots * 4;
  gt.pay = reinterpret_cast<int32_t*>(lds_raw + off);
  off += kGroupSlots * kMaxDims * 4;
  long long* lds_stats = reinterpret_cast<long long*>(lds_raw + off);
  off += kStats * 8;
  unsigned long long* lds_samp = reinterpret_cast<unsigned long long*>(lds_raw + off);  // [kMaxDims + 1]

  // ---- stage the LDS-resident filters, clear the group table
  for (int d = 0; d < a.n_dims; d++) {
    const ProbeDim& D = a.dim[d];
    if (D.bloom_lds >= 0) {
      const int nw = D.bloom_words;
      uint32_t* dst = lds_filter + D.bloom_lds;
      if (((D.bloom_lds | nw) & 3) == 0 && (((uintptr_t)D.bloom) & 15) == 0) {
        const uint4* src4 = reinterpret_cast<const uint4*>(D.bloom);
        uint4* dst4 = reinterpret_cast<uint4*>(dst);
        for (int i = threadIdx.x; i < nw / 4; i += WG) dst4[i] = src4[i];
      } else {
        for (int i = threadIdx.x; i < nw; i += WG) dst[i] = D.bloom[i];
      }
    }
  }
  for (int i = threadIdx.x; i < kGroupSlots; i += WG) {
    gt.key[i] = kEmptyGroup;
    gt.cnt[i] = 0;
    gt.sum[i] = 0;
    gt.prod[i] = 0;
  }
  for (int i = threadIdx.x; i < kStats; i += WG) lds_stats[i] = 0;
  for (int i = threadIdx.x; i <= kMaxDims; i += WG) lds_samp[i] = 0;
  __syncthreads();

  WaveStats st;
#pragma unroll
  for (int i = 0; i < kStats; i++) st.v[i] = 0;
  WaveCtx w;
  w.nd = a.n_dims;
  w.q_row = q_base;
  w.q_key = q_base + kQueue;
  w.q_pay = q_base + 2 * kQueue;
  w.qcount = 0;
  w.last1 = w.last2 = w.last3 = w.last_meas = -1;
  w.ld0 = a.load_order[0];
  w.ld1 = a.load_order[1];
  w.ld2 = a.load_order[2];
  w.ld3 = a.load_order[3];
  st.meas_rows = st.meas_lines = 0;

  const int64_t total_waves = (int64_t)gridDim.x * NW;
  const int64_t gw = (int64_t)blockIdx.x * NW + wave;
  int oob = 0;

  // ---- adaptive: sample one segment per wave with every filter --------------
  if (a.adaptive) {
    const int64_t row0 = gw * kSegRows + lane * 4;
    const bool has = gw * kSegRows < a.sample_rows;
    v4i32 kk[kMaxDims];
    uint32_t pass[kMaxDims];
    uint32_t valid = 0;
#pragma unroll
    for (int d = 0; d < kMaxDims; d++) {
      kk[d] = v4i32{0, 0, 0, 0};
      pass[d] = 0;
    }
    if (has) {
#pragma unroll
      for (int j = 0; j < 4; j++) valid |= (uint32_t)(row0 + j < a.n) << j;
#pragma unroll
      for (int d = 0; d < kMaxDims; d++) {
        if (d >= w.nd) continue;
        kk[d] = WaveLoadVec4<false>(a.dim[d].fk, row0, a.n);
        uint32_t m = 0, bad = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const int32_t key = vget(kk[d], j);
          const bool vj = (valid >> j) & 1u;
          bad |= (uint32_t)(vj && key < a.dim[d].keys_min) << j;
          m |= (uint32_t)(vj && bloom_probe_rt(a.dim[d], lds_filter, key)) << j;
        }
        pass[d] = m;
        st.v[ST_OOB] += popc_ballot4(bad);
        const int c = popc_ballot4(m);
        if (lane == 0) atomicAdd(&lds_samp[d], (unsigned long long)c);
      }
      const int nv = popc_ballot4(valid);
      if (lane == 0) atomicAdd(&lds_samp[kMaxDims], (unsigned long long)nv);
      st.v[ST_SAMPLE_ROWS] += nv;
      st.v[ST_SAMPLE_LINES] += (long long)w.nd * lines_of_ballot(__ballot(valid != 0));
    }
    __syncthreads();
    if (lds_samp[kMaxDims] > 0) {
      // rank of each dimension by sampled pass count (ties: lower index first)
      unsigned long long cntd[kMaxDims];
#pragma unroll
      for (int d = 0; d < kMaxDims; d++) cntd[d] = d < w.nd ? lds_samp[d] : ~0ull;
      int rank[kMaxDims];
#pragma unroll
      for (int d = 0; d < kMaxDims; d++) {
        rank[d] = 0;
#pragma unroll
        for (int e = 0; e < kMaxDims; e++) rank[d] += (cntd[e] < cntd[d]) || (cntd[e] == cntd[d] && e < d);
      }
      // stage s loads the dimension of rank s (value selects, no indexed stores)
      int o[kMaxDims];
#pragma unroll
      for (int s2 = 0; s2 < kMaxDims; s2++) {
        o[s2] = 0;
#pragma unroll
        for (int d = 0; d < kMaxDims; d++) o[s2] += rank[d] == s2 ? d : 0;
      }
      w.ld0 = uniform(o[0]);
      w.ld1 = uniform(o[1]);
      w.ld2 = uniform(o[2]);
      w.ld3 = uniform(o[3]);
    }
    if (has) {
      // finish the sampled rows in place: every Bloom filter is already applied
      uint32_t mask = valid;
#pragma unroll
      for (int d = 0; d < kMaxDims; d++)
        if (d < w.nd) mask &= pass[d];
      st.v[ST_BLOOM_PASS] += popc_ballot4(mask);
      Quad pay[4];
#pragma unroll
      for (int j = 0; j < 4; j++) pay[j].v[0] = pay[j].v[1] = pay[j].v[2] = pay[j].v[3] = 0;
#pragma unroll
      for (int t = 0; t < kMaxDims; t++) {
        if (t >= w.nd) continue;
        const int e = a.order[t];
        const v4i32 ke = e == 0 ? kk[0] : e == 1 ? kk[1] : e == 2 ? kk[2] : kk[3];
        uint32_t nm = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          if ((mask >> j) & 1u) {
            int32_t p = 0;
            bool ok = pht_probe(a.dim[e], vget(ke, j), p);
            pay[j].set(e, p);
            nm |= (uint32_t)ok << j;
          }
        }
        mask = nm;
      }
      st.v[ST_SAMPLE_LINES] += lines_of_ballot(__ballot(mask != 0));
      const v4i32 mv = WavePredLoadVec4(a.measure, row0, a.n, mask);
#pragma unroll
      for (int j = 0; j < 4; j++) st.v[ST_OOB] += group_add(a, gt, (mask >> j) & 1u, pay[j], vget(mv, j));
    }
  }

  // ---- main loop over tiles ------------------------------------------------------
  const int ld0 = w.ld0;
  const ProbeDim& D0 = a.dim[ld0];
  const int64_t tile_rows = (int64_t)VEC * kSegRows;
  const int64_t rest = a.n - a.sample_rows;
  const int64_t n_tiles = rest > 0 ? (rest + tile_rows - 1) / tile_rows : 0;

  v4i32 cur[VEC];
  int64_t t = gw;
  if (t < n_tiles) {
#pragma unroll
    for (int v = 0; v < VEC; v++)
      cur[v] = WaveLoadVec4<true>(D0.fk, a.sample_rows + t * tile_rows + v * kSegRows + lane * 4, a.n);
  }
  while (t < n_tiles) {
    const int64_t tn = t + total_waves;
    v4i32 nxt[VEC];
    if (tn < n_tiles) {
#pragma unroll
      for (int v = 0; v < VEC; v++)
        nxt[v] = WaveLoadVec4<true>(D0.fk, a.sample_rows + tn * tile_rows + v * kSegRows + lane * 4, a.n);
    }
    const int64_t base = a.sample_rows + t * tile_rows;
    int32_t pay[VEC][4];
    uint32_t mask[VEC];
    stage0_dispatch<VEC>(a, D0, lds_filter, cur, base, pay, mask, oob);

    // stage-0 statistics: every valid row of the tile is read
    {
      int64_t nvalid = a.n - base;
      if (nvalid > tile_rows) nvalid = tile_rows;
      st.v[ST_ROWS_IN + 0] += nvalid;
      st.v[ST_LINES_IN + 0] += (nvalid + 31) / 32;
    }
    // one (not unrolled) pass per segment; the segment's registers are picked
    // by uniform selects so that no register array is indexed dynamically
    for (int v = 0; v < VEC; v++) {
      uint32_t m = mask[0];
      v4i32 kv = cur[0];
      int32_t pv[4] = {pay[0][0], pay[0][1], pay[0][2], pay[0][3]};
#pragma unroll
      for (int u = 1; u < VEC; u++) {
        if (u == v) {
          m = mask[u];
          kv = cur[u];
#pragma unroll
          for (int j = 0; j < 4; j++) pv[j] = pay[u][j];
        }
      }
      const int cnt = popc_ballot4(m);
      if (cnt == 0) continue;
      const int64_t row0 = base + v * kSegRows + lane * 4;
      if (cnt > kDenseMin) {
        process_dense(a, lds_filter, gt, w, row0, kv, pv, m, st);
        continue;
      }
      if (w.qcount + cnt > kQueue) {
        process_batch(a, lds_filter, gt, w, w.qcount, st);
        w.qcount = 0;
      }
      enqueue(w, row0, kv, pv, m, cnt);
      if (w.qcount == kQueue) {
        process_batch(a, lds_filter, gt, w, kQueue, st);
        w.qcount = 0;
      }
    }
#pragma unroll
    for (int v = 0; v < VEC; v++) cur[v] = nxt[v];
    t = tn;
  }
  if (w.qcount > 0) process_batch(a, lds_filter, gt, w, w.qcount, st);

  // ---- flush: statistics and groups ----------------------------------------------
  {
    int o = oob;
#pragma unroll
    for (int off2 = 32; off2 > 0; off2 >>= 1) o += __shfl_down(o, off2, 64);
    if (lane == 0) {
      st.v[ST_OOB] += o;
      for (int i = 0; i < kStats; i++)
        if (i != ST_ORDER && st.v[i]) atomicAdd((unsigned long long*)&lds_stats[i], (unsigned long long)st.v[i]);
      atomicAdd((unsigned long long*)&lds_stats[ST_ROWS_IN + a.n_dims], (unsigned long long)st.meas_rows);
      atomicAdd((unsigned long long*)&lds_stats[ST_LINES_IN + a.n_dims], (unsigned long long)st.meas_lines);
      if (blockIdx.x == 0 && wave == 0) {
        long long packed = 0;
        packed = (long long)(w.ld0 & 0xff) | ((long long)(w.ld1 & 0xff) << 8) | ((long long)(w.ld2 & 0xff) << 16) |
                 ((long long)(w.ld3 & 0xff) << 24);
        lds_stats[ST_ORDER] = packed;
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kGroupSlots; i += WG) {
    const int32_t g = gt.key[i];
    if (g != kEmptyGroup && gt.cnt[i] != 0) {
      Quad pay;
      for (int d = 0; d < kMaxDims; d++) pay.v[d] = gt.pay[i * kMaxDims + d];
      global_group_add(a, g, (long long)gt.sum[i], gt.cnt[i], gt.prod[i], pay, a.n_dims);
    }
  }
  for (int i = threadIdx.x; i < kStats; i += WG) a.stats[(int64_t)blockIdx.x * kStats + i] = lds_stats[i];
}

// ---- host side -------------------------------------------------------------------
size_t probe_lds_bytes(const ProbeConfig& c, int lds_filter_words) {
  size_t off = ((size_t)lds_filter_words * 4 + 15) & ~(size_t)15;
  off += (size_t)(c.wg / 64) * 3 * kQueue * 4;
  off += 3 * kGroupSlots * 8 + kGroupSlots * 4 + kGroupSlots * kMaxDims * 4;
  off += kStats * 8 + (kMaxDims + 1) * 8;
  return off;
}

#define LJ_PROBE_CONFIGS(X) \
  X(256, 1) X(256, 2) X(256, 4) X(512, 1) X(512, 2) X(512, 4) X(1024, 1) X(1024, 2) X(1024, 4)

static const void* probe_fn(const ProbeConfig& c) {
#define X(W, V) \
  if (c.wg == W && c.vec == V) return reinterpret_cast<const void*>(&probe_kernel<W, V>);
  LJ_PROBE_CONFIGS(X)
#undef X
  return nullptr;
}

hipError_t probe_occupancy(const ProbeConfig& c, size_t lds_bytes, int* blocks_per_cu) {
  const void* fn = probe_fn(c);
  if (!fn) return hipErrorInvalidValue;
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
  if (e != hipSuccess) return e;
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, fn, c.wg, lds_bytes);
}

hipError_t launch_probe(const ProbeConfig& c, const ProbeArgs& a, int grid, size_t lds_bytes, hipStream_t s) {
#define X(W, V)                                                                                  \
  if (c.wg == W && c.vec == V) {                                                                 \
    hipLaunchKernelGGL((probe_kernel<W, V>), dim3(grid), dim3(W), lds_bytes, s, a);              \
    return hipGetLastError();                                                                    \
  }
  LJ_PROBE_CONFIGS(X)
#undef X
  return hipErrorInvalidValue;
}

}  // namespace lj
