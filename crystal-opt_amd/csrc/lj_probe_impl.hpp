// lj_probe_impl.hpp -- the fused fact-table probe of the star join (kernel template).
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
//   * One kernel per stage-0 filter kind (template K0), so the streaming loop
//     of a launch holds no per-row dispatch.  Persistent grid; each wave walks
//     256*VEC-row tiles with a grid stride; lane l owns rows 4l..4l+3 of each
//     256-row segment and reads them with one 16-byte non-temporal load.  The
//     next tile's loads are issued before the current tile is processed.
//   * Stage 0 (the first FK column, read in full) probes its filter for every
//     row (a Bloom filter staged in LDS, or the PHT for NAIVE).
//   * A segment with few stage-0 survivors pushes them into a per-wave
//     64-entry LDS queue (row, key, payload) by ballot/mbcnt; a full queue is
//     processed with all 64 lanes busy (gather next FK, probe, ..., PHT_2
//     probes, measure gather, group add).  A segment with many survivors is
//     processed in place with predicated 16-byte loads.  When a wave sees a
//     high stage-0 pass rate it also prefetches the next column with the
//     first one ("eager"), so dense tiles do not wait on a dependent load.
//   * LIP_ADAPTIVE: every wave first probes ALL Bloom filters on every 4th
//     row of one 256-row sample segment; the workgroup sums the pass counts
//     in LDS and each wave orders the filters by ascending pass count (the
//     totals so far).
//   * Group-by: a 32-slot open-addressing table in LDS per workgroup holds
//     int64 sums/counts, flushed with 64-bit global atomics at the end; a slot
//     miss goes straight to the global atomics.  The first global touch of a
//     group appends it to the `touched` list.
//   * Exact per-stage statistics: rows alive and 128-byte lines holding an
//     alive row when each column is read (the algorithmic byte count).
#pragma once

#include <hip/hip_runtime.h>

#include "lipjoin/block/crystal.hpp"
#include "lj_internal.hpp"

namespace lj {

constexpr int kStageUnroll = 8;
constexpr int kOccUnroll = 8;   // table loads in flight per lane while the occupancy bitmap builds     // 16-byte filter loads in flight per thread while staging
constexpr int kFinUnroll = 8;       // statistics rows per thread read together by the finalize
constexpr int kDenseMin = 8;         // stage-0 survivors of a segment above which it is processed in place
constexpr int kGroupProbes = 32;     // slots a row tries in the LDS group table before going global
static_assert(kGroupSlotBytes == 3 * 8 + 4 + kMaxDims * 4, "group slot layout");
static_assert(kQueue + kDenseMin <= kQueueCap, "below kQueue entries, a sparse segment always fits");
static_assert((kQueueCap - kQueue) + 16 * 4 <= kQueueCap, "after a rotation, a dense part always fits");
constexpr int32_t kEmptyGroup = -1;

// Wave-uniform counters (32-bit: one wave never sees 2^32 rows); indices into
// v[] are compile-time constants after unrolling, the measure stage (index
// n_dims, a runtime value) has its own fields and is placed at flush time.
struct WaveStats {
  uint32_t v[kStats];
  uint32_t meas_rows;
  uint32_t meas_lines;
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

// nonzero bytes of a ballot over "lane holds an alive row" (8 lanes x 4 rows = one 128-byte line)
__device__ __forceinline__ int lines_of_ballot(unsigned long long b) {
  b |= b >> 4;
  b |= b >> 2;
  b |= b >> 1;
  return __popcll(b & 0x0101010101010101ull);
}

// ---- filter probes ----------------------------------------------------------
typedef const __attribute__((address_space(3))) uint32_t* LdsWords;
typedef const __attribute__((address_space(1))) uint32_t* GlobalWords;

// A column or table pointer of the query's arguments, typed as global memory
// at the load.  The arguments of a batch are in device memory, and a pointer
// loaded from there is "generic" to the compiler: a load through it would be
// a flat load, which waits for both the vector-memory and the LDS counters.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* gl(const T* p) {
  return (const __attribute__((address_space(1))) T*)p;
}

// A filter's words are read through a pointer typed with its address space
// (LDS copy or global words, a uniform choice): a pointer that could be
// either would compile to flat loads, each followed by a wait for all of the
// wave's outstanding memory operations (the stream's prefetched tiles too).
// Bits of a filter staged as a prefix: the LDS copy below bloom_lds_words,
// the global words past it.  FB false (the probe checked at its start that no
// bit is set past any staged prefix): a bit past the prefix is 0, and only
// LDS is read.  FB true: a bit past the prefix is read from the global words
// (a global read in this code makes the wave wait for its vector-memory
// counter, the stream's prefetched tiles included, so the fast body has none).
template <bool FB>
__device__ __forceinline__ bool staged_bit(const ProbeDim& D, const uint32_t* lds, uint32_t h) {
  const uint32_t P = (uint32_t)D.bloom_lds_words, w = h >> 5;
  const LdsWords f = (LdsWords)(lds + D.bloom_lds);
  if (P == (uint32_t)D.bloom_words) return (f[w] >> (h & 31)) & 1u;  // the whole filter (uniform)
  if constexpr (FB) {
    return ((w < P ? f[w] : ((GlobalWords)D.bloom)[w]) >> (h & 31)) & 1u;
  } else {
    return w < P && ((f[w < P ? w : 0u] >> (h & 31)) & 1u);
  }
}

// The 4 bits h[j] of a lane's rows, as a mask.
template <bool FB>
__device__ __forceinline__ uint32_t staged_bits4(const ProbeDim& D, const uint32_t* lds, const uint32_t (&h)[4]) {
  const uint32_t P = (uint32_t)D.bloom_lds_words;
  const LdsWords f = (LdsWords)(lds + D.bloom_lds);
  uint32_t w[4], m = 0;
  if (P == (uint32_t)D.bloom_words) {  // the whole filter (uniform)
#pragma unroll
    for (int j = 0; j < 4; j++) w[j] = f[h[j] >> 5];
#pragma unroll
    for (int j = 0; j < 4; j++) m |= __builtin_amdgcn_ubfe(w[j], h[j], 1u) << j;
    return m;
  }
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const uint32_t wi = h[j] >> 5;
    if constexpr (FB) w[j] = wi < P ? f[wi] : ((GlobalWords)D.bloom)[wi];
    else w[j] = f[wi < P ? wi : 0u];
  }
#pragma unroll
  for (int j = 0; j < 4; j++) {
    uint32_t b = __builtin_amdgcn_ubfe(w[j], h[j], 1u);
    if constexpr (!FB) b &= (uint32_t)((h[j] >> 5) < P);
    m |= b << j;
  }
  return m;
}

__device__ __forceinline__ bool bloom_probe_rt(const ProbeDim& D, const uint32_t* lds, int32_t key) {
  if (D.bh.kind == BLOOM_NONE) return true;
  if (D.bloom_lds >= 0) {
    if (D.bloom_lds_words == D.bloom_words) return bloom_test(D.bh, (LdsWords)(lds + D.bloom_lds), key);
    // a staged prefix (a modulo hash: one bit per key)
    uint32_t bits[3];
    if (bloom_bits(D.bh, key, bits) == 0) return false;
    return staged_bit<true>(D, lds, bits[0]);
  }
  return bloom_test(D.bh, (GlobalWords)D.bloom, key);
}

// ---- hash-table lookups (the kind is uniform: a scalar branch) -------------
// PHT_2 (join.cuh:108-189): one 64-bit load of the (key, value) pair, a hit
// is a nonzero pair, no key compare.  PHT_1 (join.cuh:29-106): one 32-bit
// load, a hit is a nonzero key, no payload (0).  Textbook: the pair of the
// home slot, then linear probing until the key or an empty slot.
// ht_first_slot: the slot to load first (-1: the key is outside the domain).
__device__ __forceinline__ int64_t ht_first_slot(int32_t kind, const FastMod& len, int32_t kmin, int32_t key) {
  if (kind == HT_TEXTBOOK) return key == kTbExcludedKey ? -1 : (int64_t)tb_home(key, len.pow2_mask);
  return pht_slot(key, len, kmin);
}
__device__ __forceinline__ uint64_t ht_load(int32_t kind, const int32_t* ht, int64_t s) {
  if (kind == HT_PHT1) return (uint64_t)(uint32_t)gl(ht)[s];
  return gl(reinterpret_cast<const uint64_t*>(ht))[s];
}
// Hit and payload from `raw`, the first load (slot s >= 0).
__device__ __forceinline__ bool ht_resolve(int32_t kind, const int32_t* ht, const FastMod& len, int32_t key,
                                           int64_t s, uint64_t raw, int32_t& pay) {
  if (kind == HT_TEXTBOOK) {
    const uint32_t enc = tb_encode(key), mask = len.pow2_mask;
    for (uint32_t i = 0; i <= mask; i++) {
      const uint32_t k = (uint32_t)raw;
      if (k == enc) {
        pay = (int32_t)(uint32_t)(raw >> 32);
        return true;
      }
      if (k == 0) return false;
      s = (s + 1) & mask;
      raw = gl(reinterpret_cast<const uint64_t*>(ht))[s];
    }
    return false;
  }
  if (raw == 0) return false;
  pay = (int32_t)(uint32_t)(raw >> 32);
  return true;
}

__device__ __forceinline__ bool pht_probe(const ProbeDim& D, int32_t key, int32_t& pay) {
  const int64_t s = ht_first_slot(D.ht_kind, D.ht_len, D.keys_min, key);
  if (s < 0) return false;
  return ht_resolve(D.ht_kind, D.ht, D.ht_len, key, s, ht_load(D.ht_kind, D.ht, s), pay);
}

// A key whose table probe an exact filter decides (D.exact): with x =
// key - keys_min in [0, ht_len), the filter's bit x and the table's slot x
// are the same position, and every selected dimension key is inside that
// domain (the probe checks the build's out-of-domain count), so the bit is
// set iff the key was inserted.  Key 0 is left to the table: its pair may be
// all zero, which the reference reads as a miss (the zero sentinel).
__device__ __forceinline__ bool exact_key(const ProbeDim& D, int32_t key) {
  return (uint32_t)key - (uint32_t)D.keys_min < D.ht_len.d && key != 0;
}

// A filter probe outside stage 0 (later stages, the adaptive sample).  When
// every filter of the plan has stage 0's kind and placement (all_k0), the
// probe is that kind's short test; otherwise the generic test.  Besides the
// work, this keeps the executed code small: the generic test inlined at each
// call site is thousands of instructions, and walking them misses the
// instruction cache.  SPEC: 0 tests all_k0 here, 1 assumes it (the caller
// tested it once for a whole loop), 2 assumes it is not set, 3 is the
// one-dimension in-place body (one filter: all_k0 holds).
template <int K0, int SPEC = 0>
__device__ __forceinline__ bool probe_filter(const ProbeArgs& a, const ProbeDim& D, const uint32_t* lds,
                                             int32_t key) {
  const bool spec = SPEC == 1 || SPEC == 3 || (SPEC == 0 && a.all_k0);
  if constexpr (K0 == S0_MODP2_LDS || K0 == S0_MOD_LDS || K0 == S0_MODP2_GL || K0 == S0_MOD_GL) {
    if (spec) {
      const uint32_t x = (uint32_t)key - (uint32_t)D.bh.keys_min;
      uint32_t h;
      if constexpr (K0 == S0_MODP2_LDS || K0 == S0_MODP2_GL) h = x & D.bh.mod.pow2_mask;
      else h = fastmod_u32(x, D.bh.mod);
      bool bit;
      if (D.bloom_lds >= 0) bit = staged_bit<SPEC != 1>(D, lds, h);
      else bit = (((GlobalWords)D.bloom)[h >> 5] >> (h & 31)) & 1u;
      return key >= D.bh.keys_min && bit;
    }
  } else if constexpr (K0 == S0_MULT_LDS || K0 == S0_MULT_GL) {
    if (spec) {
      if (D.bloom_lds >= 0) return bloom_test(D.bh, (LdsWords)(lds + D.bloom_lds), key);
      return bloom_test(D.bh, (GlobalWords)D.bloom, key);
    }
  }
  return bloom_probe_rt(D, lds, key);
}

// The filter probes of a lane's 4 rows (bit j of mask4: row j alive): with
// the plan's hash family, the 4 words are read together (one wait) and a
// dead row's read is harmless (its bit index is in range whatever its key).
// Returns the surviving rows.
template <int K0, int SPEC>
__device__ __forceinline__ uint32_t probe_filter4(const ProbeArgs& a, const ProbeDim& D, const uint32_t* lds,
                                                  const v4i32& kv, uint32_t mask4) {
  const bool spec = SPEC == 1 || SPEC == 3 || (SPEC == 0 && a.all_k0);
  if constexpr (K0 == S0_MODP2_LDS || K0 == S0_MOD_LDS || K0 == S0_MODP2_GL || K0 == S0_MOD_GL) {
    if (spec) {
      const int32_t kk[4] = {kv.x, kv.y, kv.z, kv.w};
      uint32_t h[4];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const uint32_t x = (uint32_t)kk[j] - (uint32_t)D.bh.keys_min;
        if constexpr (K0 == S0_MODP2_LDS || K0 == S0_MODP2_GL) h[j] = x & D.bh.mod.pow2_mask;
        else h[j] = fastmod_u32(x, D.bh.mod);
      }
      uint32_t ok = 0;
      if (D.bloom_lds >= 0) {
        ok = staged_bits4<SPEC != 1>(D, lds, h);
      } else {
        const GlobalWords f = (GlobalWords)D.bloom;
        uint32_t w[4];
#pragma unroll
        for (int j = 0; j < 4; j++) w[j] = f[h[j] >> 5];
#pragma unroll
        for (int j = 0; j < 4; j++) ok |= __builtin_amdgcn_ubfe(w[j], h[j], 1u) << j;
      }
#pragma unroll
      for (int j = 0; j < 4; j++) ok &= ~((uint32_t)(kk[j] < D.bh.keys_min) << j);
      return ok & mask4;
    }
  }
  uint32_t ok = 0;
#pragma unroll
  for (int j = 0; j < 4; j++)
    if ((mask4 >> j) & 1u) ok |= (uint32_t)probe_filter<K0, SPEC>(a, D, lds, vget(kv, j)) << j;
  return ok;
}

// Stage-0 parameters, hoisted out of the streaming loop.
struct S0 {
  const int32_t* fk;
  const uint32_t* bf;    // staged copy in LDS or the global words
  const uint32_t* gbf;   // the global words
  uint32_t lds_words;    // words of bf that are staged (a prefix); the rest is read from gbf
  bool whole;            // the whole filter is staged
  const int32_t* ht;
  BloomHash bh;
  FastMod fm;
  int32_t kmin;
  uint32_t mask;
  int32_t ht_kind;
};

template <int K0>
__device__ __forceinline__ bool s0_probe(const S0& s, int32_t key, int32_t& pay);

// Stage 0 over a lane's 4 rows, branch-free: every table word is read
// unconditionally (a masked bit index is always inside the filter; an invalid
// PHT slot reads slot 0), so the 4 lookups are in flight together and the
// result is fixed up with selects.  Returns the alive mask (bit j = row j).
template <int K0, bool FB>
__device__ __forceinline__ uint32_t s0_probe4(const S0& s, const v4i32& k, uint32_t valid4, int32_t (&pay)[4]) {
  const int32_t kk[4] = {k.x, k.y, k.z, k.w};
  uint32_t ok = 0;
  if constexpr (K0 == S0_MODP2_LDS || K0 == S0_MODP2_GL || K0 == S0_MOD_LDS || K0 == S0_MOD_GL) {
    // bit h = (key - keys_min) mod bits: word h >> 5, bit h & 31 (a power-of-two
    // size masks the word index only: the bit index is the low 5 bits either
    // way).  A key below keys_min is counted out of domain by the caller; its
    // bit is tested like any other and a hash probe later rejects it.
    uint32_t h[4], wv[4];
    // a power-of-two size of 2^(5 + wbits) bits: word (x >> 5) & (2^wbits - 1)
    const uint32_t wbits = (uint32_t)__popc(s.mask) - 5u;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const uint32_t x = (uint32_t)kk[j] - (uint32_t)s.bh.keys_min;
      if constexpr (K0 == S0_MODP2_LDS || K0 == S0_MODP2_GL) h[j] = x;
      else h[j] = fastmod_u32(x, s.bh.mod);
    }
    uint32_t wi[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if constexpr (K0 == S0_MODP2_LDS || K0 == S0_MODP2_GL) wi[j] = __builtin_amdgcn_ubfe(h[j], 5u, wbits);
      else wi[j] = h[j] >> 5;
    }
    // v_bfe_u32 reads bits [4:0] of the offset: bit h & 31 of the word
    if constexpr (K0 == S0_MODP2_LDS || K0 == S0_MOD_LDS) {
      // the staged prefix (see staged_bit): FB reads a word past it from the
      // global words; otherwise such a bit is 0.  The whole filter staged (a
      // uniform test): plain LDS reads.
      const uint32_t P = s.lds_words;
      if (s.whole) {
#pragma unroll
        for (int j = 0; j < 4; j++) wv[j] = ((LdsWords)s.bf)[wi[j]];
#pragma unroll
        for (int j = 0; j < 4; j++) ok |= __builtin_amdgcn_ubfe(wv[j], h[j], 1u) << j;
      } else {
#pragma unroll
      for (int j = 0; j < 4; j++) {
        if constexpr (FB) wv[j] = wi[j] < P ? ((LdsWords)s.bf)[wi[j]] : ((GlobalWords)s.gbf)[wi[j]];
        else wv[j] = ((LdsWords)s.bf)[wi[j] < P ? wi[j] : 0u];
      }
#pragma unroll
      for (int j = 0; j < 4; j++) {
        uint32_t b = __builtin_amdgcn_ubfe(wv[j], h[j], 1u);
        if constexpr (!FB) b &= (uint32_t)(wi[j] < P);
        ok |= b << j;
      }
      }
    } else {
#pragma unroll
      for (int j = 0; j < 4; j++) wv[j] = ((GlobalWords)s.bf)[wi[j]];
#pragma unroll
      for (int j = 0; j < 4; j++) ok |= __builtin_amdgcn_ubfe(wv[j], h[j], 1u) << j;
    }
#pragma unroll
    for (int j = 0; j < 4; j++) pay[j] = 0;
  } else if constexpr (K0 == S0_PHT) {
    if (s.ht_kind == HT_PHT2) {
      uint64_t slot[4];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int64_t sl = pht_slot(kk[j], s.fm, s.kmin);
        slot[j] = gl(reinterpret_cast<const uint64_t*>(s.ht))[sl < 0 ? 0 : sl];
        ok |= (uint32_t)(sl >= 0) << j;
      }
#pragma unroll
      for (int j = 0; j < 4; j++) {
        ok &= ~((uint32_t)(slot[j] == 0) << j);
        pay[j] = (int32_t)(uint32_t)(slot[j] >> 32);
      }
    } else {
      // PHT_1 / textbook: the first loads of the 4 rows together, then each
      // row's resolution (a textbook row may probe further slots)
      int64_t sl[4];
      uint64_t raw[4];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        sl[j] = ht_first_slot(s.ht_kind, s.fm, s.kmin, kk[j]);
        raw[j] = ht_load(s.ht_kind, s.ht, sl[j] < 0 ? 0 : sl[j]);
      }
#pragma unroll
      for (int j = 0; j < 4; j++) {
        int32_t p = 0;
        const bool hit = ((valid4 >> j) & 1u) && sl[j] >= 0 && ht_resolve(s.ht_kind, s.ht, s.fm, kk[j], sl[j], raw[j], p);
        ok |= (uint32_t)hit << j;
        pay[j] = p;
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      int32_t p = 0;
      ok |= (uint32_t)s0_probe<K0>(s, kk[j], p) << j;
      pay[j] = p;
    }
  }
  return ok & valid4;
}

template <int K0>
__device__ __forceinline__ bool s0_probe(const S0& s, int32_t key, int32_t& pay) {
  if constexpr (K0 == S0_PHT) {
    const int64_t sl = ht_first_slot(s.ht_kind, s.fm, s.kmin, key);
    return sl >= 0 && ht_resolve(s.ht_kind, s.ht, s.fm, key, sl, ht_load(s.ht_kind, s.ht, sl), pay);
  } else if constexpr (K0 == S0_NONE) {
    return true;
  } else if constexpr (K0 == S0_MODP2_LDS || K0 == S0_MODP2_GL || K0 == S0_MOD_LDS || K0 == S0_MOD_GL) {
    const uint32_t x = (uint32_t)key - (uint32_t)s.bh.keys_min;
    uint32_t h;
    if constexpr (K0 == S0_MODP2_LDS || K0 == S0_MODP2_GL) h = x & s.mask;
    else h = fastmod_u32(x, s.bh.mod);
    const uint32_t w = h >> 5;
    const uint32_t word = w < s.lds_words ? s.bf[w] : s.gbf[w];
    return key >= s.bh.keys_min && ((word >> (h & 31)) & 1u);
  } else if constexpr (K0 == S0_MULT_LDS || K0 == S0_MULT_GL) {
    return bloom_test(s.bh, s.bf, key);
  } else {  // S0_GENERIC
    if (s.bh.kind == BLOOM_NONE) return true;
    if (s.bh.kind == BLOOM_MOD || s.bh.kind == BLOOM_MOD_POW2) {
      uint32_t bits[3];
      if (bloom_bits(s.bh, key, bits) == 0) return false;
      const uint32_t w = bits[0] >> 5;
      return ((w < s.lds_words ? s.bf[w] : s.gbf[w]) >> (bits[0] & 31)) & 1u;
    }
    return bloom_test(s.bh, s.bf, key);
  }
}

// ---- per-dimension register sets (uniform dimension index; no indexed VGPRs)
struct Quad {
  int32_t v[kMaxDims];
  __device__ __forceinline__ void set(int d, int32_t x) {
    if (d == 0) v[0] = x;
    else if (d == 1) v[1] = x;
    else if (d == 2) v[2] = x;
    else v[3] = x;
  }
  __device__ __forceinline__ int32_t get(int d) const { return d == 0 ? v[0] : d == 1 ? v[1] : d == 2 ? v[2] : v[3]; }
  __device__ __forceinline__ void clear() { v[0] = v[1] = v[2] = v[3] = 0; }
};

// ---- group table in LDS --------------------------------------------------
struct GroupLds {
  int32_t* key;             // [slots]
  unsigned long long* cnt;  // [slots]
  unsigned long long* sum;  // [slots]
  unsigned long long* prod; // [slots]
  int32_t* pay;             // [slots * 4]
};

__device__ __forceinline__ void global_group_add(const ProbeArgs& a, int64_t g, long long s, unsigned long long c,
                                                 unsigned long long p, const Quad& pay, int nd) {
  if (a.g_reps > 1) {
    // this workgroup's copy; the first add of a group to any copy lists it once
    const int64_t gi = g + (int64_t)(blockIdx.x & (uint32_t)(a.g_reps - 1)) * a.g_stride;
    atomicAdd(&a.gsum[gi], (unsigned long long)s);
    if (a.want_product) atomicAdd(&a.gprod[gi], p);
    const unsigned long long old = atomicAdd(&a.gcount[gi], c);
    if (old == 0 && atomicExch(&a.gseen[g], 1u) == 0u) {
#pragma unroll
      for (int t = 0; t < kMaxDims; t++)
        if (t < nd) a.gpay[g * kMaxDims + t] = pay.v[t];
      uint32_t idx = atomicAdd(a.n_touched, 1u);
      if (idx < a.touched_cap) a.touched[idx] = (uint32_t)g;
    }
    return;
  }
  atomicAdd(&a.gsum[g], (unsigned long long)s);
  if (a.want_product) atomicAdd(&a.gprod[g], p);
  unsigned long long old = atomicAdd(&a.gcount[g], c);
  if (old == 0) {
#pragma unroll
    for (int t = 0; t < kMaxDims; t++)
      if (t < nd) a.gpay[g * kMaxDims + t] = pay.v[t];
    uint32_t idx = atomicAdd(a.n_touched, 1u);
    if (idx < a.touched_cap) a.touched[idx] = (uint32_t)g;
  }
}

// Add one row per alive lane to the workgroup's group table.  Called by every
// lane of the wave; returns the number of rows (wave-wide) whose group index
// fell outside [0, n_groups).  pay.v[d] is dimension d's payload.
__device__ __forceinline__ int group_add(const ProbeArgs& a, GroupLds& gt, bool alive, const Quad& pay, int32_t m) {
  const int nd = a.n_dims;
  const int32_t pay0 = pay.v[0];  // the product's payload is dimension 0's
  if (a.n_groups == 1) {
    // one group (single join): wave-reduce, one LDS atomic per wave
    long long s = alive ? (long long)m : 0;
    unsigned long long p = 0;
    if (a.want_product && alive) p = (unsigned long long)(long long)(int32_t)((uint32_t)m * (uint32_t)pay0);
    const unsigned long long b = __ballot(alive);
    if (b == 0) return 0;
    s = WaveSum(s);
    if (a.want_product) p = WaveSum(p);
    if (lane_id() == 0) {
      atomicAdd(&gt.sum[0], (unsigned long long)s);
      atomicAdd(&gt.cnt[0], (unsigned long long)__popcll(b));
      if (a.want_product) atomicAdd(&gt.prod[0], p);
      if (gt.key[0] == kEmptyGroup) {
        gt.key[0] = 0;
        for (int t = 0; t < kMaxDims; t++)
          if (t < nd) gt.pay[t] = pay.v[t];
      }
    }
    return 0;
  }
  int64_t g = 0;
  if (alive) {
    int64_t acc = a.goffset;
#pragma unroll
    for (int t = 0; t < kMaxDims; t++)
      if (t < nd) acc += (int64_t)pay.v[t] * a.dim[t].gmult;
    // (a 64-bit remainder is a long instruction sequence; the group index
    // almost always fits 32 bits)
    g = acc >= 0 && acc <= (int64_t)UINT32_MAX ? (int64_t)fastmod_u32((uint32_t)acc, a.gmod) : acc % a.n_groups;
  }
  const int n_bad = __popcll(__ballot(alive && g < 0));
  if (!alive || g < 0) return n_bad;
  const unsigned long long p =
      a.want_product ? (unsigned long long)(long long)(int32_t)((uint32_t)m * (uint32_t)pay0) : 0ull;
  const uint32_t smask = (1u << a.gslot_bits) - 1u;
  uint32_t slot = ((uint32_t)g * 0x9E3779B1u) >> (32 - a.gslot_bits);
  for (int probe = 0; probe < kGroupProbes; probe++) {
    int32_t k = gt.key[slot];
    if (k == kEmptyGroup) {
      const int32_t old = atomicCAS(&gt.key[slot], kEmptyGroup, (int32_t)g);
      if (old == kEmptyGroup) {
        for (int t = 0; t < kMaxDims; t++)
          if (t < nd) gt.pay[slot * kMaxDims + t] = pay.v[t];
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
    slot = (slot + 1) & smask;
  }
  global_group_add(a, g, (long long)m, 1ull, p, pay, nd);
  return n_bad;
}

// ---- line statistics for a compacted batch (rows ascending with the lane) -----
__device__ __forceinline__ int batch_lines(bool alive, int32_t row, int32_t& last_line) {
  // last_line: line of the last alive row of this stage in earlier batches
  const int32_t line = alive ? (row >> 5) : -1;
  int32_t pm = line;  // inclusive prefix max over lanes
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int32_t o = __shfl_up(pm, off, 64);
    if ((int)lane_id() >= off) pm = pm > o ? pm : o;
  }
  int32_t prev = __shfl_up(pm, 1, 64);
  if (lane_id() == 0) prev = -1;
  if (prev < 0) prev = last_line;
  const int cnt = __popcll(__ballot(alive && line != prev));
  const int32_t top = __shfl(pm, 63, 64);
  if (top >= 0) last_line = top;
  return cnt;
}

// ---- per-wave state (named scalars: accessed by compile-time stage numbers) ---
struct WaveCtx {
  int nd;
  bool exact;               // one dimension whose exact filter replaces the table probe
  bool direct;              // one dimension, one group: streamed survivors probe the table in place
  bool occ;                 // direct, with the table's occupancy bitmap built in LDS
  long long acc_s;          // exact / direct: the lane's SUM(measure), row count and
  uint32_t acc_c;           // SUM(measure * payload), added in place
  long long acc_p;
  int ld0, ld1, ld2, ld3;   // column load order (stage -> dim)
  int32_t* q_row;
  int32_t* q_key;
  int32_t* q_pay;
  int qcount;
  int32_t last1, last2, last3, last_meas;
  // the batch taken from the queue whose first gather is in flight
  int p_cnt;        // rows in it (0: none)
  int32_t p_row;    // lane's entry: row | bit 31 = stage 1 done
  int32_t p_x0, p_x1;
  int32_t p_k;      // key gathered for the entry's first remaining stage
  uint64_t p_ht;    // one dimension, LIP: the hash-table slot of the key (gathered)
  int32_t p_m;      // one dimension: the measure of the row (gathered)
  __device__ __forceinline__ int ld(int s) const { return s == 0 ? ld0 : s == 1 ? ld1 : s == 2 ? ld2 : ld3; }
  __device__ __forceinline__ int32_t last(int s) const { return s == 1 ? last1 : s == 2 ? last2 : last3; }
  __device__ __forceinline__ void set_last(int s, int32_t x) {
    if (s == 1) last1 = x;
    else if (s == 2) last2 = x;
    else last3 = x;
  }
};

// Queued rows are processed in batches of up to 64 (lane i takes entry i),
// software-pipelined: issue_batch takes the entries out of the queue into
// registers and starts the gather of each row's next FK column; the stream
// goes on, and complete_batch, called when the queue fills again (or at the
// end), finishes the rows, so the gather's latency hides behind tiles.
//
// An entry is (row, x0, x1): x is the key (LIP: the hash probes come last)
// or the payload (naive) of the stage's dimension; x0 is stage 0's, and x1
// stage 1's when bit 31 of row says the row passed stage 1 already (the
// survivors of a dense segment).  With one dimension, bit 31 says that x1 is
// the row's measure (read from the streamed measure column).
template <int K0, int SPEC>
__device__ __forceinline__ void issue_batch(const ProbeArgs& a, WaveCtx& w, int cnt) {
  constexpr bool LIP = K0 != S0_PHT;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const int i = (int)lane_id();
  const bool alive = i < cnt;
  int32_t r = 0, x0 = 0, x1 = 0, k = 0;
  if (alive) {
    r = w.q_row[i];
    x0 = w.q_key[i];
    x1 = w.q_pay[i];
  }
  const int first = r < 0 ? 2 : 1;
  if (alive && first < w.nd) {
    const int32_t* col = first == 1 ? a.dim[w.ld1].fk : a.dim[w.ld2].fk;
    k = gl(col)[(uint32_t)(r & 0x7FFFFFFF)];
  }
  // one dimension: the rows only need their hash slot (LIP) and measure; both
  // gathers go out now
  uint64_t ht = 0;
  int32_t m = 0;
  if (alive && w.nd == 1) {
    const uint32_t row = (uint32_t)(r & 0x7FFFFFFF);
    m = r < 0 ? x1 : gl(a.measure)[row];
    if (LIP) {
      const ProbeDim& D = a.dim[0];
      if (SPEC == 3 && w.exact && exact_key(D, x0)) {
        ht = 1ull;  // a hit (payload 0) that the filter decided
      } else {
        const int64_t sl = ht_first_slot(D.ht_kind, D.ht_len, D.keys_min, x0);
        if (sl >= 0) ht = ht_load(D.ht_kind, D.ht, sl);
      }
    }
  }
  w.p_ht = ht;
  w.p_m = m;
  w.p_cnt = cnt;
  w.p_row = r;
  w.p_x0 = x0;
  w.p_x1 = x1;
  w.p_k = k;
}

// The value of dimension t from values kept by stage (stage s reads
// dimension w.ld(s); the order is wave-uniform).
__device__ __forceinline__ int32_t stage_to_dim(const WaveCtx& w, const Quad& by_stage, int t) {
  return w.ld0 == t ? by_stage.v[0] : w.ld1 == t ? by_stage.v[1] : w.ld2 == t ? by_stage.v[2] : by_stage.v[3];
}

template <int K0, int SPEC>
__device__ __forceinline__ void complete_batch(const ProbeArgs& a, const uint32_t* lds, GroupLds& gt, WaveCtx& w,
                                               WaveStats& st) {
  constexpr bool LIP = K0 != S0_PHT;
  if (w.p_cnt == 0) return;
  const int i = (int)lane_id();
  bool alive = i < w.p_cnt;
  const bool done1 = alive && w.p_row < 0;
  const int32_t row = w.p_row & 0x7FFFFFFF;
  const int first = done1 ? 2 : 1;
  // keys and payloads by stage (stage s reads dimension w.ld(s)): registers
  // picked by compile-time stage numbers only; the final table probes and the
  // group add run by dimension (compile-time dimension numbers: the table
  // arguments at constant offsets), each dimension's key picked once from the
  // stages
  Quad key, pay;
  key.clear();
  pay.clear();
  if (alive) {
    if (LIP) key.v[0] = w.p_x0;
    else pay.v[0] = w.p_x0;
    if (done1 && w.nd >= 2) {
      if (LIP) key.v[1] = w.p_x1;
      else pay.v[1] = w.p_x1;
    }
  }
  const int nd = w.nd;
#pragma unroll
  for (int s = 1; s < kMaxDims; s++) {
    if (s >= nd) continue;
    const int d = w.ld(s);
    const ProbeDim& D = a.dim[d];
    const bool load = alive && s >= first;
    st.v[ST_ROWS_IN + s] += __popcll(__ballot(load));
    int32_t ls = w.last(s);
    st.v[ST_LINES_IN + s] += batch_lines(load, row, ls);
    w.set_last(s, ls);
    int32_t k = 0;
    if (load) k = s == first ? w.p_k : gl(D.fk)[(uint32_t)row];
    st.v[ST_OOB] += __popcll(__ballot(load && k < D.keys_min));
    if (load) {
      key.v[s] = k;
      if (!LIP) {
        int32_t p = 0;
        alive = pht_probe(D, k, p);
        pay.v[s] = p;
      } else {
        alive = probe_filter<K0, SPEC>(a, D, lds, k);
      }
    }
  }
  int32_t m = 0;
  Quad pd;  // payloads by dimension
  pd.clear();
  if (!LIP) {
#pragma unroll
    for (int t = 0; t < kMaxDims; t++) pd.v[t] = stage_to_dim(w, pay, t);
  }
  if (LIP) {
    const unsigned long long bloom_pass = __ballot(alive);
    st.v[ST_BLOOM_PASS] += __popcll(bloom_pass);
    if (nd == 1) {
      // the slot gathered by issue_batch (a perfect hash: a nonzero slot is a
      // hit; a textbook table goes on probing from it)
      const ProbeDim& D = a.dim[0];
      int32_t p = 0;
      if (D.ht_kind == HT_TEXTBOOK) {
        const int64_t sl = ht_first_slot(D.ht_kind, D.ht_len, D.keys_min, w.p_x0);
        alive = alive && sl >= 0 && ht_resolve(D.ht_kind, D.ht, D.ht_len, w.p_x0, sl, w.p_ht, p);
      } else {
        alive = alive && w.p_ht != 0;
        p = (int32_t)(uint32_t)(w.p_ht >> 32);
      }
      pd.v[0] = p;
      m = w.p_m;
    } else if (bloom_pass != 0) {
      // After the Bloom chain every key is known: the PHT_2 probes of all
      // dimensions and the measure gather go out together, one round trip
      // instead of one per join.  The loads are unconditional (a lane with no
      // row, or a key outside the table, reads slot 0 / row 0 and drops it),
      // so no branch separates them.  A row is alive iff every probe hits, as
      // with the probes taken one after another in join order.
      // (every slot is loaded, the one of a dimension past nd from dimension
      // 0's table: no branch on nd between the loads either)
      Quad kd;
#pragma unroll
      for (int t = 0; t < kMaxDims; t++) kd.v[t] = stage_to_dim(w, key, t);
      uint64_t slot[kMaxDims];
      int64_t sls[kMaxDims];
      bool hit = alive;
#pragma unroll
      for (int t = 0; t < kMaxDims; t++) {
        const bool on = t < nd;
        const ProbeDim& D = a.dim[on ? t : 0];
        const int64_t sl = ht_first_slot(D.ht_kind, D.ht_len, D.keys_min, kd.v[t]);
        hit = hit && (!on || sl >= 0);
        sls[t] = sl;
        slot[t] = ht_load(D.ht_kind, D.ht, alive && on && sl >= 0 ? sl : 0);
      }
      m = gl(a.measure)[alive ? (uint32_t)row : 0u];
      if (a.measure_sub) m = (int32_t)((uint32_t)m - (uint32_t)gl(a.measure_sub)[alive ? (uint32_t)row : 0u]);
#pragma unroll
      for (int t = 0; t < kMaxDims; t++) {
        if (t >= nd) continue;
        const ProbeDim& D = a.dim[t];
        int32_t p = 0;
        if (D.ht_kind == HT_TEXTBOOK) {
          if (hit) hit = ht_resolve(D.ht_kind, D.ht, D.ht_len, kd.v[t], sls[t], slot[t], p);
        } else {
          hit = hit && slot[t] != 0;
          p = (int32_t)(uint32_t)(slot[t] >> 32);
        }
        pd.v[t] = p;
      }
      alive = hit;
    }
  }
  st.meas_rows += __popcll(__ballot(alive));
  st.meas_lines += batch_lines(alive, row, w.last_meas);
  if (!LIP && alive) {
    m = nd == 1 ? w.p_m : a.measure[(uint32_t)row];
    if (nd > 1 && a.measure_sub) m = (int32_t)((uint32_t)m - (uint32_t)a.measure_sub[(uint32_t)row]);
  }
  st.v[ST_OOB] += group_add(a, gt, alive, pd, m);
  w.p_cnt = 0;
}

// Finish the batch in flight and take the first kQueue queued rows as the
// next one; the rest (fewer than kQueue) move to the front of the queue.
template <int K0, int SPEC>
__device__ __forceinline__ void rotate_batch(const ProbeArgs& a, const uint32_t* lds, GroupLds& gt, WaveCtx& w,
                                             WaveStats& st) {
  complete_batch<K0, SPEC>(a, lds, gt, w, st);
  const int take = w.qcount < kQueue ? w.qcount : kQueue;
  issue_batch<K0, SPEC>(a, w, take);
  const int rest = w.qcount - take;
  if (rest > 0) {
    // reads from [kQueue, kQueue + rest), writes to [0, rest): no overlap, and
    // a wave's LDS operations complete in order (the batch read above first)
    const int i = (int)lane_id();
    if (i < rest) {
      const int32_t r = w.q_row[kQueue + i], x0 = w.q_key[kQueue + i], x1 = w.q_pay[kQueue + i];
      w.q_row[i] = r;
      w.q_key[i] = x0;
      w.q_pay[i] = x1;
    }
  }
  w.qcount = rest;
}

// Stage 1 of a dense segment, in place (lane: rows row0..row0+3; two or
// more dimensions).  key0 / pay0: the stage-0 keys and payloads; mask: the
// stage-0 survivors.  If use_eager, `eager` holds the lane's 4 keys of stage
// 1's FK column already (streamed with stage 0's).  Later stages are left to
// the wave queue (batches of 64 rows, so the later gathers of one segment do
// not each wait a round trip).  Returns the stage-1 survivors, with their
// stage-0 / stage-1 keys (LIP) or payloads (naive) in qx0 / qx1: the two
// columns' registers themselves, so no per-dimension register set is kept.
template <int K0, int SPEC>
__device__ __forceinline__ uint32_t dense_stage1(const ProbeArgs& a, const uint32_t* lds, const WaveCtx& w,
                                                 uint32_t row0, const v4i32& key0, const int32_t (&pay0)[4],
                                                 uint32_t mask, const v4i32& eager, bool use_eager, WaveStats& st,
                                                 int32_t (&qx0)[4], int32_t (&qx1)[4]) {
  constexpr bool LIP = K0 != S0_PHT;
  const ProbeDim& D = a.dim[w.ld1];
  st.v[ST_LINES_IN + 1] += lines_of_ballot(__ballot(mask != 0));
  st.v[ST_ROWS_IN + 1] += popc_ballot4(mask);
  const v4i32 kv = use_eager ? eager : WavePredLoadVec4(gl(D.fk), row0, a.n, mask);
  const int32_t kmin = D.keys_min;
  uint32_t bad = 0;
#pragma unroll
  for (int j = 0; j < 4; j++) bad |= (uint32_t)(vget(kv, j) < kmin) << j;
  st.v[ST_OOB] += popc_ballot4(bad & mask);
  uint32_t nm = 0;
  if (LIP) {
    nm = probe_filter4<K0, SPEC>(a, D, lds, kv, mask);
#pragma unroll
    for (int j = 0; j < 4; j++) {
      qx0[j] = vget(key0, j);
      qx1[j] = vget(kv, j);
    }
  } else {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      int32_t p = 0;
      if ((mask >> j) & 1u) nm |= (uint32_t)pht_probe(D, vget(kv, j), p) << j;
      qx0[j] = pay0[j];
      qx1[j] = p;
    }
  }
  return nm;
}

// Append the alive rows of one segment to the wave queue (row order kept).
// done1: the rows passed stage 1 too (x1 holds its key / payload).
// The four ballots of a lane mask of 4 rows (bit j: row j), computed once and
// shared by the row count and the queue positions.
struct Ballot4 {
  unsigned long long b[4];
  __device__ __forceinline__ explicit Ballot4(uint32_t m) {
#pragma unroll
    for (int j = 0; j < 4; j++) b[j] = __ballot((m >> j) & 1u);
  }
  __device__ __forceinline__ int count() const {
    return __popcll(b[0]) + __popcll(b[1]) + __popcll(b[2]) + __popcll(b[3]);
  }
  __device__ __forceinline__ uint32_t below() const { return mbcnt(b[0]) + mbcnt(b[1]) + mbcnt(b[2]) + mbcnt(b[3]); }
};

__device__ __forceinline__ void enqueue(WaveCtx& w, uint32_t row0, const int32_t (&x0)[4], const int32_t (&x1)[4],
                                        uint32_t mask, bool done1, int cnt_total, const Ballot4& bal) {
  uint32_t pos = (uint32_t)w.qcount + bal.below();
#pragma unroll
  for (int j = 0; j < 4; j++) {
    if ((mask >> j) & 1u) {
      w.q_row[pos] = (int32_t)((row0 + j) | (done1 ? 0x80000000u : 0u));
      w.q_key[pos] = x0[j];
      w.q_pay[pos] = x1[j];
      pos++;
    }
  }
  w.qcount += cnt_total;
}

__device__ __forceinline__ void enqueue(WaveCtx& w, uint32_t row0, const int32_t (&x0)[4], const int32_t (&x1)[4],
                                        uint32_t mask, bool done1, int cnt_total) {
  enqueue(w, row0, x0, x1, mask, done1, cnt_total, Ballot4(mask));
}

// One tile of a column into registers: unchecked 16-byte non-temporal loads
// when the whole tile lies below n (the uniform common case), checked loads
// for the last, partial tile.
template <int VEC>
__device__ __forceinline__ void load_tile(const int32_t* col, uint32_t base, uint32_t lane4, uint32_t n,
                                          v4i32 (&out)[VEC]) {
  if (base + (uint32_t)VEC * kSegRows <= n) {
#pragma unroll
    for (int v = 0; v < VEC; v++)
      out[v] = __builtin_nontemporal_load(reinterpret_cast<const v4i32*>(col + base + v * kSegRows + lane4));
  } else {
#pragma unroll
    for (int v = 0; v < VEC; v++) out[v] = WaveLoadVec4<true>(col, (int64_t)(base + v * kSegRows + lane4), n);
  }
}

// Result collection by one workgroup (the last to finish the probe): copy the
// touched groups out in touch order, reset their dense accumulators to zero
// for the next query, reduce the per-workgroup statistics.  `part` is LDS
// scratch of WG long longs.  With host_hdr set, the header and the first
// host_groups groups also go straight to pinned host memory.
template <int WG>
__device__ __forceinline__ void finalize_block(const ProbeArgs& a, long long* part, uint32_t nblk) {
  constexpr int ROWS = WG / kStats;
  // Every load that does not depend on another goes out first (one memory
  // round trip): the groups-hit count, this thread's entry of the touched
  // list (read whether or not it is in use), the statistics of all
  // workgroups and the build counters.  The group values, which need the
  // touched entries, are the second round trip.  No barrier follows a
  // global or host store, so nothing waits for a store to complete before
  // the kernel ends.
  const uint32_t n = __hip_atomic_load(a.n_touched, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const uint32_t g_first = threadIdx.x < a.touched_cap ? a.touched[threadIdx.x] : 0u;
  // the statistics rows: the first kFinUnroll rows of each column are read
  // together (one round trip for a grid of up to kFinUnroll * ROWS
  // workgroups), issued before the counters' loads; the rest after
  const int k = threadIdx.x % kStats, r = threadIdx.x / kStats;
  long long sv[kFinUnroll];
#pragma unroll
  for (int j = 0; j < kFinUnroll; j++) {
    const int b = r + j * ROWS;
    sv[j] = k != ST_ORDER && b < (int)nblk ? a.stats[(int64_t)b * kStats + k] : 0;
  }
  long long dimc = 0;
  if (threadIdx.x < 2 * kMaxDims) {
    // the counter block's address: a select over the four (uniform) fields,
    // not a per-lane load of the argument array
    const int d = threadIdx.x / 2;
    static_assert(kMaxDims == 4, "dimension select");
    auto sp = [](const unsigned long long* q) {  // in scalar registers
      const uint64_t u = (uint64_t)(uintptr_t)q;
      const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u);
      const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
      return (const unsigned long long*)(uintptr_t)(((uint64_t)hi << 32) | lo);
    };
    const unsigned long long *c0 = sp(a.dim_counters[0]), *c1 = sp(a.dim_counters[1]),
                             *c2 = sp(a.dim_counters[2]), *c3 = sp(a.dim_counters[3]);
    const unsigned long long* dc = d == 0 ? c0 : d == 1 ? c1 : d == 2 ? c2 : c3;
    if (dc)
      for (int l = 0; l < kCounterLanes; l++) dimc += (long long)gl(dc)[2 * l + threadIdx.x % 2];
  }
  {
    long long acc = 0;
    if (k != ST_ORDER) {
#pragma unroll
      for (int j = 0; j < kFinUnroll; j++) acc += sv[j];
      for (int b = r + kFinUnroll * ROWS; b < (int)nblk; b += ROWS) acc += a.stats[(int64_t)b * kStats + k];
    } else if (r == 0) {
      acc = a.stats[ST_ORDER];  // the order used, from workgroup 0
    }
    part[r * kStats + k] = acc;
  }
  __syncthreads();  // `part` complete; every thread has read n_touched
  if (threadIdx.x == 0) *a.n_touched = 0;
  const uint32_t listed = n < a.touched_cap ? n : a.touched_cap;
  for (uint32_t i = threadIdx.x; i < listed; i += WG) {
    const uint32_t g = i == threadIdx.x ? g_first : a.touched[i];
    if ((int64_t)i < a.out_cap) {
      GroupOut o;
      o.g = g;
      unsigned long long sum = a.gsum[g], cnt = a.gcount[g], prod = a.gprod[g];
      // every copy's loads go out together (one round trip)
#pragma unroll
      for (int r = 1; r < kGroupReps; r++) {
        if (r >= a.g_reps) continue;
        const int64_t gi = (int64_t)g + r * a.g_stride;
        sum += a.gsum[gi];
        cnt += a.gcount[gi];
        prod += a.gprod[gi];
      }
      o.sum = (long long)sum;
      o.count = cnt;
      o.prod = prod;
      for (int d = 0; d < kMaxDims; d++) o.pay[d] = a.gpay[(int64_t)g * kMaxDims + d];
      a.out[i] = o;
      if (a.host_hdr && (int64_t)i < a.host_groups) a.host_out[i] = o;
    }
    for (int r = 0; r < (a.g_reps > 1 ? a.g_reps : 1); r++) {
      const int64_t gi = (int64_t)g + r * a.g_stride;
      a.gsum[gi] = 0;
      a.gcount[gi] = 0;
      a.gprod[gi] = 0;
    }
    if (a.g_reps > 1) a.gseen[g] = 0;
  }
  for (int k = threadIdx.x; k < kStats; k += WG) {
    long long acc = 0;
    for (int r = 0; r < ROWS; r++) acc += part[r * kStats + k];
    a.out_hdr[1 + k] = acc;
    if (a.host_hdr) a.host_hdr[1 + k] = acc;
  }
  if (threadIdx.x < 2 * kMaxDims) {
    a.out_hdr[1 + kStats + threadIdx.x] = dimc;
    if (a.host_hdr) a.host_hdr[1 + kStats + threadIdx.x] = dimc;
  }
  if (threadIdx.x == 0) {
    a.out_hdr[0] = n;
    if (a.host_hdr) a.host_hdr[0] = n;
  }
}

// One dimension with an exact filter (w.exact), its measure streamed with the
// keys: a segment whose stage-0 survivors all have keys the filter decides is
// finished in place -- every survivor is a hit; the lane adds SUM(measure)
// and the row count to its running totals (one group), flushed once at the
// end.  Returns false (nothing done) for a segment with a key the filter
// cannot decide (key 0, or outside the table's domain): its survivors take
// the queue, which reads the table for those keys.
__device__ __forceinline__ bool exact_in_place(const ProbeArgs& a, WaveCtx& w, WaveStats& st, const v4i32& kv,
                                               const v4i32& mv, uint32_t m) {
  const ProbeDim& D = a.dim[0];
  bool undecided = false;
#pragma unroll
  for (int j = 0; j < 4; j++) undecided = undecided || (((m >> j) & 1u) && !exact_key(D, vget(kv, j)));
  if (__ballot(undecided)) return false;
  st.meas_rows += popc_ballot4(m);
  st.meas_lines += lines_of_ballot(__ballot(m != 0));
  long long s = 0;
#pragma unroll
  for (int j = 0; j < 4; j++) s += ((m >> j) & 1u) ? (long long)vget(mv, j) : 0ll;
  w.acc_s += s;
  w.acc_c += (uint32_t)__popc(m);
  return true;
}

// One dimension, one group, a perfect-hash table (direct), the measure
// streamed with the keys: a segment's survivors of the filter probe the table
// here -- the slots of a lane's 4 rows gathered together (a dead row reads
// slot 0 and drops it, so no branch separates the loads) -- and their
// measures and products go to the lane's sums.  No queue: the same hits
// (a nonzero slot), sums and statistics as the queued rows' complete_batch.
// With the table's occupancy bitmap in LDS (w.occ), the bit decides the
// hit and only the hits' slots are gathered, for their payloads (products
// over a PHT_2 table); the others read slot 0, one shared line.  Then the
// measure need not stream: without it (have_meas false) the hits' measures
// are gathered the same way (row 0 for the others).  Returns the segment's
// hits (wave total).
__device__ __forceinline__ int direct_in_place(const ProbeArgs& a, const uint32_t* lds, WaveCtx& w, WaveStats& st,
                                               const v4i32& kv, const v4i32& mv, uint32_t m, uint32_t row0,
                                               bool have_meas) {
  const ProbeDim& D = a.dim[0];
  int64_t sl[4];
  uint64_t raw[4];
  uint32_t hit = 0;
  if (w.occ) {
    const uint32_t* occ = lds + a.occ_lds;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      sl[j] = ht_first_slot(D.ht_kind, D.ht_len, D.keys_min, vget(kv, j));
      const bool on = ((m >> j) & 1u) && sl[j] >= 0;
      const uint32_t x = on ? (uint32_t)sl[j] : 0u;
      hit |= (uint32_t)(on && ((occ[x >> 5] >> (x & 31)) & 1u)) << j;
      raw[j] = 0;
    }
    if (a.want_product && D.ht_kind == HT_PHT2 && __ballot(hit != 0)) {
#pragma unroll
      for (int j = 0; j < 4; j++) raw[j] = ht_load(D.ht_kind, D.ht, ((hit >> j) & 1u) ? sl[j] : 0);
    }
  } else {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      sl[j] = ht_first_slot(D.ht_kind, D.ht_len, D.keys_min, vget(kv, j));
      raw[j] = ht_load(D.ht_kind, D.ht, ((m >> j) & 1u) && sl[j] >= 0 ? sl[j] : 0);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) hit |= (uint32_t)(((m >> j) & 1u) && sl[j] >= 0 && raw[j] != 0) << j;
  }
  int32_t mg[4] = {0, 0, 0, 0};
  if (!have_meas && __ballot(hit != 0)) {
#pragma unroll
    for (int j = 0; j < 4; j++) mg[j] = gl(a.measure)[((hit >> j) & 1u) ? row0 + j : 0u];
  }
  long long s = 0, p = 0;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const bool h = (hit >> j) & 1u;
    const int32_t mj = have_meas ? vget(mv, j) : mg[j];
    s += h ? (long long)mj : 0ll;
    p += h ? (long long)(int32_t)((uint32_t)mj * (uint32_t)(raw[j] >> 32)) : 0ll;
  }
  st.meas_rows += popc_ballot4(hit);
  st.meas_lines += lines_of_ballot(__ballot(hit != 0));
  w.acc_s += s;
  w.acc_c += (uint32_t)__popc(hit);
  if (a.want_product) w.acc_p += p;
  return popc_ballot4(hit);
}

// The occupancy bitmap of a table of ns slots (T: the slot type): a wave
// takes 64 slots at a time, lane i slot 64c + i, and the ballot of the
// nonzero ones is words 2c and 2c+1.  kOccUnroll chunks' loads go out
// before their ballots.  Slot indexes past the end are clamped (the loads
// stay unconditional) and their bits cleared.
template <typename T, int NW>
__device__ __forceinline__ void occupancy_build(const T* ht, uint32_t ns, uint32_t* occ, int wave) {
  const uint32_t lane = lane_id();
  const uint32_t nch = (ns + 63) / 64;
  for (uint32_t c0 = (uint32_t)wave; c0 < nch; c0 += NW * kOccUnroll) {
    T v[kOccUnroll];
#pragma unroll
    for (int u = 0; u < kOccUnroll; u++) {
      const uint32_t s = (c0 + u * NW) * 64 + lane;
      v[u] = gl(ht)[s < ns ? s : ns - 1];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < kOccUnroll; u++) {
      const uint32_t c = c0 + u * NW;
      const unsigned long long b = __ballot(c * 64 + lane < ns && v[u] != 0);
      if (c < nch && lane < 2) occ[2 * c + lane] = lane ? (uint32_t)(b >> 32) : (uint32_t)b;
    }
  }
}

struct TileCtx {
  uint32_t n;        // fact rows
  uint32_t start;    // first row of the tile space (after the adaptive sample)
  uint32_t lane4;    // 4 * lane
  const int32_t* eager_col;
};

// One tile: stage 0 on the registers `k` (VEC segments of 256 rows), then its
// survivors.  CHECKED masks rows at or beyond n.  `ke` holds the next column
// of the same rows when have_eager.  Returns the stage-0 survivor count (the
// table hits, for segments probed in place through the occupancy bitmap).
template <int VEC, int K0, bool CHECKED, int SPEC>
__device__ __forceinline__ int tile_step(const ProbeArgs& a, const S0& s0, const uint32_t* lds, GroupLds& gt,
                                         WaveCtx& w, WaveStats& st, const TileCtx& tc, uint32_t base,
                                         const v4i32 (&k)[VEC], const v4i32 (&ke)[VEC], bool have_eager, int& oob) {
  constexpr uint32_t TILE = (uint32_t)VEC * kSegRows;
  constexpr bool LIP = K0 != S0_PHT;  // a hash-table first stage is the naive plan
  uint32_t mask[VEC];
  int32_t pay[VEC][4];
#pragma unroll
  for (int v = 0; v < VEC; v++) {
    const uint32_t row0 = base + v * kSegRows + tc.lane4;
    uint32_t valid4 = 0xFu;
    if (CHECKED) {
      valid4 = 0;
#pragma unroll
      for (int j = 0; j < 4; j++) valid4 |= (uint32_t)(row0 + j < tc.n) << j;
    }
    // keys below keys_min: counted here (wave-wide, on the scalar unit); the
    // rows are counted one by one only when some lane's smallest key is below
    const int32_t km = s0.kmin;
    int32_t kmin4 = min(min(k[v].x, k[v].y), min(k[v].z, k[v].w));
    if (CHECKED) {
      const int32_t big = INT32_MAX;
      kmin4 = min(min((valid4 & 1u) ? k[v].x : big, (valid4 & 2u) ? k[v].y : big),
                  min((valid4 & 4u) ? k[v].z : big, (valid4 & 8u) ? k[v].w : big));
    }
    if (__ballot(kmin4 < km))
      st.v[ST_OOB] += __popcll(__ballot((valid4 & 1u) && k[v].x < km)) + __popcll(__ballot((valid4 & 2u) && k[v].y < km)) +
                      __popcll(__ballot((valid4 & 4u) && k[v].z < km)) + __popcll(__ballot((valid4 & 8u) && k[v].w < km));
    mask[v] = s0_probe4<K0, SPEC != 1>(s0, k[v], valid4, pay[v]);
  }
  const uint32_t nvalid = CHECKED ? (tc.n - base < TILE ? tc.n - base : TILE) : TILE;
  st.v[ST_ROWS_IN + 0] += nvalid;
  st.v[ST_LINES_IN + 0] += (nvalid + 31) / 32;
  int total = 0;
  // one (not unrolled) pass per segment; the segment's registers are picked
  // by uniform selects so that no register array is indexed dynamically
  for (int v = 0; v < VEC; v++) {
    uint32_t m = mask[0];
    v4i32 kv = k[0], ev = ke[0];
    int32_t pv[4] = {pay[0][0], pay[0][1], pay[0][2], pay[0][3]};
#pragma unroll
    for (int u = 1; u < VEC; u++) {
      if (u == v) {
        m = mask[u];
        kv = k[u];
        ev = ke[u];
#pragma unroll
        for (int j = 0; j < 4; j++) pv[j] = pay[u][j];
      }
    }
    const Ballot4 bal(m);
    const int cnt = bal.count();
    total += cnt;
    if (cnt == 0) continue;
    // callers keep qcount < kQueue between tiles; later segments of a tile re-check
    if (v > 0 && w.qcount >= kQueue) rotate_batch<K0, SPEC>(a, lds, gt, w, st);
    const uint32_t row0 = base + v * kSegRows + tc.lane4;
    int32_t qx0[4], qx1[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      qx0[j] = LIP ? vget(kv, j) : pv[j];
      qx1[j] = 0;
    }
    if (SPEC == 3 && have_eager && w.exact && exact_in_place(a, w, st, kv, ev, m)) {
      st.v[ST_BLOOM_PASS] += cnt;  // every survivor passed the only filter
      continue;
    }
    if (SPEC == 3 && w.direct && (have_eager || w.occ)) {
      const int hits = direct_in_place(a, lds, w, st, kv, ev, m, row0, have_eager);
      st.v[ST_BLOOM_PASS] += cnt;
      if (w.occ) total += hits - cnt;  // with the bitmap, the measure streams by hits
      continue;
    }
    if (cnt > kDenseMin || have_eager) {
      // with two or more dimensions stage 1 runs here on the streamed column;
      // with one, every survivor goes to the queue (its hash-slot gather is
      // issued a batch ahead) carrying its measure when the column streams
      uint32_t qm = m;
      const bool d1 = w.nd >= 2 || have_eager;
      if (w.nd == 1 && have_eager) {
#pragma unroll
        for (int j = 0; j < 4; j++) qx1[j] = vget(ev, j);
      }
      if (w.nd >= 2) qm = dense_stage1<K0, SPEC>(a, lds, w, row0, kv, pv, m, ev, have_eager, st, qx0, qx1);
      const int cq = popc_ballot4(qm);
      if (cq == 0) continue;
      // the stage-1 survivors (up to 4 per lane) go in up to four parts of
      // 16 lanes each (at most 64 rows), in row order
      const int parts = cq > kDenseMin ? 4 : 1;
#pragma unroll 1
      for (int part = 0; part < parts; part++) {
        const uint32_t pm = parts == 1 ? qm : (((int)tc.lane4 >> 6) == part ? qm : 0u);
        const int c = parts == 1 ? cq : popc_ballot4(pm);
        if (c == 0) continue;
        if (w.qcount + c > kQueueCap) rotate_batch<K0, SPEC>(a, lds, gt, w, st);
        enqueue(w, row0, qx0, qx1, pm, d1, c);
      }
      continue;
    }
    // a sparse segment adds at most kDenseMin rows: it fits (qcount < kQueue here)
    enqueue(w, row0, qx0, qx1, m, false, cnt, bal);
  }
  return total;
}

// Adaptive LIP: a wave's first tile is its sample.  Every FK column of the
// tile is loaded and every Bloom filter applied to every 4th row; the per-filter
// pass counts go into the workgroup's LDS totals.  The wave then orders its
// filters by the totals seen so far (its own sample at least; no barrier),
// most selective first, and streams the next column with the first one when
// the chosen first filter passes more than a quarter of the sampled rows.  The
// tile itself is then processed like any other (its columns are in L2).
template <int VEC, int K0, int SPEC>
__device__ __forceinline__ void sample_tile(const ProbeArgs& a, const uint32_t* lds_filter, WaveCtx& w,
                                            WaveStats& st, unsigned long long* lds_samp, uint32_t base,
                                            uint32_t lane4, int64_t n64, const v4i32 (&pre)[kMaxDims], bool& eager) {
  uint32_t cnt[kMaxDims] = {0, 0, 0, 0};
  uint32_t n_valid = 0;
#pragma unroll
  for (int v = 0; v < VEC; v++) {
    const uint32_t row0 = base + v * kSegRows + lane4;
    uint32_t valid = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) valid |= (uint32_t)((int64_t)row0 + j < n64) << j;
    // segment 0's columns were loaded while the filters staged (pre)
#pragma unroll
    for (int d = 0; d < kMaxDims; d++) {
      if (d >= w.nd) continue;
      const v4i32 kk = v == 0 ? pre[d] : WaveLoadVec4<false>(gl(a.dim[d].fk), row0, n64);
      // the pass counts come from every 4th row (a lane's first row): the
      // ranking needs rates, and a quarter of the probes keeps the sample short
      const bool pass = (valid & 1u) && probe_filter<K0, SPEC>(a, a.dim[d], lds_filter, kk.x);
      const uint32_t c = (uint32_t)__popcll(__ballot(pass));
      cnt[0] += d == 0 ? c : 0u;
      cnt[1] += d == 1 ? c : 0u;
      cnt[2] += d == 2 ? c : 0u;
      cnt[3] += d == 3 ? c : 0u;
    }
    const int nv = popc_ballot4(valid);
    n_valid += (uint32_t)__popcll(__ballot(valid & 1u));  // rows probed
    st.v[ST_SAMPLE_ROWS] += nv;
    st.v[ST_SAMPLE_LINES] += (long long)w.nd * lines_of_ballot(__ballot(valid != 0));
  }
  // workgroup totals so far, including this wave's sample: the pass counts
  // of the 4 filters as 16-bit fields of one 64-bit LDS word (a workgroup
  // samples fewer than 2^16 rows) and the rows sampled, two atomics a wave
  uint32_t tot[kMaxDims + 1] = {0, 0, 0, 0, 0};
  {
    unsigned long long packed = 0, old = 0;
    uint32_t rows_old = 0;
#pragma unroll
    for (int d = 0; d < kMaxDims; d++) packed |= (unsigned long long)(d < w.nd ? cnt[d] : 0u) << (16 * d);
    if (lane_id() == 0) {
      old = atomicAdd(&lds_samp[0], packed);
      rows_old = atomicAdd(reinterpret_cast<uint32_t*>(&lds_samp[1]), n_valid);
    }
    const unsigned long long now = old + packed;
#pragma unroll
    for (int d = 0; d < kMaxDims; d++)
      tot[d] = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((now >> (16 * d)) & 0xFFFFu));
    tot[kMaxDims] = (uint32_t)__builtin_amdgcn_readfirstlane((int)(rows_old + n_valid));
  }
  if (tot[kMaxDims] == 0) return;
  // rank of each dimension by pass count (ties: lower index first); with a
  // stage-0 mask, the dimensions outside it rank after every one inside it
  uint32_t cntd[kMaxDims];
#pragma unroll
  for (int d = 0; d < kMaxDims; d++) {
    const uint32_t late = (a.s0_mask != 0u && !((a.s0_mask >> d) & 1u)) ? 0x80000000u : 0u;
    cntd[d] = d < w.nd ? (tot[d] | late) : 0xFFFFFFFFu;
  }
  int rank[kMaxDims];
#pragma unroll
  for (int d = 0; d < kMaxDims; d++) {
    rank[d] = 0;
#pragma unroll
    for (int e = 0; e < kMaxDims; e++) rank[d] += (cntd[e] < cntd[d]) || (cntd[e] == cntd[d] && e < d);
  }
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
  const uint32_t c0 = (o[0] == 0 ? cntd[0] : o[0] == 1 ? cntd[1] : o[0] == 2 ? cntd[2] : cntd[3]) & 0x7FFFFFFFu;
  eager = (uint64_t)c0 * 4 > tot[kMaxDims];
}

// Tiles handed out inside a workgroup.  Workgroup b owns the full tiles
// r * stride + b * NW + i (rounds r, i < NW), the same set a static
// wave-per-tile assignment would give it; its waves take them in order from
// an LDS counter, so a wave that runs ahead (SIMD arbitration favours some
// waves) takes more tiles and the workgroup's waves finish together.  Slot j
// of the workgroup maps to tile (j / NW) * stride + b * NW + j % NW; slots
// 0..NW-1 are the waves' first tiles (taken without the counter).
template <int NW>
struct TileSched {
  uint32_t* next;    // LDS: slots taken after the first round
  uint32_t stride;   // waves in the grid
  uint32_t wg_base;  // b * NW
  // Issue the take; the value is read (readfirstlane) by tile() later, so the
  // LDS atomic's latency hides behind the tile in flight.  ds_append adds the
  // number of active lanes to the counter and returns its value before to
  // every lane: called with the whole wave active (uniform code only), one
  // take advances it by 64, with no exec masking or lane count around it.
  __device__ __forceinline__ uint32_t issue() const {
    return (uint32_t)__builtin_amdgcn_ds_append((__attribute__((address_space(3))) int*)next);
  }
  __device__ __forceinline__ uint32_t tile(uint32_t raw) const {
    const uint32_t j = ((uint32_t)__builtin_amdgcn_readfirstlane((int)raw) >> 6) + NW;
    return (j / NW) * stride + wg_base + (j % NW);
  }
};

// Full tiles from the workgroup's pool until it is empty, with kPrefetch
// tiles in flight ahead of the one being processed.  Every iteration issues
// the same loads (a tile past the end is clamped to tile 0), so the wait for
// the current tile is a fixed vmcnt and the prefetches stay in flight while
// it is processed.  EAGER also streams the next column.  Tiles come in
// increasing order per wave (the counter only grows).
template <int VEC, bool EAGER>
__device__ __forceinline__ void load_full_tile(const S0& s0, const TileCtx& tc, uint32_t t, v4i32 (&k)[VEC],
                                               v4i32 (&e)[VEC]) {
  constexpr uint32_t TILE = (uint32_t)VEC * kSegRows;
#pragma unroll
  for (int v = 0; v < VEC; v++) {
    const uint32_t r = tc.start + t * TILE + v * kSegRows + tc.lane4;
    if (EAGER) e[v] = __builtin_nontemporal_load(gl(reinterpret_cast<const v4i32*>(tc.eager_col + r)));
    else e[v] = v4i32{0, 0, 0, 0};
    k[v] = __builtin_nontemporal_load(gl(reinterpret_cast<const v4i32*>(s0.fk + r)));
  }
}

template <int VEC, int K0, bool EAGER, int NW, int SPEC>
__device__ __forceinline__ void stream_loop(const ProbeArgs& a, const S0& s0, const uint32_t* lds, GroupLds& gt,
                                            WaveCtx& w, WaveStats& st, const TileCtx& tc, const TileSched<NW>& sch,
                                            uint32_t n_full, int& oob) {
  constexpr uint32_t TILE = (uint32_t)VEC * kSegRows;
  uint32_t t = sch.tile(sch.issue());
  if (t >= n_full) return;
  v4i32 cur[VEC], cure[VEC];
  load_full_tile<VEC, EAGER>(s0, tc, t, cur, cure);
  uint32_t tn = sch.tile(sch.issue());
  if (kPrefetch == 1) {
    for (;;) {
      v4i32 nxt[VEC], nxte[VEC];
      load_full_tile<VEC, EAGER>(s0, tc, tn < n_full ? tn : 0u, nxt, nxte);
      const uint32_t raw = sch.issue();
      // keep the prefetch ahead of the first use of `cur` (the wait below is
      // then vmcnt(#prefetch loads) instead of vmcnt(0))
      __builtin_amdgcn_sched_barrier(0);
      tile_step<VEC, K0, false, SPEC>(a, s0, lds, gt, w, st, tc, tc.start + t * TILE, cur, cure, EAGER, oob);
      if (w.qcount >= kQueue) rotate_batch<K0, SPEC>(a, lds, gt, w, st);
      if (tn >= n_full) break;
      t = tn;
      tn = sch.tile(raw);
#pragma unroll
      for (int v = 0; v < VEC; v++) {
        cur[v] = nxt[v];
        cure[v] = nxte[v];
      }
    }
  } else {
    v4i32 nx1[VEC], nx1e[VEC];
    load_full_tile<VEC, EAGER>(s0, tc, tn < n_full ? tn : 0u, nx1, nx1e);
    uint32_t raw2 = sch.issue();
    for (;;) {
      const uint32_t tnn = sch.tile(raw2);
      v4i32 nx2[VEC], nx2e[VEC];
      load_full_tile<VEC, EAGER>(s0, tc, tnn < n_full ? tnn : 0u, nx2, nx2e);
      raw2 = sch.issue();
      __builtin_amdgcn_sched_barrier(0);
      tile_step<VEC, K0, false, SPEC>(a, s0, lds, gt, w, st, tc, tc.start + t * TILE, cur, cure, EAGER, oob);
      if (w.qcount >= kQueue) rotate_batch<K0, SPEC>(a, lds, gt, w, st);
      if (tn >= n_full) break;
      t = tn;
      tn = tnn;
#pragma unroll
      for (int v = 0; v < VEC; v++) {
        cur[v] = nx1[v];
        cure[v] = nx1e[v];
        nx1[v] = nx2[v];
        nx1e[v] = nx2e[v];
      }
    }
  }
}

// The probe of one query by workgroup `bid` of the `nblk` workgroups that
// serve it (the whole grid for a single query; a slice of it in a batch).
template <int WG, int VEC, int K0, int SPEC>
__device__ __forceinline__ void probe_body(const ProbeArgs& a, const uint32_t bid, const uint32_t nblk) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  constexpr int NW = WG / 64;
  static_assert(NW * VEC * kSegRows < (1 << 16), "adaptive sample totals are 16-bit fields");
  const int wave = uniform(threadIdx.x >> 6);  // known wave-uniform: keeps per-wave branches scalar
  const int lane = (int)lane_id();

  // LDS layout: the fixed-size parts first, at constant offsets (the wave
  // queues, statistics, sample totals, tile counter), then the group table
  // (its size is the launch's), then the staged filters (probe_lds_bytes).
  constexpr size_t kQueueBytes = (size_t)NW * 3 * kQueueCap * 4;
  int32_t* q_base = reinterpret_cast<int32_t*>(lds_raw) + wave * (3 * kQueueCap);
  long long* lds_stats = reinterpret_cast<long long*>(lds_raw + kQueueBytes);
  unsigned long long* lds_samp = reinterpret_cast<unsigned long long*>(lds_raw + kQueueBytes + kStats * 8);  // [kMaxDims + 1]
  uint32_t* lds_next = reinterpret_cast<uint32_t*>(lds_raw + kQueueBytes + kStats * 8 + (kMaxDims + 1) * 8);  // tile counter
  size_t off = kQueueBytes + kStats * 8 + (kMaxDims + 1) * 8 + 16;
  const int gslots = 1 << a.gslot_bits;
  GroupLds gt;
  gt.cnt = reinterpret_cast<unsigned long long*>(lds_raw + off);
  gt.sum = gt.cnt + gslots;
  gt.prod = gt.sum + gslots;
  off += 3 * (size_t)gslots * 8;
  gt.key = reinterpret_cast<int32_t*>(lds_raw + off);
  off += (size_t)gslots * 4;
  gt.pay = reinterpret_cast<int32_t*>(lds_raw + off);
  off += (size_t)gslots * kMaxDims * 4;
  off = (off + 15) & ~(size_t)15;
  uint32_t* lds_filter = reinterpret_cast<uint32_t*>(lds_raw + off);

  if (bid == 0 && threadIdx.x == 0) a.out_hdr[kClockSlot + 1] = (long long)wall_clock64();

  // the wave's first tile of its first column is read while the filters stage
  constexpr uint32_t TILE0 = (uint32_t)VEC * kSegRows;
  const uint32_t gw0 = bid * NW + wave;
  const int64_t n_first = a.n;
  v4i32 k_first[VEC];
#pragma unroll
  for (int v = 0; v < VEC; v++)
    k_first[v] = (int64_t)gw0 * TILE0 < n_first && !a.adaptive
                     ? WaveLoadVec4<true>(gl(a.dim[a.load_order[0]].fk), (int64_t)(gw0 * TILE0 + v * kSegRows + lane * 4), n_first)
                     : v4i32{0, 0, 0, 0};
  // adaptive: every column of the sampled segment
  v4i32 k_samp[kMaxDims];
#pragma unroll
  for (int d = 0; d < kMaxDims; d++)
    k_samp[d] = K0 != S0_PHT && a.adaptive && d < a.n_dims && (int64_t)gw0 * TILE0 < n_first
                    ? WaveLoadVec4<false>(gl(a.dim[d].fk), (int64_t)(gw0 * TILE0 + lane * 4), n_first)
                    : v4i32{0, 0, 0, 0};

  // ---- stage the LDS-resident filters, clear the group table
  for (int d = 0; d < a.n_dims; d++) {
    const ProbeDim& D = a.dim[d];
    if (D.bloom_lds >= 0) {
      const int nw = D.bloom_lds_words;  // the staged prefix
      uint32_t* dst = lds_filter + D.bloom_lds;
      if (((D.bloom_lds | nw) & 3) == 0 && (((uintptr_t)D.bloom) & 15) == 0) {
        // kStageUnroll 16-byte loads in flight per thread before their LDS
        // stores (a loop of load, wait, store pays one memory round trip per
        // 16 KB of filter)
        // (address-space-typed pointers: no flat accesses)
        const __attribute__((address_space(1))) v4i32* src4 = gl(reinterpret_cast<const v4i32*>(D.bloom));
        __attribute__((address_space(3))) v4i32* dst4 = (__attribute__((address_space(3))) v4i32*)dst;
        const int n4 = nw / 4;
        for (int i0 = 0; i0 < n4; i0 += WG * kStageUnroll) {
          // Indexes past the end are clamped to the last word, which such
          // lanes copy again (the same value): loads and stores stay
          // unconditional, so the loads are not sunk into a branch each
          // beside its store, and the barrier keeps them all ahead.
          v4i32 v[kStageUnroll];
#pragma unroll
          for (int u = 0; u < kStageUnroll; u++) {
            const int i = i0 + u * WG + (int)threadIdx.x;
            v[u] = src4[i < n4 ? i : n4 - 1];
          }
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int u = 0; u < kStageUnroll; u++) {
            const int i = i0 + u * WG + (int)threadIdx.x;
            dst4[i < n4 ? i : n4 - 1] = v[u];
          }
        }
      } else {
        for (int i = threadIdx.x; i < nw; i += WG) dst[i] = D.bloom[i];
      }
    }
  }
  // the bitmap repays its build when enough rows pass the filter: for a
  // dimension built with the query, judged from the keys its build selected
  bool use_occ = SPEC == 3 && a.occ_words > 0;
  if (use_occ && a.occ_min_sel > 0 && a.dim_counters[0]) {
    unsigned long long sel = 0;
    for (int l = 0; l < kCounterLanes; l++) sel += a.dim_counters[0][2 * l];
    use_occ = sel >= (unsigned long long)a.occ_min_sel;
  }
  if (use_occ) {
    const ProbeDim& D = a.dim[0];
    if (D.ht_kind == HT_PHT1)
      occupancy_build<uint32_t, NW>(reinterpret_cast<const uint32_t*>(D.ht), D.ht_len.d, lds_filter + a.occ_lds, wave);
    else
      occupancy_build<uint64_t, NW>(reinterpret_cast<const uint64_t*>(D.ht), D.ht_len.d, lds_filter + a.occ_lds, wave);
  }
  for (int i = threadIdx.x; i < gslots; i += WG) {
    gt.key[i] = kEmptyGroup;
    gt.cnt[i] = 0;
    gt.sum[i] = 0;
    gt.prod[i] = 0;
  }
  for (int i = threadIdx.x; i < kStats; i += WG) lds_stats[i] = 0;
  for (int i = threadIdx.x; i <= kMaxDims; i += WG) lds_samp[i] = 0;
  if (threadIdx.x == 0) *lds_next = 0;
  __syncthreads();

  WaveStats st;
#pragma unroll
  for (int i = 0; i < kStats; i++) st.v[i] = 0;
  st.meas_rows = st.meas_lines = 0;
  WaveCtx w;
  w.nd = a.n_dims;
  w.q_row = q_base;
  w.q_key = q_base + kQueueCap;
  w.q_pay = q_base + 2 * kQueueCap;
  w.qcount = 0;
  w.last1 = w.last2 = w.last3 = w.last_meas = -1;
  w.p_cnt = 0;
  w.p_row = w.p_x0 = w.p_x1 = w.p_k = w.p_m = 0;
  w.p_ht = 0;
  w.ld0 = a.load_order[0];
  w.ld1 = a.load_order[1];
  w.ld2 = a.load_order[2];
  w.ld3 = a.load_order[3];
  w.acc_s = 0;
  w.acc_c = 0;
  w.acc_p = 0;
  w.exact = SPEC == 3 && a.n_dims == 1 && a.dim[0].exact != 0;
  if (w.exact && a.dim_counters[0]) {
    // a dimension built with the query: a selected key outside the table's
    // domain (below keys_min, or aliased from keys_min + ht_len on) would
    // break the bit-slot match
    unsigned long long o = a.dim_counters[0][kCounterOutside];
    for (int l = 0; l < kCounterLanes; l++) o += a.dim_counters[0][2 * l + 1];
    w.exact = o == 0;
  }
  w.direct = SPEC == 3 && !w.exact && a.n_dims == 1 && a.n_groups == 1 && a.direct1 != 0;
  w.occ = w.direct && use_occ;

  const uint32_t total_waves = nblk * NW;
  const uint32_t gw = bid * NW + wave;
  const int64_t n64 = a.n;
  int oob = 0;
  bool eager = false;  // prefetch the next column with the first one

  // ---- first tile: the adaptive sample --------------------------------------------
  constexpr uint32_t TILE = (uint32_t)VEC * kSegRows;
  const uint32_t n_rows = (uint32_t)n64;
  const uint32_t n_full = n_rows / TILE;
  const uint32_t n_tiles = (n_rows + TILE - 1) / TILE;
  if constexpr (K0 != S0_PHT) {
    if (a.adaptive && gw < n_tiles)
      sample_tile<VEC, K0, SPEC>(a, lds_filter, w, st, lds_samp, gw * TILE, (uint32_t)lane * 4, n64, k_samp, eager);
  }

  // ---- streaming loop over tiles -------------------------------------------------
  const ProbeDim& D0 = a.dim[w.ld0];
  S0 s0;
  s0.fk = D0.fk;
  s0.bf = (K0 == S0_MODP2_LDS || K0 == S0_MOD_LDS || K0 == S0_MULT_LDS || (K0 == S0_GENERIC && D0.bloom_lds >= 0))
              ? lds_filter + D0.bloom_lds
              : D0.bloom;
  s0.gbf = D0.bloom;
  s0.lds_words = D0.bloom_lds >= 0 ? (uint32_t)D0.bloom_lds_words : 0u;
  s0.whole = D0.bloom_lds >= 0 && D0.bloom_lds_words == D0.bloom_words;
  s0.ht = D0.ht;
  s0.bh = D0.bh;
  s0.fm = D0.ht_len;
  s0.kmin = D0.keys_min;
  s0.mask = D0.bh.mod.pow2_mask;
  s0.ht_kind = D0.ht_kind;
  const int32_t* eager_col = w.nd > 1 ? a.dim[w.ld1].fk : a.measure;

  TileCtx tc;
  tc.n = n_rows;
  tc.start = 0;
  tc.lane4 = (uint32_t)lane * 4;
  tc.eager_col = eager_col;
  // The wave's first tile and the table's last, partial tile go through the
  // checked path; the first tile's pass rate (or the adaptive sample) decides
  // whether the full-tile loop prefetches the next column too.
  const uint32_t tail_tile = (n_full < n_tiles && n_full >= total_waves && gw == n_full % total_waves)
                                 ? n_full : 0xFFFFFFFFu;
  for (int pass = 0; pass < 2; pass++) {
    const uint32_t tt = pass == 0 ? gw : tail_tile;
    if (tt < n_tiles) {
      const uint32_t base = tc.start + tt * TILE;
      v4i32 k[VEC], ke[VEC];
#pragma unroll
      for (int v = 0; v < VEC; v++) {
        // the first tile of a static order was loaded before the staging; the
        // adaptive sample loaded segment 0 of every column (copied by value
        // from the chosen one: no indexed register array)
        if (pass == 0 && !a.adaptive) {
          k[v] = k_first[v];
        } else if (K0 != S0_PHT && pass == 0 && v == 0) {
          // a select over values (no lvalue conditional: that would index the array)
          const int l0 = w.ld0;
          const v4i32 s0v = k_samp[0], s1v = k_samp[1], s2v = k_samp[2], s3v = k_samp[3];
          k[v] = l0 == 0 ? s0v : l0 == 1 ? s1v : l0 == 2 ? s2v : s3v;
        } else {
          k[v] = WaveLoadVec4<true>(gl(s0.fk), (int64_t)(base + v * kSegRows + tc.lane4), n64);
        }
        ke[v] = k[v];
      }
      const int tot = tile_step<VEC, K0, true, SPEC>(a, s0, lds_filter, gt, w, st, tc, base, k, ke, false, oob);
      if (w.qcount >= kQueue) rotate_batch<K0, SPEC>(a, lds_filter, gt, w, st);
      const uint32_t nv = tc.n - base < TILE ? tc.n - base : TILE;
      // (the exact semi-join finishes a streamed segment in place, so it
      // streams from a lower pass rate than the queue's gathers would pay)
      if (pass == 0 && !a.adaptive)
        eager = (uint32_t)tot * (SPEC == 3 && w.exact    ? (uint32_t)a.exact_eager_div
                                 : SPEC == 3 && w.direct ? (uint32_t)(w.occ ? a.occ_eager_div
                                                                                      : a.direct_eager_div)
                                                         : 4u) > nv;
    }
    if (pass == 0) {
      TileSched<NW> sch;
      sch.next = lds_next;
      sch.stride = total_waves;
      sch.wg_base = bid * NW;
      if (eager) stream_loop<VEC, K0, true, NW, SPEC>(a, s0, lds_filter, gt, w, st, tc, sch, n_full, oob);
      else stream_loop<VEC, K0, false, NW, SPEC>(a, s0, lds_filter, gt, w, st, tc, sch, n_full, oob);
    }
  }
  // the batch in flight, then the rows still queued
#pragma unroll 1
  for (int it = 0; it < 2; it++) rotate_batch<K0, SPEC>(a, lds_filter, gt, w, st);

  // ---- flush: statistics and groups ----------------------------------------------
  if (SPEC == 3 && (w.exact || w.direct)) {
    const long long es = WaveSum(w.acc_s);
    const unsigned long long ec = WaveSum((unsigned long long)w.acc_c);
    const long long ep = a.want_product ? WaveSum(w.acc_p) : 0ll;
    if (lane == 0 && ec != 0) {
      atomicAdd(&gt.sum[0], (unsigned long long)es);
      atomicAdd(&gt.cnt[0], ec);
      if (a.want_product) atomicAdd(&gt.prod[0], (unsigned long long)ep);
      if (gt.key[0] == kEmptyGroup) {
        gt.key[0] = 0;
        for (int d = 0; d < kMaxDims; d++) gt.pay[d] = 0;
      }
    }
  }
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
      if (bid == 0 && wave == 0)
        lds_stats[ST_ORDER] = (long long)(w.ld0 & 0xff) | ((long long)(w.ld1 & 0xff) << 8) |
                              ((long long)(w.ld2 & 0xff) << 16) | ((long long)(w.ld3 & 0xff) << 24);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < gslots; i += WG) {
    const int32_t g = gt.key[i];
    if (g != kEmptyGroup && gt.cnt[i] != 0) {
      Quad pay;
      for (int d = 0; d < kMaxDims; d++) pay.v[d] = gt.pay[i * kMaxDims + d];
      global_group_add(a, g, (long long)gt.sum[i], gt.cnt[i], gt.prod[i], pay, a.n_dims);
    }
  }
  for (int i = threadIdx.x; i < kStats; i += WG) a.stats[(int64_t)bid * kStats + i] = lds_stats[i];

  // ---- the last workgroup to finish collects the result ---------------------------
  // Each workgroup publishes its writes with ONE agent-scope release (the
  // ticket); only the last one acquires.  A full __threadfence here would also
  // invalidate the XCD's L2 for every finishing workgroup and drop the filters
  // and hash tables that the still-running workgroups on that XCD are reading.
  __syncthreads();
  uint32_t* ticket = reinterpret_cast<uint32_t*>(lds_samp);
  if (threadIdx.x == 0) *ticket = __hip_atomic_fetch_add(a.done, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  if (*ticket == nblk - 1) {
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    finalize_block<WG>(a, reinterpret_cast<long long*>(lds_raw), nblk);
    if (threadIdx.x == 0) {
      *a.done = 0;
      a.out_hdr[kClockSlot + 2] = (long long)wall_clock64();
    }
  }
}

// One copy of the probe per filter-kind case, chosen once per launch: with
// every filter of stage 0's kind (all_k0) the later stages and the sample
// run that kind's short test; otherwise the generic one.  Each copy is a
// contiguous code path (the instruction cache sees only the one that runs).
template <int WG, int VEC, int K0>
__device__ __forceinline__ void probe_dispatch(const ProbeArgs& a, uint32_t bid, uint32_t nblk) {
  constexpr bool has_spec = K0 == S0_MODP2_LDS || K0 == S0_MOD_LDS || K0 == S0_MODP2_GL || K0 == S0_MOD_GL ||
                            K0 == S0_MULT_LDS || K0 == S0_MULT_GL;
  // A filter staged as a prefix with bits set past it (selected dimension
  // keys outside the table's domain, counted by the build): the generic body,
  // whose filter reads fall back to the global words past the prefix.
  bool past = false;
  if (a.check_past) {
#pragma unroll
    for (int d = 0; d < kMaxDims; d++)
      if (d < a.n_dims && a.dim[d].past_count && a.dim[d].bloom_lds >= 0 &&
          a.dim[d].bloom_lds_words < a.dim[d].bloom_words)
        past = past || *a.dim[d].past_count != 0ull;
  }
  constexpr bool is_mod = K0 == S0_MODP2_LDS || K0 == S0_MOD_LDS || K0 == S0_MODP2_GL || K0 == S0_MOD_GL;
  if constexpr (is_mod) {
    // one dimension summed in place (a semi-join through an exact filter, or
    // a one-group join over a perfect-hash table whose filter passes enough
    // rows): a body of its own (its in-place sums cost registers the other
    // bodies do not pay for)
    bool direct = a.direct1 != 0;
    if (direct && a.direct_min_sel > 0 && a.dim_counters[0]) {
      // a dimension built with the query: its selected keys, from the build's counters
      unsigned long long sel = 0;
      for (int l = 0; l < kCounterLanes; l++) sel += a.dim_counters[0][2 * l];
      direct = sel >= (unsigned long long)a.direct_min_sel;
    }
    if (a.n_dims == 1 && (a.dim[0].exact || direct)) {
      probe_body<WG, VEC, K0, 3>(a, bid, nblk);
      return;
    }
  }
  if constexpr (has_spec) {
    if (a.all_k0 && !past) probe_body<WG, VEC, K0, 1>(a, bid, nblk);
    else probe_body<WG, VEC, K0, 2>(a, bid, nblk);
  } else {
    probe_body<WG, VEC, K0, 2>(a, bid, nblk);
  }
}

template <int WG, int VEC, int K0>
__global__ __launch_bounds__(WG) void probe_kernel(ProbeArgs args) {
  // Read the arguments in place from the kernarg segment: indexing the
  // by-value copy with a (uniform) dimension index would copy it to scratch.
#if defined(__HIP_DEVICE_COMPILE__)
  (void)args;
  const ProbeArgs& a = *(const ProbeArgs*)(__builtin_amdgcn_kernarg_segment_ptr());
#else
  const ProbeArgs& a = args;
#endif
  probe_dispatch<WG, VEC, K0>(a, blockIdx.x, gridDim.x);
}

// A batch of queries in one launch: query q is served by workgroups
// [wg_begin[q], wg_begin[q+1]) (wg_begin[n_q] == gridDim.x); each workgroup
// finds its query by binary search and runs the same probe on the query's
// arguments, read from device memory.
__device__ __forceinline__ int query_of_block(const uint32_t* __restrict__ wg_begin, int32_t n_q) {
  int lo = 0, hi = n_q - 1;
  while (lo < hi) {  // last q with wg_begin[q] <= blockIdx.x (uniform)
    const int mid = (lo + hi + 1) >> 1;
    if (wg_begin[mid] <= blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <int WG, int VEC, int K0>
__global__ __launch_bounds__(WG) void probe_multi_kernel(const ProbeArgs* __restrict__ qs,
                                                         const uint32_t* __restrict__ wg_begin, int32_t n_q) {
  // __restrict__: the arguments do not alias what the probe writes, so their
  // fields load once instead of after every store
  const int q = query_of_block(wg_begin, n_q);
  const uint32_t b0 = wg_begin[q];
  probe_dispatch<WG, VEC, K0>(qs[q], blockIdx.x - b0, wg_begin[q + 1] - b0);
}

// Host-side table of instantiated kernels, filled by the instantiation units
// (a build may name other shapes with -DLJ_PROBE_SHAPES(X,K)=...).
#ifndef LJ_PROBE_SHAPES
#define LJ_PROBE_SHAPES(X, K) X(256, 1, K) X(512, 1, K) X(1024, 1, K)
#endif

}  // namespace lj
