// lj_internal.hpp -- kernel argument blocks shared by the kernels and the host
// code of liblipjoin.so (not part of the public ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lipjoin/block/hash.hpp"
#include "lipjoin/block/pred.hpp"

namespace lj {

constexpr int kMaxDims = 4;
constexpr int kMaxDimsConst = kMaxDims;
constexpr int kGroupSlots = 32;      // LDS-privatised groups per workgroup (the smallest table)
constexpr int kGroupSlotsMax = 4096; // the largest table (sized from the LDS that staging leaves)
constexpr int kGroupSlotBytes = 3 * 8 + 4 + kMaxDimsConst * 4;  // sum, count, product, key, payloads
constexpr int kHotFilterWords = 4096; // filters up to 16 KB are staged in LDS wherever they sit in the chain
// Group spaces up to kRepGroups keep kGroupReps copies of the global group
// accumulators (copy = workgroup index mod 8, i.e. one per XCD under the
// round-robin dispatch), so that groups overflowing the LDS table are not all
// added into the same few cache lines by every CU.  The bound keeps the
// finalize (one workgroup reads every copy of every touched group) short.
constexpr int kGroupReps = 8;
constexpr int64_t kRepGroups = 1 << 13;
constexpr int kQueue = 64;           // survivor rows per batch (one per lane)
constexpr int kQueueCap = kQueue + 8; // survivor queue entries per wave
constexpr int kStats = 16;           // per-workgroup statistics words
constexpr int kSegRows = 256;        // rows of one wave segment (64 lanes x 4)
constexpr int kPrefetch = 1;         // full tiles a wave keeps in flight ahead of the one it processes
// result header words: [0] groups hit, [1..kStats] statistics, then 2 counters
// per dimension, then 3 kernel wall-clock stamps (start, build done, end)
constexpr int kClockSlot = 1 + kStats + 2 * kMaxDims;

// statistics word indices
enum StatIdx : int {
  ST_ROWS_IN = 0,        // [0..4]
  ST_LINES_IN = 5,       // [5..9]
  ST_BLOOM_PASS = 10,
  ST_SAMPLE_ROWS = 11,
  ST_SAMPLE_LINES = 12,
  ST_OOB = 13,
  ST_ORDER = 14          // order of workgroup 0, 8 bits per stage
};

struct GroupOut;

struct BuildDim {
  const int32_t* keys;
  const int32_t* attr;
  const int32_t* payload;
  int64_t n;
  PredProgram pred;
  int32_t* ht;
  FastMod ht_len;          // textbook: pow2_mask = slots - 1
  int32_t keys_min;
  int32_t ht_kind;         // HtKind
  uint32_t* bloom;         // nullptr: no filter
  BloomHash bh;
  unsigned long long* counters;  // kCounterLanes x {selected rows, keys outside the domain}, ..., kCounterPast
  int64_t block_begin;     // first workgroup of this dimension in the build launch
  int64_t past_bit;        // modulo filter: the first bit past its staged prefix (0: not counted)
};

// A dimension's counters are spread over kCounterLanes pairs (workgroup w adds
// to pair w % kCounterLanes) so that the build's per-workgroup atomics do not
// all serialize on one address; readers sum the pairs.
constexpr int kCounterLanes = 16;
constexpr int kCounterWords = 2 * kCounterLanes + 3;  // + popcount scratch, kCounterPast, kCounterOutside
constexpr int kCounterPast = 2 * kCounterLanes + 1;   // selected keys whose filter bit is past the staged prefix
constexpr int kCounterOutside = 2 * kCounterLanes + 2;  // selected PHT keys at or past keys_min + ht_len

struct ProbeDim {
  const int32_t* fk;      // fact FK column of this dimension
  const uint32_t* bloom;  // global Bloom words (nullptr: no filter)
  const int32_t* ht;      // the hash table (PHT_2 / textbook pairs, PHT_1 keys)
  BloomHash bh;
  FastMod ht_len;          // textbook: pow2_mask = slots - 1
  int32_t keys_min;       // PHT's HASH Z argument (textbook: INT32_MIN + 1, its domain's bound)
  int32_t bloom_lds;      // word offset of the staged copy in LDS, -1 = global
  int32_t bloom_words;
  int32_t ht_kind;        // HtKind
  int32_t bloom_lds_words;// words staged: a prefix of the filter; a word past it is read from
                          // global memory (only out-of-domain keys reach it: see plan_probe)
  int64_t gmult;
  const unsigned long long* past_count;  // filter bits set past the staged prefix (device
                                         // counter of a dimension built with the query), or null
  int32_t exact;          // LIP, one dimension, one group, no product: the filter is a modulo
                          // filter of at least ht_len bits over the table's keys_min, so a
                          // passing key in [keys_min, keys_min + ht_len), other than 0, hits
                          // the table (see exact_key); its payload is not needed
  int32_t reserved0;
};

struct ProbeArgs {
  ProbeDim dim[kMaxDims];
  const int32_t* measure;
  const int32_t* measure_sub;  // nullptr, or subtracted from measure per row (two or more dimensions)
  int64_t n;               // fact rows (< 2^31)
  int32_t n_dims;
  int32_t method;          // 0 NAIVE, 1 LIP
  int32_t adaptive;        // LIP: order chosen from the sampling pass
  int32_t want_product;    // also sum (int32)(measure * payload of dim 0)
  int32_t order[kMaxDims];        // hash-probe (join) order
  int32_t load_order[kMaxDims];   // column load order (NAIVE: order, LIP: Bloom order)
  int32_t lds_filter_words;
  int32_t all_k0;          // LIP: every dimension's filter has stage 0's hash family (the
                           // placement, LDS or global, may differ: a uniform branch)
  uint32_t s0_mask;        // adaptive: dimensions the sample may put first (their filters have
                           // stage 0's kind); 0 = any
  int64_t goffset;
  int64_t n_groups;
  FastMod gmod;            // n_groups as a 32-bit divisor (n_groups <= 2^26)
  unsigned long long* gsum;
  unsigned long long* gcount;
  unsigned long long* gprod;
  int32_t* gpay;           // [n_groups * 4]
  uint32_t* touched;       // first-touched groups, in touch order
  uint32_t* n_touched;
  uint32_t touched_cap;
  int32_t gslot_bits;      // log2 of the workgroup's LDS group-table slots
  long long* stats;        // [gridDim.x * kStats]
  // result: the last workgroup to finish copies the groups out and resets the
  // dense accumulators
  uint32_t* done;          // workgroups finished (reset to 0 by the last one)
  long long* out_hdr;      // [0] groups hit, [1..kStats] statistics, then 2 counters per dimension
  GroupOut* out;
  int64_t out_cap;
  const unsigned long long* dim_counters[kMaxDims];
  // replicated accumulators (single queries with n_groups <= kRepGroups):
  // gsum/gcount/gprod hold g_reps copies g_stride apart; gseen[g] marks a
  // group as listed in `touched`.  g_reps 0 or 1: one copy, gseen unused.
  uint32_t* gseen;
  int64_t g_stride;
  int32_t g_reps;
  int32_t check_past;      // some dimension passes past_count: the probe reads it at its start
  int32_t k0;              // the query's stage-0 kind (S0Kind; read by probe_mixed_kernel)
  int32_t exact_eager_div;  // an exact one-dimension semi-join streams its measure with the keys when
                            // stage 0 passes more than 1 / exact_eager_div of its first tile's rows
  int32_t direct1;         // one dimension, one group, a perfect-hash table: while the measure
                           // streams, the survivors' table slots are gathered and summed in place
  int32_t direct_eager_div;  // ... the measure streams above 1 / direct_eager_div pass
  int32_t occ_lds;         // direct1 with occ_words > 0: the table's occupancy bitmap (bit s set iff
  int32_t occ_words;       // slot s is nonzero), built in LDS at this word offset by each workgroup
  int32_t direct_min_sel;  // direct1, a dimension built with the query: the in-place body runs when
                           // its build selected at least this many keys (0: always)
  int32_t occ_eager_div;   // with the bitmap: the measure streams above 1 / occ_eager_div of the
                           // first tile's rows hitting (below, the hits' measures are gathered)
  int32_t occ_min_sel;     // occ_words > 0, a dimension built with the query: the bitmap is built and
                           // used when its build selected at least this many keys (0: always)
  // if set (a single query): the header and the first host_groups groups are
  // also written to this pinned host memory, so no copy follows the probe
  long long* host_hdr;
  GroupOut* host_out;
  int64_t host_groups;
};


struct BuildArgs {
  BuildDim dim[kMaxDims];
  int32_t n_dims;
  int32_t pad0;
  long long* clock;        // if set: wall-clock stamp at the start of the build
};

struct FillRegion {
  uint32_t* ptr;
  int64_t words;
  uint32_t value;
  int32_t pad0;
};

constexpr int kMaxFill = 24;
struct FillArgs {
  FillRegion r[kMaxFill];
  int32_t n;
  int32_t pad0;
};

// One group of the finalized result (device layout).
struct GroupOut {
  long long g;
  long long sum;
  unsigned long long count;
  unsigned long long prod;
  int32_t pay[kMaxDims];
};


// Word copies on a stream (up to two ranges in one launch); a source may be
// pinned host memory through its device address.
struct CopyRange {
  uint32_t* dst;
  const uint32_t* src;
  int64_t words;
};
struct CopyArgs {
  CopyRange r[2];
  int32_t n;
  int32_t pad0;
};

// Host launchers (lj_kernels_*.hip).
hipError_t launch_fill(const FillArgs& a, hipStream_t s);
hipError_t launch_copy(const CopyArgs& a, hipStream_t s);
hipError_t launch_build(BuildArgs& a, hipStream_t s);
hipError_t launch_gen(int32_t kind, uint64_t key, int64_t lo, int64_t hi, int64_t first, int64_t n,
                      int32_t* out, hipStream_t s);
hipError_t launch_popcount(const uint32_t* words, int64_t n, unsigned long long* out, hipStream_t s);

// Probe kernels: one instantiation per (workgroup size, segments per wave
// tile, stage-0 filter kind).
enum S0Kind : int {
  S0_PHT = 0,        // NAIVE: PHT_2 probe
  S0_NONE = 1,       // LIP dimension without a Bloom filter
  S0_MODP2_LDS = 2,  // HASH1 over a power-of-two bit count, filter staged in LDS
  S0_MODP2_GL = 3,   //   ... filter in global memory (L2/MALL)
  S0_MOD_LDS = 4,    // HASH1, any bit count
  S0_MOD_GL = 5,
  S0_MULT_LDS = 6,   // multiplicative family (standard or blocked, k = 1..3)
  S0_MULT_GL = 7,
  S0_GENERIC = 8,    // runtime dispatch (adaptive order over mixed filter kinds)
  S0_NKINDS = 9
};
// ProbeConfig.k0 of a launch whose queries have different stage-0 kinds
// (probe_mixed_kernel: each query's ProbeArgs.k0 picks its body)
constexpr int kS0Mixed = 100;

struct ProbeConfig {
  int wg;      // threads per workgroup (256, 512, 1024)
  int vec;     // 256-row segments per wave tile (1, 2)
  int k0;      // S0Kind
};
typedef void (*ProbeKernelFn)(ProbeArgs);
struct ProbeEntry {
  int wg, vec, k0;
  const void* fn;        // probe_kernel (one query)
  const void* fn_multi;  // probe_multi_kernel (a batch of queries)
};
int probe_entries_pht(const ProbeEntry** out);
int probe_entries_mixed(const ProbeEntry** out);
int probe_entries_modp2(const ProbeEntry** out);
int probe_entries_mod(const ProbeEntry** out);
int probe_entries_mult(const ProbeEntry** out);
int probe_entries_generic(const ProbeEntry** out);

size_t probe_lds_bytes(const ProbeConfig& c, int lds_filter_words, int group_slots);
hipError_t probe_occupancy(const ProbeConfig& c, size_t lds_bytes, int* blocks_per_cu);
bool probe_has_entry(const ProbeConfig& c);  // a kernel is built for this configuration
// A batch: qs[n_q] and wg_begin[n_q + 1] in device memory; grid == wg_begin[n_q].
hipError_t launch_probe_multi(const ProbeConfig& c, const ProbeArgs* qs, const uint32_t* wg_begin, int n_q, int grid,
                              size_t lds_bytes, hipStream_t s);
// Every dimension of dims[0..n) (device memory; block_begin set): total_blocks workgroups.
hipError_t launch_build_multi(const BuildDim* dims, int n, int64_t total_blocks, hipStream_t s);
int64_t build_blocks(int64_t rows);
hipError_t launch_probe(const ProbeConfig& c, const ProbeArgs& a, int grid, size_t lds_bytes, hipStream_t s);

}  // namespace lj
