// lj_api.hip -- host side of the C-ABI in include/lipjoin.h.
//
// Orchestration mirrors the reference drivers' runQuery / hashJoin
// (src/ssb-jo/q33/0-lip.cu:205-302, src/ops/bloom_join.cu:121-156): allocate and
// clear the tables, build every dimension, run the fused probe, collect the
// groups.  Differences by design: a query is two kernel launches (one build
// of all its dimensions; the probe, whose last workgroup collects the groups)
// and one small copy back; tables come from a per-context caching allocator and
// are returned to it (never leaked) and cleared there, behind the query, so
// the next query starts on zeroed tables; the dense group arrays are kept zero
// by the result collection instead of a 7 MB memset per query; every failure
// is a status code.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "lipjoin.h"
#include "lj_gen.hpp"
#include "lj_internal.hpp"

using namespace lj;

// ------------------------------------------------------------------ context --
struct lj_dim {
  lj_dim_desc desc;
  int32_t* ht = nullptr;
  int64_t ht_words = 0;
  uint32_t* bloom = nullptr;
  int64_t bloom_words = 0;
  unsigned long long* counters = nullptr;  // device [0] selected, [1] oob, [2] popcount scratch
  BloomHash bh{};
  FastMod ht_len{};
  int64_t n_selected = -1;
  int64_t n_oob = -1;
  int64_t n_past = -1;     // resident: filter bits set past the staged prefix (-1: not known)
  int64_t n_outside = -1;  // resident: selected keys from keys_min + ht_len on (-1: not known)
  float time_build_ms = 0.f;
  // tables taken from the zero pool (cleared behind their last use) go back
  // there; tables cleared inside the query's own window (lj_hash_join, the
  // reference's timed memsets) go back to the plain free list
  bool zero_pool = true;
};

constexpr int kEvents = 4;
constexpr int64_t kMaxGroupSpace = 1ll << 26;  // a query's group space (n_groups) at most
constexpr int kEvDone = 1;

struct lj_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  uint64_t budget = 0;
  uint64_t in_use = 0;  // bytes handed out
  uint64_t cached = 0;  // bytes parked in the free lists
  std::string err;
  std::multimap<size_t, void*> free_blocks;
  std::multimap<size_t, void*> zero_blocks;  // free blocks known to be all zero (cleared on release)
  std::unordered_map<void*, size_t> live;
  int n_cu = 0;
  int max_lds = 65536;
  // ev[0] before the build, ev[3] between build and probe, ev[1] (kEvDone)
  // after the probe (which writes the result to pinned host memory itself)
  hipEvent_t ev[kEvents]{};
  // group state (kept zero between queries by the finalize kernel)
  int64_t g_cap = 0;
  unsigned long long* gsum = nullptr;
  unsigned long long* gcount = nullptr;
  unsigned long long* gprod = nullptr;
  int32_t* gpay = nullptr;
  uint32_t* gseen = nullptr;       // [g_cap] when g_reps > 1
  int32_t g_reps = 1;              // copies of gsum/gcount/gprod, g_cap apart
  uint32_t* touched = nullptr;
  uint32_t touched_cap = 0;
  uint32_t* n_touched = nullptr;  // [0] groups hit; [32] workgroups finished (probe kernel)
  void* out_buf = nullptr;       // [kHdrBytes header][GroupOut x touched_cap]
  GroupOut* out = nullptr;        // inside out_buf
  long long* out_hdr = nullptr;   // inside out_buf
  long long* stats = nullptr;
  int stats_cap = 0;
  void* host_buf = nullptr;  // pinned: header + first kHostGroups groups
  void* host_buf_dev = nullptr;  // the same memory, as the device addresses it
  // tuning (environment overrides for sweeps: LJ_PROBE_WG, LJ_PROBE_VEC, LJ_LDS_POLICY, LJ_GROUP_SLOTS)
  int probe_wg = 0;
  int probe_vec = 0;
  int exact_filter = 1;    // let an exact filter decide the table probe (LJ_EXACT_FILTER=0: off)
  int lds_policy = -1;
  int group_slots = 0;           // LJ_GROUP_SLOTS: the LDS group table's size, 0 = from the group space
  int filter_prefix = 1;         // LJ_FILTER_PREFIX=0: stage whole modulo filters
  int partial_prefix = 1;        // LJ_PARTIAL_PREFIX=0: never stage part of a filter that does not fit
  int exact_eager_div = 32;      // LJ_EXACT_EAGER: an exact semi-join streams its measure above 1/this pass
  int direct1 = 1;               // LJ_DIRECT=0: one-dimension one-group joins queue every survivor
                                 // (2: the in-place body whatever the filter's estimated pass rate)
  int direct_eager_div = 4;      // LJ_DIRECT_EAGER: ... and are summed in place above 1/this pass
  int direct_occ = 1;            // LJ_DIRECT_OCC: the table's occupancy bitmap in LDS for them
                                 // (0 never, 1 when the fact rows repay it, 2 whenever it fits)
  int occ_eager_div = 8;         // LJ_DIRECT_OCC_EAGER: with it, the measure streams above 1/this hitting
  double clock_khz = 100000.0;  // device wall clock (kernel phase timestamps)
  // Bloom autotuner cost model, per fact row, relative to the scan of an LDS filter
  // cost model of the Bloom autotuner, per fact row, in units of a scan with
  // the filter in LDS (defaults: typical values lj_tune_calibrate measures)
  double tune_scan_gl = 2.7;     // scan with the first filter in L2 instead of LDS (up to kTuneL2Bits)
  double tune_scan_big = 2.7;    // scan with a larger filter (L2 misses, served from the MALL)
  double tune_scan_mod = 0.45;   // scan with a modulo filter in LDS
  double tune_scan_blk = 0.65;   // scan with a blocked multiplicative filter in LDS
  double tune_probe_extra = 0.23;// each further bit probe of a standard k > 1 filter outside LDS (x scan)
  double tune_line = 0.36;       // a 128-byte line of the next column that holds a passing row
  double tune_gather = 5.7;      // a passing row when more than a quarter pass (the measure streams), L2 filter
  double tune_gather_lds = 9.3;  // the same with the filter in LDS (nothing else to overlap the gathers with)
  double tune_gather_q = 5.3;    // a passing row when fewer pass (queue, hash-slot and measure gathers), L2 filter
  double tune_gather_q_lds = 7.5;   // the same with the filter in LDS (its short scan hides no gather)
  int tune_calibrated = 0;
  std::map<std::pair<int, size_t>, int> occupancy;  // (config key, LDS bytes) -> workgroups per CU
  std::multimap<size_t, void*> pinned_free;         // pinned host blocks for reuse (batches)
};

static const int kHostGroups = 4096;   // groups the pinned host buffer holds
static const int kFirstGroups = 64;    // groups copied back together with the header
static const size_t kHdrBytes = 256;   // [0] groups hit, [1..16] statistics, [17..24] dimension counters,
                                       // [25..27] kernel phase clocks
static thread_local std::string g_create_err;

static constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// A query's buffers in a batch or a serving launch, in three regions:
//   zero  : gsum | gcount | gprod [ng each] | n_touched, done (kept zero
//           between runs: the finalize clears what the query touched)
//   plain : gpay [ng * 4] | touched [tcap] | statistics rows [rows * kStats]
//   out   : result header | groups [tcap]
struct QBufLayout {
  size_t zero = 0, plain = 0, out = 0;
};
static QBufLayout qbuf_layout(int64_t ng, int64_t tcap, int64_t stats_rows) {
  QBufLayout l;
  l.zero = align256((size_t)ng * 8) * 3 + 256;
  l.plain = align256((size_t)ng * kMaxDims * 4) + align256((size_t)tcap * 4) + align256((size_t)stats_rows * kStats * 8);
  l.out = align256(kHdrBytes + sizeof(GroupOut) * (size_t)tcap);
  return l;
}
static void qbuf_bind(ProbeArgs& a, int64_t ng, int64_t tcap, char* z, char* pl, char* od) {
  const size_t g8 = align256((size_t)ng * 8);
  a.gsum = reinterpret_cast<unsigned long long*>(z);
  a.gcount = reinterpret_cast<unsigned long long*>(z + g8);
  a.gprod = reinterpret_cast<unsigned long long*>(z + 2 * g8);
  a.n_touched = reinterpret_cast<uint32_t*>(z + 3 * g8);
  a.done = a.n_touched + 32;
  a.gpay = reinterpret_cast<int32_t*>(pl);
  a.touched = reinterpret_cast<uint32_t*>(pl + align256((size_t)ng * kMaxDims * 4));
  a.stats = reinterpret_cast<long long*>(pl + align256((size_t)ng * kMaxDims * 4) + align256((size_t)tcap * 4));
  a.touched_cap = (uint32_t)tcap;
  a.out_hdr = reinterpret_cast<long long*>(od);
  a.out = reinterpret_cast<GroupOut*>(od + kHdrBytes);
  a.out_cap = tcap;
}

// ---------------------------------------------------- no-allocation scratch --
// lj_star_join_noalloc runs on `scratch`, a file-scope device array: the
// runtime gives every GPU its own copy when the library loads, and each GPU
// holds it until the process ends.  Its regions are those of qbuf_layout for
// kNaGroups groups, kNaListed listed groups and kNaRows statistics rows
// (20.3 MB); the zero region is cleared once per device by lj_create and is
// kept zero by the finalize (and by a clear after a list overflow).  Host
// code never uses the array's name as an address: it uses the address that
// hipGetSymbolAddress returned for the device (g_na_ptr).
constexpr int64_t kNaGroups = 1 << 19;    // group space (Q3.3's 437,500 fits)
constexpr int64_t kNaListed = 4096;       // groups a query may hit (the pinned result buffer's size)
constexpr int64_t kNaRows = 2048;         // workgroups of one probe launch
constexpr size_t kNaZeroBytes = align256((size_t)kNaGroups * 8) * 3 + 256;
constexpr size_t kNaPlainBytes =
    align256((size_t)kNaGroups * kMaxDims * 4) + align256((size_t)kNaListed * 4) + align256((size_t)kNaRows * kStats * 8);
constexpr size_t kNaOutBytes = align256(kHdrBytes + sizeof(GroupOut) * (size_t)kNaListed);
constexpr size_t kNaScratchBytes = kNaZeroBytes + kNaPlainBytes + kNaOutBytes;
constexpr int kNaMaxDevices = 64;
static_assert(kNaListed == kHostGroups, "the result goes through the context's pinned buffer");

static __device__ __align__(256) unsigned char scratch[kNaScratchBytes];

static std::mutex g_na_mu;
static char* g_na_ptr[kNaMaxDevices];   // scratch's address on each device (null: not set up)

static lj_status fail(lj_ctx* c, lj_status s, const char* fmt, ...) __attribute__((format(printf, 3, 4)));
static lj_status fail(lj_ctx* c, lj_status s, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  else g_create_err = buf;
  return s;
}

#define HIP_TRY(c, expr)                                                                           \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess) return fail(c, LJ_EHIP, "%s: %s", #expr, hipGetErrorString(e_));        \
  } while (0)

// Wait for an event by polling: a query is a few hundred microseconds of
// device work, and a blocking wait's wake-up would add tens of microseconds.
static hipError_t spin_wait(hipEvent_t e) {
  for (;;) {
    const hipError_t q = hipEventQuery(e);
    if (q != hipErrorNotReady) return q;
    std::this_thread::yield();
  }
}

// ---------------------------------------------------------- caching allocator --
static size_t round_block(size_t bytes) {
  size_t b = 256;
  while (b < bytes) b <<= 1;
  return b;
}

static lj_status dev_alloc(lj_ctx* c, size_t bytes, void** out) {
  const size_t sz = round_block(bytes ? bytes : 1);
  auto it = c->free_blocks.find(sz);
  if (it != c->free_blocks.end()) {
    *out = it->second;
    c->free_blocks.erase(it);
    c->cached -= sz;
    c->live[*out] = sz;
    c->in_use += sz;
    return LJ_OK;
  }
  if (c->budget && c->in_use + sz > c->budget)
    return fail(c, LJ_EBUDGET, "allocation of %zu bytes exceeds the HBM budget (%llu in use of %llu)", sz,
                (unsigned long long)c->in_use, (unsigned long long)c->budget);
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, sz);
  if (e != hipSuccess) {
    // give the cache back and retry once
    (void)hipGetLastError();
    (void)hipStreamSynchronize(c->stream);  // parked blocks may still be cleared by queued fills
    for (auto& kv : c->free_blocks) (void)hipFree(kv.second);
    for (auto& kv : c->zero_blocks) (void)hipFree(kv.second);
    c->free_blocks.clear();
    c->zero_blocks.clear();
    c->cached = 0;
    e = hipMalloc(&p, sz);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return fail(c, LJ_ENOMEM, "hipMalloc(%zu) failed: %s", sz, hipGetErrorString(e));
    }
  }
  c->live[p] = sz;
  c->in_use += sz;
  *out = p;
  return LJ_OK;
}

static void dev_release(lj_ctx* c, void* p) {
  if (!p) return;
  auto it = c->live.find(p);
  if (it == c->live.end()) return;
  c->free_blocks.emplace(it->second, p);
  c->cached += it->second;
  c->in_use -= it->second;
  c->live.erase(it);
}

// Blocks whose every byte is zero.  A released table is cleared by a fill
// queued on the context's stream behind the work that used it, so the next
// query finds it clean without a memset on its own critical path.
static lj_status dev_alloc_zero(lj_ctx* c, size_t bytes, void** out) {
  const size_t sz = round_block(bytes ? bytes : 1);
  auto it = c->zero_blocks.find(sz);
  if (it != c->zero_blocks.end()) {
    *out = it->second;
    c->zero_blocks.erase(it);
    c->cached -= sz;
    c->live[*out] = sz;
    c->in_use += sz;
    return LJ_OK;
  }
  lj_status s = dev_alloc(c, sz, out);
  if (s != LJ_OK) return s;
  FillArgs f{};
  f.r[0].ptr = static_cast<uint32_t*>(*out);
  f.r[0].words = (int64_t)(sz / 4);
  f.n = 1;
  HIP_TRY(c, launch_fill(f, c->stream));
  return LJ_OK;
}

// Queue the clear of a block and park it in the zero pool.  used_bytes: the
// prefix its user could have written (the block came zero from the pool, so
// the rest of it still is); 0 = the whole block.
static void dev_release_zero(lj_ctx* c, void* p, FillArgs& f, size_t used_bytes = 0) {
  if (!p) return;
  auto it = c->live.find(p);
  if (it == c->live.end()) return;
  if (f.n == kMaxFill) {
    (void)launch_fill(f, c->stream);
    f.n = 0;
  }
  const size_t clear = used_bytes && used_bytes < it->second ? used_bytes : it->second;
  f.r[f.n].ptr = static_cast<uint32_t*>(p);
  f.r[f.n].words = (int64_t)((clear + 3) / 4);
  f.r[f.n].value = 0;
  f.n++;
  c->zero_blocks.emplace(it->second, p);
  c->cached += it->second;
  c->in_use -= it->second;
  c->live.erase(it);
}

template <typename T>
static lj_status dev_alloc_zero_t(lj_ctx* c, int64_t count, T** out) {
  void* p = nullptr;
  lj_status s = dev_alloc_zero(c, (size_t)count * sizeof(T), &p);
  *out = static_cast<T*>(p);
  return s;
}

// Pinned host memory (power-of-two blocks, kept by the context for reuse).
static lj_status host_alloc(lj_ctx* c, size_t bytes, void** out) {
  const size_t sz = round_block(bytes ? bytes : 1);
  auto it = c->pinned_free.find(sz);
  if (it != c->pinned_free.end()) {
    *out = it->second;
    c->pinned_free.erase(it);
    return LJ_OK;
  }
  HIP_TRY(c, hipHostMalloc(out, sz, hipHostMallocDefault));
  return LJ_OK;
}

static void host_release(lj_ctx* c, void* p, size_t bytes) {
  if (p) c->pinned_free.emplace(round_block(bytes ? bytes : 1), p);
}

template <typename T>
static lj_status dev_alloc_t(lj_ctx* c, int64_t count, T** out) {
  void* p = nullptr;
  lj_status s = dev_alloc(c, (size_t)count * sizeof(T), &p);
  *out = static_cast<T*>(p);
  return s;
}

// -------------------------------------------------------------- ABI basics --
extern "C" int32_t lj_abi_version(void) { return LJ_ABI_VERSION; }

extern "C" int64_t lj_abi_struct_size(int32_t which) {
  switch (which) {
    case 0: return sizeof(lj_dim_desc);
    case 1: return sizeof(lj_dim_info);
    case 2: return sizeof(lj_probe_stats);
    case 3: return sizeof(lj_join_desc);
    case 4: return sizeof(lj_join_result);
    case 5: return sizeof(lj_star_desc);
    case 6: return sizeof(lj_group);
    case 7: return sizeof(lj_star_result);
    case 8: return sizeof(lj_batch_opts);
    case 9: return sizeof(lj_bloom_choice);
    case 10: return sizeof(lj_serve_cfg);
    case 11: return sizeof(lj_serve_stats);
    case 12: return sizeof(lj_tune_model);
  }
  return -1;
}

extern "C" const char* lj_status_str(lj_status s) {
  switch (s) {
    case LJ_OK: return "ok";
    case LJ_EINVAL: return "invalid argument";
    case LJ_ENOMEM: return "out of device memory";
    case LJ_EBUDGET: return "HBM budget exceeded";
    case LJ_EHIP: return "HIP runtime error";
    case LJ_EUNSUPPORTED: return "unsupported";
    case LJ_EDOMAIN: return "key outside the hash domain";
    case LJ_ETIMEOUT: return "timeout";
    case LJ_ESTATE: return "invalid state";
  }
  return "unknown status";
}

extern "C" const char* lj_last_error(const lj_ctx* c) { return c ? c->err.c_str() : g_create_err.c_str(); }

static int env_int(const char* name, int dflt) {
  const char* v = getenv(name);
  return v && *v ? atoi(v) : dflt;
}

extern "C" lj_status lj_create(int32_t device, uint64_t hbm_budget_bytes, void* hip_stream, lj_ctx** out) {
  if (!out) return fail(nullptr, LJ_EINVAL, "out is NULL");
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0)
    return fail(nullptr, LJ_EHIP, "no HIP device available (%s)", hipGetErrorString(e));
  if (device < 0 || device >= ndev) return fail(nullptr, LJ_EINVAL, "device %d out of range [0,%d)", device, ndev);
  lj_ctx* c = new lj_ctx();
  c->device = device;
  c->budget = hbm_budget_bytes;
  auto bail = [&](lj_status s) {
    g_create_err = c->err;
    delete c;
    return s;
  };
  if (hipSetDevice(device) != hipSuccess) return bail(fail(c, LJ_EHIP, "hipSetDevice failed"));
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return bail(fail(c, LJ_EHIP, "hipGetDeviceProperties failed"));
  c->n_cu = prop.multiProcessorCount;
  {
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
      c->clock_khz = khz;
  }
  c->max_lds = (int)prop.sharedMemPerBlock;
  if (hip_stream) {
    c->stream = static_cast<hipStream_t>(hip_stream);
  } else {
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
      return bail(fail(c, LJ_EHIP, "hipStreamCreate failed"));
    c->own_stream = true;
  }
  // ev[1] follows the probe and is the one the host waits for: it keeps the
  // system-scope release (the result is in pinned host memory).  The others only time phases; without the
  // system fence, recording them does not write back and invalidate the
  // caches between the build and the probe (the probe would re-read the
  // filters and tables it finds in L2 from memory).
  for (int i = 0; i < kEvents; i++)
    if (hipEventCreateWithFlags(&c->ev[i], i == kEvDone ? hipEventDefault : hipEventDisableSystemFence) != hipSuccess)
      return bail(fail(c, LJ_EHIP, "hipEventCreate failed"));
  if (hipHostMalloc(&c->host_buf, kHdrBytes + sizeof(GroupOut) * kHostGroups) != hipSuccess)
    return bail(fail(c, LJ_EHIP, "hipHostMalloc failed"));
  if (hipHostGetDevicePointer(&c->host_buf_dev, c->host_buf, 0) != hipSuccess)
    return bail(fail(c, LJ_EHIP, "hipHostGetDevicePointer failed"));
  c->probe_wg = env_int("LJ_PROBE_WG", 0);
  c->probe_vec = env_int("LJ_PROBE_VEC", 0);
  c->exact_filter = env_int("LJ_EXACT_FILTER", 1);
  c->lds_policy = env_int("LJ_LDS_POLICY", -1);
  c->group_slots = env_int("LJ_GROUP_SLOTS", 0);
  c->filter_prefix = env_int("LJ_FILTER_PREFIX", 1);
  c->partial_prefix = env_int("LJ_PARTIAL_PREFIX", 1);
  c->exact_eager_div = std::max(1, env_int("LJ_EXACT_EAGER", 32));
  c->direct1 = env_int("LJ_DIRECT", 1);
  c->direct_eager_div = std::max(1, env_int("LJ_DIRECT_EAGER", 4));
  c->direct_occ = env_int("LJ_DIRECT_OCC", 1);
  c->occ_eager_div = std::max(1, env_int("LJ_DIRECT_OCC_EAGER", 8));
  // lj_star_join_noalloc's scratch on this device: its address, once per
  // device (the device is current), and the one clear of its zero region
  if (device < kNaMaxDevices) {
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[device]) {
      void* sp = nullptr;
      const hipError_t e2 = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
      if (e2 != hipSuccess || !sp)
        return bail(fail(c, LJ_EHIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e2)));
      FillArgs f{};
      f.r[0].ptr = static_cast<uint32_t*>(sp);
      f.r[0].words = (int64_t)(kNaZeroBytes / 4);
      f.n = 1;
      if (launch_fill(f, c->stream) != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess)
        return bail(fail(c, LJ_EHIP, "clearing the no-allocation scratch failed"));
      g_na_ptr[device] = static_cast<char*>(sp);
    }
  }
  *out = c;
  return LJ_OK;
}

extern "C" void lj_destroy(lj_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (auto& kv : c->live) (void)hipFree(kv.first);
  for (auto& kv : c->free_blocks) (void)hipFree(kv.second);
  for (auto& kv : c->zero_blocks) (void)hipFree(kv.second);
  for (auto& ev : c->ev)
    if (ev) (void)hipEventDestroy(ev);
  if (c->host_buf) (void)hipHostFree(c->host_buf);
  for (auto& kv : c->pinned_free) (void)hipHostFree(kv.second);
  if (c->own_stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

extern "C" lj_status lj_sync(lj_ctx* c) {
  if (!c) return LJ_EINVAL;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return LJ_OK;
}

extern "C" uint64_t lj_bytes_in_use(const lj_ctx* c) { return c ? c->in_use : 0; }

// ------------------------------------------------------------- dimensions --
static int ilog2_u64(uint64_t v) {
  int r = -1;
  while (v) {
    v >>= 1;
    r++;
  }
  return r;
}

static lj_status check_dim_desc(lj_ctx* c, const lj_dim_desc* d) {
  if (!d) return fail(c, LJ_EINVAL, "dimension descriptor is NULL");
  if (d->n_rows < 0 || d->n_rows >= (1ll << 31)) return fail(c, LJ_EINVAL, "dimension n_rows %lld out of range", (long long)d->n_rows);
  if (d->n_rows > 0 && !d->keys) return fail(c, LJ_EINVAL, "dimension keys is NULL");
  if (d->pred.n_steps < 0 || d->pred.n_steps > LJ_MAX_PRED_STEPS)
    return fail(c, LJ_EINVAL, "predicate has %d steps (max %d)", d->pred.n_steps, LJ_MAX_PRED_STEPS);
  for (int i = 0; i < d->pred.n_steps; i++) {
    if (d->pred.steps[i].comb < 0 || d->pred.steps[i].comb > 2 || d->pred.steps[i].cmp < 0 || d->pred.steps[i].cmp > 4)
      return fail(c, LJ_EINVAL, "predicate step %d has an unknown operator", i);
  }
  if (d->n_rows > 0 && d->pred.n_steps > 0 && !d->attr) return fail(c, LJ_EINVAL, "predicate needs the attr column");
  if (d->ht_kind != LJ_HT_PHT2 && d->ht_kind != LJ_HT_PHT1 && d->ht_kind != LJ_HT_TEXTBOOK)
    return fail(c, LJ_EINVAL, "ht_kind %d unknown", d->ht_kind);
  if (d->n_rows > 0 && d->ht_kind != LJ_HT_PHT1 && !d->payload && !d->attr)
    return fail(c, LJ_EINVAL, "the table needs a payload or attr column");
  if (d->ht_len < 1 || d->ht_len > (1ll << 30)) return fail(c, LJ_EINVAL, "ht_len %lld out of range [1, 2^30]", (long long)d->ht_len);
  if (d->ht_kind == LJ_HT_TEXTBOOK && (d->ht_len & (d->ht_len - 1)))
    return fail(c, LJ_EINVAL, "a textbook table needs a power-of-two ht_len (got %lld)", (long long)d->ht_len);
  if (d->bloom_hash == LJ_BLOOM_MOD) {
    if (d->bloom_k != 1) return fail(c, LJ_EINVAL, "MOD Bloom hash supports k = 1 only");
    if (d->bloom_bits < 32 || d->bloom_bits > (1ll << 31)) return fail(c, LJ_EINVAL, "bloom_bits %lld out of range", (long long)d->bloom_bits);
  } else if (d->bloom_hash == LJ_BLOOM_MULT) {
    if (d->bloom_k < 1 || d->bloom_k > 3) return fail(c, LJ_EINVAL, "bloom_k %d not in 1..3", d->bloom_k);
    if (d->bloom_bits < 64 || d->bloom_bits > (1ll << 31) || (d->bloom_bits & (d->bloom_bits - 1)))
      return fail(c, LJ_EINVAL, "MULT Bloom needs a power-of-two bloom_bits in [64, 2^31]");
    if (d->bloom_layout != LJ_BLOOM_STANDARD && d->bloom_layout != LJ_BLOOM_BLOCKED)
      return fail(c, LJ_EINVAL, "bloom_layout %d unknown", d->bloom_layout);
  } else if (d->bloom_hash != LJ_BLOOM_NONE) {
    return fail(c, LJ_EINVAL, "bloom_hash %d unknown", d->bloom_hash);
  }
  return LJ_OK;
}

static void dim_hashes(lj_dim* D) {
  const lj_dim_desc& d = D->desc;
  D->ht_len = make_fastmod((uint32_t)d.ht_len);
  D->ht_words = d.ht_kind == LJ_HT_PHT1 ? d.ht_len : 2 * d.ht_len;
  BloomHash& b = D->bh;
  b = BloomHash{};
  b.kind = BLOOM_NONE;
  b.keys_min = d.keys_min;
  D->bloom_words = 0;
  if (d.bloom_hash == LJ_BLOOM_MOD) {
    D->bloom_words = (d.bloom_bits + 31) / 32;
    const uint64_t m = (uint64_t)D->bloom_words * 32;
    b.k = 1;
    b.mod = make_fastmod((uint32_t)m);
    b.kind = (m & (m - 1)) == 0 ? BLOOM_MOD_POW2 : BLOOM_MOD;
  } else if (d.bloom_hash == LJ_BLOOM_MULT) {
    D->bloom_words = d.bloom_bits / 32;
    const int lg = ilog2_u64((uint64_t)d.bloom_bits);
    b.k = d.bloom_k;
    if (d.bloom_layout == LJ_BLOOM_BLOCKED) {
      b.kind = BLOOM_MULT_BLOCKED;
      b.shift = (uint32_t)(32 - (lg - 5));
    } else {
      b.kind = BLOOM_MULT_STD;
      b.shift = (uint32_t)(32 - lg);
    }
  }
}

// Return a dimension's tables to the zero pool; the clears are queued on the
// stream (after whatever already used the tables).
// Release dimensions to the zero pool: one clearing launch for all of them.
static void dims_free(lj_ctx* c, lj_dim* const* dims, int n) {
  FillArgs f{};
  for (int i = 0; i < n; i++) {
    lj_dim* D = dims[i];
    if (!D) continue;
    if (D->zero_pool) {
      dev_release_zero(c, D->ht, f, (size_t)D->ht_words * 4);
      dev_release_zero(c, D->bloom, f, (size_t)D->bloom_words * 4);
      dev_release_zero(c, D->counters, f, (size_t)kCounterWords * 8);
    } else {
      dev_release(c, D->ht);
      dev_release(c, D->bloom);
      dev_release(c, D->counters);
    }
    delete D;
  }
  if (f.n) (void)launch_fill(f, c->stream);
}

static void dim_free(lj_ctx* c, lj_dim* D) { dims_free(c, &D, 1); }

// The counters of a dimension summed over their lanes: {selected, outside the domain}.
static void sum_counters(const unsigned long long* h, unsigned long long out[2]) {
  out[0] = out[1] = 0;
  for (int l = 0; l < kCounterLanes; l++) {
    out[0] += h[2 * l];
    out[1] += h[2 * l + 1];
  }
}

// Allocate the tables of a dimension (no launches): all-zero tables from the
// zero pool, or (zero_pool false) uncleared ones that the caller clears.
static lj_status dim_alloc(lj_ctx* c, const lj_dim_desc* desc, lj_dim** out, bool zero_pool = true) {
  lj_status s = check_dim_desc(c, desc);
  if (s != LJ_OK) return s;
  lj_dim* D = new lj_dim();
  D->desc = *desc;
  D->zero_pool = zero_pool;
  if (!D->desc.payload) D->desc.payload = D->desc.attr;
  dim_hashes(D);
  if (zero_pool) {
    if ((s = dev_alloc_zero_t(c, D->ht_words, &D->ht)) != LJ_OK ||
        (D->bloom_words && (s = dev_alloc_zero_t(c, D->bloom_words, &D->bloom)) != LJ_OK) ||
        (s = dev_alloc_zero_t(c, kCounterWords, &D->counters)) != LJ_OK) {
      dim_free(c, D);
      return s;
    }
  } else if ((s = dev_alloc_t(c, D->ht_words, &D->ht)) != LJ_OK ||
             (D->bloom_words && (s = dev_alloc_t(c, D->bloom_words, &D->bloom)) != LJ_OK) ||
             (s = dev_alloc_t(c, kCounterWords, &D->counters)) != LJ_OK) {
    dim_free(c, D);
    return s;
  }
  *out = D;
  return LJ_OK;
}

// 32-bit words of a dimension's counter block (kCounterWords 64-bit counters).
constexpr int64_t kCounterFillWords = (int64_t)kCounterWords * 2;

static void add_fill(FillArgs& f, void* p, int64_t words) {
  if (!p || words <= 0) return;
  f.r[f.n].ptr = static_cast<uint32_t*>(p);
  f.r[f.n].words = words;
  f.r[f.n].value = 0;
  f.n++;
}

// Words of a dimension's filter that the probe stages in LDS.  A modulo
// filter sets bit (key - keys_min) mod m, so the keys of the table's domain
// [keys_min, keys_min + ht_len) set bits below ht_len only: that prefix is
// staged (the build counts the selected keys whose bit is past it; with any,
// the probe reads the filter's global words for bits past the prefix).
// Other hashes spread their bits: the whole filter.
static int64_t filter_prefix_words(const lj_ctx* c, const lj_dim* D) {
  if (!D->bloom_words) return 0;
  if (!c->filter_prefix || (D->bh.kind != BLOOM_MOD && D->bh.kind != BLOOM_MOD_POW2) ||
      D->desc.ht_kind == LJ_HT_TEXTBOOK)
    return D->bloom_words;
  const int64_t bits = std::min<int64_t>(D->desc.ht_len, D->bloom_words * 32);
  return std::min<int64_t>(D->bloom_words, ((bits + 31) / 32 + 3) & ~3ll);
}

// Queue fill + build of up to 4 allocated dimensions.
static void set_build_dim(const lj_ctx* c, const lj_dim* D, BuildDim& bd) {
  bd.keys = D->desc.keys;
  bd.attr = D->desc.attr;
  bd.payload = D->desc.payload;
  bd.n = D->desc.n_rows;
  bd.pred.n_steps = D->desc.pred.n_steps;
  for (int k = 0; k < 4; k++) {
    bd.pred.comb[k] = D->desc.pred.steps[k].comb;
    bd.pred.cmp[k] = D->desc.pred.steps[k].cmp;
    bd.pred.value[k] = D->desc.pred.steps[k].value;
  }
  bd.ht = D->ht;
  bd.ht_len = D->ht_len;
  bd.keys_min = D->desc.keys_min;
  bd.ht_kind = D->desc.ht_kind;
  bd.bloom = D->bloom;
  bd.bh = D->bh;
  bd.counters = D->counters;
  const int64_t pw = filter_prefix_words(c, D);
  bd.past_bit = pw < D->bloom_words ? pw * 32 : 0;
}

// Build of allocated (all-zero) dimensions: one launch for all of them.
static lj_status dims_launch(lj_ctx* c, lj_dim* const* dims, int n, long long* clock) {
  BuildArgs b{};
  for (int i = 0; i < n; i++) set_build_dim(c, dims[i], b.dim[b.n_dims++]);
  b.clock = clock;
  HIP_TRY(c, launch_build(b, c->stream));
  return LJ_OK;
}

extern "C" lj_status lj_dim_build(lj_ctx* c, const lj_dim_desc* desc, lj_dim** out) {
  if (!c || !out) return fail(c, LJ_EINVAL, "NULL argument");
  *out = nullptr;
  HIP_TRY(c, hipSetDevice(c->device));
  lj_dim* D = nullptr;
  lj_status s = dim_alloc(c, desc, &D);
  if (s != LJ_OK) return s;
  HIP_TRY(c, hipEventRecord(c->ev[0], c->stream));
  if ((s = dims_launch(c, &D, 1, nullptr)) != LJ_OK) {
    dim_free(c, D);
    return s;
  }
  HIP_TRY(c, hipEventRecord(c->ev[1], c->stream));
  unsigned long long hc[kCounterWords], h[2];
  HIP_TRY(c, hipMemcpyAsync(hc, D->counters, sizeof hc, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  sum_counters(hc, h);
  D->n_past = (int64_t)hc[kCounterPast];
  D->n_outside = (int64_t)hc[kCounterOutside];
  HIP_TRY(c, hipEventElapsedTime(&D->time_build_ms, c->ev[0], c->ev[1]));
  D->n_selected = (int64_t)h[0];
  D->n_oob = (int64_t)h[1];
  *out = D;
  if (D->n_oob > 0)
    return fail(c, LJ_EDOMAIN, "%lld selected keys are below keys_min (tables built without them)", (long long)D->n_oob);
  return LJ_OK;
}

extern "C" lj_status lj_dim_get_info(lj_ctx* c, const lj_dim* D, lj_dim_info* info) {
  if (!c || !D || !info) return fail(c, LJ_EINVAL, "NULL argument");
  HIP_TRY(c, hipSetDevice(c->device));
  memset(info, 0, sizeof *info);
  info->n_selected = D->n_selected;
  info->n_oob = D->n_oob;
  info->ht_words = D->ht_words;
  info->bloom_words = D->bloom_words;
  info->time_build_ms = D->time_build_ms;
  if (D->bloom_words) {
    unsigned long long* cnt = D->counters + 2 * kCounterLanes;
    HIP_TRY(c, hipMemsetAsync(cnt, 0, 8, c->stream));
    HIP_TRY(c, launch_popcount(D->bloom, D->bloom_words, cnt, c->stream));
    unsigned long long h = 0;
    HIP_TRY(c, hipMemcpyAsync(&h, cnt, 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    info->bloom_bits_set = (int64_t)h;
  }
  return LJ_OK;
}

extern "C" lj_status lj_dim_export(lj_ctx* c, const lj_dim* D, int32_t* ht_host, int64_t ht_words,
                                   uint32_t* bloom_host, int64_t bloom_words) {
  if (!c || !D) return fail(c, LJ_EINVAL, "NULL argument");
  HIP_TRY(c, hipSetDevice(c->device));
  if (ht_host) {
    if (ht_words != D->ht_words) return fail(c, LJ_EINVAL, "ht_words %lld != %lld", (long long)ht_words, (long long)D->ht_words);
    HIP_TRY(c, hipMemcpyAsync(ht_host, D->ht, (size_t)ht_words * 4, hipMemcpyDeviceToHost, c->stream));
  }
  if (bloom_host) {
    if (bloom_words != D->bloom_words)
      return fail(c, LJ_EINVAL, "bloom_words %lld != %lld", (long long)bloom_words, (long long)D->bloom_words);
    HIP_TRY(c, hipMemcpyAsync(bloom_host, D->bloom, (size_t)bloom_words * 4, hipMemcpyDeviceToHost, c->stream));
  }
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return LJ_OK;
}

extern "C" void lj_dim_release(lj_ctx* c, lj_dim* D) {
  if (!c || !D) return;
  (void)hipStreamSynchronize(c->stream);
  dim_free(c, D);
}

// ------------------------------------------------------------------ queries --
static lj_status ensure_groups(lj_ctx* c, int64_t n_groups) {
  if (n_groups <= c->g_cap) return LJ_OK;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  dev_release(c, c->gsum);
  dev_release(c, c->gcount);
  dev_release(c, c->gprod);
  dev_release(c, c->gpay);
  dev_release(c, c->gseen);
  dev_release(c, c->touched);
  dev_release(c, c->out_buf);
  c->out_buf = nullptr;
  c->gseen = nullptr;
  c->g_cap = 0;
  c->g_reps = 1;
  const int64_t cap = n_groups;
  const int reps = cap <= kRepGroups ? kGroupReps : 1;
  const uint32_t tcap = (uint32_t)std::min<int64_t>(n_groups, 1 << 20);
  lj_status s;
  if ((s = dev_alloc_t(c, cap * reps, &c->gsum)) != LJ_OK || (s = dev_alloc_t(c, cap * reps, &c->gcount)) != LJ_OK ||
      (s = dev_alloc_t(c, cap * reps, &c->gprod)) != LJ_OK ||
      (s = dev_alloc_t(c, cap * kMaxDims, &c->gpay)) != LJ_OK ||
      (reps > 1 && (s = dev_alloc_t(c, cap, &c->gseen)) != LJ_OK) ||
      (s = dev_alloc_t(c, (int64_t)tcap, &c->touched)) != LJ_OK ||
      (s = dev_alloc(c, kHdrBytes + sizeof(GroupOut) * tcap, &c->out_buf)) != LJ_OK)
    return s;
  c->out_hdr = static_cast<long long*>(c->out_buf);
  c->out = reinterpret_cast<GroupOut*>(static_cast<char*>(c->out_buf) + kHdrBytes);
  if (!c->n_touched && (s = dev_alloc_t(c, 64, &c->n_touched)) != LJ_OK) return s;  // [0] touched, [32] done
  FillArgs f{};
  add_fill(f, c->gsum, cap * reps * 2);
  add_fill(f, c->gcount, cap * reps * 2);
  add_fill(f, c->gprod, cap * reps * 2);
  if (reps > 1) add_fill(f, c->gseen, cap);
  add_fill(f, c->n_touched, 64);
  HIP_TRY(c, launch_fill(f, c->stream));
  c->g_cap = cap;
  c->g_reps = reps;
  c->touched_cap = tcap;
  return LJ_OK;
}

static lj_status ensure_stats(lj_ctx* c, int blocks) {
  if (blocks <= c->stats_cap) return LJ_OK;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  dev_release(c, c->stats);
  c->stats = nullptr;
  c->stats_cap = 0;
  lj_status s = dev_alloc_t(c, (int64_t)blocks * kStats, &c->stats);
  if (s != LJ_OK) return s;
  c->stats_cap = blocks;
  return LJ_OK;
}

static bool is_perm(const int32_t* o, int n) {
  int seen = 0;
  for (int i = 0; i < n; i++) {
    if (o[i] < 0 || o[i] >= n || (seen >> o[i]) & 1) return false;
    seen |= 1 << o[i];
  }
  return true;
}

struct QueryOut {
  std::vector<GroupOut> groups;
  long long stats[kStats];
  float t_fill = 0, t_build = 0, t_probe = 0, t_total = 0;
  int64_t dim_selected[kMaxDims];
  int64_t dim_oob[kMaxDims];
};

// Workgroups per CU of a probe configuration, cached per context (cache
// false: a cached value is used, a new one is not stored).
static lj_status occupancy_of(lj_ctx* c, const ProbeConfig& cfg, size_t lds_bytes, int* per_cu, bool cache = true) {
  const auto key = std::make_pair(cfg.wg * 100000 + cfg.vec * 1000 + cfg.k0, lds_bytes);
  auto it = c->occupancy.find(key);
  if (it == c->occupancy.end()) {
    HIP_TRY(c, probe_occupancy(cfg, lds_bytes, per_cu));
    if (cache) c->occupancy[key] = *per_cu;
  } else {
    *per_cu = it->second;
  }
  if (*per_cu < 1)
    return fail(c, LJ_EUNSUPPORTED, "probe configuration wg=%d vec=%d with %zu B LDS does not fit", cfg.wg, cfg.vec,
                lds_bytes);
  return LJ_OK;
}

// ---- query planning (shared by single queries and batches) ----------------------
struct Plan {
  int nd = 0;
  int32_t bloom_order[LJ_MAX_DIMS] = {0, 1, 2, 3};
  lj_dim* dims[LJ_MAX_DIMS] = {nullptr, nullptr, nullptr, nullptr};
  lj_dim* owned[LJ_MAX_DIMS] = {nullptr, nullptr, nullptr, nullptr};
  int n_owned = 0;
  ProbeConfig cfg{};
  ProbeArgs a{};     // everything but the per-query buffers
  size_t lds_bytes = 0;
  int per_cu = 0;
};

static lj_status validate_query(lj_ctx* c, const lj_star_desc* q, Plan& p) {
  if (!q) return fail(c, LJ_EINVAL, "query descriptor is NULL");
  const int nd = q->n_dims;
  if (nd < 1 || nd > LJ_MAX_DIMS) return fail(c, LJ_EINVAL, "n_dims %d not in 1..%d", nd, LJ_MAX_DIMS);
  if (q->n_fact < 0 || q->n_fact >= (1ll << 31) - 4096) return fail(c, LJ_EINVAL, "n_fact %lld out of range", (long long)q->n_fact);
  if (q->method < LJ_METHOD_NAIVE || q->method > LJ_METHOD_LIP_ADAPTIVE) return fail(c, LJ_EINVAL, "unknown method %d", q->method);
  if (!is_perm(q->order, nd)) return fail(c, LJ_EINVAL, "order is not a permutation of 0..%d", nd - 1);
  bool bo_default = true;
  for (int i = 0; i < nd; i++) bo_default = bo_default && q->bloom_order[i] < 0;
  for (int i = 0; i < nd; i++) p.bloom_order[i] = bo_default ? q->order[i] : q->bloom_order[i];
  if (!is_perm(p.bloom_order, nd)) return fail(c, LJ_EINVAL, "bloom_order is not a permutation");
  if (q->n_groups < 1 || q->n_groups > kMaxGroupSpace) return fail(c, LJ_EINVAL, "n_groups %lld not in [1, 2^26]", (long long)q->n_groups);
  if (q->n_fact > 0) {
    if (!q->fact_measure || ((uintptr_t)q->fact_measure & 15)) return fail(c, LJ_EINVAL, "fact_measure must be a 16-byte aligned device pointer");
    if ((uintptr_t)q->fact_measure_sub & 15) return fail(c, LJ_EINVAL, "fact_measure_sub must be 16-byte aligned");
    if (q->fact_measure_sub && q->n_dims < 2)
      return fail(c, LJ_EUNSUPPORTED, "fact_measure_sub needs two or more dimensions");
    for (int d = 0; d < nd; d++)
      if (!q->fact_fk[d] || ((uintptr_t)q->fact_fk[d] & 15))
        return fail(c, LJ_EINVAL, "fact_fk[%d] must be a 16-byte aligned device pointer", d);
  }
  p.nd = nd;
  return LJ_OK;
}

// Resident dimensions, or tables allocated for a build inside the call (all
// zero, or with zero_pool false uncleared: the caller queues their clear).
static lj_status acquire_dims(lj_ctx* c, const lj_star_desc* q, Plan& p, bool zero_pool = true) {
  lj_status s = LJ_OK;
  for (int d = 0; d < p.nd && s == LJ_OK; d++) {
    if (q->resident[d]) {
      p.dims[d] = const_cast<lj_dim*>(q->resident[d]);
    } else {
      s = dim_alloc(c, &q->dims[d], &p.dims[d], zero_pool);
      if (s == LJ_OK) p.owned[p.n_owned++] = p.dims[d];
    }
  }
  if (s != LJ_OK) {
    dims_free(c, p.owned, p.n_owned);
    p.n_owned = 0;
  }
  return s;
}

// Workgroup shape, LDS-resident filters, stage-0 kernel kind, occupancy.
static lj_status plan_probe(lj_ctx* c, const lj_star_desc* q, int want_product, int wg, Plan& p,
                            bool cache_occupancy = true) {
  const int nd = p.nd;
  const bool lip = q->method != LJ_METHOD_NAIVE;
  const bool adaptive = q->method == LJ_METHOD_LIP_ADAPTIVE;
  ProbeConfig& cfg = p.cfg;
  cfg.wg = wg ? wg : (c->probe_wg ? c->probe_wg : 1024);
  cfg.vec = c->probe_vec ? c->probe_vec : 1;
  cfg.k0 = S0_GENERIC;
  const int lds_policy = c->lds_policy >= 0 ? c->lds_policy : 0;
  ProbeArgs& a = p.a;
  a = ProbeArgs{};
  int lds_words = 0;
  const int64_t lds_budget_words =
      ((int64_t)c->max_lds - (int64_t)probe_lds_bytes(cfg, 0, kGroupSlots) - 16) / 4;
  for (int d = 0; d < nd; d++) {
    ProbeDim& P = a.dim[d];
    lj_dim* D = p.dims[d];
    P.fk = q->fact_fk[d];
    P.bloom = D->bloom;
    P.ht = D->ht;
    P.bh = lip ? D->bh : BloomHash{};
    if (!lip) P.bh.kind = BLOOM_NONE;
    P.ht_len = D->ht_len;
    P.ht_kind = D->desc.ht_kind;
    // the fact keys outside the table's domain: below keys_min (perfect
    // hash), the excluded key INT32_MIN (textbook)
    P.keys_min = D->desc.ht_kind == LJ_HT_TEXTBOOK ? kTbExcludedKey + 1 : D->desc.keys_min;
    P.bloom_lds = -1;
    P.bloom_words = (int32_t)D->bloom_words;
    P.past_count = nullptr;
    P.bloom_lds_words = (int32_t)D->bloom_words;
    P.gmult = q->group_mult[d];
  }
  // A one-dimension LIP join whose result needs no payload (one group, no
  // product) with a modulo filter of at least ht_len bits over the table's
  // keys_min: the filter decides the table probe of every key inside the
  // table's domain (exact_key in the probe), provided every selected key lies
  // in that domain -- known here for a resident dimension, checked by the
  // probe on the build's counters for one built with the query.
  if (lip && nd == 1 && !want_product && q->n_groups == 1) {
    ProbeDim& P = a.dim[0];
    const lj_dim* D = p.dims[0];
    const bool mod = P.bh.kind == BLOOM_MOD || P.bh.kind == BLOOM_MOD_POW2;
    const bool pht = D->desc.ht_kind == LJ_HT_PHT1 || D->desc.ht_kind == LJ_HT_PHT2;
    const bool known_in = !q->resident[0] || (D->n_oob == 0 && D->n_outside == 0);
    if (c->exact_filter && mod && pht && P.bloom && P.bh.keys_min == P.keys_min && P.bh.mod.d >= P.ht_len.d &&
        known_in)
      P.exact = 1;
  }
  // A one-dimension, one-group LIP join over a perfect-hash table: where the
  // filter passes enough rows for the measure to stream, the survivors probe
  // the table and are summed in place instead of queued (a modulo filter:
  // the bodies of the other kinds keep the queue).  Its body pays from about
  // a fifth of the rows passing: a modulo filter holding S selected keys
  // passes about 1 - exp(-S / min(bits, ht_len)) of uniform fact keys, so
  // S >= -ln(4/5) min(bits, ht_len).  Known here for a resident dimension;
  // the probe checks it on the build's counters for one built with the query.
  if (lip && nd == 1 && q->n_groups == 1 && c->direct1) {
    const lj_dim* D = p.dims[0];
    const int32_t k = D->desc.ht_kind;
    const ProbeDim& P = a.dim[0];
    if ((k == LJ_HT_PHT1 || k == LJ_HT_PHT2) && (P.bh.kind == BLOOM_MOD || P.bh.kind == BLOOM_MOD_POW2)) {
      const double dom = (double)std::min<uint32_t>(P.bh.mod.d, P.ht_len.d);
      const int64_t min_sel = c->direct1 >= 2 ? 0 : (int64_t)std::ceil(0.2231 * dom);
      if (q->resident[0]) {
        a.direct1 = D->n_selected >= min_sel;
      } else {
        a.direct1 = 1;
        a.direct_min_sel = (int32_t)min_sel;
      }
    }
  }
  // Whole filters, or their prefixes (filter_prefix_words): prefixes free
  // LDS for a larger group table when the group space is small enough for
  // every group to be hit (at most kRepGroups), and let more filters be
  // staged when the whole ones do not all fit.  Otherwise whole filters: no
  // prefix check at the probe's start, the small group table (a large one
  // costs a 437,500-group space with 24 groups hit more than it saves).
  const bool small_space = q->n_groups > 1 && q->n_groups <= kRepGroups;
  bool use_prefix = small_space;
  if (lip && !use_prefix) {
    int64_t need = 0;
    for (int d = 0; d < nd; d++)
      if (a.dim[d].bh.kind != BLOOM_NONE && a.dim[d].bloom) need += (a.dim[d].bloom_words + 3) & ~3ll;
    use_prefix = need > lds_budget_words;
  }
  if (lip && use_prefix) {
    // a resident dimension with bits past its prefix stages the whole
    // filter; one built with the query passes its counter to the probe
    // (which checks it once, at its start)
    for (int d = 0; d < nd; d++) {
      ProbeDim& P = a.dim[d];
      lj_dim* D = p.dims[d];
      int64_t pw = filter_prefix_words(c, D);
      if (pw < D->bloom_words) {
        if (q->resident[d] && D->n_past != 0) pw = D->bloom_words;
        else if (!q->resident[d]) {
          P.past_count = D->counters + kCounterPast;
          a.check_past = 1;
        }
      }
      P.bloom_lds_words = (int32_t)pw;
    }
  }
  if (lip) {
    // Filters staged in LDS (their staged prefix), in load order.  Policy 0:
    // every filter that fits.  Policy 1: the stage-0 filter only.  Policy 2:
    // a static plan's stage-0 filter and every filter of at most
    // kHotFilterWords, larger ones later in the chain probed from L2 (the
    // adaptive order stages the small filters, or the largest that fits if
    // none is small).  Measured on Q3.3 and SSB flights 2-4: policy 0 is the
    // fastest (a 128 KiB filter probed from L2 by a fifth of the rows costs
    // more than the group table its LDS would hold saves).
    for (int st = 0; st < nd; st++) {
      const int d = adaptive ? st : p.bloom_order[st];
      ProbeDim& P = a.dim[d];
      if (P.bh.kind == BLOOM_NONE || !P.bloom) continue;
      const int64_t w4 = (P.bloom_lds_words + 3) & ~3ll;
      if (lds_words + w4 > lds_budget_words) continue;
      if (lds_policy == 1 && !adaptive && st > 0) continue;
      if (lds_policy == 2 && w4 > kHotFilterWords && (adaptive || st > 0)) continue;
      P.bloom_lds = lds_words;
      lds_words += (int)w4;
    }
    if (adaptive && lds_words == 0) {
      int best = -1;
      for (int d = 0; d < nd; d++) {
        const ProbeDim& P = a.dim[d];
        const int64_t w4 = (P.bloom_lds_words + 3) & ~3ll;
        if (P.bh.kind != BLOOM_NONE && P.bloom && w4 <= lds_budget_words &&
            (best < 0 || w4 > ((a.dim[best].bloom_lds_words + 3) & ~3ll)))
          best = d;
      }
      if (best >= 0) {
        a.dim[best].bloom_lds = 0;
        lds_words = (int)((a.dim[best].bloom_lds_words + 3) & ~3ll);
      }
    }
  }
  // A one-dimension semi-join through an exact filter (P.exact) that does
  // not fit in LDS: its first words, as many as the LDS holds, are staged,
  // and the probe reads a word past them from the global filter (the bodies
  // that fall back to global words: all_k0 is cleared below).  Such a probe
  // is bound by the rate of L2 requests of its filter reads, and the staged
  // share of them leaves that stream (2^28 fact rows over a 2^22-bit filter:
  // 1.38 -> 0.99 ms).  A join that probes the table for each survivor gains
  // while few rows pass (10 %: 1.58 -> 1.41 ms) and loses once the gathers
  // bound it (50 %: 2.42 -> 2.91 ms; the global reads past the prefix stall
  // the streamed columns' loads), and its pass rate is not known when it is
  // planned: it keeps the whole filter in global memory.  Only
  // when the fact rows repay each workgroup's staging copy kPrefixRepay
  // times over: the probes moved to LDS, n * staged / words, against the
  // copy, a 128-byte line per 32 staged words on each of about n_cu
  // workgroups.
  bool partial = false;
  if (lip && !adaptive && nd == 1 && c->partial_prefix && lds_policy == 0 && a.dim[0].exact) {
    ProbeDim& P = a.dim[0];
    const bool mod = P.bh.kind == BLOOM_MOD || P.bh.kind == BLOOM_MOD_POW2;
    const int64_t pw = lds_budget_words & ~3ll;
    constexpr double kPrefixRepay = 4.0;
    if (mod && P.bloom && P.bloom_lds < 0 && pw >= 1024 && pw < P.bloom_words &&
        (double)q->n_fact >= kPrefixRepay * (double)c->n_cu * (double)P.bloom_words / 32.0) {
      P.bloom_lds = 0;
      P.bloom_lds_words = (int32_t)pw;
      lds_words = (int)pw;
      partial = true;
    }
  }
  // stage-0 kernel specialisation: the kind of the first filter (every
  // candidate's kind for the adaptive order; GENERIC when they differ)
  auto kind_of = [&](int d) -> int {
    const ProbeDim& P = a.dim[d];
    const bool l = P.bloom_lds >= 0;
    switch (P.bh.kind) {
      case BLOOM_NONE: return S0_NONE;
      case BLOOM_MOD_POW2: return l ? S0_MODP2_LDS : S0_MODP2_GL;
      case BLOOM_MOD: return l ? S0_MOD_LDS : S0_MOD_GL;
      default: return l ? S0_MULT_LDS : S0_MULT_GL;
    }
  };
  if (!lip) {
    cfg.k0 = S0_PHT;
  } else if (!adaptive) {
    cfg.k0 = kind_of(p.bloom_order[0]);
  } else {
    cfg.k0 = kind_of(0);
    for (int d = 1; d < nd; d++)
      if (kind_of(d) != cfg.k0) cfg.k0 = S0_GENERIC;
    if (cfg.k0 == S0_GENERIC) {
      // filters of mixed kinds or placements (some do not fit in LDS): stage 0
      // is kept to the staged filters when they share one kind, so the kernel
      // stays specialised; the sample orders those first, the others after
      int k = -1;
      bool one = true;
      for (int d = 0; d < nd; d++)
        if (a.dim[d].bloom_lds >= 0) {
          if (k < 0) k = kind_of(d);
          else if (kind_of(d) != k) one = false;
        }
      if (k >= 0 && one) {
        cfg.k0 = k;
        for (int d = 0; d < nd; d++)
          if (kind_of(d) == k) a.s0_mask |= 1u << d;
      }
    }
  }
  // every filter of stage 0's hash family (in LDS or not): the kernel
  // probes later stages with that family's short test
  auto family = [](int k) { return k == S0_MODP2_LDS || k == S0_MODP2_GL ? 1 : k == S0_MOD_LDS || k == S0_MOD_GL ? 2
                                   : k == S0_MULT_LDS || k == S0_MULT_GL ? 3 : 0; };
  a.all_k0 = lip && family(cfg.k0) != 0 ? 1 : 0;
  for (int d = 0; d < nd; d++)
    if (lip && family(kind_of(d)) != family(cfg.k0)) a.all_k0 = 0;
  if (partial) a.all_k0 = 0;
  // A table probed in place (direct1, behind a modulo filter: the in-place
  // body) whose occupancy bitmap fits in the LDS left: each workgroup builds
  // the bitmap from the table at its start (ht_len slots read from L2), and
  // the probe then decides hit or miss with an LDS read and gathers only the
  // hits' payloads.  Repaid when the rows passing the filter, n * pass, are
  // at least kOccRepay x CUs x slots (about 8 bytes read per slot by each
  // workgroup, against a slot gather saved per passing row; measured: 2^25
  // rows over 2^18 slots, a 0.5 ratio of rows to CUs x slots, gain at every
  // pass rate from 35 % up).  The pass
  // rate is estimated from the selected keys S as for direct1: known here
  // for a resident dimension, checked by the probe on the build's counters
  // (occ_min_sel) for one built with the query.
  if (a.direct1 && !partial && c->direct_occ > 0 &&
      (a.dim[0].bh.kind == BLOOM_MOD || a.dim[0].bh.kind == BLOOM_MOD_POW2)) {
    constexpr double kOccRepay = 0.125;
    const ProbeDim& P = a.dim[0];
    const int64_t ns = P.ht_len.d;
    const int64_t ow = (((ns + 63) / 64) * 2 + 3) & ~3ll;
    const double pass_min = kOccRepay * (double)c->n_cu * (double)ns / std::max<double>(1.0, (double)q->n_fact);
    const double dom = (double)std::min<uint32_t>(P.bh.mod.d, P.ht_len.d);
    const int64_t min_sel = c->direct_occ >= 2 ? 0 : pass_min < 1.0 ? (int64_t)std::ceil(-std::log1p(-pass_min) * dom)
                                                                     : INT64_MAX;
    bool want = min_sel < INT64_MAX;
    if (want && q->resident[0]) want = p.dims[0]->n_selected >= min_sel;
    if (want && lds_words + ow <= lds_budget_words) {
      a.occ_lds = lds_words;
      a.occ_words = (int32_t)ow;
      a.occ_min_sel = q->resident[0] ? 0 : (int32_t)std::min<int64_t>(min_sel, INT32_MAX);
      lds_words += (int)ow;
    }
  }
  // Without the bitmap, probing in place pays where the slot gathers hit L2
  // (a 4 MB table: 6.8 -> 4.6 ms); against a 32 MB table, gathered from the
  // MALL, the queue's gathers issued a batch ahead are 1-3 % faster.
  constexpr int64_t kDirectTableBytes = 8ll << 20;
  if (a.direct1 && a.occ_words == 0 && c->direct1 < 2 && p.dims[0]->ht_words * 4 > kDirectTableBytes)
    a.direct1 = 0;
  // the LDS group table: 32 slots, or for a small group space up to twice
  // its size (at most kGroupSlotsMax) in the LDS that staging leaves.  A
  // 1024-thread workgroup is alone on its CU (the probe's registers allow 16
  // waves per CU), so the rest of the CU's LDS is free; smaller workgroups
  // keep the small table.
  int gslots = kGroupSlots;
  const int env_slots = c->group_slots;
  if ((small_space && cfg.wg == 1024) || (q->n_groups > 1 && env_slots > 0)) {
    int64_t want = std::min<int64_t>(kGroupSlotsMax, 2 * q->n_groups);
    if (env_slots > 0) want = env_slots;
    const int64_t left = (int64_t)c->max_lds - (int64_t)probe_lds_bytes(cfg, lds_words, 0) - 16;
    while ((int64_t)gslots * 2 <= want && (int64_t)gslots * 2 * kGroupSlotBytes <= left) gslots *= 2;
  }
  a.gslot_bits = ilog2_u64((uint64_t)gslots);
  p.lds_bytes = probe_lds_bytes(cfg, lds_words, gslots);
  lj_status s = occupancy_of(c, cfg, p.lds_bytes, &p.per_cu, cache_occupancy);
  if (s != LJ_OK) return s;
  a.measure = q->fact_measure;
  a.measure_sub = q->fact_measure_sub;
  a.n = q->n_fact;
  a.n_dims = nd;
  a.method = lip ? 1 : 0;
  a.adaptive = adaptive ? 1 : 0;
  a.want_product = want_product;
  a.k0 = cfg.k0;
  a.exact_eager_div = c->exact_eager_div;
  a.direct_eager_div = c->direct_eager_div;
  a.occ_eager_div = c->occ_eager_div;
  for (int i = 0; i < LJ_MAX_DIMS; i++) {
    a.order[i] = i < nd ? q->order[i] : i;
    a.load_order[i] = i < nd ? (lip ? p.bloom_order[i] : q->order[i]) : i;
  }
  if (a.s0_mask) {
    // the order a wave starts from (and keeps if it samples no row): the
    // dimensions allowed at stage 0 first, each group in its given order
    int32_t lo[LJ_MAX_DIMS];
    int k = 0;
    for (int pass = 0; pass < 2; pass++)
      for (int i = 0; i < nd; i++) {
        const bool ok = (a.s0_mask >> a.load_order[i]) & 1u;
        if (ok == (pass == 0)) lo[k++] = a.load_order[i];
      }
    for (int i = 0; i < nd; i++) a.load_order[i] = lo[i];
  }
  a.lds_filter_words = lds_words;
  a.goffset = q->group_offset;
  a.n_groups = q->n_groups;
  a.gmod = make_fastmod((uint32_t)q->n_groups);
  for (int d = 0; d < nd; d++) a.dim_counters[d] = q->resident[d] ? nullptr : p.dims[d]->counters;
  return LJ_OK;
}

// Parse one query's result region (header + groups, host copy).
static lj_status parse_result(lj_ctx* c, const lj_star_desc* q, const Plan& p, const long long* hdr,
                              const GroupOut* groups, int64_t groups_avail, QueryOut& qo) {
  const int nd = p.nd;
  const int64_t n_touched = hdr[0];
  memcpy(qo.stats, hdr + 1, sizeof qo.stats);
  qo.groups.assign(groups, groups + std::min<int64_t>(n_touched, groups_avail));
  std::sort(qo.groups.begin(), qo.groups.end(), [](const GroupOut& x, const GroupOut& y) { return x.g < y.g; });
  int64_t build_oob = 0;
  for (int d = 0; d < nd; d++) {
    if (q->resident[d]) {
      qo.dim_selected[d] = p.dims[d]->n_selected;
      qo.dim_oob[d] = p.dims[d]->n_oob;
    } else {
      qo.dim_selected[d] = (int64_t)hdr[1 + kStats + 2 * d];
      qo.dim_oob[d] = (int64_t)hdr[2 + kStats + 2 * d];
    }
    build_oob += qo.dim_oob[d];
  }
  if (build_oob > 0 || qo.stats[ST_OOB] > 0)
    return fail(c, LJ_EDOMAIN, "%lld dimension keys and %lld fact keys/groups fell outside their hash domains",
                (long long)build_oob, (long long)qo.stats[ST_OOB]);
  return LJ_OK;
}

// The common path of lj_star_join and lj_hash_join.  With fill_in_window
// (lj_hash_join) the query's tables are cleared inside its own timed window,
// as the reference's hashJoin times its memsets (bloom_join.cu:135-137), and
// that clear is t_fill; otherwise they come from the zero pool, cleared by a
// fill queued behind the query that last used them (t_fill = 0).
static lj_status run_query(lj_ctx* c, const lj_star_desc* q, int want_product, QueryOut& qo,
                           bool fill_in_window = false) {
  Plan p;
  lj_status s = validate_query(c, q, p);
  if (s != LJ_OK) return s;
  HIP_TRY(c, hipSetDevice(c->device));
  if ((s = acquire_dims(c, q, p, !fill_in_window)) != LJ_OK) return s;
  auto cleanup = [&]() {
    dims_free(c, p.owned, p.n_owned);
    p.n_owned = 0;
  };
  if ((s = ensure_groups(c, q->n_groups)) != LJ_OK) {
    cleanup();
    return s;
  }
  // ---- two launches: the build of the query's own dimensions, then the probe
  // (which also collects the result groups in its last workgroup).  The build
  // goes out first; the probe is planned while it runs.
  const bool built_here = p.n_owned > 0;
  HIP_TRY(c, hipEventRecord(c->ev[0], c->stream));
  const bool timed_fill = fill_in_window && built_here;
  if (timed_fill) {
    FillArgs f{};
    for (int i = 0; i < p.n_owned; i++) {
      add_fill(f, p.owned[i]->ht, p.owned[i]->ht_words);
      add_fill(f, p.owned[i]->bloom, p.owned[i]->bloom_words);
      add_fill(f, p.owned[i]->counters, kCounterFillWords);
    }
    HIP_TRY(c, launch_fill(f, c->stream));
    HIP_TRY(c, hipEventRecord(c->ev[2], c->stream));
  }
  if (built_here) {
    if ((s = dims_launch(c, p.owned, p.n_owned, c->out_hdr + kClockSlot)) != LJ_OK) {
      cleanup();
      return s;
    }
  }
  if ((s = plan_probe(c, q, want_product, 0, p)) != LJ_OK) {
    cleanup();
    return s;
  }
  const ProbeConfig& cfg = p.cfg;
  ProbeArgs& a = p.a;
  const int64_t tile_rows = (int64_t)cfg.vec * kSegRows;
  const int nw = cfg.wg / 64;
  int64_t grid = (int64_t)c->n_cu * p.per_cu;
  const int64_t need = (q->n_fact + tile_rows * nw - 1) / (tile_rows * nw);
  if (grid > need) grid = need;
  if (grid < 1) grid = 1;
  if ((s = ensure_stats(c, (int)grid)) != LJ_OK) {
    cleanup();
    return s;
  }
  a.gsum = c->gsum;
  a.gcount = c->gcount;
  a.gprod = c->gprod;
  a.gpay = c->gpay;
  if (c->g_reps > 1) {
    // the allocation holds g_reps copies of g_cap >= n_groups entries
    a.g_reps = c->g_reps;
    a.g_stride = c->g_cap;
    a.gseen = c->gseen;
  }
  a.touched = c->touched;
  a.n_touched = c->n_touched;
  a.touched_cap = c->touched_cap;
  a.stats = c->stats;
  a.done = c->n_touched + 32;
  a.out_hdr = c->out_hdr;
  a.out = c->out;
  a.out_cap = c->touched_cap;
  // the header and the first groups go straight to the pinned host buffer
  const int64_t first = std::min<int64_t>(kFirstGroups, c->touched_cap);
  a.host_hdr = static_cast<long long*>(c->host_buf_dev);
  a.host_out = reinterpret_cast<GroupOut*>(static_cast<char*>(c->host_buf_dev) + kHdrBytes);
  a.host_groups = first;

  HIP_TRY(c, hipEventRecord(c->ev[3], c->stream));
  HIP_TRY(c, launch_probe(cfg, a, (int)grid, p.lds_bytes, c->stream));
  HIP_TRY(c, hipEventRecord(c->ev[kEvDone], c->stream));

  // ---- results: the header (groups hit, statistics, dimension counters) and
  // the first groups are in the pinned host buffer once the probe is done
  long long* hdr = static_cast<long long*>(c->host_buf);
  GroupOut* hg = reinterpret_cast<GroupOut*>(static_cast<char*>(c->host_buf) + kHdrBytes);
  // the built tables go back to the zero pool now: their clears queue behind the probe
  cleanup();
  HIP_TRY(c, spin_wait(c->ev[kEvDone]));
  const int64_t n_touched = hdr[0];
  if (n_touched > (int64_t)c->touched_cap) {
    // the finalize zeroed only the listed groups: clear the rest so that the
    // next query on this context starts from all-zero accumulators
    FillArgs f{};
    add_fill(f, c->gsum, c->g_cap * c->g_reps * 2);
    add_fill(f, c->gcount, c->g_cap * c->g_reps * 2);
    add_fill(f, c->gprod, c->g_cap * c->g_reps * 2);
    if (c->g_reps > 1) add_fill(f, c->gseen, c->g_cap);
    HIP_TRY(c, launch_fill(f, c->stream));
    return fail(c, LJ_EUNSUPPORTED, "%lld groups were hit; at most %u can be listed", (long long)n_touched, c->touched_cap);
  }
  std::vector<GroupOut> big;
  const GroupOut* groups = hg;
  if (n_touched > first) {
    if (n_touched <= kHostGroups) {
      HIP_TRY(c, hipMemcpyAsync(hg + first, c->out + first, sizeof(GroupOut) * (size_t)(n_touched - first),
                                hipMemcpyDeviceToHost, c->stream));
      HIP_TRY(c, hipStreamSynchronize(c->stream));
    } else {
      big.resize((size_t)n_touched);
      HIP_TRY(c, hipMemcpy(big.data(), c->out, sizeof(GroupOut) * (size_t)n_touched, hipMemcpyDeviceToHost));
      groups = big.data();
    }
  }
  // phases from HIP events on the query's stream: [fill] [build launch] [probe launch]
  qo.t_fill = 0.f;
  if (timed_fill) {
    HIP_TRY(c, hipEventElapsedTime(&qo.t_fill, c->ev[0], c->ev[2]));
    HIP_TRY(c, hipEventElapsedTime(&qo.t_build, c->ev[2], c->ev[3]));
  } else {
    HIP_TRY(c, hipEventElapsedTime(&qo.t_build, c->ev[0], c->ev[3]));
  }
  HIP_TRY(c, hipEventElapsedTime(&qo.t_probe, c->ev[3], c->ev[1]));
  HIP_TRY(c, hipEventElapsedTime(&qo.t_total, c->ev[0], c->ev[1]));
  if (!built_here) qo.t_build = 0.f;
  return parse_result(c, q, p, hdr, groups, n_touched, qo);
}

static void fill_stats(const QueryOut& qo, int nd, lj_probe_stats* ps) {
  memset(ps, 0, sizeof *ps);
  for (int s = 0; s <= LJ_MAX_DIMS; s++) {
    ps->rows_in[s] = qo.stats[ST_ROWS_IN + s];
    ps->lines_in[s] = qo.stats[ST_LINES_IN + s];
  }
  ps->rows_bloom_pass = qo.stats[ST_BLOOM_PASS];
  ps->sample_rows = qo.stats[ST_SAMPLE_ROWS];
  ps->sample_lines = qo.stats[ST_SAMPLE_LINES];
  ps->n_oob = qo.stats[ST_OOB];
  for (int s = 0; s < LJ_MAX_DIMS; s++) ps->order_used[s] = s < nd ? (int32_t)((qo.stats[ST_ORDER] >> (8 * s)) & 0xff) : -1;
}

// The caller's result from n groups in ascending g and the statistics.
static void groups_to_result(const lj_star_desc* q, const GroupOut* groups, size_t n, lj_star_result* out) {
  out->n_groups = (int64_t)n;
  out->n_rows = 0;
  out->sum_total = 0;
  for (size_t i = 0; i < n; i++) {
    const GroupOut& g = groups[i];
    out->n_rows += (int64_t)g.count;
    out->sum_total += g.sum;
    if ((int64_t)i < out->groups_cap && out->groups) {
      lj_group& o = out->groups[i];
      o.g = g.g;
      for (int d = 0; d < LJ_MAX_DIMS; d++) o.payload[d] = d < q->n_dims ? g.pay[d] : 0;
      o.sum = g.sum;
      o.count = (int64_t)g.count;
      o.sum_i32 = (int32_t)(uint32_t)(uint64_t)g.sum;
      o.reserved0 = 0;
    }
  }
}

static void to_star_result(const lj_star_desc* q, const QueryOut& qo, lj_star_result* out) {
  groups_to_result(q, qo.groups.data(), qo.groups.size(), out);
  out->time_query_ms = qo.t_total;
  out->time_build_ms = qo.t_fill + qo.t_build;
  out->time_probe_ms = qo.t_probe;
  out->status = LJ_OK;
  fill_stats(qo, q->n_dims, &out->stats);
}

extern "C" lj_status lj_star_join(lj_ctx* c, const lj_star_desc* q, lj_star_result* out) {
  if (!c || !q || !out) return fail(c, LJ_EINVAL, "NULL argument");
  QueryOut qo;
  lj_status s = run_query(c, q, 0, qo);
  if (s != LJ_OK && s != LJ_EDOMAIN) {
    out->status = s;
    return s;
  }
  to_star_result(q, qo, out);
  out->status = s;
  return s;
}

// lj_star_join over resident dimensions with no allocation: the probe's
// buffers are this device's `scratch` (its address from lj_create), the
// result comes through the context's pinned buffer, the groups are sorted in
// place there.  Nothing here allocates device or host memory.
extern "C" lj_status lj_star_join_noalloc(lj_ctx* c, const lj_star_desc* q, lj_star_result* out) {
  if (!c || !q || !out) return fail(c, LJ_EINVAL, "NULL argument");
  Plan p;
  lj_status s = validate_query(c, q, p);
  if (s != LJ_OK) return s;
  for (int d = 0; d < q->n_dims; d++)
    if (!q->resident[d])
      return fail(c, LJ_EINVAL, "lj_star_join_noalloc needs resident dimensions (dimension %d is not)", d);
  if (q->n_groups > kNaGroups)
    return fail(c, LJ_EINVAL, "n_groups %lld exceeds the no-allocation limit of %lld", (long long)q->n_groups,
                (long long)kNaGroups);
  char* base = c->device < kNaMaxDevices ? g_na_ptr[c->device] : nullptr;
  if (!base) return fail(c, LJ_ESTATE, "no scratch was set up for device %d", c->device);
  HIP_TRY(c, hipSetDevice(c->device));
  if ((s = acquire_dims(c, q, p)) != LJ_OK) return s;  // resident dimensions only: nothing allocated
  if ((s = plan_probe(c, q, 0, 0, p, false)) != LJ_OK) return s;
  const ProbeConfig& cfg = p.cfg;
  ProbeArgs& a = p.a;
  qbuf_bind(a, kNaGroups, kNaListed, base, base + kNaZeroBytes, base + kNaZeroBytes + kNaPlainBytes);
  a.host_hdr = static_cast<long long*>(c->host_buf_dev);
  a.host_out = reinterpret_cast<GroupOut*>(static_cast<char*>(c->host_buf_dev) + kHdrBytes);
  a.host_groups = kNaListed;
  const int64_t tile = (int64_t)cfg.vec * kSegRows * (cfg.wg / 64);
  const int64_t need = std::max<int64_t>(1, (q->n_fact + tile - 1) / tile);
  const int64_t grid = std::min<int64_t>({(int64_t)c->n_cu * p.per_cu, need, kNaRows});
  HIP_TRY(c, hipEventRecord(c->ev[3], c->stream));
  HIP_TRY(c, launch_probe(cfg, a, (int)grid, p.lds_bytes, c->stream));
  HIP_TRY(c, hipEventRecord(c->ev[kEvDone], c->stream));
  HIP_TRY(c, spin_wait(c->ev[kEvDone]));
  const long long* hdr = static_cast<const long long*>(c->host_buf);
  GroupOut* hg = reinterpret_cast<GroupOut*>(static_cast<char*>(c->host_buf) + kHdrBytes);
  const int64_t n_touched = hdr[0];
  if (n_touched > kNaListed) {
    // the finalize zeroed only the listed groups: clear the whole region, and
    // wait for the clear, so that the next call (on any stream) finds it zero
    FillArgs f{};
    f.r[0].ptr = reinterpret_cast<uint32_t*>(base);
    f.r[0].words = (int64_t)(kNaZeroBytes / 4);
    f.n = 1;
    HIP_TRY(c, launch_fill(f, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return fail(c, LJ_EUNSUPPORTED, "%lld groups were hit; the no-allocation path lists at most %lld",
                (long long)n_touched, (long long)kNaListed);
  }
  std::sort(hg, hg + n_touched, [](const GroupOut& x, const GroupOut& y) { return x.g < y.g; });
  groups_to_result(q, hg, (size_t)n_touched, out);
  QueryOut qo;  // statistics and timings only (its group vector stays empty)
  memcpy(qo.stats, hdr + 1, sizeof qo.stats);
  HIP_TRY(c, hipEventElapsedTime(&qo.t_probe, c->ev[3], c->ev[kEvDone]));
  out->time_query_ms = qo.t_probe;
  out->time_build_ms = 0.f;
  out->time_probe_ms = qo.t_probe;
  fill_stats(qo, q->n_dims, &out->stats);
  int64_t build_oob = 0;
  for (int d = 0; d < q->n_dims; d++) build_oob += p.dims[d]->n_oob;
  s = LJ_OK;
  if (build_oob > 0 || qo.stats[ST_OOB] > 0)
    s = fail(c, LJ_EDOMAIN, "%lld dimension keys and %lld fact keys/groups fell outside their hash domains",
             (long long)build_oob, (long long)qo.stats[ST_OOB]);
  out->status = s;
  return s;
}

extern "C" lj_status lj_hash_join(lj_ctx* c, const lj_join_desc* j, lj_join_result* out) {
  if (!c || !j || !out) return fail(c, LJ_EINVAL, "NULL argument");
  if (j->method != LJ_METHOD_NAIVE && j->method != LJ_METHOD_LIP && j->method != LJ_METHOD_LIP_ADAPTIVE)
    return fail(c, LJ_EINVAL, "unknown method %d", j->method);
  lj_star_desc q;
  memset(&q, 0, sizeof q);
  q.fact_fk[0] = j->fact_fk;
  q.fact_measure = j->fact_measure;
  q.n_fact = j->n_fact;
  q.n_dims = 1;
  q.method = j->method;
  q.order[0] = 0;
  for (int i = 0; i < LJ_MAX_DIMS; i++) q.bloom_order[i] = -1;
  q.dims[0] = j->dim;
  q.n_groups = 1;
  if (j->flags & ~LJ_JOIN_SUM_ONLY) return fail(c, LJ_EINVAL, "unknown flags 0x%x", (unsigned)j->flags);
  QueryOut qo;
  lj_status s = run_query(c, &q, (j->flags & LJ_JOIN_SUM_ONLY) ? 0 : 1, qo, true);
  if (s != LJ_OK && s != LJ_EDOMAIN) return s;
  memset(out, 0, sizeof *out);
  out->n_selected = qo.dim_selected[0];
  if (!qo.groups.empty()) {
    out->n_match = (int64_t)qo.groups[0].count;
    out->sum_measure = qo.groups[0].sum;
    out->sum_product = qo.groups[0].prod;
  }
  out->n_bloom_pass = j->method == LJ_METHOD_NAIVE ? j->n_fact : qo.stats[ST_BLOOM_PASS];
  out->time_build = qo.t_build;
  out->time_probe = qo.t_probe;
  out->time_extra = qo.t_fill;
  out->time_total = qo.t_total;
  fill_stats(qo, 1, &out->stats);
  return s;
}

// ---------------------------------------------------------------- generators --
extern "C" lj_status lj_gen_host(int32_t kind, uint64_t seed, uint64_t stream, int64_t lo, int64_t hi, int64_t first,
                                 int64_t n, int32_t* out) {
  if (n < 0 || (n > 0 && !out) || kind < 0 || kind > 4) return LJ_EINVAL;
  if (kind == GEN_UNIFORM && (hi < lo || hi - lo + 1 > (1ll << 32))) return LJ_EINVAL;
  if (kind == GEN_DATEKEY && (hi < 1 || hi > (1ll << 32))) return LJ_EINVAL;
  const uint64_t key = stream_key(seed, stream);
  unsigned nt = std::thread::hardware_concurrency();
  if (nt == 0) nt = 1;
  if (nt > 32) nt = 32;
  if (n < (1 << 16)) nt = 1;
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; t++) {
    const int64_t b = n * t / nt, e = n * (t + 1) / nt;
    auto work = [=]() {
      for (int64_t i = b; i < e; i++) out[i] = gen_value(kind, key, lo, hi, first + i);
    };
    if (nt == 1) work();
    else th.emplace_back(work);
  }
  for (auto& x : th) x.join();
  return LJ_OK;
}

extern "C" lj_status lj_gen_device(lj_ctx* c, int32_t kind, uint64_t seed, uint64_t stream, int64_t lo, int64_t hi,
                                   int64_t first, int64_t n, int32_t* out_dev) {
  if (!c) return LJ_EINVAL;
  if (n < 0 || (n > 0 && !out_dev) || kind < 0 || kind > 4) return fail(c, LJ_EINVAL, "bad generator arguments");
  if (kind == GEN_UNIFORM && (hi < lo || hi - lo + 1 > (1ll << 32))) return fail(c, LJ_EINVAL, "bad uniform range");
  if (kind == GEN_DATEKEY && (hi < 1 || hi > (1ll << 32))) return fail(c, LJ_EINVAL, "bad day range");
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, launch_gen(kind, stream_key(seed, stream), lo, hi, first, n, out_dev, c->stream));
  return LJ_OK;
}

extern "C" lj_status lj_gen_shuffle_host(uint64_t seed, uint64_t stream, int64_t n, int32_t* out) {
  if (n < 0 || (n > 0 && !out)) return LJ_EINVAL;
  const uint64_t key = stream_key(seed, stream);
  // Durstenfeld's Fisher-Yates: swap out[i] with out[j], j uniform in [0, i]
  for (int64_t i = n - 1; i > 0; i--) {
    const int64_t j = (int64_t)bounded_at(key, (uint64_t)i, (uint64_t)(i + 1));
    const int32_t t = out[i];
    out[i] = out[j];
    out[j] = t;
  }
  return LJ_OK;
}

// RAND_RANGE(N) of generator.h:32 over the C library's rand(), truncated to
// int as `int j = RAND_RANGE(i)` does (generator.h:56).
static int32_t ref_rand_range(int64_t n) { return (int32_t)((double)rand() / ((double)RAND_MAX + 1) * (double)n); }

// random_unique_gen (generator.h:69-79): 1..n, then knuth_shuffle (:51-60).
static void ref_unique_shuffled(int32_t* a, int64_t n, std::vector<int32_t>& scratch) {
  int32_t* out = a;
  if (!out) {
    scratch.resize((size_t)n);
    out = scratch.data();
  }
  for (int64_t i = 0; i < n; i++) out[i] = (int32_t)(i + 1);
  for (int64_t i = n - 1; i > 0; i--) {
    const int32_t j = ref_rand_range(i);
    const int32_t t = out[i];
    out[i] = out[j];
    out[j] = t;
  }
}

extern "C" lj_status lj_gen_reference_host(int64_t n_dim, int64_t n_fact, int32_t* dim_keys, int32_t* dim_vals,
                                           int32_t* fact_keys, int32_t* fact_vals) {
  if (n_dim < 1 || n_dim >= (1ll << 31) || n_fact < 0 || n_fact >= (1ll << 31))
    return fail(nullptr, LJ_EINVAL, "n_dim %lld / n_fact %lld out of range", (long long)n_dim, (long long)n_fact);
  std::vector<int32_t> scratch;
  srand(0);  // check_seed (generator.h:38-45)
  // create_relation_pk (generator.h:88-106)
  ref_unique_shuffled(dim_keys, n_dim, scratch);
  if (dim_vals)
    for (int64_t i = 0; i < n_dim; i++) dim_vals[i] = (int32_t)i;
  // create_relation_fk (generator.h:108-136)
  const int64_t iters = n_fact / n_dim, rem = n_fact % n_dim;
  for (int64_t i = 0; i < iters; i++) ref_unique_shuffled(fact_keys ? fact_keys + n_dim * i : nullptr, n_dim, scratch);
  if (rem > 0) ref_unique_shuffled(fact_keys ? fact_keys + n_dim * iters : nullptr, rem, scratch);
  if (fact_vals)
    for (int64_t i = 0; i < n_fact; i++) fact_vals[i] = (int32_t)i;
  return LJ_OK;
}

// The reference's star-schema inputs (src/ops/utils/star.cuh:67-120,
// star_table_init_host): srand(0); the fact table's create_relation_pk (keys
// 1..n_fact shuffled, values 0..n_fact-1); then per dimension, in order, its
// create_relation_pk (keys 1..n_rows shuffled, values 0..n_rows-1) and the
// fact's foreign keys by create_relation_fk_only (generator.h:139-165:
// successive shuffled blocks of 1..n_rows and a shuffled tail).
extern "C" lj_status lj_gen_reference_star_host(int64_t n_fact, int32_t n_dims, const int64_t* dim_rows,
                                                int32_t* fact_keys, int32_t* fact_vals, int32_t* const* dim_keys,
                                                int32_t* const* dim_vals, int32_t* const* fact_fks) {
  if (n_fact < 0 || n_fact >= (1ll << 31) || n_dims < 0 || n_dims > LJ_MAX_DIMS || (n_dims > 0 && !dim_rows))
    return fail(nullptr, LJ_EINVAL, "n_fact %lld / n_dims %d out of range", (long long)n_fact, n_dims);
  for (int d = 0; d < n_dims; d++)
    if (dim_rows[d] < 1 || dim_rows[d] >= (1ll << 31))
      return fail(nullptr, LJ_EINVAL, "dimension %d: %lld rows", d, (long long)dim_rows[d]);
  std::vector<int32_t> scratch;
  srand(0);  // check_seed (generator.h:38-45)
  ref_unique_shuffled(fact_keys, n_fact, scratch);
  if (fact_vals)
    for (int64_t i = 0; i < n_fact; i++) fact_vals[i] = (int32_t)i;
  for (int d = 0; d < n_dims; d++) {
    const int64_t nr = dim_rows[d];
    int32_t* dk = dim_keys ? dim_keys[d] : nullptr;
    int32_t* dv = dim_vals ? dim_vals[d] : nullptr;
    int32_t* fk = fact_fks ? fact_fks[d] : nullptr;
    ref_unique_shuffled(dk, nr, scratch);
    if (dv)
      for (int64_t i = 0; i < nr; i++) dv[i] = (int32_t)i;
    const int64_t iters = n_fact / nr, rem = n_fact % nr;
    for (int64_t i = 0; i < iters; i++) ref_unique_shuffled(fk ? fk + nr * i : nullptr, nr, scratch);
    if (rem > 0) ref_unique_shuffled(fk ? fk + nr * iters : nullptr, rem, scratch);
  }
  return LJ_OK;
}

// ------------------------------------------------------------------- batches --
// n queries in one build launch and one probe launch per stage-0 kernel kind;
// query q's workgroups are [wg_begin[q], wg_begin[q+1]) of its kind's grid.
namespace {
struct BatchQuery {
  lj_star_desc desc;
  Plan p;
  int64_t grid = 0;
  int64_t tcap = 0;
  size_t zero_off = 0;   // gsum | gcount | gprod | n_touched+done (kept zero between runs)
  size_t plain_off = 0;  // gpay | touched | stats
  size_t out_off = 0;    // header | groups
  int group = 0;
};
struct BatchKind {
  ProbeConfig cfg{};
  size_t lds = 0;
  int per_cu = 0;
  std::vector<int> qs;
  int64_t grid = 0;
  ProbeArgs* d_qs = nullptr;
  uint32_t* d_wgb = nullptr;
};
}  // namespace

struct lj_batch {
  lj_ctx* c = nullptr;
  lj_batch_opts opts{};
  std::vector<BatchQuery> q;
  std::vector<BatchKind> kinds;
  BuildDim* d_build = nullptr;
  int n_build = 0;
  int64_t build_blocks = 0;
  std::vector<FillArgs> clears;   // the built tables, cleared before every run but the first
  char* zero_buf = nullptr;
  char* plain_buf = nullptr;
  char* out_dev = nullptr;
  char* out_host = nullptr;
  size_t out_bytes = 0;
  void* args_buf = nullptr;       // ProbeArgs[] and wg_begin[] of every kind, BuildDim[]
  void* args_host = nullptr;      // pinned staging of args_buf
  size_t args_bytes = 0;
  hipStream_t stream = nullptr;
  bool own_stream = true;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int runs = 0;
};

static void batch_free(lj_batch* b) {
  if (!b) return;
  lj_ctx* c = b->c;
  if (b->stream && b->own_stream) (void)hipStreamSynchronize(b->stream);
  else if (b->stream && b->e1 && b->runs > 0) (void)hipEventSynchronize(b->e1);  // a shared stream: only this batch
  else if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (auto& bq : b->q) dims_free(c, bq.p.owned, bq.p.n_owned);
  if (b->zero_buf) {
    FillArgs f{};
    dev_release_zero(c, b->zero_buf, f);
    (void)launch_fill(f, c->stream);
  }
  dev_release(c, b->plain_buf);
  dev_release(c, b->out_dev);
  dev_release(c, b->args_buf);
  host_release(c, b->out_host, b->out_bytes);
  host_release(c, b->args_host, b->args_bytes);
  if (b->e0) (void)hipEventDestroy(b->e0);
  if (b->e1) (void)hipEventDestroy(b->e1);
  if (b->stream && b->own_stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

// Plans and allocates a batch and queues the upload of its arguments on
// `stream` (the batch's own stream when NULL); nothing waits for the device.
static lj_status batch_build(lj_ctx* c, const lj_star_desc* descs, int32_t n, const lj_batch_opts* opts,
                             hipStream_t stream, lj_batch** out) {
  *out = nullptr;
  if (n < 1 || n > (1 << 16)) return fail(c, LJ_EINVAL, "batch of %d queries (1..65536)", n);
  lj_batch* b = new lj_batch();
  b->c = c;
  if (stream) {
    b->stream = stream;
    b->own_stream = false;
  }
  if (opts) b->opts = *opts;
  if (b->opts.groups_cap <= 0) b->opts.groups_cap = 64;
  const int wg = b->opts.probe_wg;
  if (wg != 0 && wg != 256 && wg != 512 && wg != 1024) {
    delete b;
    return fail(c, LJ_EINVAL, "probe_wg %d (0, 256, 512 or 1024)", wg);
  }
  auto bail = [&](lj_status s) {
    batch_free(b);
    return s;
  };
  b->q.resize((size_t)n);
  // plan every query; group them by stage-0 kernel kind
  for (int i = 0; i < n; i++) {
    BatchQuery& bq = b->q[(size_t)i];
    bq.desc = descs[i];
    lj_status s = validate_query(c, &bq.desc, bq.p);
    if (s == LJ_OK) s = acquire_dims(c, &bq.desc, bq.p);
    if (s == LJ_OK) s = plan_probe(c, &bq.desc, 0, wg, bq.p);
    if (s != LJ_OK) {
      c->err = "query " + std::to_string(i) + ": " + c->err;
      return bail(s);
    }
    int k = -1;
    for (size_t j = 0; j < b->kinds.size(); j++)
      if (b->kinds[j].cfg.wg == bq.p.cfg.wg && b->kinds[j].cfg.vec == bq.p.cfg.vec && b->kinds[j].cfg.k0 == bq.p.cfg.k0)
        k = (int)j;
    if (k < 0) {
      b->kinds.push_back(BatchKind{});
      k = (int)b->kinds.size() - 1;
      b->kinds[(size_t)k].cfg = bq.p.cfg;
    }
    BatchKind& K = b->kinds[(size_t)k];
    K.lds = std::max(K.lds, bq.p.lds_bytes);
    K.qs.push_back(i);
    bq.group = k;
  }
  // queries of several stage-0 kinds at the default shape: one launch of the
  // mixed-kind probe (each workgroup runs its query's specialised body)
  if (b->kinds.size() > 1 && b->opts.launch_per_kind == 0) {
    bool one_shape = true;
    for (const BatchKind& K : b->kinds) one_shape = one_shape && K.cfg.wg == 1024 && K.cfg.vec == 1;
    if (one_shape) {
      BatchKind M{};
      M.cfg = ProbeConfig{1024, 1, kS0Mixed};
      for (int i = 0; i < n; i++) {
        M.lds = std::max(M.lds, b->q[(size_t)i].p.lds_bytes);
        M.qs.push_back(i);
        b->q[(size_t)i].group = 0;
      }
      b->kinds.clear();
      b->kinds.push_back(M);
    }
  }
  // workgroups: each query gets as many as one query alone would (the
  // resident slots of the kind, fewer if it has fewer tiles), in query order.
  // The dispatcher starts workgroups in order, so the queries run nearly one
  // after another -- each query's filters and hash table keep the caches to
  // themselves -- with the next query filling the previous one's tail and no
  // launch gap between them.
  for (BatchKind& K : b->kinds) {
    lj_status s = occupancy_of(c, K.cfg, K.lds, &K.per_cu);
    if (s != LJ_OK) return bail(s);
    const int64_t slots = (int64_t)c->n_cu * K.per_cu;
    const int64_t tile_rows = (int64_t)K.cfg.vec * kSegRows * (K.cfg.wg / 64);
    for (int i : K.qs) {
      BatchQuery& bq = b->q[(size_t)i];
      const int64_t need = std::max<int64_t>(1, (bq.desc.n_fact + tile_rows - 1) / tile_rows);
      bq.grid = std::min(slots, need);
      K.grid += bq.grid;
    }
    if (K.grid >= (1ll << 31)) return bail(fail(c, LJ_EUNSUPPORTED, "batch grid too large"));
  }
  // per-query buffers, carved from three allocations
  size_t zero_bytes = 0, plain_bytes = 0, out_bytes = 0;
  for (BatchQuery& bq : b->q) {
    const int64_t ng = bq.desc.n_groups;
    bq.tcap = std::min<int64_t>(ng, b->opts.groups_cap);
    const QBufLayout l = qbuf_layout(ng, bq.tcap, bq.grid);
    bq.zero_off = zero_bytes;
    zero_bytes += l.zero;
    bq.plain_off = plain_bytes;
    plain_bytes += l.plain;
    bq.out_off = out_bytes;
    out_bytes += l.out;
  }
  lj_status s;
  void* zp = nullptr;
  void* pp = nullptr;
  void* op = nullptr;
  if ((s = dev_alloc_zero(c, zero_bytes, &zp)) != LJ_OK) return bail(s);
  b->zero_buf = static_cast<char*>(zp);
  if ((s = dev_alloc(c, plain_bytes, &pp)) != LJ_OK) return bail(s);
  b->plain_buf = static_cast<char*>(pp);
  if ((s = dev_alloc(c, out_bytes, &op)) != LJ_OK) return bail(s);
  b->out_dev = static_cast<char*>(op);
  b->out_bytes = out_bytes;
  void* oh = nullptr;
  if ((s = host_alloc(c, out_bytes, &oh)) != LJ_OK) return bail(s);
  b->out_host = static_cast<char*>(oh);
  // the argument arrays of every kind and the build descriptors
  size_t args_bytes = 0;
  std::vector<size_t> qs_off(b->kinds.size()), wgb_off(b->kinds.size());
  for (size_t k = 0; k < b->kinds.size(); k++) {
    qs_off[k] = args_bytes;
    args_bytes += align256(sizeof(ProbeArgs) * b->kinds[k].qs.size());
    wgb_off[k] = args_bytes;
    args_bytes += align256(sizeof(uint32_t) * (b->kinds[k].qs.size() + 1));
  }
  std::vector<BuildDim> bds;
  for (BatchQuery& bq : b->q)
    for (int i = 0; i < bq.p.n_owned; i++) {
      BuildDim bd{};
      set_build_dim(c, bq.p.owned[i], bd);
      bd.block_begin = b->build_blocks;
      b->build_blocks += build_blocks(bd.n);
      bds.push_back(bd);
    }
  b->n_build = (int)bds.size();
  const size_t build_off = args_bytes;
  args_bytes += align256(sizeof(BuildDim) * std::max<size_t>(bds.size(), 1));
  if ((s = dev_alloc(c, args_bytes, &b->args_buf)) != LJ_OK) return bail(s);
  b->args_bytes = args_bytes;
  if ((s = host_alloc(c, args_bytes, &b->args_host)) != LJ_OK) return bail(s);
  char* args = static_cast<char*>(b->args_buf);
  char* host_args = static_cast<char*>(b->args_host);
  memset(host_args, 0, args_bytes);
  for (size_t k = 0; k < b->kinds.size(); k++) {
    BatchKind& K = b->kinds[k];
    K.d_qs = reinterpret_cast<ProbeArgs*>(args + qs_off[k]);
    K.d_wgb = reinterpret_cast<uint32_t*>(args + wgb_off[k]);
    ProbeArgs* hq = reinterpret_cast<ProbeArgs*>(host_args + qs_off[k]);
    uint32_t* hw = reinterpret_cast<uint32_t*>(host_args + wgb_off[k]);
    uint32_t begin = 0;
    for (size_t j = 0; j < K.qs.size(); j++) {
      BatchQuery& bq = b->q[(size_t)K.qs[j]];
      ProbeArgs a = bq.p.a;
      qbuf_bind(a, bq.desc.n_groups, bq.tcap, b->zero_buf + bq.zero_off, b->plain_buf + bq.plain_off,
                b->out_dev + bq.out_off);
      hq[j] = a;
      hw[j] = begin;
      begin += (uint32_t)bq.grid;
    }
    hw[K.qs.size()] = begin;
  }
  if (!bds.empty()) memcpy(host_args + build_off, bds.data(), sizeof(BuildDim) * bds.size());
  b->d_build = reinterpret_cast<BuildDim*>(args + build_off);
  if (!b->stream) HIP_TRY(c, hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  HIP_TRY(c, hipEventCreateWithFlags(&b->e0, hipEventDefault));
  HIP_TRY(c, hipEventCreateWithFlags(&b->e1, hipEventDefault));
  // after everything queued on the context's stream (zero-pool clears)
  HIP_TRY(c, hipEventRecord(b->e0, c->stream));
  HIP_TRY(c, hipStreamWaitEvent(b->stream, b->e0, 0));
  HIP_TRY(c, hipMemcpyAsync(args, host_args, args_bytes, hipMemcpyHostToDevice, b->stream));
  // the tables to clear before a repeated run
  FillArgs f{};
  for (BatchQuery& bq : b->q)
    for (int i = 0; i < bq.p.n_owned; i++) {
      lj_dim* D = bq.p.owned[i];
      void* ps[3] = {D->ht, D->bloom, D->counters};
      int64_t ws[3] = {D->ht_words, D->bloom_words, kCounterFillWords};
      for (int t = 0; t < 3; t++) {
        if (!ps[t] || ws[t] <= 0) continue;
        if (f.n == kMaxFill) {
          b->clears.push_back(f);
          f = FillArgs{};
        }
        add_fill(f, ps[t], ws[t]);
      }
    }
  if (f.n) b->clears.push_back(f);
  *out = b;
  return LJ_OK;
}

extern "C" lj_status lj_batch_create(lj_ctx* c, const lj_star_desc* descs, int32_t n, const lj_batch_opts* opts,
                                     lj_batch** out) {
  if (!c || !out || (n > 0 && !descs)) return fail(c, LJ_EINVAL, "NULL argument");
  *out = nullptr;
  HIP_TRY(c, hipSetDevice(c->device));
  lj_status s = batch_build(c, descs, n, opts, nullptr, out);
  if (s != LJ_OK) return s;
  HIP_TRY(c, hipStreamSynchronize((*out)->stream));  // the arguments are on the device
  return LJ_OK;
}

// The launches of one run (recorded once when use_graph).
static lj_status batch_enqueue(lj_batch* b, bool clear) {
  lj_ctx* c = b->c;
  if (clear)
    for (const FillArgs& f : b->clears) HIP_TRY(c, launch_fill(f, b->stream));
  HIP_TRY(c, launch_build_multi(b->d_build, b->n_build, b->build_blocks, b->stream));
  for (const BatchKind& K : b->kinds)
    HIP_TRY(c, launch_probe_multi(K.cfg, K.d_qs, K.d_wgb, (int)K.qs.size(), (int)K.grid, K.lds, b->stream));
  HIP_TRY(c, hipMemcpyAsync(b->out_host, b->out_dev, b->out_bytes, hipMemcpyDeviceToHost, b->stream));
  return LJ_OK;
}

// Queues one run (graph replay or the launches); e0/e1 bracket it.
static lj_status batch_launch(lj_batch* b) {
  lj_ctx* c = b->c;
  lj_status s;
  if (b->opts.use_graph) {
    if (!b->exec) {
      HIP_TRY(c, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      s = batch_enqueue(b, true);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(b->stream, &g);
      if (s != LJ_OK) {
        if (g) (void)hipGraphDestroy(g);
        return s;
      }
      HIP_TRY(c, e);
      b->graph = g;
      HIP_TRY(c, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    HIP_TRY(c, hipEventRecord(b->e0, b->stream));
    HIP_TRY(c, hipGraphLaunch(b->exec, b->stream));
    HIP_TRY(c, hipEventRecord(b->e1, b->stream));
  } else {
    HIP_TRY(c, hipEventRecord(b->e0, b->stream));
    if ((s = batch_enqueue(b, b->runs > 0)) != LJ_OK) return s;
    HIP_TRY(c, hipEventRecord(b->e1, b->stream));
  }
  b->runs++;
  return LJ_OK;
}

// After e1 has completed: the results of the last run.  status (optional)
// receives each query's own status; the return value is the first failure.
static lj_status batch_collect(lj_batch* b, lj_star_result* results, lj_status* status = nullptr) {
  lj_ctx* c = b->c;
  lj_status s;
  float ms = 0.f;
  HIP_TRY(c, hipEventElapsedTime(&ms, b->e0, b->e1));
  lj_status first_err = LJ_OK;
  FillArgs f{};
  for (size_t i = 0; i < b->q.size(); i++) {
    BatchQuery& bq = b->q[i];
    const long long* hdr = reinterpret_cast<const long long*>(b->out_host + bq.out_off);
    const GroupOut* groups = reinterpret_cast<const GroupOut*>(b->out_host + bq.out_off + kHdrBytes);
    QueryOut qo;
    s = parse_result(c, &bq.desc, bq.p, hdr, groups, bq.tcap, qo);
    if (hdr[0] > bq.tcap) {
      // the finalize zeroed only the listed groups: clear the query's whole
      // accumulator region before the next run
      if (f.n == kMaxFill) {
        HIP_TRY(c, launch_fill(f, b->stream));
        f = FillArgs{};
      }
      add_fill(f, b->zero_buf + bq.zero_off, (int64_t)(align256((size_t)bq.desc.n_groups * 8) * 3 / 4));
      if (s == LJ_OK)
        s = fail(c, LJ_EUNSUPPORTED, "query %zu hit %lld groups; groups_cap is %lld", i, hdr[0], (long long)bq.tcap);
    }
    qo.t_total = qo.t_probe = ms;
    qo.t_build = qo.t_fill = 0.f;
    to_star_result(&bq.desc, qo, &results[i]);
    results[i].status = s;
    if (status) status[i] = s;
    if (s != LJ_OK && first_err == LJ_OK) first_err = s;
  }
  if (f.n) HIP_TRY(c, launch_fill(f, b->stream));
  return first_err;
}

extern "C" lj_status lj_batch_run(lj_batch* b, lj_star_result* results) {
  if (!b) return fail(nullptr, LJ_EINVAL, "NULL batch");
  lj_ctx* c = b->c;
  if (!results) return fail(c, LJ_EINVAL, "NULL results");
  HIP_TRY(c, hipSetDevice(c->device));
  lj_status s = batch_launch(b);
  if (s != LJ_OK) return s;
  HIP_TRY(c, spin_wait(b->e1));
  return batch_collect(b, results);
}

extern "C" void lj_batch_destroy(lj_batch* b) { batch_free(b); }

// ------------------------------------------------------------ Bloom autotuner --
// Filters up to this size (2 MiB) stay in an XCD's L2 next to the hash table's
// hot lines; larger ones are mostly served from the MALL.
constexpr int64_t kTuneL2Bits = 1ll << 24;

// False-positive rate of a filter of m bits holding n keys.
static double fpr_of(int hash, int layout, int k, double n, double m, double dim_rows) {
  if (n <= 0) return 0.0;
  if (hash == LJ_BLOOM_MOD) {
    // bit (key - min) % m over keys dense in [min, min + dim_rows): exact when
    // m covers the range; else a bit is set if any of its ~dim_rows/m keys is
    if (m >= dim_rows) return 0.0;
    const double sel = std::min(1.0, n / dim_rows);
    return 1.0 - std::pow(1.0 - sel, dim_rows / m);
  }
  if (layout == LJ_BLOOM_STANDARD) return std::pow(1.0 - std::exp(-(double)k * n / m), (double)k);
  // blocked: k bits in one 32-bit word of m / 32; a word holds Poisson(lambda)
  // keys.  The sum runs over lambda +- 12 standard deviations with each term
  // in log space (exp(-lambda) alone underflows for a heavily loaded filter).
  const double words = m / 32.0, lambda = n / words;
  const double sd = std::sqrt(lambda), ll = std::log(lambda);
  const int64_t j0 = (int64_t)std::max(0.0, std::floor(lambda - 12.0 * sd - 12.0));
  const int64_t j1 = (int64_t)std::ceil(lambda + 12.0 * sd + 12.0);
  double acc = 0.0;
  for (int64_t j = j0; j <= j1; j++) {
    const double p = std::exp(-lambda + (double)j * ll - std::lgamma((double)j + 1.0));
    const double filled = 1.0 - std::pow(1.0 - 1.0 / 32.0, (double)k * (double)j);
    acc += p * std::pow(filled, (double)k);
  }
  return std::min(1.0, acc);
}

// Fraction of 128-byte lines (32 rows) of a column that hold at least one
// row passing with probability p (rows pass independently).
static double lines_of(double p) { return 1.0 - std::pow(1.0 - p, 32.0); }

extern "C" lj_status lj_tune_bloom(lj_ctx* c, int64_t n_selected, int64_t n_dim_rows, double fact_pass,
                                   int32_t flags, lj_bloom_choice* out) {
  if (!c || !out) return fail(c, LJ_EINVAL, "NULL argument");
  if (flags & ~LJ_TUNE_SEMI_JOIN) return fail(c, LJ_EINVAL, "unknown flags 0x%x", (unsigned)flags);
  const bool semi = (flags & LJ_TUNE_SEMI_JOIN) != 0 && c->exact_filter;
  if (n_selected < 0 || n_dim_rows < 1 || n_selected > n_dim_rows || !(fact_pass >= 0.0 && fact_pass <= 1.0))
    return fail(c, LJ_EINVAL, "n_selected %lld, n_dim_rows %lld, fact_pass %g", (long long)n_selected,
                (long long)n_dim_rows, fact_pass);
  ProbeConfig cfg{1024, 1, S0_MULT_LDS};
  const int64_t lds_bits = ((int64_t)c->max_lds - (int64_t)probe_lds_bytes(cfg, 0, kGroupSlots) - 16) * 8;
  lj_bloom_choice best{};
  best.cost = 1e300;
  struct Cand { int hash, layout, k; };
  const Cand cands[] = {{LJ_BLOOM_MOD, LJ_BLOOM_STANDARD, 1},  {LJ_BLOOM_MULT, LJ_BLOOM_STANDARD, 1},
                        {LJ_BLOOM_MULT, LJ_BLOOM_STANDARD, 2}, {LJ_BLOOM_MULT, LJ_BLOOM_STANDARD, 3},
                        {LJ_BLOOM_MULT, LJ_BLOOM_BLOCKED, 2},  {LJ_BLOOM_MULT, LJ_BLOOM_BLOCKED, 3}};
  for (int lg = 16; lg <= 26; lg++) {
    const double m = std::ldexp(1.0, lg);
    for (const Cand& cd : cands) {
      const bool in_lds = (int64_t)m <= lds_bits;
      const double fpr = fpr_of(cd.hash, cd.layout, cd.k, (double)n_selected, m, (double)n_dim_rows);
      // a semi-join's exact modulo filter decides the table probe: no gathers
      const bool exact = semi && cd.hash == LJ_BLOOM_MOD && m >= (double)n_dim_rows;
      // in units of a scan with a standard multiplicative filter in LDS
      const double lds_unit = cd.hash == LJ_BLOOM_MOD ? c->tune_scan_mod
                              : cd.layout == LJ_BLOOM_BLOCKED ? c->tune_scan_blk : 1.0;
      double unit = in_lds ? lds_unit : m <= (double)kTuneL2Bits ? c->tune_scan_gl : c->tune_scan_big;
      // and, too large for LDS, it is staged in part (plan_probe): that share
      // of its probes costs what an LDS filter's do
      if (exact && !in_lds && c->partial_prefix) {
        const double f = (double)lds_bits / m;
        unit = f * c->tune_scan_mod + (1.0 - f) * unit;
      }
      // (further bit probes of a standard filter cost nothing measurable in
      // LDS; outside it each is another L2 request for the rows still alive)
      double scan = unit;
      if (cd.layout == LJ_BLOOM_STANDARD && cd.k > 1 && !in_lds) scan += (cd.k - 1) * c->tune_probe_extra * unit;
      const double pass = fact_pass + (1.0 - fact_pass) * fpr;
      // The probe streams the measure with the keys when more than a quarter
      // of a tile's rows pass: the cost per passing row is its hash-slot
      // gather (smaller beside an L2 filter's scan, which it overlaps).
      // Otherwise the passing rows are queued and gather the measure too.
      // Each wave decides on its first tile, so between 20 % and 30 % the
      // two costs are blended (no step for a filter to slip across).
      double cost;
      if (exact) {
        cost = scan + lines_of(pass) * c->tune_line;
      } else {
        const double streamed = scan + pass * (in_lds ? c->tune_gather_lds : c->tune_gather);
        const double queued = scan + lines_of(pass) * c->tune_line + pass * (in_lds ? c->tune_gather_q_lds
                                                                                    : c->tune_gather_q);
        const double w = std::min(1.0, std::max(0.0, (pass - 0.2) / 0.1));
        cost = (1.0 - w) * queued + w * streamed;
      }
      if (cost < best.cost * (1.0 - 1e-9)) {
        best.bloom_hash = cd.hash;
        best.bloom_layout = cd.layout;
        best.bloom_k = cd.k;
        best.in_lds = in_lds ? 1 : 0;
        best.bloom_bits = (int64_t)m;
        best.fpr = fpr;
        best.cost = cost;
      }
    }
  }
  *out = best;
  return LJ_OK;
}

extern "C" lj_status lj_tune_get_model(const lj_ctx* c, lj_tune_model* out) {
  if (!c || !out) return LJ_EINVAL;
  *out = lj_tune_model{};
  out->scan_l2 = c->tune_scan_gl;
  out->scan_big = c->tune_scan_big;
  out->probe_extra = c->tune_probe_extra;
  out->line = c->tune_line;
  out->gather = c->tune_gather;
  out->scan_mod = c->tune_scan_mod;
  out->scan_blocked = c->tune_scan_blk;
  out->gather_queue = c->tune_gather_q;
  out->gather_lds = c->tune_gather_lds;
  out->gather_queue_lds = c->tune_gather_q_lds;
  out->calibrated = c->tune_calibrated;
  return LJ_OK;
}

extern "C" lj_status lj_tune_set_model(lj_ctx* c, const lj_tune_model* in) {
  if (!c || !in) return fail(c, LJ_EINVAL, "NULL argument");
  auto finite = [](double x) { return x == x && x < 1e300 && x > -1e300; };
  if (!finite(in->scan_l2) || !finite(in->scan_big) || !finite(in->probe_extra) || !finite(in->line) ||
      !finite(in->gather) || !finite(in->scan_mod) || !finite(in->scan_blocked) || !finite(in->gather_queue) ||
      !finite(in->gather_lds) || !finite(in->gather_queue_lds) || in->scan_l2 < 1.0 ||
      in->scan_big < in->scan_l2 || in->probe_extra < 0.0 || in->line < 0.0 || in->gather < 0.0 ||
      in->gather_queue < 0.0 || in->gather_lds < 0.0 || in->gather_queue_lds < 0.0 ||
      !(in->scan_mod > 0.0) || !(in->scan_blocked > 0.0))
    return fail(c, LJ_EINVAL,
                "tune model: scan_l2 >= 1, scan_big >= scan_l2, scan_mod and scan_blocked > 0, the others >= 0, "
                "all finite");
  c->tune_scan_gl = in->scan_l2;
  c->tune_scan_big = in->scan_big;
  c->tune_probe_extra = in->probe_extra;
  c->tune_line = in->line;
  c->tune_gather = in->gather;
  c->tune_scan_mod = in->scan_mod;
  c->tune_scan_blk = in->scan_blocked;
  c->tune_gather_q = in->gather_queue;
  c->tune_gather_lds = in->gather_lds;
  c->tune_gather_q_lds = in->gather_queue_lds;
  c->tune_calibrated = 1;
  return LJ_OK;
}

// (a calibration run's filter: the blocked multiplicative layout, k = 2)
constexpr int kCalBlocked = 100;

// Calibration: a dimension of 2^22 dense keys and a fact table of 2^25 rows
// (device-generated), nine LIP joins, each timed over all rows and over a
// quarter of them (their difference is the cost per row): 2^20-bit filters in
// LDS that nothing passes (standard multiplicative, the unit; modulo;
// blocked), a 2^22-bit (L2) and a 2^25-bit filter that nothing passes, the
// exact 2^22-bit filter that 20 % and 40 % pass, and the 2^20-bit modulo
// filter in LDS over 3 % and 10 % of the keys (4 keys a bit: about 14 % and
// 41 % pass; survivor costs are measured against the empty filter of the
// same kind).

extern "C" lj_status lj_tune_calibrate(lj_ctx* c) {
  if (!c) return fail(c, LJ_EINVAL, "NULL context");
  HIP_TRY(c, hipSetDevice(c->device));
  const int64_t nd = 1 << 22, nf = 1 << 25;
  int32_t *keys = nullptr, *attr = nullptr, *fk = nullptr, *meas = nullptr;
  lj_status s;
  if ((s = dev_alloc_t(c, nd, &keys)) != LJ_OK) return s;
  auto release = [&]() {
    (void)hipStreamSynchronize(c->stream);
    dev_release(c, keys);
    dev_release(c, attr);
    dev_release(c, fk);
    dev_release(c, meas);
  };
  if ((s = dev_alloc_t(c, nd, &attr)) != LJ_OK || (s = dev_alloc_t(c, nf, &fk)) != LJ_OK ||
      (s = dev_alloc_t(c, nf, &meas)) != LJ_OK) {
    release();
    return s;
  }
  const uint64_t seed = 0x1A7E5EED7Full;
  if (launch_gen(LJ_GEN_SEQUENCE, mix64(seed), 0, 0, 0, nd, keys, c->stream) != hipSuccess ||
      launch_gen(LJ_GEN_UNIFORM, stream_key(seed, 1), 0, 9999, 0, nd, attr, c->stream) != hipSuccess ||
      launch_gen(LJ_GEN_UNIFORM, stream_key(seed, 2), 0, nd - 1, 0, nf, fk, c->stream) != hipSuccess ||
      launch_gen(LJ_GEN_UNIFORM, stream_key(seed, 3), 1, 10000, 0, nf, meas, c->stream) != hipSuccess) {
    release();
    return fail(c, LJ_EHIP, "generator launch failed");
  }
  auto run1 = [&](int hash, int64_t bits, int32_t below, int64_t n_fact, float* ms) -> lj_status {
    lj_star_desc q{};
    q.fact_fk[0] = fk;
    q.fact_measure = meas;
    q.n_fact = n_fact;
    q.n_dims = 1;
    q.method = LJ_METHOD_LIP;
    for (int i = 0; i < LJ_MAX_DIMS; i++) {
      q.order[i] = i;
      q.bloom_order[i] = -1;
    }
    lj_dim_desc& d = q.dims[0];
    d.keys = keys;
    d.attr = attr;
    d.n_rows = nd;
    d.pred.n_steps = 1;
    d.pred.steps[0].comb = LJ_COMB_SET;
    d.pred.steps[0].cmp = LJ_CMP_LT;
    d.pred.steps[0].value = below;
    d.ht_kind = LJ_HT_PHT2;
    d.ht_len = nd;
    d.bloom_hash = hash;
    d.bloom_k = hash == kCalBlocked ? 2 : 1;
    if (hash == kCalBlocked) {
      d.bloom_hash = LJ_BLOOM_MULT;
      d.bloom_layout = LJ_BLOOM_BLOCKED;
    }
    d.bloom_bits = bits;
    q.n_groups = 1;
    float best = 1e30f;
    for (int r = 0; r < 4; r++) {
      QueryOut qo;
      lj_status st = run_query(c, &q, 0, qo);
      if (st != LJ_OK) return st;
      if (r > 0) best = std::min(best, qo.t_probe);
    }
    *ms = best;
    return LJ_OK;
  };
  // the cost per fact row: the difference between a run over all nf rows and
  // one over a quarter of them (the launch, staging and finalize cancel out)
  auto run = [&](int hash, int64_t bits, int32_t below, float* per_row) -> lj_status {
    float t_all = 0, t_quarter = 0;
    lj_status st = run1(hash, bits, below, nf, &t_all);
    if (st == LJ_OK) st = run1(hash, bits, below, nf / 4, &t_quarter);
    *per_row = (t_all - t_quarter) / (0.75f * (float)nf);
    return st;
  };
  float t_lds = 0, t_mod = 0, t_blk = 0, t_l2 = 0, t_big = 0, t_p20 = 0, t_p40 = 0, t_m14 = 0, t_m41 = 0;
  // the survivor costs are measured with every passing row reading the table
  // (these one-group joins would otherwise let the exact filter decide)
  const int exact_saved = c->exact_filter;
  c->exact_filter = 0;
  if ((s = run(LJ_BLOOM_MULT, 1 << 20, 0, &t_lds)) != LJ_OK || (s = run(LJ_BLOOM_MOD, 1 << 20, 0, &t_mod)) != LJ_OK ||
      (s = run(kCalBlocked, 1 << 20, 0, &t_blk)) != LJ_OK || (s = run(LJ_BLOOM_MOD, nd, 0, &t_l2)) != LJ_OK ||
      (s = run(LJ_BLOOM_MOD, 1 << 25, 0, &t_big)) != LJ_OK || (s = run(LJ_BLOOM_MOD, nd, 2000, &t_p20)) != LJ_OK ||
      (s = run(LJ_BLOOM_MOD, nd, 4000, &t_p40)) != LJ_OK ||
      (s = run(LJ_BLOOM_MOD, 1 << 20, 300, &t_m14)) != LJ_OK || (s = run(LJ_BLOOM_MOD, 1 << 20, 1000, &t_m41)) != LJ_OK) {
    c->exact_filter = exact_saved;
    release();
    return s;
  }
  c->exact_filter = exact_saved;
  release();
  if (!(t_lds > 0 && t_mod > 0 && t_blk > 0 && t_l2 > 0 && t_big > 0))
    return fail(c, LJ_EHIP, "calibration timings are not positive");
  c->tune_scan_gl = std::max(1.0, (double)t_l2 / t_lds);
  c->tune_scan_big = std::max(c->tune_scan_gl, (double)t_big / t_lds);
  c->tune_scan_mod = std::min(1.0, (double)t_mod / t_lds);
  c->tune_scan_blk = std::min(1.0, (double)t_blk / t_lds);
  // t_p20 / t_p40 scan the same 2^22-bit (512 KiB) filter as t_l2 and pass
  // 20 % and 40 % of the rows.  Up to a quarter the survivors' measure is
  // gathered: e(p) = lines_of(p) * line + p * gather_q gives gather_q at 20 %
  // (line stays the model's: the lines term barely changes between the pass
  // rates a short calibration can time, so it cannot be fitted from them;
  // 0.36 fits a 2^28-row sweep from 1 % to 19 %), and gather_q_lds from the
  // LDS filter that passes about 14 % (the queued cost per row falls as more
  // rows pass; the filters the tuner weighs against each other pass 5-20 %).
  // Above a quarter the measure streams
  // with the keys: gather from 40 %, gather_lds from the LDS filter that
  // passes about 41 %.
  const double e20 = std::max(0.0, ((double)t_p20 - t_l2) / t_lds);
  const double e40 = std::max(0.0, ((double)t_p40 - t_l2) / t_lds);
  const double gather_q = (e20 - lines_of(0.20) * c->tune_line) / 0.20;
  const auto p_mod20 = [&](double sel) {  // the 2^20-bit modulo filter over sel of the keys
    return sel + (1.0 - sel) * fpr_of(LJ_BLOOM_MOD, LJ_BLOOM_STANDARD, 1, sel * (double)nd, (double)(1 << 20), (double)nd);
  };
  const double p14 = p_mod20(0.03), p41 = p_mod20(0.1);
  const double e14 = std::max(0.0, ((double)t_m14 - t_mod) / t_lds);
  const double e41 = std::max(0.0, ((double)t_m41 - t_mod) / t_lds);
  c->tune_gather_q = std::max(1.0, gather_q);
  c->tune_gather_q_lds = std::max(1.0, (e14 - lines_of(p14) * c->tune_line) / p14);
  c->tune_gather = std::max(1.0, e40 / 0.4);
  c->tune_gather_lds = std::max(1.0, e41 / p41);
  c->tune_calibrated = 1;
  return LJ_OK;
}

// ------------------------------------------------------------------- serving --
// The loop keeps everything a launch needs: per launch slot a stream, two
// events and the argument buffers (pinned staging + device copy); per request
// a buffer set from a pool (group accumulators kept zero by the finalize,
// statistics rows, the device result region and a pinned host result region
// that the finalize writes directly).  A launch of one request is the
// single-query probe (arguments in the kernarg segment, no upload); a launch
// of several uploads their arguments once and runs the batched probe.
namespace {
using Clock = std::chrono::steady_clock;

// Statistics rows per request buffer: the most workgroups one query is given
// (n_cu x the most workgroups of a configuration per CU).
constexpr int kServeWgPerCu = 8;
// The loop spins this long for new work before sleeping on the condition.
constexpr auto kServeIdleSpin = std::chrono::milliseconds(2);

struct ReqBuf {
  int64_t ng = 0;       // group-space class (a power of two >= n_groups)
  int64_t tcap = 0;     // groups listed
  QBufLayout l;
  char* dev = nullptr;  // zero | plain | out
  char* host = nullptr; // pinned: header | groups [tcap]
  char* host_dev = nullptr;  // the same memory, as the device addresses it
  size_t host_bytes = 0;
  uint64_t bytes = 0;   // device bytes (the budget's unit)
  bool fresh = true;    // zero region not cleared yet
};

struct ServeReq {
  uint64_t id = 0;
  lj_star_desc desc{};
  Clock::time_point t_submit;
  uint64_t bytes = 0;
  // admission
  Plan p;
  ReqBuf* buf = nullptr;
  int64_t grid = 0;
};

struct ServeDone {
  lj_status st = LJ_OK;
  std::string err;
  lj_star_result r{};
  std::vector<lj_group> groups;
};

struct ServeSlot {
  hipStream_t stream = nullptr;  // one of the server's streams
  hipEvent_t e0 = nullptr, e1 = nullptr;
  char* args_dev = nullptr;   // ProbeArgs [max_batch] | wg_begin [max_batch + S0_NKINDS]
  char* args_host = nullptr;  // pinned staging of args_dev
  char* args_host_dev = nullptr;  // the same memory, as the device addresses it
  size_t args_bytes = 0;
  size_t wgb_off = 0;
  bool busy = false;
  std::vector<ServeReq> reqs;
  uint64_t bytes = 0;
};

int64_t pow2_at_least(int64_t x) {
  int64_t v = 1;
  while (v < x) v <<= 1;
  return v;
}
}  // namespace

struct lj_server {
  lj_ctx* c = nullptr;
  lj_serve_cfg cfg{};
  std::thread th;
  std::mutex mu;
  std::condition_variable cv_work;   // the loop waits for requests
  std::condition_variable cv_done;   // pollers wait for results
  std::deque<ServeReq> pending;
  std::atomic<uint64_t> n_pending{0};  // pending.size(), read without the lock by the idle spin
  std::unordered_map<uint64_t, ServeDone> done;
  uint64_t next_id = 1;
  bool stopping = false;
  lj_serve_stats stats{};
  std::vector<ServeSlot> slots;
  std::vector<hipStream_t> streams;
  std::multimap<int64_t, ReqBuf*> free_bufs;  // idle request buffers by group class
  std::vector<ReqBuf*> all_bufs;
  int64_t stats_rows = 0;
};

// Device bytes of a request buffer for n_groups (the serving budget's unit).
static uint64_t serve_bytes(const lj_server* S, int64_t n_groups) {
  const int64_t ng = pow2_at_least(n_groups);
  const int64_t tcap = std::min<int64_t>(ng, S->cfg.groups_cap);
  const QBufLayout l = qbuf_layout(ng, tcap, S->stats_rows);
  return l.zero + l.plain + l.out;
}

static void reqbuf_free(lj_ctx* c, ReqBuf* b) {
  dev_release(c, b->dev);
  host_release(c, b->host, b->host_bytes);
  delete b;
}

// An idle buffer of n_groups' class, or a new one (its zero region is cleared
// on the stream of its first launch).
static lj_status reqbuf_get(lj_server* S, int64_t n_groups, ReqBuf** out) {
  lj_ctx* c = S->c;
  const int64_t ng = pow2_at_least(n_groups);
  auto it = S->free_bufs.find(ng);
  if (it != S->free_bufs.end()) {
    *out = it->second;
    S->free_bufs.erase(it);
    return LJ_OK;
  }
  ReqBuf* b = new ReqBuf();
  b->ng = ng;
  b->tcap = std::min<int64_t>(ng, S->cfg.groups_cap);
  b->l = qbuf_layout(ng, b->tcap, S->stats_rows);
  b->bytes = b->l.zero + b->l.plain + b->l.out;
  void* d = nullptr;
  lj_status s = dev_alloc(c, b->bytes, &d);
  if (s != LJ_OK) {
    delete b;
    return s;
  }
  b->dev = static_cast<char*>(d);
  b->host_bytes = kHdrBytes + sizeof(GroupOut) * (size_t)b->tcap;
  void* h = nullptr;
  void* hd = nullptr;
  if (host_alloc(c, b->host_bytes, &h) != LJ_OK || hipHostGetDevicePointer(&hd, h, 0) != hipSuccess) {
    host_release(c, h, b->host_bytes);
    b->host = nullptr;
    reqbuf_free(c, b);
    return fail(c, LJ_EHIP, "pinned result buffer allocation failed");
  }
  b->host = static_cast<char*>(h);
  b->host_dev = static_cast<char*>(hd);
  S->all_bufs.push_back(b);
  *out = b;
  return LJ_OK;
}

// The request's arguments: its plan with its buffers; the header and groups
// also go to the pinned host region.
static ProbeArgs serve_args(const ServeReq& rq) {
  ProbeArgs a = rq.p.a;
  ReqBuf* b = rq.buf;
  qbuf_bind(a, b->ng, b->tcap, b->dev, b->dev + b->l.zero, b->dev + b->l.zero + b->l.plain);
  a.host_hdr = reinterpret_cast<long long*>(b->host_dev);
  a.host_out = reinterpret_cast<GroupOut*>(b->host_dev + kHdrBytes);
  a.host_groups = b->tcap;
  return a;
}

static void serve_finish(lj_server* S, const ServeReq& rq, lj_status st, const char* err, const lj_star_result* r,
                         const lj_group* groups, Clock::time_point t_done) {
  ServeDone d;
  d.st = st;
  if (err) d.err = err;
  if (r) {
    d.r = *r;
    const int64_t n = std::min<int64_t>(r->n_groups, r->groups_cap);
    d.groups.assign(groups, groups + std::max<int64_t>(n, 0));
  }
  d.r.groups = nullptr;
  d.r.status = st;
  d.r.time_query_ms = (float)std::chrono::duration<double, std::milli>(t_done - rq.t_submit).count();
  S->done[rq.id] = std::move(d);
}

// The slot's first n_q ProbeArgs and n_w workgroup offsets, from its pinned
// staging into its device buffer, ordered on the slot's stream (a copy
// kernel: see copy_kernel).
static hipError_t upload_args(const ServeSlot& sl, size_t n_q, size_t n_w) {
  static_assert(sizeof(ProbeArgs) % 4 == 0, "ProbeArgs in words");
  CopyArgs ca{};
  ca.r[0].dst = reinterpret_cast<uint32_t*>(sl.args_dev);
  ca.r[0].src = reinterpret_cast<const uint32_t*>(sl.args_host_dev);
  ca.r[0].words = (int64_t)(n_q * sizeof(ProbeArgs) / 4);
  ca.r[1].dst = reinterpret_cast<uint32_t*>(sl.args_dev + sl.wgb_off);
  ca.r[1].src = reinterpret_cast<const uint32_t*>(sl.args_host_dev + sl.wgb_off);
  ca.r[1].words = (int64_t)n_w;
  ca.n = 2;
  return launch_copy(ca, sl.stream);
}

// Plan the admitted requests and queue their launch on the slot's stream.
// Requests that fail to plan finish at once with their own status (out).
static lj_status serve_launch(lj_server* S, ServeSlot& sl, std::vector<std::pair<ServeReq, lj_status>>& failed) {
  lj_ctx* c = S->c;
  std::vector<ServeReq> ok;
  for (ServeReq& rq : sl.reqs) {
    lj_status s = validate_query(c, &rq.desc, rq.p);
    if (s == LJ_OK) s = acquire_dims(c, &rq.desc, rq.p);
    if (s == LJ_OK) s = plan_probe(c, &rq.desc, 0, 0, rq.p);
    if (s == LJ_OK) s = reqbuf_get(S, rq.desc.n_groups, &rq.buf);
    if (s != LJ_OK) {
      failed.emplace_back(std::move(rq), s);
      continue;
    }
    ok.push_back(std::move(rq));
  }
  sl.reqs = std::move(ok);
  if (sl.reqs.empty()) return LJ_OK;
  // clear the zero regions of new buffers first (stream order)
  FillArgs f{};
  for (ServeReq& rq : sl.reqs)
    if (rq.buf->fresh) {
      if (f.n == kMaxFill) {
        HIP_TRY(c, launch_fill(f, sl.stream));
        f = FillArgs{};
      }
      add_fill(f, rq.buf->dev, (int64_t)(rq.buf->l.zero / 4));
      rq.buf->fresh = false;
    }
  if (f.n) HIP_TRY(c, launch_fill(f, sl.stream));
  HIP_TRY(c, hipEventRecord(sl.e0, sl.stream));
  if (sl.reqs.size() == 1) {
    ServeReq& rq = sl.reqs[0];
    const ProbeConfig& cfg = rq.p.cfg;
    const int64_t tile = (int64_t)cfg.vec * kSegRows * (cfg.wg / 64);
    const int64_t need = std::max<int64_t>(1, (rq.desc.n_fact + tile - 1) / tile);
    rq.grid = std::min<int64_t>({(int64_t)c->n_cu * rq.p.per_cu, need, S->stats_rows});
    const ProbeArgs a = serve_args(rq);
    HIP_TRY(c, launch_probe(cfg, a, (int)rq.grid, rq.p.lds_bytes, sl.stream));
  } else {
    // group by kernel configuration; one upload, one launch per configuration
    std::vector<std::vector<size_t>> kinds;
    std::vector<ProbeConfig> kcfg;
    std::vector<size_t> klds;
    for (size_t i = 0; i < sl.reqs.size(); i++) {
      const ProbeConfig& cf = sl.reqs[i].p.cfg;
      size_t k = 0;
      while (k < kcfg.size() && !(kcfg[k].wg == cf.wg && kcfg[k].vec == cf.vec && kcfg[k].k0 == cf.k0)) k++;
      if (k == kcfg.size()) {
        kcfg.push_back(cf);
        kinds.emplace_back();
        klds.push_back(0);
      }
      kinds[k].push_back(i);
      klds[k] = std::max(klds[k], sl.reqs[i].p.lds_bytes);
    }
    // (one launch per kind: an admitted group is one or two requests at the
    // rates where latency matters, and the per-kind kernels are faster than
    // the mixed-kind one)
    ProbeArgs* hq = reinterpret_cast<ProbeArgs*>(sl.args_host);
    uint32_t* hw = reinterpret_cast<uint32_t*>(sl.args_host + sl.wgb_off);
    std::vector<size_t> q0(kinds.size()), w0(kinds.size());
    std::vector<int64_t> kgrid(kinds.size(), 0);
    size_t qi = 0, wi = 0;
    for (size_t k = 0; k < kinds.size(); k++) {
      int per_cu = 0;
      lj_status s = occupancy_of(c, kcfg[k], klds[k], &per_cu);
      if (s != LJ_OK) return s;
      const int64_t tile = (int64_t)kcfg[k].vec * kSegRows * (kcfg[k].wg / 64);
      q0[k] = qi;
      w0[k] = wi;
      uint32_t begin = 0;
      for (size_t i : kinds[k]) {
        ServeReq& rq = sl.reqs[i];
        const int64_t need = std::max<int64_t>(1, (rq.desc.n_fact + tile - 1) / tile);
        rq.grid = std::min<int64_t>({(int64_t)c->n_cu * per_cu, need, S->stats_rows});
        hq[qi++] = serve_args(rq);
        hw[wi++] = begin;
        begin += (uint32_t)rq.grid;
      }
      hw[wi++] = begin;
      kgrid[k] = begin;
    }
    HIP_TRY(c, upload_args(sl, qi, wi));
    const ProbeArgs* dq = reinterpret_cast<const ProbeArgs*>(sl.args_dev);
    const uint32_t* dw = reinterpret_cast<const uint32_t*>(sl.args_dev + sl.wgb_off);
    for (size_t k = 0; k < kinds.size(); k++)
      HIP_TRY(c, launch_probe_multi(kcfg[k], dq + q0[k], dw + w0[k], (int)kinds[k].size(), (int)kgrid[k], klds[k],
                                    sl.stream));
  }
  HIP_TRY(c, hipEventRecord(sl.e1, sl.stream));
  return LJ_OK;
}

// A completed launch: every request's result from its pinned host region.
static void serve_complete(lj_server* S, ServeSlot& sl) {
  lj_ctx* c = S->c;
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, sl.e0, sl.e1);
  const Clock::time_point t_done = Clock::now();
  bool dirty = false;
  std::vector<ServeDone> outs(sl.reqs.size());
  for (size_t j = 0; j < sl.reqs.size(); j++) {
    ServeReq& rq = sl.reqs[j];
    ReqBuf* b = rq.buf;
    const long long* hdr = reinterpret_cast<const long long*>(b->host);
    const GroupOut* groups = reinterpret_cast<const GroupOut*>(b->host + kHdrBytes);
    QueryOut qo;
    lj_status s = parse_result(c, &rq.desc, rq.p, hdr, groups, b->tcap, qo);
    if (hdr[0] > b->tcap) {
      // the finalize zeroed only the listed groups: clear the rest before reuse
      FillArgs f{};
      add_fill(f, b->dev, (int64_t)(b->l.zero / 4));
      (void)launch_fill(f, sl.stream);
      dirty = true;
      if (s == LJ_OK)
        s = fail(c, LJ_EUNSUPPORTED, "request hit %lld groups; groups_cap is %lld", hdr[0], (long long)b->tcap);
    }
    qo.t_total = qo.t_probe = ms;
    qo.t_build = qo.t_fill = 0.f;
    ServeDone& d = outs[j];
    d.groups.resize((size_t)b->tcap);
    d.r.groups = d.groups.data();
    d.r.groups_cap = b->tcap;
    to_star_result(&rq.desc, qo, &d.r);
    d.r.status = s;
    d.st = s;
    if (s != LJ_OK) d.err = c->err;
  }
  if (dirty) (void)hipStreamSynchronize(sl.stream);  // the clears, before the buffers are reused
  std::lock_guard<std::mutex> lk(S->mu);
  for (size_t j = 0; j < sl.reqs.size(); j++) {
    ServeReq& rq = sl.reqs[j];
    serve_finish(S, rq, outs[j].st, outs[j].st == LJ_OK ? nullptr : outs[j].err.c_str(), &outs[j].r,
                 outs[j].groups.data(), t_done);
    S->free_bufs.emplace(rq.buf->ng, rq.buf);
    rq.buf = nullptr;
  }
  S->stats.completed += sl.reqs.size();
  S->stats.bytes_in_flight -= sl.bytes;
  sl.reqs.clear();
  sl.bytes = 0;
  sl.busy = false;
}

static void serve_loop(lj_server* S) {
  lj_ctx* c = S->c;
  (void)hipSetDevice(c->device);
  for (;;) {
    // 1. completions (the loop alone drives the context: no lock for HIP work)
    bool any_busy = false, completed = false;
    for (ServeSlot& sl : S->slots) {
      if (!sl.busy) continue;
      if (hipEventQuery(sl.e1) != hipSuccess) {
        any_busy = true;
        continue;
      }
      serve_complete(S, sl);
      completed = true;
    }
    if (completed) S->cv_done.notify_all();
    // 2. admission: every waiting request the budget admits, into one launch on a free slot
    ServeSlot* slot = nullptr;
    for (ServeSlot& sl : S->slots)
      if (!sl.busy) slot = &sl;
    if (!any_busy && S->n_pending.load(std::memory_order_acquire) == 0) {
      // idle: spin for the next request before sleeping on the condition
      const Clock::time_point until = Clock::now() + kServeIdleSpin;
      while (S->n_pending.load(std::memory_order_acquire) == 0 && Clock::now() < until) std::this_thread::yield();
    }
    std::vector<ServeReq> take;
    uint64_t take_bytes = 0;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      if (!any_busy && S->pending.empty()) {
        if (S->stopping) break;
        S->cv_work.wait(lk, [&] { return S->stopping || !S->pending.empty(); });
        if (S->pending.empty()) break;
      }
      if (slot) {
        const uint64_t budget = S->cfg.hbm_budget;
        while (!S->pending.empty() && (int)take.size() < S->cfg.max_batch) {
          ServeReq& rq = S->pending.front();
          if (budget && rq.bytes > budget) {  // can never be admitted
            serve_finish(S, rq, LJ_EBUDGET, "request needs more device memory than the serving budget", nullptr,
                         nullptr, Clock::now());
            S->stats.rejected++;
            S->pending.pop_front();
            S->cv_done.notify_all();
            continue;
          }
          if (budget && S->stats.bytes_in_flight + take_bytes + rq.bytes > budget) break;
          take_bytes += rq.bytes;
          take.push_back(std::move(rq));
          S->pending.pop_front();
        }
        S->n_pending.store(S->pending.size(), std::memory_order_release);
        if (!take.empty()) {
          S->stats.bytes_in_flight += take_bytes;
          S->stats.peak_bytes_in_flight = std::max(S->stats.peak_bytes_in_flight, S->stats.bytes_in_flight);
          S->stats.launches++;
          S->stats.max_batch_seen = std::max<uint64_t>(S->stats.max_batch_seen, take.size());
        }
      }
    }
    if (!take.empty()) {
      slot->reqs = std::move(take);
      slot->bytes = take_bytes;
      slot->busy = true;
      std::vector<std::pair<ServeReq, lj_status>> failed;
      lj_status st = serve_launch(S, *slot, failed);
      if (st != LJ_OK) {
        // the launch itself failed: every request of it fails with that status
        for (ServeReq& rq : slot->reqs) failed.emplace_back(std::move(rq), st);
        slot->reqs.clear();
      }
      if (!failed.empty()) {
        const std::string err = c->err;
        std::lock_guard<std::mutex> lk(S->mu);
        for (auto& fr : failed) {
          if (fr.first.buf) S->free_bufs.emplace(fr.first.buf->ng, fr.first.buf);
          serve_finish(S, fr.first, fr.second, err.c_str(), nullptr, nullptr, Clock::now());
          S->stats.rejected++;
          S->stats.bytes_in_flight -= fr.first.bytes;
          slot->bytes -= fr.first.bytes;
        }
        S->cv_done.notify_all();
      }
      if (slot->reqs.empty()) slot->busy = false;
      continue;
    }
    if (any_busy) std::this_thread::yield();  // poll the launches in flight
  }
}

static void server_free(lj_server* S) {
  lj_ctx* c = S->c;
  for (hipStream_t st : S->streams) (void)hipStreamSynchronize(st);
  for (ServeSlot& sl : S->slots) {
    if (sl.e0) (void)hipEventDestroy(sl.e0);
    if (sl.e1) (void)hipEventDestroy(sl.e1);
    dev_release(c, sl.args_dev);
    if (sl.args_host) (void)hipHostFree(sl.args_host);
  }
  for (hipStream_t st : S->streams) (void)hipStreamDestroy(st);
  for (ReqBuf* b : S->all_bufs) reqbuf_free(c, b);
  delete S;
}

// One launch of each batched probe kernel (the 1024-thread kinds the loop
// launches) on an empty query: no fact rows, one workgroup, the buffers of
// a one-group request.  A kernel's first launch in a process spends
// milliseconds loading it (6.9 ms measured), which the first requests of
// that kind in a batched launch, and every request queued behind them,
// would otherwise wait for.
static lj_status serve_warm(lj_server* S) {
  lj_ctx* c = S->c;
  ServeSlot& sl = S->slots[0];
  ReqBuf* b = nullptr;
  lj_status s = reqbuf_get(S, 1, &b);
  if (s != LJ_OK) return s;
  auto done = [&](lj_status st) {
    (void)hipStreamSynchronize(sl.stream);
    S->free_bufs.emplace(b->ng, b);
    return st;
  };
  if (b->fresh) {
    FillArgs f{};
    add_fill(f, b->dev, (int64_t)(b->l.zero / 4));
    if (launch_fill(f, sl.stream) != hipSuccess) return done(fail(c, LJ_EHIP, "fill launch failed"));
    b->fresh = false;
  }
  ProbeArgs a{};
  qbuf_bind(a, b->ng, b->tcap, b->dev, b->dev + b->l.zero, b->dev + b->l.zero + b->l.plain);
  a.n = 0;
  a.n_dims = 1;
  a.n_groups = 1;
  a.gmod = make_fastmod(1);
  a.gslot_bits = ilog2_u64((uint64_t)kGroupSlots);
  a.exact_eager_div = c->exact_eager_div;
  a.direct_eager_div = c->direct_eager_div;
  a.occ_eager_div = c->occ_eager_div;
  for (int i = 0; i < LJ_MAX_DIMS; i++) {
    a.order[i] = a.load_order[i] = i;
    a.dim[i].bloom_lds = -1;
  }
  ProbeArgs* hq = reinterpret_cast<ProbeArgs*>(sl.args_host);
  uint32_t* hw = reinterpret_cast<uint32_t*>(sl.args_host + sl.wgb_off);
  hq[0] = a;
  hw[0] = 0;
  hw[1] = 1;
  if (upload_args(sl, 1, 2) != hipSuccess) return done(fail(c, LJ_EHIP, "argument upload failed"));
  const ProbeArgs* dq = reinterpret_cast<const ProbeArgs*>(sl.args_dev);
  const uint32_t* dw = reinterpret_cast<const uint32_t*>(sl.args_dev + sl.wgb_off);
  for (int k0 = 0; k0 < S0_NKINDS; k0++) {
    const ProbeConfig cfg{1024, 1, k0};
    if (!probe_has_entry(cfg)) continue;
    const size_t lds = probe_lds_bytes(cfg, 0, kGroupSlots);
    int per_cu = 0;
    if ((s = occupancy_of(c, cfg, lds, &per_cu)) != LJ_OK) return done(s);
    if (launch_probe_multi(cfg, dq, dw, 1, 1, lds, sl.stream) != hipSuccess)
      return done(fail(c, LJ_EHIP, "warm-up launch failed"));
  }
  return done(hipStreamSynchronize(sl.stream) == hipSuccess ? LJ_OK : fail(c, LJ_EHIP, "warm-up failed"));
}

extern "C" lj_status lj_serve_start(lj_ctx* c, const lj_serve_cfg* cfg, lj_server** out) {
  if (!c || !out) return fail(c, LJ_EINVAL, "NULL argument");
  *out = nullptr;
  HIP_TRY(c, hipSetDevice(c->device));
  lj_server* S = new lj_server();
  S->c = c;
  if (cfg) S->cfg = *cfg;
  if (S->cfg.max_batch <= 0) S->cfg.max_batch = 64;
  if (S->cfg.max_inflight <= 0) S->cfg.max_inflight = 4;
  if (S->cfg.groups_cap <= 0) S->cfg.groups_cap = 64;
  if (S->cfg.max_batch > 4096 || S->cfg.max_inflight > 8) {
    delete S;
    return fail(c, LJ_EINVAL, "max_batch <= 4096 and max_inflight <= 8");
  }
  if (S->cfg.prewarm_groups < 0 || S->cfg.prewarm_groups > kMaxGroupSpace) {
    delete S;
    return fail(c, LJ_EINVAL, "prewarm_groups %lld outside [0, %lld]", (long long)S->cfg.prewarm_groups,
                (long long)kMaxGroupSpace);
  }
  if (S->cfg.streams <= 0) S->cfg.streams = 1;
  S->cfg.streams = std::min(S->cfg.streams, S->cfg.max_inflight);
  S->stats_rows = (int64_t)c->n_cu * kServeWgPerCu;
  for (int i = 0; i < S->cfg.streams; i++) {
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
      server_free(S);
      return fail(c, LJ_EHIP, "hipStreamCreate failed");
    }
    S->streams.push_back(st);
  }
  S->slots.resize((size_t)S->cfg.max_inflight);
  const size_t nb = (size_t)S->cfg.max_batch;
  for (size_t i = 0; i < S->slots.size(); i++) {
    ServeSlot& sl = S->slots[i];
    sl.stream = S->streams[i % S->streams.size()];
    sl.wgb_off = align256(sizeof(ProbeArgs) * nb);
    sl.args_bytes = sl.wgb_off + align256(sizeof(uint32_t) * (nb + S0_NKINDS));
    void* d = nullptr;
    lj_status s = LJ_OK;
    if (hipEventCreateWithFlags(&sl.e0, hipEventDefault) != hipSuccess ||
        hipEventCreateWithFlags(&sl.e1, hipEventDefault) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&sl.args_host), sl.args_bytes, hipHostMallocDefault) != hipSuccess ||
        hipHostGetDevicePointer(reinterpret_cast<void**>(&sl.args_host_dev), sl.args_host, 0) != hipSuccess)
      s = fail(c, LJ_EHIP, "serving slot setup failed");
    if (s == LJ_OK) s = dev_alloc(c, sl.args_bytes, &d);
    sl.args_dev = static_cast<char*>(d);
    if (s != LJ_OK) {
      server_free(S);
      return s;
    }
  }
  // request buffers of the announced group-space class, allocated and
  // cleared now: the loop then takes them from its free list instead of
  // allocating device and pinned host memory on the first requests (a
  // millisecond or more each, which every request behind them waits for)
  if (S->cfg.prewarm_groups > 0) {
    // (no more than the serving budget can hold in flight at once)
    int64_t nb_warm = (int64_t)S->cfg.max_inflight * S->cfg.max_batch;
    if (S->cfg.hbm_budget)
      nb_warm = std::min<int64_t>(nb_warm, (int64_t)(S->cfg.hbm_budget / serve_bytes(S, S->cfg.prewarm_groups)));
    std::vector<ReqBuf*> warm;
    FillArgs f{};
    lj_status s = LJ_OK;
    for (int64_t i = 0; i < nb_warm && s == LJ_OK; i++) {
      ReqBuf* b = nullptr;
      s = reqbuf_get(S, S->cfg.prewarm_groups, &b);
      if (s != LJ_OK) break;
      warm.push_back(b);
      if (f.n == kMaxFill) {
        if (launch_fill(f, c->stream) != hipSuccess) s = fail(c, LJ_EHIP, "fill launch failed");
        f = FillArgs{};
      }
      add_fill(f, b->dev, (int64_t)(b->l.zero / 4));
      b->fresh = false;
    }
    if (s == LJ_OK && f.n && launch_fill(f, c->stream) != hipSuccess) s = fail(c, LJ_EHIP, "fill launch failed");
    for (ReqBuf* b : warm) S->free_bufs.emplace(b->ng, b);
    if (s != LJ_OK) {
      (void)hipStreamSynchronize(c->stream);
      server_free(S);
      return s;
    }
  }
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (lj_status s = serve_warm(S); s != LJ_OK) {
    server_free(S);
    return s;
  }
  S->th = std::thread(serve_loop, S);
  *out = S;
  return LJ_OK;
}

extern "C" lj_status lj_serve_submit(lj_server* S, const lj_star_desc* q, uint64_t* id) {
  if (!S || !q || !id) return fail(nullptr, LJ_EINVAL, "NULL argument");
  // checked here, on the caller's thread (its message: lj_last_error(NULL) on
  // this thread), so that one bad request never shares a launch
  Plan p;
  lj_status st = validate_query(nullptr, q, p);
  if (st != LJ_OK) return st;
  for (int d = 0; d < q->n_dims; d++)
    if (!q->resident[d]) return fail(nullptr, LJ_EINVAL, "serving needs resident dimensions (dimension %d is not)", d);
  ServeReq rq;
  rq.desc = *q;
  rq.t_submit = Clock::now();
  rq.bytes = serve_bytes(S, q->n_groups);
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stopping) return fail(nullptr, LJ_ESTATE, "the server is stopping");
    rq.id = S->next_id++;
    *id = rq.id;
    S->pending.push_back(std::move(rq));
    S->n_pending.store(S->pending.size(), std::memory_order_release);
    S->stats.submitted++;
  }
  S->cv_work.notify_one();
  return LJ_OK;
}

extern "C" lj_status lj_serve_poll(lj_server* S, uint64_t id, lj_star_result* out, int32_t timeout_ms) {
  if (!S || !out) return LJ_EINVAL;
  std::unique_lock<std::mutex> lk(S->mu);
  auto ready = [&] { return S->done.count(id) != 0; };
  if (!ready()) {
    if (timeout_ms == 0) return LJ_ETIMEOUT;
    if (timeout_ms < 0) S->cv_done.wait(lk, ready);
    else if (!S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) return LJ_ETIMEOUT;
  }
  ServeDone& d = S->done[id];
  lj_group* g = out->groups;
  const int64_t cap = out->groups_cap;
  *out = d.r;
  out->groups = g;
  out->groups_cap = cap;
  if (g)
    for (int64_t i = 0; i < std::min<int64_t>(cap, (int64_t)d.groups.size()); i++) g[i] = d.groups[(size_t)i];
  const lj_status st = d.st;
  S->done.erase(id);
  return st;
}

extern "C" lj_status lj_serve_get_stats(lj_server* S, lj_serve_stats* out) {
  if (!S || !out) return LJ_EINVAL;
  std::lock_guard<std::mutex> lk(S->mu);
  *out = S->stats;
  return LJ_OK;
}

extern "C" lj_status lj_serve_stop(lj_server* S) {
  if (!S) return LJ_EINVAL;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stopping = true;
  }
  S->cv_work.notify_all();
  if (S->th.joinable()) S->th.join();
  server_free(S);
  return LJ_OK;
}
