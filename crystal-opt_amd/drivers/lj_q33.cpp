// lj_q33 -- SSB Q3.3 driver over the C-ABI, a drop-in for the reference
// binaries bin/ssb-jo/q33/{0..5}-{naive,lip} and 3-lip-best-bloom
// (src/ssb-jo/q33/*.cu main(), :307-371): same --t flag, same output
// ("Result:" rows "year c_city s_city revenue" in group order, "Res Count",
// "Time Taken Total", and the {"query":33,"time_query":ms} line per trial that
// eval/jo_q33.py parses).
//
//   --plan=<i>-naive | <i>-lip (i = 0..5, join orders of q33/{i}-*.cu),
//          3-lip-best-bloom, lip-adaptive          (default 0-lip)
//   --data_dir=<dir> --sf=<1..256>  read the reference's columnar SSB files
//          (raw little-endian int32, named by ssb_utils.h:106-153, e.g.
//          LINEORDER5 = lo_orderdate); otherwise synthetic seeded columns
//   --n=<fact rows>  synthetic fact size (default 120,000,000)
//   --i32  print the reference's int32 (wrapping) sums instead of int64
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "lipjoin.h"

static const char* arg_str(int argc, char** argv, const char* name) {
  const size_t ln = strlen(name);
  for (int i = 1; i < argc; i++)
    if (!strncmp(argv[i], "--", 2) && !strncmp(argv[i] + 2, name, ln) && argv[i][2 + ln] == '=') return argv[i] + 3 + ln;
  return nullptr;
}
static bool arg_flag(int argc, char** argv, const char* name) {
  for (int i = 1; i < argc; i++)
    if (!strncmp(argv[i], "--", 2) && !strcmp(argv[i] + 2, name)) return true;
  return false;
}

struct Sizes {
  int sf;
  long long lo, s, c;
};
// ssb_utils.h:8-83 (D_LEN is 2556 at every scale factor)
static const Sizes kSizes[] = {{1, 6001171, 2000, 30000},        {2, 11998051, 4000, 60000},
                               {4, 23996670, 8000, 120000},      {8, 47989129, 16000, 240000},
                               {16, 95988758, 32000, 480000},    {20, 119994746, 40000, 600000},
                               {32, 192000754, 64000, 960000},   {64, 384016864, 128000, 1920000},
                               {128, 768047048, 256000, 3840000}, {256, 1536094096, 512000, 7680000}};

static bool load_column(const std::string& path, long long n, std::vector<int32_t>& out) {
  std::ifstream f(path, std::ios::in | std::ios::binary);
  if (!f) return false;
  out.resize(n);
  f.read(reinterpret_cast<char*>(out.data()), n * 4);
  return (long long)f.gcount() == n * 4;
}

#define CHECK_LJ(ctx, x)                                                                   \
  do {                                                                                     \
    lj_status s_ = (x);                                                                    \
    if (s_ != LJ_OK) {                                                                     \
      fprintf(stderr, "%s failed: %s (%s)\n", #x, lj_status_str(s_), lj_last_error(ctx)); \
      return 1;                                                                            \
    }                                                                                      \
  } while (0)
#define CHECK_HIP(x)                                                           \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) {                                                    \
      fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));          \
      return 1;                                                                \
    }                                                                          \
  } while (0)

enum { S = 0, C = 1, D = 2 };

static void set_city_pred(lj_pred& p) {  // q33/0-lip.cu:132-135
  p.n_steps = 2;
  p.steps[0] = {LJ_COMB_SET, LJ_CMP_EQ, 231};
  p.steps[1] = {LJ_COMB_OR, LJ_CMP_EQ, 235};
}

static int64_t round_up_pow2(int64_t n) {
  int64_t p = 1;
  while (p < n) p <<= 1;
  return p;
}

int main(int argc, char** argv) {
  if (arg_flag(argc, argv, "help")) {
    printf("%s [--t=<num trials>] [--plan=<i>-naive|<i>-lip|3-lip-best-bloom|lip-adaptive] "
           "[--data_dir=<dir> --sf=<sf>] [--n=<fact rows>] [--i32] [--device=<id>]\n",
           argv[0]);
    return 0;
  }
  int num_trials = arg_str(argc, argv, "t") ? atoi(arg_str(argc, argv, "t")) : 3;
  const char* plan = arg_str(argc, argv, "plan") ? arg_str(argc, argv, "plan") : "0-lip";
  const int device = arg_str(argc, argv, "device") ? atoi(arg_str(argc, argv, "device")) : 0;
  const bool print_i32 = arg_flag(argc, argv, "i32");
  static const int orders[6][3] = {{S, C, D}, {S, D, C}, {D, S, C}, {D, C, S}, {C, D, S}, {C, S, D}};
  int method = LJ_METHOD_LIP;
  int order[3], bloom_order[3] = {-1, -1, -1};
  if (!strcmp(plan, "3-lip-best-bloom")) {
    memcpy(order, orders[3], sizeof order);
    memcpy(bloom_order, orders[0], sizeof bloom_order);
  } else if (!strcmp(plan, "lip-adaptive")) {
    memcpy(order, orders[0], sizeof order);
    method = LJ_METHOD_LIP_ADAPTIVE;
  } else if (plan[0] >= '0' && plan[0] <= '5' && plan[1] == '-') {
    memcpy(order, orders[plan[0] - '0'], sizeof order);
    method = !strcmp(plan + 2, "naive") ? LJ_METHOD_NAIVE : LJ_METHOD_LIP;
  } else {
    fprintf(stderr, "unknown plan %s\n", plan);
    return 2;
  }

  long long lo_len = 120000000, s_len = 40000, c_len = 600000;
  const long long d_len = 2556;
  const char* data_dir = arg_str(argc, argv, "data_dir");
  if (data_dir) {
    int sf = arg_str(argc, argv, "sf") ? atoi(arg_str(argc, argv, "sf")) : 20;
    bool ok = false;
    for (const Sizes& z : kSizes)
      if (z.sf == sf) {
        lo_len = z.lo, s_len = z.s, c_len = z.c;
        ok = true;
      }
    if (!ok) {
      fprintf(stderr, "unknown scale factor %d\n", sf);
      return 2;
    }
  } else if (arg_str(argc, argv, "n")) {
    lo_len = atoll(arg_str(argc, argv, "n"));
  }

  lj_ctx* ctx = nullptr;
  CHECK_LJ(nullptr, lj_create(device, 0, nullptr, &ctx));
  // column order: lo_orderdate lo_custkey lo_suppkey lo_revenue d_datekey d_year s_suppkey s_city c_custkey c_city
  const char* files[10] = {"LINEORDER5", "LINEORDER2", "LINEORDER4", "LINEORDER12", "DDATE0",
                           "DDATE4",     "SUPPLIER0",  "SUPPLIER3",  "CUSTOMER0",   "CUSTOMER3"};
  const long long lens[10] = {lo_len, lo_len, lo_len, lo_len, d_len, d_len, s_len, s_len, c_len, c_len};
  int32_t* dev[10];
  for (int i = 0; i < 10; i++) CHECK_HIP(hipMalloc(&dev[i], sizeof(int32_t) * (lens[i] ? lens[i] : 1)));
  if (data_dir) {
    std::vector<int32_t> h;
    for (int i = 0; i < 10; i++) {
      if (!load_column(std::string(data_dir) + "/" + files[i], lens[i], h)) {
        fprintf(stderr, "cannot read %s/%s (%lld int32 values)\n", data_dir, files[i], lens[i]);
        return 1;
      }
      CHECK_HIP(hipMemcpy(dev[i], h.data(), 4 * lens[i], hipMemcpyHostToDevice));
    }
    printf("** LOADED DATA **\n");
  } else {
    const uint64_t seed = 0x1A7E5EED03ull;
    CHECK_LJ(ctx, lj_gen_device(ctx, LJ_GEN_DATEKEY, seed, 12, 0, d_len, 0, lo_len, dev[0]));
    CHECK_LJ(ctx, lj_gen_device(ctx, LJ_GEN_UNIFORM, seed, 10, 1, c_len, 0, lo_len, dev[1]));
    CHECK_LJ(ctx, lj_gen_device(ctx, LJ_GEN_UNIFORM, seed, 11, 1, s_len, 0, lo_len, dev[2]));
    CHECK_LJ(ctx, lj_gen_device(ctx, LJ_GEN_UNIFORM, seed, 14, 1, 1000000, 0, lo_len, dev[3]));
    CHECK_LJ(ctx, lj_gen_device(ctx, LJ_GEN_DATEKEY_SEQ, seed, 0, 0, 0, 0, d_len, dev[4]));
    CHECK_LJ(ctx, lj_gen_device(ctx, LJ_GEN_YEAR_OF_DAY, seed, 0, 0, 0, 0, d_len, dev[5]));
    CHECK_LJ(ctx, lj_gen_device(ctx, LJ_GEN_SEQUENCE, seed, 0, 1, 0, 0, s_len, dev[6]));
    CHECK_LJ(ctx, lj_gen_device(ctx, LJ_GEN_UNIFORM, seed, 2, 0, 249, 0, s_len, dev[7]));
    CHECK_LJ(ctx, lj_gen_device(ctx, LJ_GEN_SEQUENCE, seed, 0, 1, 0, 0, c_len, dev[8]));
    CHECK_LJ(ctx, lj_gen_device(ctx, LJ_GEN_UNIFORM, seed, 1, 0, 249, 0, c_len, dev[9]));
    CHECK_LJ(ctx, lj_sync(ctx));
    printf("** GENERATED DATA **\n");
  }
  printf("** LOADED DATA TO GPU **\n");

  lj_star_desc q;
  memset(&q, 0, sizeof q);
  q.fact_fk[S] = dev[2];
  q.fact_fk[C] = dev[1];
  q.fact_fk[D] = dev[0];
  q.fact_measure = dev[3];
  q.n_fact = lo_len;
  q.n_dims = 3;
  q.method = method;
  for (int i = 0; i < 4; i++) {
    q.order[i] = i < 3 ? order[i] : i;
    q.bloom_order[i] = i < 3 ? bloom_order[i] : -1;
  }
  const int32_t d_val_len = 19981230 - 19920101 + 1;
  lj_dim_desc& ds = q.dims[S];
  ds.keys = dev[6], ds.attr = dev[7], ds.n_rows = s_len, ds.ht_len = s_len;
  set_city_pred(ds.pred);
  lj_dim_desc& dc = q.dims[C];
  dc.keys = dev[8], dc.attr = dev[9], dc.n_rows = c_len, dc.ht_len = c_len;
  set_city_pred(dc.pred);
  lj_dim_desc& dd = q.dims[D];
  dd.keys = dev[4], dd.attr = dev[5], dd.n_rows = d_len, dd.ht_len = d_val_len, dd.keys_min = 19920101;
  dd.pred.n_steps = 2;  // q33/0-lip.cu:192-195: GTE 1992, then LTE 1997 overwrites
  dd.pred.steps[0] = {LJ_COMB_SET, LJ_CMP_GTE, 1992};
  dd.pred.steps[1] = {LJ_COMB_SET, LJ_CMP_LTE, 1997};
  if (method != LJ_METHOD_NAIVE) {
    ds.bloom_hash = dc.bloom_hash = dd.bloom_hash = LJ_BLOOM_MOD;
    ds.bloom_k = dc.bloom_k = dd.bloom_k = 1;
    ds.bloom_bits = round_up_pow2(s_len);   // q33/0-lip.cu:221-223
    dc.bloom_bits = round_up_pow2(c_len);
    dd.bloom_bits = round_up_pow2(d_val_len);
  }
  q.group_mult[S] = 250 * 7;  // q33/0-lip.cu:102-104
  q.group_mult[C] = 7;
  q.group_mult[D] = 1;
  q.group_offset = -1992;
  q.n_groups = (1998 - 1992 + 1) * 250 * 250;

  std::vector<lj_group> groups(1 << 16);
  for (int t = 0; t < num_trials; t++) {
    auto st = std::chrono::high_resolution_clock::now();
    lj_star_result r;
    memset(&r, 0, sizeof r);
    r.groups = groups.data();
    r.groups_cap = (int64_t)groups.size();
    CHECK_LJ(ctx, lj_star_join(ctx, &q, &r));
    auto fin = std::chrono::high_resolution_clock::now();
    printf("Result:\n");
    for (int64_t i = 0; i < r.n_groups && i < r.groups_cap; i++) {
      const lj_group& g = groups[i];
      if (print_i32)
        printf("%d %d %d %d\n", g.payload[D], g.payload[C], g.payload[S], g.sum_i32);
      else
        printf("%d %d %d %lld\n", g.payload[D], g.payload[C], g.payload[S], (long long)g.sum);
    }
    printf("Res Count: %lld\n", (long long)r.n_groups);
    printf("Time Taken Total: %g\n", std::chrono::duration<double>(fin - st).count() * 1000);
    printf("{\"query\":33,\"time_query\":%g}\n", r.time_query_ms);
  }
  for (int i = 0; i < 10; i++) (void)hipFree(dev[i]);
  lj_destroy(ctx);
  return 0;
}
