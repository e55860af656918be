// lj_join -- single FK-PK join driver over the C-ABI, a drop-in for the
// reference binaries bin/ops/join and bin/ops/bloom_join:
//   src/ops/join.cu:163-238        (no --b: plain join)
//   src/ops/bloom_join.cu:171-249  (--b=<bits>: Bloom pre-filter, LIP method)
// Same flags (--n --d --t --s --b --device --help) and the same JSON line per
// trial ({"num_dim",...,"time_join_total"}).
//
// Inputs: by default the reference's own generator semantics
// (src/ops/utils/generator.h:37-136): srand(0); dimension keys 1..d in a
// shuffled order (j uniform in [0, i) per step, RAND_RANGE), values 0..d-1;
// fact keys as successive shuffled blocks of 1..d (and a shuffled 1..(n%d)
// tail), fact values 0..n-1; predicate dim_val <= num_select.  The product
// sum of join.cu:96 is reported as "sum".  --gen=seeded uses the counter-based
// generator instead (dimension 0..d-1 keys, attr uniform).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "lipjoin.h"

static bool arg_int(int argc, char** argv, const char* name, long long* v) {
  const size_t ln = strlen(name);
  for (int i = 1; i < argc; i++)
    if (!strncmp(argv[i], "--", 2) && !strncmp(argv[i] + 2, name, ln) && argv[i][2 + ln] == '=') {
      *v = atoll(argv[i] + 3 + ln);
      return true;
    }
  return false;
}

static bool arg_flag(int argc, char** argv, const char* name) {
  for (int i = 1; i < argc; i++)
    if (!strncmp(argv[i], "--", 2) && !strcmp(argv[i] + 2, name)) return true;
  return false;
}

static const char* arg_str(int argc, char** argv, const char* name) {
  const size_t ln = strlen(name);
  for (int i = 1; i < argc; i++)
    if (!strncmp(argv[i], "--", 2) && !strncmp(argv[i] + 2, name, ln) && argv[i][2 + ln] == '=') return argv[i] + 3 + ln;
  return nullptr;
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

int main(int argc, char** argv) {
  long long num_fact = 1ll << 30, num_dim = 16ll << 20, num_trials = 3, device = 0, bits = 0;
  arg_int(argc, argv, "n", &num_fact);
  arg_int(argc, argv, "d", &num_dim);
  arg_int(argc, argv, "t", &num_trials);
  long long num_select = num_dim / 128;
  arg_int(argc, argv, "s", &num_select);
  const bool bloom = arg_int(argc, argv, "b", &bits);
  arg_int(argc, argv, "device", &device);
  const char* gen = arg_str(argc, argv, "gen");
  const bool seeded = gen && !strcmp(gen, "seeded");
  const char* method_s = arg_str(argc, argv, "method");
  if (arg_flag(argc, argv, "help")) {
    printf("%s [--n=<num fact>] [--d=<num dim>] [--t=<num trials>] [--s=<num select>] "
           "[--b=<num bloom filter bits>] [--method=naive|lip|lip-adaptive] [--gen=reference|seeded] "
           "[--device=<device-id>] [--v]\n",
           argv[0]);
    return 0;
  }
  int32_t method = bloom ? LJ_METHOD_LIP : LJ_METHOD_NAIVE;
  if (method_s) method = !strcmp(method_s, "naive") ? LJ_METHOD_NAIVE : !strcmp(method_s, "lip") ? LJ_METHOD_LIP : LJ_METHOD_LIP_ADAPTIVE;
  printf("Running with %lld fact, %lld dim, num_select: %lld, num_bloom_filter_bits: %lld, %lld trials\n", num_fact,
         num_dim, num_select, bits, num_trials);

  lj_ctx* ctx = nullptr;
  CHECK_LJ(nullptr, lj_create((int32_t)device, 0, nullptr, &ctx));
  int32_t *d_dim_key, *d_dim_val, *d_fact_fkey, *d_fact_val;
  CHECK_HIP(hipMalloc(&d_dim_key, sizeof(int32_t) * num_dim));
  CHECK_HIP(hipMalloc(&d_dim_val, sizeof(int32_t) * num_dim));
  CHECK_HIP(hipMalloc(&d_fact_fkey, sizeof(int32_t) * num_fact));
  CHECK_HIP(hipMalloc(&d_fact_val, sizeof(int32_t) * num_fact));
  if (!seeded) {
    // create_relation_pk / create_relation_fk after srand(0) (generator.h:37-136)
    std::vector<int32_t> k(num_dim), v(num_dim), fk(num_fact), fv(num_fact);
    CHECK_LJ(nullptr, lj_gen_reference_host(num_dim, num_fact, k.data(), v.data(), fk.data(), fv.data()));
    CHECK_HIP(hipMemcpy(d_dim_key, k.data(), 4 * num_dim, hipMemcpyHostToDevice));
    CHECK_HIP(hipMemcpy(d_dim_val, v.data(), 4 * num_dim, hipMemcpyHostToDevice));
    CHECK_HIP(hipMemcpy(d_fact_fkey, fk.data(), 4 * num_fact, hipMemcpyHostToDevice));
    CHECK_HIP(hipMemcpy(d_fact_val, fv.data(), 4 * num_fact, hipMemcpyHostToDevice));
  } else {
    std::vector<int32_t> k(num_dim);
    for (long long i = 0; i < num_dim; i++) k[i] = (int32_t)i;
    CHECK_LJ(ctx, lj_gen_shuffle_host(0x1A7E5EED02ull, 0, num_dim, k.data()));
    CHECK_HIP(hipMemcpy(d_dim_key, k.data(), 4 * num_dim, hipMemcpyHostToDevice));
    CHECK_LJ(ctx, lj_gen_device(ctx, LJ_GEN_UNIFORM, 0x1A7E5EED02ull, 1, 0, 9999, 0, num_dim, d_dim_val));
    CHECK_LJ(ctx, lj_gen_device(ctx, LJ_GEN_UNIFORM, 0x1A7E5EED02ull, 2, 0, num_dim - 1, 0, num_fact, d_fact_fkey));
    CHECK_LJ(ctx, lj_gen_device(ctx, LJ_GEN_UNIFORM, 0x1A7E5EED02ull, 3, 1, 10000, 0, num_fact, d_fact_val));
    CHECK_LJ(ctx, lj_sync(ctx));
  }

  lj_join_desc j;
  memset(&j, 0, sizeof j);
  j.fact_fk = d_fact_fkey;
  j.fact_measure = d_fact_val;
  j.n_fact = num_fact;
  j.method = method;
  j.dim.keys = d_dim_key;
  j.dim.attr = d_dim_val;
  j.dim.payload = d_dim_val;
  j.dim.n_rows = num_dim;
  j.dim.pred.n_steps = 1;
  j.dim.pred.steps[0].comb = LJ_COMB_SET;
  j.dim.pred.steps[0].cmp = seeded ? LJ_CMP_LT : LJ_CMP_LTE;   // bloom_join.cu:61: dim_val <= num_select
  j.dim.pred.steps[0].value = (int32_t)num_select;
  j.dim.ht_kind = LJ_HT_PHT2;
  j.dim.ht_len = num_dim;
  if (bloom) {
    j.dim.bloom_hash = LJ_BLOOM_MOD;
    j.dim.bloom_k = 1;
    j.dim.bloom_bits = bits;
  }
  for (long long t = 0; t < num_trials; t++) {
    lj_join_result r;
    CHECK_LJ(ctx, lj_hash_join(ctx, &j, &r));
    printf("{\"num_dim\":%lld,\"num_fact\":%lld,\"num_select\":%lld,\"num_bloom_filter_bits\":%lld,"
           "\"time_build\":%g,\"time_probe\":%g,\"time_extra\":%g,\"time_join_total\":%g,"
           "\"n_match\":%lld,\"sum\":%llu,\"sum_measure\":%lld}\n",
           num_dim, num_fact, num_select, bits, r.time_build, r.time_probe, r.time_extra, r.time_total,
           (long long)r.n_match, (unsigned long long)r.sum_product, (long long)r.sum_measure);
  }
  (void)hipFree(d_dim_key);
  (void)hipFree(d_dim_val);
  (void)hipFree(d_fact_fkey);
  (void)hipFree(d_fact_val);
  lj_destroy(ctx);
  return 0;
}
