/*
 * oracle.c -- CPU restatement of the reference's star-join hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker and the CPU baseline).  The
 * shipped library never links or calls it.  Semantics follow the reference
 * files cited per function; the fact-table loops are plain row-at-a-time C,
 * optionally split over pthreads by row ranges (results are integer sums, so
 * the split does not change them).
 */
#include "oracle.h"

#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ssb_utils.h:85-96 (roundUpToPowerOfTwo) */
uint32_t orc_round_up_pow2(uint32_t n) {
  if (n == 0) return 1;
  n--;
  n |= n >> 1;
  n |= n >> 2;
  n |= n >> 4;
  n |= n >> 8;
  n |= n >> 16;
  n++;
  return n;
}

/* crystal/pred.cuh:35-356.  BlockPred* overwrites, BlockPredAnd* ANDs,
 * BlockPredOr* ORs the functor result into the flags. */
static int cmp_eval(int32_t cmp, int32_t a, int32_t c) {
  switch (cmp) {
    case ORC_CMP_LT: return a < c;   /* LessThan      pred.cuh:166-177 */
    case ORC_CMP_GT: return a > c;   /* GreaterThan   pred.cuh:180-191 */
    case ORC_CMP_LTE: return a <= c; /* LessThanEq    pred.cuh:194-205 */
    case ORC_CMP_GTE: return a >= c; /* GreaterThanEq pred.cuh:208-219 */
    case ORC_CMP_EQ: return a == c;  /* Eq            pred.cuh:222-233 */
  }
  return 0;
}

void orc_predicate(const int32_t* attr, int64_t n, const orc_pred_step* steps,
                   int32_t n_steps, uint8_t* flags) {
  for (int64_t i = 0; i < n; i++) {
    int f = 1; /* InitFlags, pred.cuh:25-33 */
    for (int32_t s = 0; s < n_steps; s++) {
      int r = cmp_eval(steps[s].cmp, attr[i], steps[s].value);
      if (steps[s].comb == ORC_COMB_SET) f = r;
      else if (steps[s].comb == ORC_COMB_AND) f = f && r;
      else f = f || r;
    }
    flags[i] = (uint8_t)f;
  }
}

/* crystal/join.cuh:27  HASH(X,Y,Z) ((X-Z) % Y), C truncating remainder. */
int64_t orc_hash_pht(int32_t key, int64_t len, int32_t keys_min) {
  int64_t x = (int64_t)key - (int64_t)keys_min;
  if (len <= 0) return -1;
  int64_t h = x % len;
  return h < 0 ? -1 : h;
}

/* join.cuh:191-258 */
void orc_pht1_build(const int32_t* keys, const uint8_t* flags, int64_t n,
                    int32_t* ht, int64_t len, int32_t keys_min, int64_t* n_oob) {
  int64_t oob = 0;
  for (int64_t i = 0; i < n; i++) {
    if (flags && !flags[i]) continue;
    int64_t h = orc_hash_pht(keys[i], len, keys_min);
    if (h < 0) { oob++; continue; }
    if (ht[h] == 0) ht[h] = keys[i]; /* atomicCAS(&ht[h], 0, key) */
  }
  if (n_oob) *n_oob += oob;
}

/* join.cuh:29-106 */
int orc_pht1_probe(const int32_t* ht, int64_t len, int32_t keys_min, int32_t key) {
  int64_t h = orc_hash_pht(key, len, keys_min);
  if (h < 0) return 0;
  return ht[h] != 0;
}

/* join.cuh:260-335 */
void orc_pht2_build(const int32_t* keys, const int32_t* vals, const uint8_t* flags,
                    int64_t n, int32_t* ht, int64_t len, int32_t keys_min,
                    int64_t* n_oob) {
  int64_t oob = 0;
  for (int64_t i = 0; i < n; i++) {
    if (flags && !flags[i]) continue;
    int64_t h = orc_hash_pht(keys[i], len, keys_min);
    if (h < 0) { oob++; continue; }
    if (ht[2 * h] == 0) ht[2 * h] = keys[i]; /* atomicCAS(&ht[h<<1], 0, key)  join.cuh:275 */
    ht[2 * h + 1] = vals[i];                 /* unconditional store          join.cuh:276 */
  }
  if (n_oob) *n_oob += oob;
}

/* join.cuh:108-189: slot = *(uint64_t*)&ht[h<<1]; hit iff slot != 0, payload = slot>>32 */
int orc_pht2_probe(const int32_t* ht, int64_t len, int32_t keys_min, int32_t key,
                   int32_t* payload) {
  int64_t h = orc_hash_pht(key, len, keys_min);
  if (h < 0) return 0;
  uint64_t slot = (uint64_t)(uint32_t)ht[2 * h] | ((uint64_t)(uint32_t)ht[2 * h + 1] << 32);
  if (slot == 0) return 0;
  if (payload) *payload = (int32_t)(uint32_t)(slot >> 32);
  return 1;
}

/* ---- textbook open addressing (project-defined, see oracle.h) ------------ */
int64_t orc_tb_home(int32_t key, int64_t cap) {
  int lg = 0;
  while (((int64_t)1 << lg) < cap) lg++;
  if (lg == 0) return 0;
  return (int64_t)((uint32_t)((uint32_t)key * 0x9E3779B1u) >> (32 - lg));
}

void orc_tb_build(const int32_t* keys, const int32_t* vals, const uint8_t* flags, int64_t n,
                  int32_t* ht, int64_t cap, int64_t* n_oob) {
  int64_t oob = 0;
  for (int64_t i = 0; i < n; i++) {
    if (flags && !flags[i]) continue;
    if (keys[i] == INT32_MIN) { oob++; continue; }
    const int32_t enc = (int32_t)((uint32_t)keys[i] ^ 0x80000000u);
    int64_t s = orc_tb_home(keys[i], cap), j = 0;
    for (; j < cap; j++, s = (s + 1) & (cap - 1)) {
      if (ht[2 * s] == 0) ht[2 * s] = enc;
      if (ht[2 * s] == enc) { ht[2 * s + 1] = vals[i]; break; }
    }
    if (j == cap) oob++; /* table full */
  }
  if (n_oob) *n_oob += oob;
}

int orc_tb_probe(const int32_t* ht, int64_t cap, int32_t key, int32_t* payload) {
  if (key == INT32_MIN) return 0;
  const int32_t enc = (int32_t)((uint32_t)key ^ 0x80000000u);
  int64_t s = orc_tb_home(key, cap);
  for (int64_t j = 0; j < cap; j++, s = (s + 1) & (cap - 1)) {
    if (ht[2 * s] == enc) {
      if (payload) *payload = ht[2 * s + 1];
      return 1;
    }
    if (ht[2 * s] == 0) return 0;
  }
  return 0;
}

/* ---- Bloom filters ------------------------------------------------------- */
/* Odd 32-bit multipliers for the project-defined multiplicative family:
 * 0x9E3779B1 is Knuth's golden-ratio multiplier (TAOCP 6.4); the others are
 * the 32-bit xxHash primes.  All odd, so x -> x*A mod 2^32 is a bijection. */
static const uint32_t kMult[4] = {0x9E3779B1u, 0x85EBCA77u, 0xC2B2AE3Du, 0x27D4EB2Fu};

static int ilog2_u64(uint64_t v) {
  int r = -1;
  while (v) { v >>= 1; r++; }
  return r;
}

int64_t orc_bloom_nwords(const orc_bloom* b) {
  if (b->hash == ORC_BLOOM_MOD) return (b->nbits + 31) / 32; /* bloom_join.cu:128 */
  if (b->hash == ORC_BLOOM_MULT) return b->nbits / 32;
  return 0;
}

int orc_bloom_bits(const orc_bloom* b, int32_t key, int64_t* bits_out) {
  if (b->hash == ORC_BLOOM_MOD) {
    /* bloom.cuh:7  HASH1(X,Y,Z) ((X-Z) % (Y)) with Y = words*32 (bloom.cuh:24) */
    int64_t m = orc_bloom_nwords(b) * 32;
    int64_t x = (int64_t)key - (int64_t)b->keys_min;
    if (m <= 0) return 0;
    int64_t h = x % m;
    if (h < 0) return 0;
    bits_out[0] = h;
    return 1;
  }
  if (b->hash == ORC_BLOOM_MULT) {
    /* Project-defined: x = key - keys_min (mod 2^32), h_i = (x*A_i mod 2^32) >> (32 - log2 m). */
    uint32_t x = (uint32_t)key - (uint32_t)b->keys_min;
    int lg = ilog2_u64((uint64_t)b->nbits);
    if (b->layout == ORC_LAYOUT_STANDARD) {
      for (int i = 0; i < b->k; i++) bits_out[i] = (int64_t)((uint32_t)(x * kMult[i]) >> (32 - lg));
    } else {
      /* BLOCKED: all k bits in one 32-bit word; word from A_0, bits from A_1..A_k. */
      int lgw = lg - 5;
      uint32_t w = (uint32_t)(x * kMult[0]) >> (32 - lgw);
      for (int i = 0; i < b->k; i++)
        bits_out[i] = (int64_t)w * 32 + (int64_t)((uint32_t)(x * kMult[i + 1]) >> 27);
    }
    return b->k;
  }
  return 0;
}

/* bloom.cuh:10-88 (MOD) and the project-defined MULT family */
void orc_bloom_build(const orc_bloom* b, const int32_t* keys, const uint8_t* flags,
                     int64_t n, uint32_t* words, int64_t* n_oob) {
  int64_t oob = 0, bits[4];
  for (int64_t i = 0; i < n; i++) {
    if (flags && !flags[i]) continue;
    int k = orc_bloom_bits(b, keys[i], bits);
    if (k == 0) { oob++; continue; }
    for (int j = 0; j < k; j++) words[bits[j] >> 5] |= 1u << (bits[j] & 31);
  }
  if (n_oob) *n_oob += oob;
}

/* bloom.cuh:90-174: flag = (bf[h>>5] & (1<<(h&31))) != 0 for every hash */
int orc_bloom_probe(const orc_bloom* b, const uint32_t* words, int32_t key) {
  int64_t bits[4];
  int k = orc_bloom_bits(b, key, bits);
  if (k == 0) return 0;
  for (int j = 0; j < k; j++)
    if (!(words[bits[j] >> 5] & (1u << (bits[j] & 31)))) return 0;
  return 1;
}

/* ---- dimension build shared by both drivers ------------------------------ */
typedef struct {
  int32_t* ht;     /* 2 * ht_len int32 */
  uint32_t* bloom; /* nwords */
  int64_t n_selected, n_oob;
} built_dim;

/* build_kernel (join.cu:43-63, bloom_join.cu:43-65) and build_hashtable_{s,c,d}
 * (q33/0-lip.cu:114-203): predicate -> Bloom (if any) -> PHT_2(key -> payload);
 * PHT_1 (q21/0-lip.cu:135) or the textbook table by the dimension's kind. */
static int build_dim(const orc_dim* d, built_dim* out) {
  memset(out, 0, sizeof(*out));
  uint8_t* flags = (uint8_t*)malloc(d->n > 0 ? (size_t)d->n : 1);
  out->ht = (int32_t*)calloc((size_t)(2 * d->ht_len + 2), sizeof(int32_t));
  int64_t nw = orc_bloom_nwords(&d->bloom);
  out->bloom = (uint32_t*)calloc((size_t)(nw + 1), sizeof(uint32_t));
  if (!flags || !out->ht || !out->bloom) return -1;
  orc_predicate(d->attr, d->n, d->pred, d->n_pred, flags);
  for (int64_t i = 0; i < d->n; i++) out->n_selected += flags[i];
  if (d->bloom.hash != ORC_BLOOM_NONE) orc_bloom_build(&d->bloom, d->keys, flags, d->n, out->bloom, &out->n_oob);
  const int32_t* vals = d->payload ? d->payload : d->attr;
  if (d->ht_kind == ORC_HT_PHT1) orc_pht1_build(d->keys, flags, d->n, out->ht, d->ht_len, d->keys_min, &out->n_oob);
  else if (d->ht_kind == ORC_HT_TEXTBOOK) orc_tb_build(d->keys, vals, flags, d->n, out->ht, d->ht_len, &out->n_oob);
  else orc_pht2_build(d->keys, vals, flags, d->n, out->ht, d->ht_len, d->keys_min, &out->n_oob);
  free(flags);
  return 0;
}

/* The hash-table probe of a dimension of any kind (PHT_1 has no payload: 0). */
static int probe_dim(const orc_dim* d, const built_dim* b, int32_t key, int32_t* pay) {
  if (d->ht_kind == ORC_HT_PHT1) {
    *pay = 0;
    return orc_pht1_probe(b->ht, d->ht_len, d->keys_min, key);
  }
  if (d->ht_kind == ORC_HT_TEXTBOOK) return orc_tb_probe(b->ht, d->ht_len, key, pay);
  return orc_pht2_probe(b->ht, d->ht_len, d->keys_min, key, pay);
}

/* A fact key outside the table's domain (the reference indexes out of bounds). */
static int key_oob(const orc_dim* d, int32_t key) {
  if (d->ht_kind == ORC_HT_TEXTBOOK) return key == INT32_MIN;
  return orc_hash_pht(key, d->ht_len, d->keys_min) < 0;
}

static void free_dim(built_dim* b) {
  free(b->ht);
  free(b->bloom);
}

/* ---- single join --------------------------------------------------------- */
typedef struct {
  const orc_dim* dim;
  const built_dim* bd;
  const int32_t* fk;
  const int32_t* measure;
  int64_t lo, hi;
  int32_t method;
  orc_join_result r;
} sj_task;

/* probe_kernel (join.cu:65-108; bloom_join.cu:67-112), row-at-a-time. */
static void* sj_worker(void* arg) {
  sj_task* t = (sj_task*)arg;
  const orc_dim* d = t->dim;
  int use_bloom = (t->method == ORC_METHOD_LIP) && d->bloom.hash != ORC_BLOOM_NONE;
  int line_alive = 0;
  for (int64_t i = t->lo; i < t->hi; i++) {
    if ((i & 31) == 0) {
      t->r.lines_fk++;
      line_alive = 0;
    }
    int32_t key = t->fk[i];
    if (use_bloom && !orc_bloom_probe(&d->bloom, t->bd->bloom, key)) continue;
    t->r.n_bloom_pass++;
    if (key_oob(d, key)) t->r.n_oob++;
    int32_t pay = 0;
    if (!probe_dim(d, t->bd, key, &pay)) continue;
    t->r.n_match++;
    int32_t m = t->measure[i];
    t->r.sum_measure += m;
    /* join.cu:96: `sum += vals[ITEM] * join_vals[ITEM]`: a 32-bit int product
     * (wrapping) converted to unsigned long long (sign extension). */
    int32_t prod = (int32_t)((uint32_t)m * (uint32_t)pay);
    t->r.sum_product += (uint64_t)(int64_t)prod;
    if (!line_alive) {
      t->r.lines_measure++;
      line_alive = 1;
    }
  }
  return NULL;
}

static int64_t split_point(int64_t n, int32_t parts, int32_t p) {
  /* chunk boundaries rounded to 32 rows so that line statistics do not double count */
  int64_t lines = (n + 31) / 32;
  int64_t b = (lines * p / parts) * 32;
  return b > n ? n : b;
}

int orc_single_join(const orc_dim* dim, const int32_t* fact_fk, const int32_t* fact_measure,
                    int64_t n_fact, int32_t method, int32_t n_threads, orc_join_result* out) {
  built_dim bd;
  memset(out, 0, sizeof(*out));
  if (build_dim(dim, &bd) != 0) return -1;
  if (n_threads < 1) n_threads = 1;
  sj_task* tasks = (sj_task*)calloc((size_t)n_threads, sizeof(sj_task));
  pthread_t* th = (pthread_t*)calloc((size_t)n_threads, sizeof(pthread_t));
  for (int32_t p = 0; p < n_threads; p++) {
    tasks[p].dim = dim;
    tasks[p].bd = &bd;
    tasks[p].fk = fact_fk;
    tasks[p].measure = fact_measure;
    tasks[p].lo = split_point(n_fact, n_threads, p);
    tasks[p].hi = split_point(n_fact, n_threads, p + 1);
    tasks[p].method = method;
    if (n_threads > 1) pthread_create(&th[p], NULL, sj_worker, &tasks[p]);
    else sj_worker(&tasks[p]);
  }
  for (int32_t p = 0; p < n_threads; p++) {
    if (n_threads > 1) pthread_join(th[p], NULL);
    out->n_bloom_pass += tasks[p].r.n_bloom_pass;
    out->n_match += tasks[p].r.n_match;
    out->sum_measure += tasks[p].r.sum_measure;
    out->sum_product += tasks[p].r.sum_product;
    out->n_oob += tasks[p].r.n_oob;
    out->lines_fk += tasks[p].r.lines_fk;
    out->lines_measure += tasks[p].r.lines_measure;
  }
  out->n_selected = bd.n_selected;
  out->n_oob += bd.n_oob;
  free(tasks);
  free(th);
  free_dim(&bd);
  return 0;
}

/* ---- star join ----------------------------------------------------------- */
typedef struct {
  const orc_dim* dims;
  const built_dim* bd;
  int32_t n_dims;
  const orc_fact* fact;
  const int32_t* order;
  const int32_t* bloom_order;
  int32_t method;
  const int64_t* gmult;
  int64_t goffset, n_groups;
  int64_t lo, hi;
  int64_t* gsum;
  int64_t* gcount;
  int32_t* gpay;
  int64_t* glast; /* row index of the last write to gpay (for the merge) */
  int64_t lines[ORC_MAX_DIMS + 1];
  int64_t rows[ORC_MAX_DIMS + 2];
  int64_t n_oob;
} star_task;

/* probe<128,4> of q33/{i}-naive.cu:28-91 and q33/{i}-lip.cu:28-112, row at a time:
 * NAIVE: per dim in join order, (Pred)Load FK -> PHT_2 probe;
 * LIP:   per dim in Bloom order, (Pred)Load FK -> Bloom probe; then PHT_2 per dim
 *        in join order on the FKs already loaded; then the measure and the group add. */
static void* star_worker(void* arg) {
  star_task* t = (star_task*)arg;
  const int32_t nd = t->n_dims;
  const int32_t* load_order = (t->method == ORC_METHOD_LIP) ? t->bloom_order : t->order;
  int line_alive[ORC_MAX_DIMS + 1];
  for (int64_t i = t->lo; i < t->hi; i++) {
    if ((i & 31) == 0)
      for (int s = 0; s <= nd; s++) line_alive[s] = 0;
    int32_t key[ORC_MAX_DIMS], pay[ORC_MAX_DIMS];
    int alive = 1;
    for (int s = 0; s < nd && alive; s++) {
      int32_t d = load_order[s];
      t->rows[s]++;
      if (!line_alive[s]) { t->lines[s]++; line_alive[s] = 1; }
      key[d] = t->fact->fk[d][i];
      if (key_oob(&t->dims[d], key[d])) t->n_oob++;
      if (t->method == ORC_METHOD_LIP) {
        if (t->dims[d].bloom.hash != ORC_BLOOM_NONE)
          alive = orc_bloom_probe(&t->dims[d].bloom, t->bd[d].bloom, key[d]);
      } else {
        alive = probe_dim(&t->dims[d], &t->bd[d], key[d], &pay[d]);
      }
    }
    if (alive && t->method == ORC_METHOD_LIP) {
      t->rows[nd + 1]++; /* passed every Bloom filter */
      for (int s = 0; s < nd && alive; s++) {
        int32_t d = t->order[s];
        alive = probe_dim(&t->dims[d], &t->bd[d], key[d], &pay[d]);
      }
    }
    if (!alive) continue;
    t->rows[nd]++;
    if (!line_alive[nd]) { t->lines[nd]++; line_alive[nd] = 1; }
    int32_t m = t->fact->measure[i];
    /* q33/0-lip.cu:102-104: (s*250*7 + c*7 + (year-1992)) % (7*250*250) */
    int64_t acc = t->goffset;
    for (int d = 0; d < nd; d++) acc += (int64_t)pay[d] * t->gmult[d];
    int64_t g = acc % t->n_groups;
    if (g < 0) { t->n_oob++; continue; }
    t->gsum[g] += m;
    t->gcount[g] += 1;
    for (int d = 0; d < nd; d++) t->gpay[g * ORC_MAX_DIMS + d] = pay[d];
    t->glast[g] = i;
  }
  return NULL;
}

int64_t orc_star_join(const orc_dim* dims, int32_t n_dims, const orc_fact* fact,
                      const int32_t* order, const int32_t* bloom_order, int32_t method,
                      const int64_t* gmult, int64_t goffset, int64_t n_groups,
                      int64_t* gsum, int64_t* gcount, int32_t* gpay, int64_t* lines_out,
                      int64_t* rows_out, int64_t* n_oob, int32_t n_threads) {
  if (n_dims < 1 || n_dims > ORC_MAX_DIMS || n_groups < 1) return -1;
  built_dim bd[ORC_MAX_DIMS];
  int64_t build_oob = 0;
  for (int d = 0; d < n_dims; d++) {
    if (build_dim(&dims[d], &bd[d]) != 0) return -1;
    build_oob += bd[d].n_oob;
  }
  if (!bloom_order) bloom_order = order;
  if (n_threads < 1) n_threads = 1;
  star_task* tasks = (star_task*)calloc((size_t)n_threads, sizeof(star_task));
  pthread_t* th = (pthread_t*)calloc((size_t)n_threads, sizeof(pthread_t));
  for (int32_t p = 0; p < n_threads; p++) {
    star_task* t = &tasks[p];
    t->dims = dims;
    t->bd = bd;
    t->n_dims = n_dims;
    t->fact = fact;
    t->order = order;
    t->bloom_order = bloom_order;
    t->method = method;
    t->gmult = gmult;
    t->goffset = goffset;
    t->n_groups = n_groups;
    t->lo = split_point(fact->n, n_threads, p);
    t->hi = split_point(fact->n, n_threads, p + 1);
    if (p == 0) {
      t->gsum = gsum;
      t->gcount = gcount;
      t->gpay = gpay;
      memset(gsum, 0, sizeof(int64_t) * (size_t)n_groups);
      memset(gcount, 0, sizeof(int64_t) * (size_t)n_groups);
      memset(gpay, 0, sizeof(int32_t) * (size_t)n_groups * ORC_MAX_DIMS);
    } else {
      t->gsum = (int64_t*)calloc((size_t)n_groups, sizeof(int64_t));
      t->gcount = (int64_t*)calloc((size_t)n_groups, sizeof(int64_t));
      t->gpay = (int32_t*)calloc((size_t)n_groups * ORC_MAX_DIMS, sizeof(int32_t));
    }
    t->glast = (int64_t*)malloc(sizeof(int64_t) * (size_t)n_groups);
    for (int64_t g = 0; g < n_groups; g++) t->glast[g] = -1;
    if (n_threads > 1) pthread_create(&th[p], NULL, star_worker, t);
    else star_worker(t);
  }
  int64_t oob = build_oob;
  for (int32_t p = 0; p < n_threads; p++) {
    if (n_threads > 1) pthread_join(th[p], NULL);
  }
  for (int s = 0; s <= ORC_MAX_DIMS + 1; s++) {
    if (lines_out && s <= ORC_MAX_DIMS) lines_out[s] = 0;
    if (rows_out) rows_out[s] = 0;
  }
  for (int32_t p = 0; p < n_threads; p++) {
    star_task* t = &tasks[p];
    for (int s = 0; s <= n_dims; s++) {
      if (lines_out) lines_out[s] += t->lines[s];
      if (rows_out) rows_out[s] += t->rows[s];
    }
    if (rows_out) rows_out[n_dims + 1] += t->rows[n_dims + 1];
    oob += t->n_oob;
    if (p > 0) {
      /* later row ranges are later writers: their payloads win (plain stores) */
      for (int64_t g = 0; g < n_groups; g++) {
        gsum[g] += t->gsum[g];
        gcount[g] += t->gcount[g];
        if (t->glast[g] >= 0)
          for (int d = 0; d < ORC_MAX_DIMS; d++) gpay[g * ORC_MAX_DIMS + d] = t->gpay[g * ORC_MAX_DIMS + d];
      }
      free(t->gsum);
      free(t->gcount);
      free(t->gpay);
    }
    free(t->glast);
  }
  int64_t hit = 0;
  for (int64_t g = 0; g < n_groups; g++) hit += gcount[g] != 0;
  if (n_oob) *n_oob = oob;
  free(tasks);
  free(th);
  for (int d = 0; d < n_dims; d++) free_dim(&bd[d]);
  return hit;
}
