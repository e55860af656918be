/*
 * oracle.h -- CPU restatement of the reference's star-join hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code: it is used by tests/ as the parity checker, by
 * __graft_entry__.smoke() as the checker of one small GPU run, and by
 * bench.py's cpu_baseline leg as the timed CPU baseline ("port").
 *
 * Every function cites the reference file:line whose semantics it restates.
 * The reference is header-only CUDA (the crystal/ headers) plus drivers; it cannot be
 * compiled in this image (no nvcc, CUB 1.6.4 is fetched by wget), so this file
 * is a from-scratch C restatement, pinned by the known-answer tests and
 * brute-force checks under tests/ (see each test's docstring).
 *
 * Parity status per function (what pins it):
 *   orc_hash_pht / orc_bloom MOD     : closed forms + bloom.cu KAT (tests/test_oracle_kat.py)
 *   orc_predicate                    : truth-table tests of SET/AND/OR x 5 comparators
 *   orc_pht1 / orc_pht2 build+probe  : direct-index textbook join on dense keys, sentinel quirk tests
 *   orc_bloom MULT / BLOCKED         : build-defined hash (not in the reference): hand-computed
 *                                      known values + no-false-negative / order-independence invariants
 *   orc_single_join                  : closed-form sums (all-selected, none-selected) + brute force
 *   orc_star_join                    : nested-loop SQL brute force (q33.sql semantics) on small inputs
 *   orc_tb_build / orc_tb_probe      : project-defined (not in the reference): a dictionary of the
 *                                      selected keys (last value wins), tables with collisions, a
 *                                      full table, the excluded key
 *   line statistics                  : brute-force recount in tests
 */
#ifndef LIPJOIN_ORACLE_H
#define LIPJOIN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- predicates: crystal/pred.cuh:25-356 ---------------------------------- */
enum { ORC_COMB_SET = 0, ORC_COMB_AND = 1, ORC_COMB_OR = 2 };
enum { ORC_CMP_LT = 0, ORC_CMP_GT = 1, ORC_CMP_LTE = 2, ORC_CMP_GTE = 3, ORC_CMP_EQ = 4 };

typedef struct {
  int32_t comb;  /* ORC_COMB_* : BlockPred (overwrite) / BlockPredAnd / BlockPredOr */
  int32_t cmp;   /* ORC_CMP_*  : LessThan / GreaterThan / LessThanEq / GreaterThanEq / Eq */
  int32_t value; /* the functor's `compare` */
} orc_pred_step;

/* flags[i] = result of applying the steps in order to attr[i]; n_steps == 0
 * means InitFlags (all ones, crystal/pred.cuh:25-33). */
void orc_predicate(const int32_t* attr, int64_t n, const orc_pred_step* steps,
                   int32_t n_steps, uint8_t* flags);

/* ---- perfect hash tables: crystal/join.cuh -------------------------------- */
/* HASH(X,Y,Z) ((X-Z) % Y)  (crystal/join.cuh:27).  Returns -1 when the
 * reference would index out of bounds (X < Z gives a negative C remainder). */
int64_t orc_hash_pht(int32_t key, int64_t len, int32_t keys_min);

/* BlockBuildSelectivePHT_1 (join.cuh:191-258): CAS(ht[s], 0, key). */
void orc_pht1_build(const int32_t* keys, const uint8_t* flags, int64_t n,
                    int32_t* ht, int64_t len, int32_t keys_min, int64_t* n_oob);
/* BlockProbeAndPHT_1 (join.cuh:29-106): hit iff ht[s] != 0, no key compare. */
int orc_pht1_probe(const int32_t* ht, int64_t len, int32_t keys_min, int32_t key);

/* BlockBuildSelectivePHT_2 (join.cuh:260-335): CAS(ht[2s], 0, key) then an
 * unconditional ht[2s+1] = val (last writer wins). */
void orc_pht2_build(const int32_t* keys, const int32_t* vals, const uint8_t* flags,
                    int64_t n, int32_t* ht, int64_t len, int32_t keys_min,
                    int64_t* n_oob);
/* BlockProbeAndPHT_2 (join.cuh:108-189): 64-bit slot (key | val<<32) != 0 is a
 * hit and yields payload = val; no key compare. */
int orc_pht2_probe(const int32_t* ht, int64_t len, int32_t keys_min, int32_t key,
                   int32_t* payload);

/* ---- textbook open addressing (not in the reference) ---------------------
 * The hash table with key compare that the reference's perfect hash replaces
 * (join.cuh:108-189 compares no keys).  cap slots (a power of two) of int32
 * (stored key, value) pairs, stored key = key ^ 0x80000000 so that an
 * all-zero table is empty (the key INT32_MIN is outside the domain); home
 * slot = Knuth's multiplicative hash (TAOCP vol. 3 sec. 6.4)
 * (uint32)(key * 0x9E3779B1) >> (32 - log2 cap); linear probing.  A
 * repeated key keeps its slot and takes the last value. */
int64_t orc_tb_home(int32_t key, int64_t cap);
void orc_tb_build(const int32_t* keys, const int32_t* vals, const uint8_t* flags, int64_t n,
                  int32_t* ht, int64_t cap, int64_t* n_oob);
/* hit iff the key is stored; payload = its value */
int orc_tb_probe(const int32_t* ht, int64_t cap, int32_t key, int32_t* payload);

/* ---- Bloom filters: crystal/bloom.cuh ------------------------------------ */
enum { ORC_BLOOM_NONE = 0, ORC_BLOOM_MOD = 1, ORC_BLOOM_MULT = 2 };
enum { ORC_LAYOUT_STANDARD = 0, ORC_LAYOUT_BLOCKED = 1 };

typedef struct {
  int32_t hash;     /* ORC_BLOOM_* */
  int32_t layout;   /* ORC_LAYOUT_* (MULT only) */
  int32_t k;        /* 1 for MOD; 1..3 for MULT */
  int32_t keys_min; /* HASH1's Z */
  int64_t nbits;    /* MOD: any (rounded up to words*32); MULT: power of two >= 64 */
} orc_bloom;

int64_t orc_bloom_nwords(const orc_bloom* b);
/* MOD = BlockBuildBloomFilter (bloom.cuh:10-88): h = HASH1(key, words*32, min),
 * atomicOr(&bf[h>>5], 1<<(h&31)).  MULT is defined by this project (Knuth
 * multiplicative hashing, TAOCP vol. 3 sec. 6.4), see orc_bloom_bits(). */
void orc_bloom_build(const orc_bloom* b, const int32_t* keys, const uint8_t* flags,
                     int64_t n, uint32_t* words, int64_t* n_oob);
/* BlockProbeBloomFilter (bloom.cuh:90-174): pass iff all k bits are set. */
int orc_bloom_probe(const orc_bloom* b, const uint32_t* words, int32_t key);
/* The k bit positions of `key` (returns k, or 0 when the key is out of domain). */
int orc_bloom_bits(const orc_bloom* b, int32_t key, int64_t* bits_out);

/* ---- whole-query drivers -------------------------------------------------- */
enum { ORC_METHOD_NAIVE = 0, ORC_METHOD_LIP = 1 };

enum { ORC_HT_PHT2 = 0, ORC_HT_PHT1 = 1, ORC_HT_TEXTBOOK = 2 };

typedef struct {
  const int32_t* keys;
  const int32_t* attr;    /* predicate column (NULL only when n_pred == 0) */
  const int32_t* payload; /* PHT_2 value column (NULL -> attr) */
  int64_t n;
  const orc_pred_step* pred;
  int32_t n_pred;
  int32_t keys_min;
  int64_t ht_len;         /* slots (TEXTBOOK: the capacity, a power of two) */
  orc_bloom bloom;        /* hash == NONE -> no Bloom filter */
  int32_t ht_kind;        /* ORC_HT_*: PHT_2 (join.cuh:108-189,260-335), PHT_1 (no payload,
                             join.cuh:29-106,191-258) or TEXTBOOK */
  int32_t reserved0;
} orc_dim;

/* Single FK-PK join with an optional Bloom pre-probe
 * (src/ops/join.cu:43-108, src/ops/bloom_join.cu:43-112). */
typedef struct {
  int64_t n_selected;   /* dim rows passing the predicate */
  int64_t n_bloom_pass; /* fact rows passing the Bloom probe (== n_fact when NAIVE) */
  int64_t n_match;      /* fact rows passing the PHT_2 probe */
  int64_t sum_measure;  /* int64 SUM(measure) of matches */
  uint64_t sum_product; /* join.cu:96 / bloom_join.cu:100: sum += (int)(val*join_val) into u64 */
  int64_t n_oob;        /* keys outside the hash domain (reference UB; here: miss) */
  int64_t lines_fk;     /* 32-row lines of the FK column read (all) */
  int64_t lines_measure;/* 32-row lines holding >= 1 match (measure column) */
} orc_join_result;

int orc_single_join(const orc_dim* dim, const int32_t* fact_fk,
                    const int32_t* fact_measure, int64_t n_fact, int32_t method,
                    int32_t n_threads, orc_join_result* out);

/* Star join with dense group-by (src/ssb-jo/q33/{0..5}-{naive,lip}.cu).
 * group g = (sum_d payload_d * gmult[d] + goffset) % n_groups (q33/0-lip.cu:102-104).
 * gsum / gcount are dense [n_groups]; gpay is [n_groups * 4] and holds the
 * payloads of the last row written to each group (the reference's plain stores,
 * q33/0-lip.cu:105-107).  `order` is the join (hash-probe) order; `bloom_order`
 * is the LIP Bloom order (3-lip-best-bloom.cu decouples them).
 * Stage s < n_dims is the s-th FK column loaded (join order for NAIVE, Bloom
 * order for LIP); stage n_dims is the measure column.  lines_out[0..n_dims]
 * (size ORC_MAX_DIMS+1) = 32-row lines holding >= 1 row alive when the stage
 * loads its column; rows_out[0..n_dims+1] (size ORC_MAX_DIMS+2) = rows alive
 * entering each stage, and rows_out[n_dims+1] = LIP rows passing every Bloom
 * filter.  Returns the number of non-empty groups. */
#define ORC_MAX_DIMS 4
typedef struct {
  const int32_t* fk[ORC_MAX_DIMS];
  const int32_t* measure;
  int64_t n;
} orc_fact;

int64_t orc_star_join(const orc_dim* dims, int32_t n_dims, const orc_fact* fact,
                      const int32_t* order, const int32_t* bloom_order,
                      int32_t method, const int64_t* gmult, int64_t goffset,
                      int64_t n_groups, int64_t* gsum, int64_t* gcount,
                      int32_t* gpay, int64_t* lines_out, int64_t* rows_out,
                      int64_t* n_oob, int32_t n_threads);

/* ssb_utils.h:85-96 */
uint32_t orc_round_up_pow2(uint32_t n);

#ifdef __cplusplus
}
#endif
#endif
