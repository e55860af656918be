/*
 * lipjoin.h -- C-ABI of the MI355X (gfx950) star-join engine.
 *
 * Plain C99 types only.  Device pointers are raw HIP device addresses owned by
 * the caller (the reference's loadToGPU columns); tables, Bloom filters and
 * scratch are owned by the context.  Every entry point returns an lj_status and
 * never exits the process; lj_last_error() gives the message of the last
 * failure on a context.  A context is driven by one host thread at a time.
 *
 * Reference interfaces replaced (yliang412/crystal-opt):
 *   lj_hash_join   <- TimeKeeper hashJoin(...)   src/ops/join.cu:117-148,
 *                                                 src/ops/bloom_join.cu:121-156
 *   lj_star_join   <- float runQuery(...)        src/ssb-jo/q33/0-lip.cu:205-302,
 *                                                 src/ssb-jo/q33/0-naive.cu:178-257
 *   lj_dim_build   <- build_kernel / build_hashtable_{s,c,d} + their memsets
 *                     (bloom_join.cu:43-65,128-141; q33/0-lip.cu:114-203,218-257)
 *   lj_dim_export  <- the device tables those kernels fill (no reference
 *                     accessor: the reference never copies them back)
 *   lj_gen_*       <- create_relation_pk / create_relation_fk
 *                     (src/ops/utils/generator.h:88-136) and loadColumn
 *                     (src/ssb/ssb_utils.h:161-171), replaced by seeded
 *                     counter-based generators
 * The device-side operator API (the crystal/ headers) is mirrored by the headers in
 * include/lipjoin/block/.
 */
#ifndef LIPJOIN_H
#define LIPJOIN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LJ_ABI_VERSION 5
#define LJ_MAX_DIMS 4
#define LJ_MAX_PRED_STEPS 4

typedef enum {
  LJ_OK = 0,
  LJ_EINVAL = 1,        /* bad descriptor / argument */
  LJ_ENOMEM = 2,        /* device allocation failed */
  LJ_EBUDGET = 3,       /* request exceeds the context's HBM budget */
  LJ_EHIP = 4,          /* HIP runtime error */
  LJ_EUNSUPPORTED = 5,  /* valid but not implemented combination */
  LJ_EDOMAIN = 6,       /* keys fell outside the hash domain (reference: UB) */
  LJ_ETIMEOUT = 7,
  LJ_ESTATE = 8         /* call not valid in the current state */
} lj_status;

typedef struct lj_ctx lj_ctx;
typedef struct lj_dim lj_dim;

/* ---- predicates: crystal/pred.cuh ---------------------------------------- */
typedef enum { LJ_COMB_SET = 0, LJ_COMB_AND = 1, LJ_COMB_OR = 2 } lj_comb;
typedef enum { LJ_CMP_LT = 0, LJ_CMP_GT = 1, LJ_CMP_LTE = 2, LJ_CMP_GTE = 3, LJ_CMP_EQ = 4 } lj_cmp;

/* One BlockPred / BlockPredAnd / BlockPredOr call (pred.cuh:35-356).  Steps
 * run in order; SET overwrites the flag (so GTE followed by SET LTE keeps only
 * the LTE test, as build_hashtable_d does, q33/0-lip.cu:192-195). */
typedef struct {
  int32_t comb;   /* lj_comb */
  int32_t cmp;    /* lj_cmp */
  int32_t value;  /* comparison constant */
} lj_pred_step;

typedef struct {
  int32_t n_steps;                      /* 0 = InitFlags: every row selected */
  lj_pred_step steps[LJ_MAX_PRED_STEPS];
} lj_pred;

/* ---- dimension tables ----------------------------------------------------- */
typedef enum {
  LJ_HT_PHT2 = 0,     /* crystal PHT_2: (key,val) int32 pairs, slot (key-min)%len,
                         0 = empty, no key compare (join.cuh:108-189,260-335) */
  LJ_HT_PHT1 = 1,     /* crystal PHT_1: int32 keys only, slot (key-min)%len, hit iff the
                         slot is nonzero, no payload (reported as 0) (join.cuh:29-106,
                         191-258; the Q2.x / Q4.x suppliers, q21/0-lip.cu:77,135) */
  LJ_HT_TEXTBOOK = 2  /* open addressing with key compare (not in the reference):
                         ht_len slots (a power of two) of (key ^ 0x80000000, val) int32
                         pairs, home slot (uint32)(key * 0x9E3779B1) >> (32 - log2 ht_len),
                         linear probing; keys_min is not used by the table (the key
                         INT32_MIN is outside its domain).  Agrees with PHT_2 on dense,
                         nonzero, in-domain keys; differs on key 0 with value 0 and on
                         fact keys outside the dimension's key range */
} lj_ht_kind;

typedef enum {
  LJ_BLOOM_NONE = 0,
  LJ_BLOOM_MOD = 1,   /* HASH1 (bloom.cuh:7): bit (key-min) % (words*32), k = 1 */
  LJ_BLOOM_MULT = 2   /* multiplicative family, k in 1..3, bits a power of two */
} lj_bloom_hash;

typedef enum {
  LJ_BLOOM_STANDARD = 0, /* k independent bit positions over the whole array */
  LJ_BLOOM_BLOCKED = 1   /* MULT only: all k bits inside one 32-bit word */
} lj_bloom_layout;

typedef struct {
  const int32_t* keys;     /* device, n_rows */
  const int32_t* attr;     /* device, predicate column (may be NULL if pred.n_steps == 0) */
  const int32_t* payload;  /* device, PHT_2 / textbook value column; NULL -> attr */
  int64_t n_rows;
  lj_pred pred;
  int32_t ht_kind;         /* lj_ht_kind */
  int32_t keys_min;        /* HASH's / HASH1's Z argument */
  int64_t ht_len;          /* slots (the reference's num_slots) */
  int32_t bloom_hash;      /* lj_bloom_hash */
  int32_t bloom_layout;    /* lj_bloom_layout */
  int32_t bloom_k;         /* hashes per key */
  int32_t reserved0;
  int64_t bloom_bits;      /* MOD: rounded up to a multiple of 32 (bloom_join.cu:128);
                              MULT: power of two in [64, 2^31] */
} lj_dim_desc;

typedef struct {
  int64_t n_selected;    /* dimension rows that passed the predicate */
  int64_t n_oob;         /* selected keys whose hash fell outside the table */
  int64_t ht_words;      /* int32 words of the PHT (2 * ht_len) */
  int64_t bloom_words;   /* uint32 words of the Bloom filter (0 if none) */
  int64_t bloom_bits_set;/* popcount of the Bloom filter */
  float time_build_ms;   /* device time of the table fill + build launches */
} lj_dim_info;

/* ---- context -------------------------------------------------------------- */
/* hip_stream: a hipStream_t to run on (NULL -> the context creates its own
 * default-priority stream).  hbm_budget_bytes: cap on context-owned device
 * memory (0 -> no cap). */
lj_status lj_create(int32_t device, uint64_t hbm_budget_bytes, void* hip_stream, lj_ctx** out);
void lj_destroy(lj_ctx* ctx);
const char* lj_last_error(const lj_ctx* ctx);
const char* lj_status_str(lj_status s);
int32_t lj_abi_version(void);
/* sizeof of the public structs, for binding checks: 0 lj_dim_desc, 1 lj_dim_info,
 * 2 lj_probe_stats, 3 lj_join_desc, 4 lj_join_result, 5 lj_star_desc,
 * 6 lj_group, 7 lj_star_result, 8 lj_batch_opts, 9 lj_bloom_choice,
 * 10 lj_serve_cfg, 11 lj_serve_stats, 12 lj_tune_model; -1 for an unknown
 * index. */
int64_t lj_abi_struct_size(int32_t which);
/* Wait for all work queued by the context. */
lj_status lj_sync(lj_ctx* ctx);
/* Device memory currently held by the context (tables, filters, scratch). */
uint64_t lj_bytes_in_use(const lj_ctx* ctx);

/* Build a device-resident dimension (predicate -> Bloom filter -> PHT_2).
 * The lj_dim stays valid until lj_dim_release. */
lj_status lj_dim_build(lj_ctx* ctx, const lj_dim_desc* desc, lj_dim** out);
lj_status lj_dim_get_info(lj_ctx* ctx, const lj_dim* dim, lj_dim_info* info);
/* Copy the table (2*ht_len int32; PHT_1: ht_len) and/or the Bloom words to
 * host memory; either pointer may be NULL.  Sizes are checked against the
 * table sizes.  (A textbook table's layout depends on the order of the
 * inserts under collisions; its contents as a map do not.) */
lj_status lj_dim_export(lj_ctx* ctx, const lj_dim* dim, int32_t* ht_host, int64_t ht_words,
                        uint32_t* bloom_host, int64_t bloom_words);
void lj_dim_release(lj_ctx* ctx, lj_dim* dim);

/* ---- queries ---------------------------------------------------------------- */
typedef enum {
  LJ_METHOD_NAIVE = 0,        /* hash probes in join order (q33/{i}-naive.cu) */
  LJ_METHOD_LIP = 1,          /* all Bloom probes in bloom_order, then hash probes (q33/{i}-lip.cu) */
  LJ_METHOD_LIP_ADAPTIVE = 2  /* LIP with the Bloom order chosen on the device from sampled pass rates */
} lj_method;

/* Per-stage statistics reported by the fused probe (exact, all integers).
 * Stage s is the s-th FK column loaded (join order for NAIVE, Bloom order for
 * LIP); stage n_dims is the measure column. */
typedef struct {
  int64_t rows_in[LJ_MAX_DIMS + 1];   /* rows alive when stage s loads its column */
  int64_t lines_in[LJ_MAX_DIMS + 1];  /* 32-row (128 B) lines holding >= 1 such row */
  int64_t rows_bloom_pass;            /* LIP: rows that passed every Bloom filter */
  int64_t sample_rows;                /* adaptive: rows of the sampled tiles (each wave's first
                                         tile; all columns read; not in rows_in/lines_in) */
  int64_t sample_lines;               /* lines read for the sampled tiles, all columns */
  int32_t order_used[LJ_MAX_DIMS];    /* the order stage 0..n-1 used (adaptive: by the first
                                         wave of workgroup 0; each wave orders its own) */
  int64_t n_oob;                      /* FK / group indices outside their domains */
} lj_probe_stats;

/* Single FK-PK join, optional Bloom pre-filter (join.cu / bloom_join.cu). */
typedef struct {
  const int32_t* fact_fk;      /* device */
  const int32_t* fact_measure; /* device */
  int64_t n_fact;
  lj_dim_desc dim;             /* built inside the call (hashJoin semantics) */
  int32_t method;              /* NAIVE: no Bloom; LIP: Bloom then PHT_2 */
  int32_t flags;               /* 0, or LJ_JOIN_SUM_ONLY */
} lj_join_desc;

/* lj_join_desc.flags: report SUM(measure) and the counts only; sum_product
 * (the reference's vals * join_vals aggregate, join.cu:96) is then 0.  With no
 * payload needed, a LIP join whose modulo filter has at least ht_len bits lets
 * the filter decide the table probe of every in-domain key (a semi-join:
 * identical counts and sums, no hash-table reads for those keys). */
#define LJ_JOIN_SUM_ONLY 1

typedef struct {
  int64_t n_selected;     /* dimension rows selected */
  int64_t n_bloom_pass;   /* fact rows that passed the Bloom filter (n_fact for NAIVE) */
  int64_t n_match;        /* fact rows that matched the PHT */
  int64_t sum_measure;    /* int64 SUM(measure) over matches */
  uint64_t sum_product;   /* reference aggregate: sum of (int32)(measure*payload) as u64 (join.cu:96) */
  float time_build;       /* TimeKeeper fields (ms), join.cu:110-115 */
  float time_probe;
  float time_extra;       /* the clear of the table, filter and counters inside the join's
                             window (the reference's timed memsets, bloom_join.cu:135-137) */
  float time_total;       /* clear + build + probe */
  lj_probe_stats stats;
} lj_join_result;

lj_status lj_hash_join(lj_ctx* ctx, const lj_join_desc* desc, lj_join_result* out);

/* Star join with a dense group-by (Q3.3 shape, q33/{0..5}-{naive,lip}.cu).
 * group g = (sum_d payload_d * group_mult[d] + group_offset) % n_groups, computed
 * exactly in 64-bit (q33/0-lip.cu:102-104).  Dimensions come either from
 * dims[d] (built inside the call: runQuery semantics) or from resident[d]. */
typedef struct {
  const int32_t* fact_fk[LJ_MAX_DIMS]; /* device FK column per dimension */
  const int32_t* fact_measure;          /* device */
  int64_t n_fact;
  int32_t n_dims;
  int32_t method;                       /* lj_method */
  int32_t order[LJ_MAX_DIMS];           /* join (hash-probe) order, a permutation of 0..n_dims-1 */
  int32_t bloom_order[LJ_MAX_DIMS];     /* LIP Bloom order; all -1 -> same as order */
  lj_dim_desc dims[LJ_MAX_DIMS];
  const lj_dim* resident[LJ_MAX_DIMS];  /* non-NULL -> use this built dimension */
  int64_t group_mult[LJ_MAX_DIMS];
  int64_t group_offset;
  int64_t n_groups;                     /* dense group space, <= 2^26 */
  int32_t flags;                        /* 0 */
  int32_t reserved0;
  const int32_t* fact_measure_sub;      /* device, or NULL.  Non-NULL: each row adds
                                           (int32)(measure - measure_sub), as Q4.x's
                                           revenue - supplycost (q41/0-lip.cu:127-128);
                                           needs n_dims >= 2 and 16-byte alignment */
} lj_star_desc;

typedef struct {
  int64_t g;                       /* dense group index */
  int32_t payload[LJ_MAX_DIMS];    /* the dimension payloads of the group (year, c_city, ...) */
  int64_t sum;                     /* int64 SUM(measure) */
  int64_t count;                   /* rows in the group */
  int32_t sum_i32;                 /* reference int32 atomicAdd value (wraps), q33/0-lip.cu:108 */
  int32_t reserved0;
} lj_group;

typedef struct {
  lj_group* groups;        /* caller-provided host array, ascending g */
  int64_t groups_cap;
  int64_t n_groups;        /* non-empty groups (may exceed groups_cap: then truncated) */
  int64_t n_rows;          /* rows aggregated */
  int64_t sum_total;
  float time_query_ms;     /* device time from the build launch to the end of the probe
                              (builds + probe + finalize).  The reference's runQuery window
                              (q33/0-lip.cu:216-278) also holds its allocations and memsets;
                              here a query's tables come from the context's zero pool,
                              cleared by a fill queued behind the query that last used them,
                              so no clear runs inside this window.  lj_hash_join clears its
                              tables inside its window and reports that as time_extra. */
  float time_build_ms;
  float time_probe_ms;
  int32_t status;          /* this query's own lj_status (batches and serving report one
                              per query; lj_star_join: its return value) */
  lj_probe_stats stats;
} lj_star_result;

lj_status lj_star_join(lj_ctx* ctx, const lj_star_desc* desc, lj_star_result* out);

/* lj_star_join_noalloc -- lj_star_join that allocates nothing.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Every dimension must be resident
 * (lj_dim_build); the caller passes the fact columns and the result struct.
 * The probe's buffers (group accumulators, statistics rows, the group list,
 * the result region) are a fixed device array of 20.3 MB that each GPU holds
 * from the library's load until the process ends (set up by lj_create).
 * Limits: n_groups <= LJ_NOALLOC_MAX_GROUPS, at most LJ_NOALLOC_MAX_LISTED
 * groups hit (else LJ_EINVAL / LJ_EUNSUPPORTED, and the array is cleared).
 * All calls on one GPU share the array: calls that overlap on one GPU race. */
#define LJ_NOALLOC_MAX_GROUPS (1 << 19)
#define LJ_NOALLOC_MAX_LISTED 4096
lj_status lj_star_join_noalloc(lj_ctx* ctx, const lj_star_desc* desc, lj_star_result* out);

/* ---- batches of queries ---------------------------------------------------------
 * n independent star joins (any mix of sizes, predicates and Bloom filters)
 * run as ONE build launch over all their dimensions and ONE probe launch per
 * stage-0 kernel kind (one in total when the filters share a kind), the
 * workgroups of the grid partitioned over the queries in proportion to their
 * fact rows.  With use_graph, the launches (table clears, builds, probes and
 * the result copy) are recorded once as a HIP graph and replayed by every
 * later lj_batch_run.  Results are identical to n calls of lj_star_join.
 * No reference counterpart: the reference runs one query per launch
 * sequence; this is the batched mode of the engine. */
typedef struct lj_batch lj_batch;

typedef struct {
  int32_t use_graph;       /* record the launches as a HIP graph on the first run */
  int32_t probe_wg;        /* workgroup size 256/512/1024; 0: the context's choice */
  int64_t groups_cap;      /* result groups kept per query on the device (>= 1; 0: 64) */
  int32_t launch_per_kind; /* 0: one probe launch for the whole batch (queries of different
                              stage-0 kinds share it at the 1024-thread shape); 1: one probe
                              launch per stage-0 kind */
  int32_t reserved0;
  int64_t reserved[1];
} lj_batch_opts;

/* descs[0..n): every dimension built inside the batch (dims[d]) or resident.
 * Device columns must stay valid until lj_batch_destroy. */
lj_status lj_batch_create(lj_ctx* ctx, const lj_star_desc* descs, int32_t n, const lj_batch_opts* opts,
                          lj_batch** out);
/* Runs the batch; results[i] is filled like lj_star_join's (groups ascending g,
 * at most groups_cap of them); time_query_ms is the whole batch's time. */
lj_status lj_batch_run(lj_batch* b, lj_star_result* results);
/* (results[i].status holds query i's own status; the return value is the
 * first failure.  A query that hits more groups than groups_cap fails with
 * LJ_EUNSUPPORTED and its accumulators are cleared before the next run.) */
void lj_batch_destroy(lj_batch* b);

/* ---- serving ---------------------------------------------------------------------
 * A request loop over device-resident dimensions (lj_dim_build): callers
 * submit star joins (every dims[d] resident, fact columns anywhere in device
 * memory, e.g. slices of a resident pool) from any thread and poll for the
 * results.  The loop runs on its own host thread and drives the context
 * alone (do not call other lj_* functions on the context until lj_serve_stop).
 * Continuous batching: requests that arrive while launches are in flight
 * wait; whenever one of max_inflight launch slots is free, every waiting
 * request that the HBM budget admits (up to max_batch) goes into ONE launch
 * (the batched probe, whose queries' workgroups overlap on the device; the
 * launch queues behind the one in flight, or runs beside it with
 * cfg.streams > 1).  The budget bounds the per-request device memory in flight
 * (statistics rows, group accumulators, result regions).  Results are
 * identical to lj_star_join of the same query.  No reference counterpart:
 * the reference runs one query per process invocation. */
typedef struct lj_server lj_server;

typedef struct {
  uint64_t hbm_budget;     /* bytes of per-request device memory in flight; 0: unlimited */
  int32_t max_batch;       /* requests per launch; 0: 64 */
  int32_t max_inflight;    /* launch slots (launches in flight); 0: 4 */
  int64_t groups_cap;      /* result groups kept per request; 0: 64 */
  int32_t streams;         /* streams the launch slots use: 0 or 1 = one stream, so a new
                              launch queues behind the one in flight (no two launches
                              compete for the device's caches); k > 1: slot i runs on
                              stream i % k and launches overlap on the device */
  int32_t reserved0;
  int64_t prewarm_groups;  /* > 0: at start, allocate and clear max_inflight * max_batch
                              request buffers for queries of up to this many groups, so
                              the first requests do not wait for allocations; 0: none */
} lj_serve_cfg;

typedef struct {
  uint64_t submitted;
  uint64_t completed;
  uint64_t rejected;       /* failed admission or validation */
  uint64_t launches;
  uint64_t bytes_in_flight;
  uint64_t peak_bytes_in_flight;
  uint64_t max_batch_seen;
  uint64_t reserved;
} lj_serve_stats;

lj_status lj_serve_start(lj_ctx* ctx, const lj_serve_cfg* cfg, lj_server** out);
/* Queues a request (the descriptor is copied); *id identifies it for poll.
 * The descriptor is checked here (lj_last_error(NULL) on the calling thread
 * gives the message of a refusal), so a malformed request never joins a
 * launch; a request that fails later (planning, out-of-domain keys, more
 * groups than groups_cap) fails alone, with its own status. */
lj_status lj_serve_submit(lj_server* srv, const lj_star_desc* q, uint64_t* id);
/* Waits up to timeout_ms (0: do not wait, < 0: forever) for request id.
 * LJ_ETIMEOUT while it is not done.  When done, fills out like lj_star_join
 * (at most out->groups_cap groups, ascending g) with time_query_ms = submit to
 * result ready (host clock) and time_probe_ms = the device time of its
 * launch, returns the request's status and forgets the request. */
lj_status lj_serve_poll(lj_server* srv, uint64_t id, lj_star_result* out, int32_t timeout_ms);
lj_status lj_serve_get_stats(lj_server* srv, lj_serve_stats* out);
/* Finishes every queued request, stops the loop and frees the server
 * (results not yet polled are dropped). */
lj_status lj_serve_stop(lj_server* srv);

/* ---- Bloom filter autotuner ------------------------------------------------------
 * Picks the Bloom filter (hash, layout, k, bits) of a LIP join on one
 * dimension: n_selected keys pass the dimension predicate, the fact side has
 * n_fact rows of which a fraction fact_pass joins.  The cost model per fact
 * row is scan(hash, layout, k, where the filter lives: LDS, L2 up to 2 MiB,
 * or beyond) + P * gather (gather_lds for a filter in LDS) when the probe
 * streams the next column (P > 1/4), scan + lines(P) * line + P * gather_queue
 * (gather_queue_lds) when it queues the passing rows, blended for P in
 * [0.2, 0.3] (and no gather
 * term for an exact filter of a semi-join), with
 * P = fact_pass + (1 - fact_pass) * false-positive rate(bits, k, layout) and
 * lines(P) = 1 - (1 - P)^32 the fraction of 128-byte lines of the next column
 * that hold a passing row; the constants are the context's (defaults, or
 * measured by lj_tune_calibrate). */
typedef struct {
  int32_t bloom_hash;      /* lj_bloom_hash */
  int32_t bloom_layout;    /* lj_bloom_layout */
  int32_t bloom_k;
  int32_t in_lds;          /* 1 if the filter fits the probe's LDS */
  int64_t bloom_bits;
  double fpr;              /* predicted false-positive rate */
  double cost;             /* predicted cost per fact row (scan of an LDS filter = 1) */
} lj_bloom_choice;

lj_status lj_tune_bloom(lj_ctx* ctx, int64_t n_selected, int64_t n_dim_rows, double fact_pass,
                        int32_t flags, lj_bloom_choice* out);
/* lj_tune_bloom flags: the join needs no payload (one group, no product), so
 * an exact modulo filter (bits >= n_dim_rows, dense keys) decides the table
 * probe and its passing rows cost no hash-table gather. */
#define LJ_TUNE_SEMI_JOIN 1
/* Measures the model's constants on this device (a few probes of synthetic
 * data, about a second) and keeps them in the context. */
lj_status lj_tune_calibrate(lj_ctx* ctx);
/* The model's constants, per fact row in units of a scan with the filter in
 * LDS.  Read them after lj_tune_calibrate to keep a measurement; set them to
 * reuse it on another context of the same device (no counterpart in the
 * reference, whose Bloom sizes are fixed by its drivers' arguments). */
typedef struct {
  double scan_l2;          /* scan with the first filter in L2 (up to 2 MiB), >= 1 */
  double scan_big;         /* scan with a larger filter, >= scan_l2 */
  double probe_extra;      /* each further bit probe of a standard k > 1 filter outside LDS,
                              times its scan, >= 0 */
  double line;             /* a 128-byte line of the next column that holds a passing row, >= 0 */
  double gather;           /* a passing row when more than a quarter pass (the measure
                              streams with the keys), filter outside LDS, >= 0 */
  double scan_mod;         /* scan with a modulo filter in LDS, > 0 */
  double scan_blocked;     /* scan with a blocked multiplicative filter in LDS, > 0 */
  double gather_queue;     /* a passing row when at most a quarter pass (hash-slot and
                              measure gathers), >= 0 */
  double gather_lds;       /* as gather, with the filter in LDS, >= 0 */
  double gather_queue_lds; /* as gather_queue, with the filter in LDS, >= 0 */
  int32_t calibrated;      /* get: 1 once measured or set */
  int32_t reserved;
} lj_tune_model;
lj_status lj_tune_get_model(const lj_ctx* ctx, lj_tune_model* out);
lj_status lj_tune_set_model(lj_ctx* ctx, const lj_tune_model* in);

/* ---- seeded synthetic columns ------------------------------------------------ */
/* Counter-based generator: value j = first + i of stream `stream` depends only
 * on (seed, stream, j), so host and device produce identical columns. */
typedef enum {
  LJ_GEN_UNIFORM = 0,      /* uniform integer in [lo, hi] */
  LJ_GEN_DATEKEY = 1,      /* yyyymmdd of day lo + U[0, hi), day 0 = 1992-01-01 */
  LJ_GEN_SEQUENCE = 2,     /* lo + j */
  LJ_GEN_YEAR_OF_DAY = 3,  /* calendar year of day lo + j */
  LJ_GEN_DATEKEY_SEQ = 4   /* yyyymmdd of day lo + j */
} lj_gen_kind;

/* Fill host memory out[0..n) (multithreaded). */
lj_status lj_gen_host(int32_t kind, uint64_t seed, uint64_t stream, int64_t lo, int64_t hi,
                      int64_t first, int64_t n, int32_t* out);
/* Same values, generated on the device into a device buffer. */
lj_status lj_gen_device(lj_ctx* ctx, int32_t kind, uint64_t seed, uint64_t stream, int64_t lo,
                        int64_t hi, int64_t first, int64_t n, int32_t* out_dev);
/* Host Fisher-Yates shuffle of out[0..n) driven by the same generator. */
lj_status lj_gen_shuffle_host(uint64_t seed, uint64_t stream, int64_t n, int32_t* out);

/* The reference's single-join inputs (src/ops/utils/generator.h:37-136, as
 * bloom_join.cu:221-222 calls them): srand(0), then create_relation_pk --
 * dimension keys 1..n_dim shuffled by its Knuth shuffle (j = RAND_RANGE(i),
 * generator.h:32,51-60), values 0..n_dim-1 -- then create_relation_fk -- fact
 * keys as successive shuffled blocks of 1..n_dim and a shuffled 1..(n_fact %
 * n_dim) tail, values 0..n_fact-1.  Uses the C library's rand() (its global
 * state: not thread-safe).  Any output pointer may be NULL (its values are
 * still drawn, so the others do not change). */
lj_status lj_gen_reference_host(int64_t n_dim, int64_t n_fact, int32_t* dim_keys, int32_t* dim_vals,
                                int32_t* fact_keys, int32_t* fact_vals);
/* The reference's star-schema inputs (src/ops/utils/star.cuh:67-120,
 * star_table_init_host, with the generators above): srand(0), the fact
 * table's keys 1..n_fact shuffled and values 0..n_fact-1, then for each of
 * n_dims dimensions in turn its keys 1..dim_rows[d] shuffled, its values
 * 0..dim_rows[d]-1 and the fact's foreign keys into it (shuffled blocks of
 * 1..dim_rows[d] and a shuffled tail, create_relation_fk_only).  Any array,
 * or entry of dim_keys / dim_vals / fact_fks, may be NULL (its values are
 * still drawn).  Uses the C library's rand() (not thread-safe). */
lj_status lj_gen_reference_star_host(int64_t n_fact, int32_t n_dims, const int64_t* dim_rows,
                                     int32_t* fact_keys, int32_t* fact_vals, int32_t* const* dim_keys,
                                     int32_t* const* dim_vals, int32_t* const* fact_fks);

#ifdef __cplusplus
}
#endif
#endif
