/*
 * proteus_amd.h -- C ABI of the MI355X (gfx950) join / aggregation hot path.
 *
 * Plain pointers and sizes only.  Device pointers ("d_") live in HBM of the
 * context's device; host pointers ("h_") are ordinary or pinned host memory.
 * Every entry point returns a pj_status; on failure pj_last_error() returns a
 * message for the calling thread.  Work is issued on the context's stream and
 * each entry point returns after its results are available unless its comment
 * says otherwise.  One context per device; calls on one context are
 * serialized by the caller.
 *
 * Each declaration cites the reference interface it replaces (paths relative
 * to the reference repository).
 */
#ifndef PROTEUS_AMD_H
#define PROTEUS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  PJ_OK = 0,
  PJ_E_CAPACITY = 1,    /* build/group input exceeded maxBuildInputSize / maxInputSize */
  PJ_E_OOM = 2,         /* device allocation failed or HBM budget exceeded   */
  PJ_E_UNSUPPORTED = 3, /* operator / expression / type not on this path     */
  PJ_E_HIP = 4,         /* HIP runtime error                                  */
  PJ_E_ARG = 5,         /* invalid argument                                   */
  PJ_E_PARSE = 6,       /* malformed plan JSON / unknown operator             */
  PJ_E_CANCELLED = 7    /* request cancelled before completion                */
} pj_status;

typedef struct pj_ctx pj_ctx;

/* ---- context, memory, errors --------------------------------------------
 * Replaces MemoryManager::mallocGpu/freeGpu/mallocPinned/freePinned
 * (core/platform/include/platform/memory/memory-manager.hpp) and the
 * gpu_run / stream helpers of core/platform/include/platform/common/gpu/
 * gpu-common.hpp:93,383-398 with status codes instead of aborts. */
const char *pj_last_error(void);
int pj_ctx_create(int device, uint64_t hbm_budget_bytes, pj_ctx **out);
int pj_ctx_destroy(pj_ctx *ctx);
int pj_ctx_sync(pj_ctx *ctx);
int pj_ctx_device_info(pj_ctx *ctx, int *num_cus, uint64_t *hbm_bytes,
                       uint64_t *lds_bytes_per_cu);
int pj_malloc(pj_ctx *ctx, uint64_t bytes, void **d_ptr);
int pj_free(pj_ctx *ctx, void *d_ptr);
int pj_malloc_pinned(pj_ctx *ctx, uint64_t bytes, void **h_ptr);
int pj_free_pinned(pj_ctx *ctx, void *h_ptr);
int pj_memcpy_h2d(pj_ctx *ctx, void *d_dst, const void *h_src, uint64_t bytes);
int pj_memcpy_d2h(pj_ctx *ctx, void *h_dst, const void *d_src, uint64_t bytes);
int pj_memset(pj_ctx *ctx, void *d_dst, int byte, uint64_t bytes);
int pj_mem_in_use(pj_ctx *ctx, uint64_t *bytes);
/* Pool counters: out[0] bytes handed out, out[1] bytes cached for reuse,
 * out[2] hipMalloc calls (cache misses), out[3] cached blocks released.
 * Host-side bookkeeping of this library; the reference's memory managers
 * (core/platform/lib/memory/memory-manager.cpp) keep no such counters. */
int pj_mem_stats(pj_ctx *ctx, uint64_t out[4]);
/* Device-to-device copy with a streaming HIP kernel (the reference's
 * gpu memcpy helper, core/platform/lib/memory/buffer-manager.cu:846,908). */
int pj_copy(pj_ctx *ctx, void *d_dst, const void *d_src, uint64_t bytes);

/* ---- per-kernel timing ---------------------------------------------------
 * Replaces the time_block / nvtx ranges around pipelines
 * (core/platform/include/platform/util/timing.hpp:98-162,
 * core/olap/lib/plan/prepared-statement.cpp:117): HIP events recorded on the
 * context stream around every launch while enabled. */
int pj_prof_enable(pj_ctx *ctx, int on);
int pj_prof_reset(pj_ctx *ctx);
/* Total milliseconds and launch count of kernels whose name starts with
 * `prefix` since the last reset (synchronizes the context). */
int pj_prof_get(pj_ctx *ctx, const char *prefix, double *total_ms,
                uint64_t *launches);

/* ---- seeded counter-based generators (device) ----------------------------
 * The synthetic-data counterpart of tools/prepare_data.py:389-431: columns
 * generated directly in HBM.  Output = f(seed, stream, row) only. */
int pj_gen_perm_i32(pj_ctx *ctx, int32_t *d_out, uint64_t n, uint64_t seed);
int pj_gen_uniform_i32(pj_ctx *ctx, int32_t *d_out, uint64_t n, uint64_t seed,
                       uint64_t stream, int64_t lo, uint64_t range);
int pj_gen_uniform_i64(pj_ctx *ctx, int64_t *d_out, uint64_t n, uint64_t seed,
                       uint64_t stream, int64_t lo, uint64_t range);
int pj_gen_unit_f32(pj_ctx *ctx, float *d_out, uint64_t n, uint64_t seed,
                    uint64_t stream);
int pj_gen_zipf_fk_i32(pj_ctx *ctx, int32_t *d_out, uint64_t n, uint64_t seed,
                       uint64_t stream, uint64_t nbuild, uint64_t range,
                       double theta, uint64_t perm_seed);
int pj_gen_zipf_i32(pj_ctx *ctx, int32_t *d_out, uint64_t n, uint64_t seed,
                    uint64_t stream, uint64_t g, double theta,
                    uint64_t perm_seed);

/* ---- join -> reduce --------------------------------------------------------
 * Aggregates of the join output that a parent reduce consumes
 * (opt::GpuReduce, core/olap/lib/operators/gpu/gpu-reduce.cpp:45-202, fed by
 * the probe loop's getParent()->consume(), hash-join-chained.cpp:557-559):
 * COUNT(*), SUM/MIN/MAX of the int32 build value and of the int32 probe value
 * over all matches, in exact int64.  Empty result: sums 0, min INT64_MAX,
 * max INT64_MIN (the monoid identities of reduce-opt.cpp:374-430). */
typedef struct {
  int64_t count;
  int64_t sum_b;
  int64_t sum_p;
  int64_t min_b, max_b, min_p, max_p;
} pj_join_agg;

/* hashjoin-chained -> reduce.  Replaces the JIT'd GpuHashJoinChained
 * build/probe (core/olap/lib/operators/hash-join-chained.cpp:262-566,
 * gpu/gpu-hash-join-chained.cpp:31-73) behind the plan operator
 * "hashjoin-chained" (core/olap/lib/plan/plan-parser.cpp:977-1006).
 * hash_bits and max_build are the operator's fields; nb > max_build returns
 * PJ_E_CAPACITY (the reference asserts, hash-join-chained.cpp:599-608).
 * d_bv / d_pv may be NULL (value 0).  The device algorithm depends on the
 * inputs, not the result: dense build keys (as many distinct keys as
 * min..max holds, up to 2^22) join through a direct table of the build
 * values; others through the open-addressing table (small probe sides) or
 * the radix-partitioned join (large ones). */
int pj_join_chained_reduce(pj_ctx *ctx, const int32_t *d_bk,
                           const int32_t *d_bv, uint64_t nb,
                           const int32_t *d_pk, const int32_t *d_pv,
                           uint64_t np, int hash_bits, uint64_t max_build,
                           pj_join_agg *out);

/* Allocation-free variant of pj_join_chained_reduce.  It has no counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under hipGraph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, results and errors, but the table and the
 * accumulator live in a static 32 MiB + 8 KiB device array of this library
 * (one per GPU, held until the process ends) instead of the context pool:
 * the table (max(2^hash_bits, 2 * next_pow2(nb)) 8-byte slots) must fit
 * 2^22 slots, i.e. nb <= 2^21 and hash_bits <= 22, else PJ_E_CAPACITY.
 * No hipMalloc / hipFree or other allocation happens in the call.  Every
 * call on a GPU shares that GPU's array, from every context on it: calls
 * that overlap on one GPU race (calls on one context are serialized). */
int pj_join_chained_reduce_noalloc(pj_ctx *ctx, const int32_t *d_bk,
                                   const int32_t *d_bv, uint64_t nb,
                                   const int32_t *d_pk, const int32_t *d_pv,
                                   uint64_t np, int hash_bits,
                                   uint64_t max_build, pj_join_agg *out);

/* hashjoin-chained with materialized output: every (build row, probe row)
 * pair the reference's probe loop hands to its parent
 * (hash-join-chained.cpp:321-566).  Writes up to `cap` pairs to d_out_b /
 * d_out_p (device, uint32 row ids), *n_out = total matches; if the total
 * exceeds cap the first cap are written and PJ_E_CAPACITY is returned. */
int pj_join_chained_pairs(pj_ctx *ctx, const int32_t *d_bk, uint64_t nb,
                          const int32_t *d_pk, uint64_t np, int hash_bits,
                          uint64_t max_build, uint32_t *d_out_b,
                          uint32_t *d_out_p, uint64_t cap, uint64_t *n_out);

/* Radix-partitioned join tuning knobs; zero fields pick the autotuned value. */
typedef struct {
  int bits1;           /* pass-1 radix bits (fan-out 2^bits1)                */
  int bits2;           /* pass-2 radix bits                                  */
  int lds_log;         /* log2 of the LDS table's head array                 */
  uint32_t probe_unit; /* max probe tuples per join work unit                */
  int grid_mul;        /* persistent workgroups per CU                       */
} pj_radix_cfg;

/* partitioned-hashjoin-chained -> reduce.  Replaces HashPartitioner
 * (core/olap/lib/operators/gpu/gpu-partitioned-hash-join-chained.cpp:47-404
 * with the kernels of gpu-partitioned-hash-join-chained-helpers.cpp:63-481)
 * and GpuPartitionedHashJoinChained's join loop (:612-1586) behind the plan
 * operator "partitioned-hashjoin-chained" (plan-parser.cpp:851-976).
 * nb > max_build or np > max_probe returns PJ_E_CAPACITY.  cfg may be NULL. */
int pj_join_partitioned_reduce(pj_ctx *ctx, const int32_t *d_bk,
                               const int32_t *d_bv, uint64_t nb,
                               const int32_t *d_pk, const int32_t *d_pv,
                               uint64_t np, uint64_t max_build,
                               uint64_t max_probe, const pj_radix_cfg *cfg,
                               pj_join_agg *out);

/* out-of-gpu-join -> reduce (plan-parser.cpp:531-850 with ZipCollect /
 * packet-zip.cpp and the CPU HashRearrange into numOfBuckets buckets): the
 * inputs h_* are in host memory and may exceed HBM.  Both sides are
 * bucketed by hashInt32(key) URem num_buckets on the device in chunks that
 * fit hbm_budget bytes (0: half of HBM), through a pinned host arena, and
 * each bucket pair is joined by the radix-partitioned join; the aggregates
 * are those of pj_join_partitioned_reduce over the whole inputs.
 * num_buckets = 0 picks the fewest buckets that fit the budget.  At most 256
 * buckets (one device bucketing pass): num_buckets > 256 returns PJ_E_ARG,
 * and a budget that would need more than 256 buckets returns PJ_E_CAPACITY,
 * as does a bucket that still does not fit (skewed keys).
 * h_bv / h_pv may be NULL (value 0). */
int pj_join_out_of_gpu_reduce(pj_ctx *ctx, const int32_t *h_bk,
                              const int32_t *h_bv, uint64_t nb,
                              const int32_t *h_pk, const int32_t *h_pv,
                              uint64_t np, uint32_t num_buckets,
                              uint64_t hbm_budget, pj_join_agg *out);

/* The configuration pj_join_partitioned_reduce picks for (nb, np) when cfg
 * is NULL (the autotuner's choice; runs a timed sweep the first time a size
 * class is seen if tuning is enabled with pj_autotune_enable). */
int pj_join_partitioned_config(pj_ctx *ctx, uint64_t nb, uint64_t np,
                               pj_radix_cfg *cfg);
int pj_autotune_enable(pj_ctx *ctx, int on);

/* ---- hash group-by ---------------------------------------------------------
 * groupby / hashgroupby-chained (plan-parser.cpp:479-530) replacing the JIT'd
 * GpuHashGroupByChained build + scan (gpu/gpu-hash-group-by-chained.cpp:
 * 77-556, hash-group-by-chained.cpp:638-853).  One int32 key; per group
 * COUNT, SUM(v), MIN(v), MAX(v) in exact int64 and SUM((double) w) in FP64.
 * d_v / d_w may be NULL.  Output columns are device arrays of capacity
 * max_groups, compacted, in unspecified group order; *n_groups is set.
 * More than max_groups distinct keys returns PJ_E_CAPACITY (the reference
 * asserts, hash-group-by-chained.cpp:915-919).  hash_bits is accepted for
 * API parity; the table geometry is chosen from max_groups. */
int pj_groupby_i32(pj_ctx *ctx, const int32_t *d_k, const int64_t *d_v,
                   const float *d_w, uint64_t n, int hash_bits,
                   uint64_t max_groups, int32_t *d_out_k, int64_t *d_cnt,
                   int64_t *d_sum, int64_t *d_min, int64_t *d_max,
                   double *d_fsum, uint64_t *n_groups);

/* ---- reduce ----------------------------------------------------------------
 * reduce (plan-parser.cpp:248-286) over one int64 column: opt::GpuReduce's
 * SUM / MIN / MAX monoids (gpu/gmonoids.cpp:122-244, gpu-reduce.cpp:155-202). */
int pj_reduce_i64(pj_ctx *ctx, const int64_t *d_v, uint64_t n, int64_t *sum,
                  int64_t *min, int64_t *max);

/* ---- hash-rearrange --------------------------------------------------------
 * hash-rearrange / hash-pack (plan-parser.cpp:1424-1457) replacing the JIT'd
 * GpuHashRearrange (gpu/gpu-hash-rearrange.cpp:87-1049): bucket =
 * hashInt32(key) URem buckets (:338-348).  Output rows are grouped by bucket:
 * bucket b occupies [h_offsets[b], h_offsets[b+1]) of d_out_k / d_out_v
 * (d_v / d_out_v may be NULL).  Order inside a bucket is unspecified. */
int pj_hash_rearrange_i32(pj_ctx *ctx, const int32_t *d_k, const int32_t *d_v,
                          uint64_t n, uint32_t buckets, int32_t *d_out_k,
                          int32_t *d_out_v, uint64_t *h_offsets);

/* ---- resident dimensions, batched queries, serving ---------------------------
 * A query is the plan hashjoin-chained -> groupby over a resident build
 * relation ("dimension"): SELECT S.grp, COUNT(*), SUM(R.value + S.val)
 * FROM S JOIN R ON S.fk = R.key GROUP BY S.grp, with S.grp dense in
 * [0, groups).  The reference runs one JIT'd pipeline per query
 * (PreparedStatement::execute, prepared-statement.cpp:41-140, build side
 * rebuilt in every open, gpu-hash-join-chained.cpp:31-73); here dimension
 * tables stay resident in HBM across queries and many queries share one
 * persistent launch over a device work queue of (query, tile) items,
 * replayed from a hipGraph per stream slot. */
typedef struct pj_dim pj_dim;
/* Resident HBM table of (key, value); keys unique and != INT32_MIN
 * (PJ_E_ARG otherwise).  d_vals may be NULL (value 0).  The input columns
 * are not kept.  Keys whose span fits the open-addressing table they would
 * take get a direct table (one slot per key of the span: 8 bytes, or 4 --
 * the value alone -- when the span is dense and above 2048 keys), others
 * open addressing; pj_dim_info reports the table's bytes. */
int pj_dim_create(pj_ctx *ctx, const int32_t *d_keys, const int32_t *d_vals,
                  uint64_t n, pj_dim **out);
int pj_dim_destroy(pj_ctx *ctx, pj_dim *dim);
int pj_dim_info(pj_ctx *ctx, const pj_dim *dim, uint64_t *rows,
                uint64_t *table_bytes);

typedef struct {
  const pj_dim *dim;       /* build side                                     */
  const int32_t *d_fk;     /* probe columns (device); with gen_seed != 0 the  */
  const int32_t *d_val;    /* batch first fills them from the seed:           */
  const int32_t *d_grp;    /*   fk uniform [0, fk_range), val [0, 2^31),      */
  uint64_t n;              /*   grp [0, groups) (pj_gen_uniform_i32 streams   */
  uint32_t groups;         /*   1, 2, 3 of gen_seed)                          */
  uint32_t flags;          /* 0                                               */
  uint64_t gen_seed;
  uint64_t fk_range;        /* 1 .. 2^31 when gen_seed != 0                    */
  int64_t *d_cnt;          /* outputs, `groups` entries each                  */
  int64_t *d_sum;
} pj_query;

/* Zero fields take the context's autotuned value (pj_batch_autotune), or
 * the default in parentheses before any tuning. */
typedef struct {
  uint32_t tile_rows;      /* probe rows per work item (65536)               */
  int wg_per_cu;           /* persistent workgroups per CU (2)               */
  int lds_dim;             /* dimensions of <= lds_dim rows are probed from  */
                           /* LDS, the rest from HBM; -1: none (2048, max)   */
} pj_batch_cfg;

/* One query, one launch sequence, synchronous (the plain path). */
int pj_query_run(pj_ctx *ctx, const pj_query *q, const pj_batch_cfg *cfg);
/* Creates n (1..8) stream slots for pj_batch_launch (default priority). */
int pj_batch_slots(pj_ctx *ctx, int n);
/* Starts a batch of 1..4096 queries on a slot and returns; the slot's
 * graph is captured on first use and replayed afterwards. */
int pj_batch_launch(pj_ctx *ctx, int slot, const pj_query *qs, uint32_t nq,
                    const pj_batch_cfg *cfg);
int pj_batch_poll(pj_ctx *ctx, int slot, int *done);
int pj_batch_wait(pj_ctx *ctx, int slot);
/* Times batches of the given queries (their outputs are overwritten) over
 * tile sizes, workgroups per CU and the LDS-vs-HBM table threshold, one
 * knob at a time, and keeps the fastest as the context's default for zero
 * fields of pj_batch_cfg; *best receives it.  Queries with gen_seed need
 * their probe columns allocated (d_fk, d_val, d_grp). */
int pj_batch_autotune(pj_ctx *ctx, const pj_query *qs, uint32_t nq, pj_batch_cfg *best);

/* Serving loop: queries arrive at arrival_s[i] seconds after the call
 * starts; mode 0 admits every waiting query whose probe inputs fit the HBM
 * budget into the next free slot (continuous batching: batches of later
 * arrivals overlap the ones already on the device; with >= 2 slots, slot 0
 * takes only light queries so they never wait behind a heavy batch); mode
 * 1 runs one query at a time with pj_query_run in arrival order.  Probe
 * inputs of queries with gen_seed != 0 and d_fk == NULL are placed in
 * arenas allocated once per call; their pointers are reset to NULL when the
 * call returns (caller-supplied columns are left as given).  done_s[i] =
 * completion time of query i (seconds after the start), -1 for a query not
 * run.  The reference has no serving loop: this is the added pattern. */
typedef struct {
  int mode;                /* 0 batched, 1 one query per launch              */
  int slots;               /* streams in flight (mode 0), 1..8               */
  uint32_t max_batch;      /* queries per batch (<= 4096)                    */
  uint64_t hbm_budget;     /* bytes of probe inputs in flight (0: no limit)  */
  uint64_t heavy_rows;     /* with >= 2 slots, slot 0 serves only queries of */
                           /* fewer probe rows (0: 2^22)                     */
  pj_batch_cfg batch;
  uint64_t max_batch_rows; /* probe rows per batch: a batch takes no further  */
                           /* query that would pass it (0: 2^26); its first   */
                           /* query always fits                              */
  const volatile int32_t *cancel; /* NULL, or a host flag: once non-zero, no  */
                           /* query is admitted, batches on the device finish, */
                           /* queries not run keep done_s = -1 and the call  */
                           /* returns PJ_E_CANCELLED                         */
} pj_serve_cfg;
typedef struct {
  uint64_t batches;
  uint64_t max_in_flight_bytes;
  double wall_s;
} pj_serve_stats;
int pj_serve_run(pj_ctx *ctx, pj_query *qs, const double *arrival_s,
                 uint32_t nq, const pj_serve_cfg *cfg, double *done_s,
                 pj_serve_stats *stats);

/* ---- bloom filters ---------------------------------------------------------
 * BloomFilterBuild / BloomFilterProbe / BloomFilterRepack
 * (core/olap/lib/operators/bloom-filter/bloom-filter-build.cpp:29-58,
 * bloom-filter-probe.cpp:29-52, bloom-filter-repack.cpp:43-470) as called by
 * RelBuilder::bloomfilter_build / _probe / _repack (relbuilder.cpp:957-975).
 * A filter has filter_size entries (a power of two, 1 .. 2^32); key k maps to
 * entry (uint32)k & (filter_size - 1), as BloomFilter::findInFilter
 * (bloom-filter.cpp:85-107).  One filter per bloom_id per context, like the
 * reference's registry (setBloomFilter / getBloomFilter, bloom-filter.cpp:36-49):
 * a build makes a fresh filter and replaces the one of the same id.
 * Probe writes the ascending row ids whose key's entry is set (d_idx holds n
 * entries); repack also gathers ncols columns of esize[c] bytes (1, 4 or 8)
 * at those rows into d_out[c] (n rows of room each).  Probing an id that was
 * never built returns PJ_E_ARG. */
int pj_bloom_build_i32(pj_ctx *ctx, uint64_t bloom_id, uint64_t filter_size,
                       const int32_t *d_keys, uint64_t n);
int pj_bloom_probe_i32(pj_ctx *ctx, uint64_t bloom_id, const int32_t *d_keys,
                       uint64_t n, uint32_t *d_idx, uint64_t *n_out);
int pj_bloom_repack(pj_ctx *ctx, uint64_t bloom_id, const int32_t *d_keys,
                    uint64_t n, int ncols, const void *const *d_cols,
                    const int *esize, void *const *d_out, uint64_t *n_out);
int pj_bloom_release(pj_ctx *ctx, uint64_t bloom_id);

/* ---- device sort ---------------------------------------------------------------
 * qsort_<fields>(ptr, N) of core/platform/lib/util/sort/gpu/sort-util.hpp:27-58
 * (the records gpu-sort.cpp packs and sorts on the device): sorts n records in
 * place in device memory, ascending and lexicographic with the first field
 * most significant; equal records keep their order.  `fields` names the
 * fields in order, 'i' = int32 and 'l' = int64 (1 to 8 of them); each field
 * sits at its natural alignment and the record is padded to its widest field
 * (a C struct of those members).  d_rows must be aligned to the widest field.
 * A stable LSD radix sort (8-bit digits; digit positions every record shares
 * are skipped).  Waits for the context. */
int pj_sort_records(pj_ctx *ctx, void *d_rows, uint64_t n, const char *fields);

/* ---- host -> HBM staging -----------------------------------------------------
 * The reference's cpu-to-gpu operator (core/olap/lib/operators/cpu-to-gpu.cpp:
 * 69-239) moves host-resident blocks of a relation to the device, and its
 * storage manager loads the files of prepared relations
 * (core/storage/lib/storage-manager.cpp:300-330; binary-block-plugin.cpp:67-100
 * reads column <rel>.<attr> as a headerless array).  Here a copy of `bytes`
 * bytes from a file at `offset` (path != NULL) or from pageable host memory
 * (h_src != NULL; exactly one of the two) to d_dst is cut into 8 MiB chunks
 * that worker threads fill into pinned buffers, each chunk DMA'd on a copy
 * stream of the context as soon as it is filled.
 *   pj_stage_begin queues the copy and returns at once with a ticket; the
 *     host memory must stay valid until pj_stage_wait returns.  A file shorter
 *     than offset + bytes is PJ_E_ARG, before anything is queued.
 *   pj_stage_wait(ticket) waits on the host until every chunk of the copy is
 *     issued, then makes the context's stream wait for the copy on the device:
 *     kernels launched on the context afterwards see the data, and kernels
 *     launched before it overlap with the DMA.  Each ticket is waited for once;
 *     a read error is reported here.
 *   pj_stage_file is begin + wait + a context synchronisation. */
int pj_stage_begin(pj_ctx *ctx, const char *path, const void *h_src,
                   uint64_t offset, uint64_t bytes, void *d_dst,
                   uint64_t *ticket);
int pj_stage_wait(pj_ctx *ctx, uint64_t ticket);
int pj_stage_file(pj_ctx *ctx, const char *path, uint64_t offset,
                  uint64_t bytes, void *d_dst);

/* ---- plan API ----------------------------------------------------------------
 * PlanExecutor::parseOperator (core/olap/lib/plan/plan-parser.cpp:238-1675)
 * + PreparedStatement::execute (prepared-statement.cpp:41-140): parse a
 * Proteus JSON physical plan and run it.  Relations named by "scan" operators
 * resolve to columns registered with pj_register_column (device or host
 * memory; a host column is staged to HBM on first use) or to CSV files
 * (plugin type "csv"/"pm-csv").  The result rows are written as text in the
 * reference's print format (one row per line, fields comma-separated) into
 * out (NUL-terminated, truncated to out_cap). */
int pj_register_column(pj_ctx *ctx, const char *rel, const char *attr,
                       const char *type, const void *ptr, uint64_t n,
                       int on_device);
int pj_clear_catalog(pj_ctx *ctx);
int pj_plan_execute(pj_ctx *ctx, const char *plan_json, const char *base_dir,
                    char *out, uint64_t out_cap);

#ifdef __cplusplus
}
#endif
#endif
