/*
 * oracle.h -- CPU restatement of Proteus's join / group-by / reduce hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only to check
 * results (or as the timed CPU baseline).  The product library never links it.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Parity pins (see tests/test_oracle.py):
 *   - or_phjc_bucket (the 32-bit murmur3 finaliser): scikit-learn's
 *     murmurhash3_32 known answers.
 *   - or_hash_i32 (hashInt32, 64-bit arithmetic): PARITY UNPINNED.  The
 *     reference holds no vectors for it (its hashing test asserts only true,
 *     tests/olap/deprecated/tests-hashing.cpp:78-110) and no library routine
 *     computes it; only special cases (0, sign extension) are checked.  No
 *     checked result depends on it: counts, checksums and per-group
 *     aggregates do not depend on bucket layout.
 *   - or_join_chained(_mt): reference sailors goldens (74,383,10), closed
 *     forms (COUNT = |S| for a permutation build), dense direct-address
 *     oracle, brute force on tiny inputs.
 *   - or_groupby_chained: dense-array oracle, brute force, first-appearance
 *     order, exact FP64 sums for u/2^24 inputs.
 *   - generators: bijection / range / distribution invariants.
 *   - or_bloom_build / or_bloom_probe: set-based semi-join on the masked keys
 *     (brute force), no false negatives, exact semi-join when the filter
 *     covers every key, a one-entry filter passing every row.
 */
#ifndef PROTEUS_ORACLE_H
#define PROTEUS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- hashing ---------------------------------------------------------- */
uint64_t or_hash_i32(int32_t k);
uint32_t or_phjc_bucket(int32_t k, int log_parts, int hash_bits);
uint32_t or_rearrange_bucket(int32_t k, uint32_t buckets);

/* ---- seeded counter-based generators ---------------------------------- */
uint64_t or_mix64(uint64_t z);
uint64_t or_rng(uint64_t seed, uint64_t stream, uint64_t i);
uint64_t or_uniform(uint64_t seed, uint64_t stream, uint64_t i, uint64_t n);
uint64_t or_perm(uint64_t seed, uint64_t i, uint64_t n);
uint64_t or_zipf(uint64_t seed, uint64_t stream, uint64_t i, uint64_t n,
                 double theta);

void or_gen_perm_i32(int32_t *out, uint64_t n, uint64_t seed, int nthreads);
void or_gen_uniform_i32(int32_t *out, uint64_t n, uint64_t seed,
                        uint64_t stream, int64_t lo, uint64_t range,
                        int nthreads);
void or_gen_uniform_i64(int64_t *out, uint64_t n, uint64_t seed,
                        uint64_t stream, int64_t lo, uint64_t range,
                        int nthreads);
void or_gen_unit_f32(float *out, uint64_t n, uint64_t seed, uint64_t stream,
                     int nthreads);
void or_gen_zipf_fk_i32(int32_t *out, uint64_t n, uint64_t seed,
                        uint64_t stream, uint64_t nbuild, uint64_t range,
                        double theta, uint64_t perm_seed, int nthreads);
void or_gen_zipf_i32(int32_t *out, uint64_t n, uint64_t seed, uint64_t stream,
                     uint64_t g, double theta, uint64_t perm_seed,
                     int nthreads);

/* ---- join -> reduce ---------------------------------------------------- */
typedef struct {
  int64_t count;   /* number of (build, probe) matches                */
  int64_t sum_b;   /* SUM over matches of (int64) build value        */
  int64_t sum_p;   /* SUM over matches of (int64) probe value        */
  int64_t min_b, max_b, min_p, max_p; /* over matches; identity if none */
} or_join_out;

#define OR_OK 0
#define OR_E_CAPACITY 1
#define OR_E_ARG 2
#define OR_E_OOM 3

int or_join_chained(const int32_t *rk, const int32_t *rv, uint64_t nr,
                    const int32_t *sk, const int32_t *sv, uint64_t ns,
                    int hash_bits, uint64_t max_build, or_join_out *out);
int or_join_chained_mt(const int32_t *rk, const int32_t *rv, uint64_t nr,
                       const int32_t *sk, const int32_t *sv, uint64_t ns,
                       int hash_bits, uint64_t max_build, int nthreads,
                       or_join_out *out);
int or_join_dense(const int32_t *rk, const int32_t *rv, uint64_t nr,
                  const int32_t *sk, const int32_t *sv, uint64_t ns,
                  int nthreads, or_join_out *out);
int or_join_pairs(const int32_t *rk, uint64_t nr, const int32_t *sk,
                  uint64_t ns, int hash_bits, uint64_t max_build,
                  uint32_t *out_b, uint32_t *out_p, uint64_t cap,
                  uint64_t *n_out);

/* ---- group-by ---------------------------------------------------------- */
typedef struct {
  int32_t *keys;
  int64_t *cnt, *sum, *min, *max;
  double *fsum;
} or_groups;

int or_groupby_chained(const int32_t *k, const int64_t *v, const float *w,
                       uint64_t n, int hash_bits, uint64_t max_groups,
                       or_groups *out, uint64_t *ngroups);
/* or_groupby_chained over nthreads slices, partial tables merged per key
 * (the reference's per-worker group-by and final merge). */
int or_groupby_chained_mt(const int32_t *k, const int64_t *v, const float *w, uint64_t n,
                          int hash_bits, uint64_t max_groups, int nthreads, or_groups *out,
                          uint64_t *ngroups);
int or_groupby_dense(const int32_t *k, const int64_t *v, const float *w,
                     uint64_t n, uint64_t g, int nthreads, or_groups *out);

/* ---- reduce ------------------------------------------------------------ */
void or_reduce_i64(const int64_t *v, uint64_t n, int64_t *sum, int64_t *min,
                   int64_t *max);

/* ---- radix partition invariants ---------------------------------------- */
void or_histogram(const int32_t *k, uint64_t n, int shift, int bits,
                  uint64_t *hist);

/* ---- bloom filters ------------------------------------------------------ */
void or_bloom_build(uint8_t *filter, uint64_t size, const int32_t *keys,
                    uint64_t n);
uint64_t or_bloom_probe(const uint8_t *filter, uint64_t size,
                        const int32_t *keys, uint64_t n, uint32_t *out_idx);

#ifdef __cplusplus
}
#endif
#endif
