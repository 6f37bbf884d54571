/*
 * oracle.c -- CPU restatement of Proteus's GPU join / group-by hot path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg, to check results.  Never linked by the product.
 *
 * Reference files restated (paths relative to the reference repository):
 *   core/olap/lib/expressions/expressions-hasher.cpp:39-58
 *   core/olap/lib/operators/hash-join-chained.cpp:233-319,321-566,568-625
 *   core/olap/lib/operators/hash-join-chained-morsel.cpp:27-129
 *   core/olap/lib/operators/hash-group-by-chained.cpp:70-143,193-210,267-633
 *   core/olap/lib/operators/reduce-opt.cpp:52-430, agg-t.cpp:28
 *   core/olap/lib/operators/gpu/gpu-partitioned-hash-join-chained.cpp:595-610
 *   core/olap/lib/operators/gpu/gpu-hash-rearrange.cpp:338-348
 *   core/olap/lib/operators/bloom-filter/bloom-filter.cpp:36-107,
 *     bloom-filter-build.cpp:29-58, bloom-filter-probe.cpp:29-52
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

/* ======================================================================= */
/* hashing                                                                  */
/* ======================================================================= */

/* hashInt32 (expressions-hasher.cpp:39-58): sign-extend to 64 bits, then the
 * murmur3 fmix steps with LOGICAL shifts, all in 64-bit arithmetic. */
uint64_t or_hash_i32(int32_t k) {
  uint64_t h = (uint64_t)(int64_t)k;
  h ^= h >> 16;
  h *= 0x85ebca6bULL;
  h ^= h >> 13;
  h *= 0xc2b2ae35ULL;
  h ^= h >> 16;
  return h;
}

static inline uint32_t fmix32u(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bU;
  h ^= h >> 13;
  h *= 0xc2b2ae35U;
  h ^= h >> 16;
  return h;
}

/* GpuPartitionedHashJoinChained::hash (gpu-partitioned-hash-join-chained.cpp:
 * 595-610): 32-bit fmix with logical shifts, >> log_parts, & (2^hash_bits-1). */
uint32_t or_phjc_bucket(int32_t k, int log_parts, int hash_bits) {
  uint32_t h = fmix32u((uint32_t)k);
  h >>= log_parts;
  return h & ((1U << hash_bits) - 1U);
}

/* GpuHashRearrange target (gpu-hash-rearrange.cpp:338-348):
 * URem(hashInt32(e), numOfBuckets). */
uint32_t or_rearrange_bucket(int32_t k, uint32_t buckets) {
  return (uint32_t)(or_hash_i32(k) % (uint64_t)buckets);
}

/* ======================================================================= */
/* generators: counter-based, identical on host and device                  */
/* ======================================================================= */

/* splitmix64 finalizer (Steele, Lea, Flood, OOPSLA 2014). */
uint64_t or_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

#define GOLDEN 0x9e3779b97f4a7c15ULL

/* i-th output of the splitmix64 sequence keyed by (seed, stream). */
uint64_t or_rng(uint64_t seed, uint64_t stream, uint64_t i) {
  uint64_t key = or_mix64(seed + GOLDEN * (stream + 1));
  return or_mix64(key + GOLDEN * (i + 1));
}

static inline uint64_t mulhi64(uint64_t a, uint64_t b) {
  return (uint64_t)(((unsigned __int128)a * b) >> 64);
}

/* uniform in [0, n): multiply-high of a 64-bit draw. */
uint64_t or_uniform(uint64_t seed, uint64_t stream, uint64_t i, uint64_t n) {
  return mulhi64(or_rng(seed, stream, i), n);
}

/* Seeded bijection of [0, 2^k): 4 rounds of (x*odd, x^=x>>s, x+=c) mod 2^k. */
static uint64_t perm_pow2(uint64_t seed, uint64_t x, int k) {
  uint64_t mask = (k >= 64) ? ~0ULL : ((1ULL << k) - 1ULL);
  int s = (k + 1) / 2;
  for (int r = 0; r < 4; ++r) {
    uint64_t m = or_rng(seed, 0xFEED0000ULL + (uint64_t)r, 0) | 1ULL;
    uint64_t c = or_rng(seed, 0xFEED0000ULL + (uint64_t)r, 1);
    x = (x * m) & mask;
    x ^= x >> s;
    x = (x + c) & mask;
  }
  return x;
}

/* Seeded bijection of [0, n) by cycle-walking the power-of-two bijection. */
uint64_t or_perm(uint64_t seed, uint64_t i, uint64_t n) {
  if (n <= 1) return 0;
  int k = 1;
  while (k < 63 && (1ULL << k) < n) ++k;
  uint64_t x = i;
  do {
    x = perm_pow2(seed, x, k);
  } while (x >= n);
  return x;
}

/* Zipf(theta) over ranks [0, n) by rejection-inversion (W. Hormann and
 * G. Derflinger, "Rejection-inversion to generate variates from monotone
 * discrete distributions", ACM TOMACS 6(3), 1996).  Attempt a draws its
 * uniform from stream (stream + a * 2^32); at most 64 attempts. */
static double zh1(double x) { /* log1p(x)/x */
  return fabs(x) > 1e-8 ? log1p(x) / x : 1.0 - x * (0.5 - x * (1.0 / 3.0 - 0.25 * x));
}
static double zh2(double x) { /* expm1(x)/x */
  return fabs(x) > 1e-8 ? expm1(x) / x : 1.0 + x * 0.5 * (1.0 + x * (1.0 / 3.0) * (1.0 + 0.25 * x));
}
static double zH(double x, double s) {
  double lx = log(x);
  return zh2((1.0 - s) * lx) * lx;
}
static double zHinv(double x, double s) {
  double t = x * (1.0 - s);
  if (t < -1.0) t = -1.0;
  return exp(zh1(t) * x);
}
static double zhh(double x, double s) { return exp(-s * log(x)); }

uint64_t or_zipf(uint64_t seed, uint64_t stream, uint64_t i, uint64_t n,
                 double theta) {
  if (n <= 1) return 0;
  if (theta <= 0.0) return or_uniform(seed, stream, i, n);
  double s = theta;
  double hx1 = zH(1.5, s) - 1.0;
  double hn = zH((double)n + 0.5, s);
  double sp = 2.0 - zHinv(zH(2.5, s) - zhh(2.0, s), s);
  uint64_t k = 1;
  for (uint64_t a = 0; a < 64; ++a) {
    uint64_t r = or_rng(seed, stream + (a << 32), i);
    double u01 = (double)((r >> 11) + 1) * (1.0 / 9007199254740992.0);
    double u = hn + u01 * (hx1 - hn);
    double x = zHinv(u, s);
    double kd = floor(x + 0.5);
    if (kd < 1.0) kd = 1.0;
    if (kd > (double)n) kd = (double)n;
    k = (uint64_t)kd;
    if (kd - x <= sp || u >= zH(kd + 0.5, s) - zhh(kd, s)) break;
  }
  return k - 1;
}

/* ---- tiny parallel-for on pthreads ------------------------------------- */
typedef void (*pf_body)(void *ctx, uint64_t lo, uint64_t hi);
typedef struct {
  pf_body f;
  void *ctx;
  uint64_t lo, hi;
} pf_task;
static void *pf_run(void *p) {
  pf_task *t = (pf_task *)p;
  t->f(t->ctx, t->lo, t->hi);
  return NULL;
}
static void parallel_for(uint64_t n, int nthreads, pf_body f, void *ctx) {
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  if (n < 65536 || nthreads == 1) {
    f(ctx, 0, n);
    return;
  }
  pthread_t th[256];
  pf_task tk[256];
  uint64_t per = (n + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; ++t) {
    tk[t].f = f;
    tk[t].ctx = ctx;
    tk[t].lo = per * t < n ? per * t : n;
    tk[t].hi = per * (t + 1) < n ? per * (t + 1) : n;
    pthread_create(&th[t], NULL, pf_run, &tk[t]);
  }
  for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
}

typedef struct {
  void *out;
  uint64_t seed, stream, n, range, nbuild, perm_seed, g;
  int64_t lo;
  double theta;
} gen_ctx;

static void gen_perm_body(void *c, uint64_t lo, uint64_t hi) {
  gen_ctx *g = (gen_ctx *)c;
  int32_t *o = (int32_t *)g->out;
  for (uint64_t i = lo; i < hi; ++i) o[i] = (int32_t)or_perm(g->seed, i, g->n);
}
void or_gen_perm_i32(int32_t *out, uint64_t n, uint64_t seed, int nthreads) {
  gen_ctx c = {out, seed, 0, n, 0, 0, 0, 0, 0, 0.0};
  parallel_for(n, nthreads, gen_perm_body, &c);
}

static void gen_u32_body(void *c, uint64_t lo, uint64_t hi) {
  gen_ctx *g = (gen_ctx *)c;
  int32_t *o = (int32_t *)g->out;
  for (uint64_t i = lo; i < hi; ++i)
    o[i] = (int32_t)(g->lo + (int64_t)or_uniform(g->seed, g->stream, i, g->range));
}
void or_gen_uniform_i32(int32_t *out, uint64_t n, uint64_t seed,
                        uint64_t stream, int64_t lo, uint64_t range,
                        int nthreads) {
  gen_ctx c = {out, seed, stream, n, range, 0, 0, 0, lo, 0.0};
  parallel_for(n, nthreads, gen_u32_body, &c);
}

static void gen_u64_body(void *c, uint64_t lo, uint64_t hi) {
  gen_ctx *g = (gen_ctx *)c;
  int64_t *o = (int64_t *)g->out;
  for (uint64_t i = lo; i < hi; ++i)
    o[i] = g->lo + (int64_t)or_uniform(g->seed, g->stream, i, g->range);
}
void or_gen_uniform_i64(int64_t *out, uint64_t n, uint64_t seed,
                        uint64_t stream, int64_t lo, uint64_t range,
                        int nthreads) {
  gen_ctx c = {out, seed, stream, n, range, 0, 0, 0, lo, 0.0};
  parallel_for(n, nthreads, gen_u64_body, &c);
}

/* float32 = u / 2^24 with u uniform in [0, 2^24): exact in f32 and f64. */
static void gen_f32_body(void *c, uint64_t lo, uint64_t hi) {
  gen_ctx *g = (gen_ctx *)c;
  float *o = (float *)g->out;
  for (uint64_t i = lo; i < hi; ++i)
    o[i] = (float)or_uniform(g->seed, g->stream, i, 1ULL << 24) * (1.0f / 16777216.0f);
}
void or_gen_unit_f32(float *out, uint64_t n, uint64_t seed, uint64_t stream,
                     int nthreads) {
  gen_ctx c = {out, seed, stream, n, 0, 0, 0, 0, 0, 0.0};
  parallel_for(n, nthreads, gen_f32_body, &c);
}

/* Probe foreign keys: a uniform draw u over [0, range); u >= nbuild is a miss
 * key u, otherwise the key is the build key at Zipf rank r (build row r holds
 * perm(perm_seed, r, nbuild)).  The miss rate equals the uniform variant's. */
static void gen_zfk_body(void *c, uint64_t lo, uint64_t hi) {
  gen_ctx *g = (gen_ctx *)c;
  int32_t *o = (int32_t *)g->out;
  for (uint64_t i = lo; i < hi; ++i) {
    uint64_t u = or_uniform(g->seed, g->stream, i, g->range);
    if (u >= g->nbuild) {
      o[i] = (int32_t)u;
    } else {
      uint64_t r = or_zipf(g->seed, g->stream + 7, i, g->nbuild, g->theta);
      o[i] = (int32_t)or_perm(g->perm_seed, r, g->nbuild);
    }
  }
}
void or_gen_zipf_fk_i32(int32_t *out, uint64_t n, uint64_t seed,
                        uint64_t stream, uint64_t nbuild, uint64_t range,
                        double theta, uint64_t perm_seed, int nthreads) {
  gen_ctx c = {out, seed, stream, n, range, nbuild, perm_seed, 0, 0, theta};
  parallel_for(n, nthreads, gen_zfk_body, &c);
}

static void gen_zg_body(void *c, uint64_t lo, uint64_t hi) {
  gen_ctx *g = (gen_ctx *)c;
  int32_t *o = (int32_t *)g->out;
  for (uint64_t i = lo; i < hi; ++i) {
    uint64_t r = or_zipf(g->seed, g->stream, i, g->g, g->theta);
    o[i] = (int32_t)or_perm(g->perm_seed, r, g->g);
  }
}
void or_gen_zipf_i32(int32_t *out, uint64_t n, uint64_t seed, uint64_t stream,
                     uint64_t g, double theta, uint64_t perm_seed,
                     int nthreads) {
  gen_ctx c = {out, seed, stream, n, 0, 0, perm_seed, g, 0, theta};
  parallel_for(n, nthreads, gen_zg_body, &c);
}

/* ======================================================================= */
/* join -> reduce                                                           */
/* ======================================================================= */

static void jo_init(or_join_out *o) {
  o->count = 0;
  o->sum_b = 0;
  o->sum_p = 0;
  o->min_b = INT64_MAX;
  o->max_b = INT64_MIN;
  o->min_p = INT64_MAX;
  o->max_p = INT64_MIN;
}
static inline void jo_match(or_join_out *o, int64_t b, int64_t p) {
  o->count += 1;
  o->sum_b = (int64_t)((uint64_t)o->sum_b + (uint64_t)b);
  o->sum_p = (int64_t)((uint64_t)o->sum_p + (uint64_t)p);
  if (b < o->min_b) o->min_b = b;
  if (b > o->max_b) o->max_b = b;
  if (p < o->min_p) o->min_p = p;
  if (p > o->max_p) o->max_p = p;
}
static void jo_merge(or_join_out *o, const or_join_out *x) {
  o->count += x->count;
  o->sum_b = (int64_t)((uint64_t)o->sum_b + (uint64_t)x->sum_b);
  o->sum_p = (int64_t)((uint64_t)o->sum_p + (uint64_t)x->sum_p);
  if (x->min_b < o->min_b) o->min_b = x->min_b;
  if (x->max_b > o->max_b) o->max_b = x->max_b;
  if (x->min_p < o->min_p) o->min_p = x->min_p;
  if (x->max_p > o->max_p) o->max_p = x->max_p;
}

/* HashJoinChained (hash-join-chained.cpp): head[2^hash_bits] = ~0
 * (open_build :568-597); build row i: idx = cnt++ (:241-252),
 * old = xchg(head[h], idx) (:254-260), packet[idx] = {old, key, value}
 * (:262-319); probe walks the chain from head[h] emitting every key match
 * (:321-566); close_build checks cnt <= maxBuildInputSize (:599-608).
 * h = hashInt32(key) URem 2^hash_bits (:233-239). */
int or_join_chained(const int32_t *rk, const int32_t *rv, uint64_t nr,
                    const int32_t *sk, const int32_t *sv, uint64_t ns,
                    int hash_bits, uint64_t max_build, or_join_out *out) {
  jo_init(out);
  if (hash_bits < 0 || hash_bits > 32) return OR_E_ARG;
  if (nr > max_build) return OR_E_CAPACITY;
  uint64_t hsz = 1ULL << hash_bits, mask = hsz - 1;
  uint32_t *head = (uint32_t *)malloc(hsz * 4);
  uint32_t *next = (uint32_t *)malloc((nr ? nr : 1) * 4);
  if (!head || !next) {
    free(head);
    free(next);
    return OR_E_OOM;
  }
  memset(head, 0xff, hsz * 4);
  uint32_t cnt = 0;
  for (uint64_t i = 0; i < nr; ++i) {
    uint64_t h = or_hash_i32(rk[i]) & mask;
    uint32_t idx = cnt++;
    next[idx] = head[h];
    head[h] = idx;
  }
  for (uint64_t j = 0; j < ns; ++j) {
    uint64_t h = or_hash_i32(sk[j]) & mask;
    for (uint32_t cur = head[h]; cur != 0xffffffffU; cur = next[cur])
      if (rk[cur] == sk[j]) jo_match(out, rv ? rv[cur] : 0, sv ? sv[j] : 0);
  }
  free(head);
  free(next);
  return OR_OK;
}

/* HashJoinChainedMorsel (hash-join-chained-morsel.cpp:27-129): one shared
 * table; every worker inserts with seq_cst fetch_add(cnt) and seq_cst
 * exchange(head[h]); probe workers walk the shared chains.  Morsels of 16Ki
 * rows are handed out by an atomic cursor. */
typedef struct {
  const int32_t *rk, *rv, *sk, *sv;
  uint64_t nr, ns, mask;
  _Atomic uint32_t *head;
  uint32_t *next;
  uint32_t *idx_key; /* build row of each table index */
  _Atomic uint32_t cnt;
  _Atomic uint64_t morsel;
  pthread_barrier_t bar;
  or_join_out part[256];
} mt_join;
typedef struct {
  mt_join *j;
  int t;
} mt_arg;

#define MORSEL 16384ULL
static void *mt_join_worker(void *p) {
  mt_arg *a = (mt_arg *)p;
  mt_join *j = a->j;
  for (;;) {
    uint64_t m = atomic_fetch_add(&j->morsel, 1);
    uint64_t lo = m * MORSEL;
    if (lo >= j->nr) break;
    uint64_t hi = lo + MORSEL < j->nr ? lo + MORSEL : j->nr;
    for (uint64_t i = lo; i < hi; ++i) {
      uint64_t h = or_hash_i32(j->rk[i]) & j->mask;
      uint32_t idx = atomic_fetch_add_explicit(&j->cnt, 1, memory_order_seq_cst);
      uint32_t old = atomic_exchange_explicit(&j->head[h], idx, memory_order_seq_cst);
      j->next[idx] = old;
      j->idx_key[idx] = (uint32_t)i;
    }
  }
  pthread_barrier_wait(&j->bar);
  if (a->t == 0) atomic_store(&j->morsel, 0);
  pthread_barrier_wait(&j->bar);
  or_join_out acc;
  jo_init(&acc);
  for (;;) {
    uint64_t m = atomic_fetch_add(&j->morsel, 1);
    uint64_t lo = m * MORSEL;
    if (lo >= j->ns) break;
    uint64_t hi = lo + MORSEL < j->ns ? lo + MORSEL : j->ns;
    for (uint64_t s = lo; s < hi; ++s) {
      int32_t k = j->sk[s];
      uint64_t h = or_hash_i32(k) & j->mask;
      for (uint32_t cur = atomic_load_explicit(&j->head[h], memory_order_relaxed);
           cur != 0xffffffffU; cur = j->next[cur]) {
        uint32_t r = j->idx_key[cur];
        if (j->rk[r] == k) jo_match(&acc, j->rv ? j->rv[r] : 0, j->sv ? j->sv[s] : 0);
      }
    }
  }
  j->part[a->t] = acc;
  return NULL;
}

int or_join_chained_mt(const int32_t *rk, const int32_t *rv, uint64_t nr,
                       const int32_t *sk, const int32_t *sv, uint64_t ns,
                       int hash_bits, uint64_t max_build, int nthreads,
                       or_join_out *out) {
  jo_init(out);
  if (hash_bits < 0 || hash_bits > 32) return OR_E_ARG;
  if (nr > max_build) return OR_E_CAPACITY;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  mt_join *j = (mt_join *)calloc(1, sizeof(mt_join));
  if (!j) return OR_E_OOM;
  uint64_t hsz = 1ULL << hash_bits;
  j->head = (_Atomic uint32_t *)malloc(hsz * 4);
  j->next = (uint32_t *)malloc((nr ? nr : 1) * 4);
  j->idx_key = (uint32_t *)malloc((nr ? nr : 1) * 4);
  if (!j->head || !j->next || !j->idx_key) {
    free((void *)j->head);
    free(j->next);
    free(j->idx_key);
    free(j);
    return OR_E_OOM;
  }
  memset((void *)j->head, 0xff, hsz * 4);
  j->rk = rk;
  j->rv = rv;
  j->sk = sk;
  j->sv = sv;
  j->nr = nr;
  j->ns = ns;
  j->mask = hsz - 1;
  atomic_init(&j->cnt, 0);
  atomic_init(&j->morsel, 0);
  pthread_barrier_init(&j->bar, NULL, (unsigned)nthreads);
  pthread_t th[256];
  mt_arg args[256];
  for (int t = 0; t < nthreads; ++t) {
    args[t].j = j;
    args[t].t = t;
    pthread_create(&th[t], NULL, mt_join_worker, &args[t]);
  }
  for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
  for (int t = 0; t < nthreads; ++t) jo_merge(out, &j->part[t]);
  pthread_barrier_destroy(&j->bar);
  free((void *)j->head);
  free(j->next);
  free(j->idx_key);
  free(j);
  return OR_OK;
}

/* Direct-address oracle: requires the build keys to be distinct and in
 * [0, nr).  A probe key matches iff 0 <= key < nr and that key was built.
 * Hash-free, so it pins the chained restatement independently. */
typedef struct {
  const int32_t *sk, *sv;
  const int32_t *val;
  const uint8_t *present;
  uint64_t nr;
  or_join_out part[256];
  int nthreads;
} dense_ctx;
typedef struct {
  dense_ctx *d;
  int t;
  uint64_t lo, hi;
} dense_arg;
static void *dense_worker(void *p) {
  dense_arg *a = (dense_arg *)p;
  dense_ctx *d = a->d;
  or_join_out acc;
  jo_init(&acc);
  for (uint64_t s = a->lo; s < a->hi; ++s) {
    int32_t k = d->sk[s];
    if (k >= 0 && (uint64_t)k < d->nr && d->present[k])
      jo_match(&acc, d->val ? d->val[k] : 0, d->sv ? d->sv[s] : 0);
  }
  d->part[a->t] = acc;
  return NULL;
}
int or_join_dense(const int32_t *rk, const int32_t *rv, uint64_t nr,
                  const int32_t *sk, const int32_t *sv, uint64_t ns,
                  int nthreads, or_join_out *out) {
  jo_init(out);
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  int32_t *val = (int32_t *)malloc((nr ? nr : 1) * 4);
  uint8_t *present = (uint8_t *)calloc(nr ? nr : 1, 1);
  if (!val || !present) {
    free(val);
    free(present);
    return OR_E_OOM;
  }
  for (uint64_t i = 0; i < nr; ++i) {
    int32_t k = rk[i];
    if (k < 0 || (uint64_t)k >= nr || present[k]) {
      free(val);
      free(present);
      return OR_E_ARG;
    }
    present[k] = 1;
    val[k] = rv ? rv[i] : 0;
  }
  dense_ctx d;
  d.sk = sk;
  d.sv = sv;
  d.val = val;
  d.present = present;
  d.nr = nr;
  pthread_t th[256];
  dense_arg args[256];
  uint64_t per = (ns + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; ++t) {
    args[t].d = &d;
    args[t].t = t;
    args[t].lo = per * t < ns ? per * t : ns;
    args[t].hi = per * (t + 1) < ns ? per * (t + 1) : ns;
    pthread_create(&th[t], NULL, dense_worker, &args[t]);
  }
  for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
  for (int t = 0; t < nthreads; ++t) jo_merge(out, &d.part[t]);
  free(val);
  free(present);
  return OR_OK;
}

/* Materialised variant of the chained join: every (build row, probe row)
 * pair the probe loop of hash-join-chained.cpp:321-566 hands to its parent. */
int or_join_pairs(const int32_t *rk, uint64_t nr, const int32_t *sk,
                  uint64_t ns, int hash_bits, uint64_t max_build,
                  uint32_t *out_b, uint32_t *out_p, uint64_t cap,
                  uint64_t *n_out) {
  *n_out = 0;
  if (hash_bits < 0 || hash_bits > 32) return OR_E_ARG;
  if (nr > max_build) return OR_E_CAPACITY;
  uint64_t hsz = 1ULL << hash_bits, mask = hsz - 1;
  uint32_t *head = (uint32_t *)malloc(hsz * 4);
  uint32_t *next = (uint32_t *)malloc((nr ? nr : 1) * 4);
  if (!head || !next) {
    free(head);
    free(next);
    return OR_E_OOM;
  }
  memset(head, 0xff, hsz * 4);
  for (uint64_t i = 0; i < nr; ++i) {
    uint64_t h = or_hash_i32(rk[i]) & mask;
    next[i] = head[h];
    head[h] = (uint32_t)i;
  }
  uint64_t n = 0;
  int rc = OR_OK;
  for (uint64_t j = 0; j < ns; ++j) {
    uint64_t h = or_hash_i32(sk[j]) & mask;
    for (uint32_t cur = head[h]; cur != 0xffffffffU; cur = next[cur])
      if (rk[cur] == sk[j]) {
        if (n < cap) {
          out_b[n] = cur;
          out_p[n] = (uint32_t)j;
        } else {
          rc = OR_E_CAPACITY;
        }
        ++n;
      }
  }
  *n_out = n;
  free(head);
  free(next);
  return rc;
}

/* ======================================================================= */
/* group-by                                                                 */
/* ======================================================================= */

/* HashGroupByChained (hash-group-by-chained.cpp): hash = hashInt32(key)
 * URem 2^hash_bits (:193-210); walk the chain from head[h] comparing keys;
 * a hit updates the aggregates, a miss appends a new entry at cnt++ whose
 * aggregates start from the row (:267-633); the scan emits entries in index
 * order, i.e. first-appearance order.  Aggregates here: COUNT = SUM(int64 1)
 * (agg-t.cpp:28), SUM/MIN/MAX(int64 v), SUM((double) w) -- Proteus float is
 * FP64 (expressionTypes.cpp:87-89). */
int or_groupby_chained(const int32_t *k, const int64_t *v, const float *w,
                       uint64_t n, int hash_bits, uint64_t max_groups,
                       or_groups *out, uint64_t *ngroups) {
  *ngroups = 0;
  if (hash_bits < 0 || hash_bits > 32) return OR_E_ARG;
  uint64_t hsz = 1ULL << hash_bits, mask = hsz - 1;
  uint32_t *head = (uint32_t *)malloc(hsz * 4);
  uint32_t *next = (uint32_t *)malloc((max_groups ? max_groups : 1) * 4);
  if (!head || !next) {
    free(head);
    free(next);
    return OR_E_OOM;
  }
  memset(head, 0xff, hsz * 4);
  uint64_t cnt = 0;
  int rc = OR_OK;
  for (uint64_t i = 0; i < n; ++i) {
    int32_t key = k[i];
    uint64_t h = or_hash_i32(key) & mask;
    uint32_t cur = head[h], last = 0xffffffffU;
    while (cur != 0xffffffffU && out->keys[cur] != key) {
      last = cur;
      cur = next[cur];
    }
    int64_t vi = v ? v[i] : 0;
    double wi = w ? (double)w[i] : 0.0;
    if (cur == 0xffffffffU) {
      if (cnt >= max_groups) {
        rc = OR_E_CAPACITY;
        break;
      }
      cur = (uint32_t)cnt++;
      next[cur] = 0xffffffffU;
      if (last == 0xffffffffU)
        head[h] = cur;
      else
        next[last] = cur;
      out->keys[cur] = key;
      out->cnt[cur] = 1;
      out->sum[cur] = vi;
      out->min[cur] = vi;
      out->max[cur] = vi;
      out->fsum[cur] = wi;
    } else {
      out->cnt[cur] += 1;
      out->sum[cur] = (int64_t)((uint64_t)out->sum[cur] + (uint64_t)vi);
      if (vi < out->min[cur]) out->min[cur] = vi;
      if (vi > out->max[cur]) out->max[cur] = vi;
      out->fsum[cur] += wi;
    }
  }
  *ngroups = cnt;
  free(head);
  free(next);
  return rc;
}

/* The same group-by with nthreads workers, as the reference runs a CPU
 * group-by under a router: every worker builds its own chained table over a
 * contiguous slice of the rows (or_groupby_chained), then one pass merges the
 * partial groups in worker order into the output table with the monoids'
 * combine (COUNT and SUM add, MIN / MAX compare, FP64 SUM adds).  Output
 * order: first appearance in the concatenation of the partials. */
typedef struct {
  const int32_t *k;
  const int64_t *v;
  const float *w;
  uint64_t lo, hi, max_groups;
  int hash_bits, rc;
  or_groups part;
  uint64_t ng;
} gbmt_arg;
static void *gbmt_worker(void *p) {
  gbmt_arg *a = (gbmt_arg *)p;
  a->rc = or_groupby_chained(a->k + a->lo, a->v ? a->v + a->lo : NULL, a->w ? a->w + a->lo : NULL,
                             a->hi - a->lo, a->hash_bits, a->max_groups, &a->part, &a->ng);
  return NULL;
}
int or_groupby_chained_mt(const int32_t *k, const int64_t *v, const float *w, uint64_t n,
                          int hash_bits, uint64_t max_groups, int nthreads, or_groups *out,
                          uint64_t *ngroups) {
  *ngroups = 0;
  if (hash_bits < 0 || hash_bits > 32) return OR_E_ARG;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 64) nthreads = 64;
  const uint64_t cap = max_groups ? max_groups : 1;
  gbmt_arg args[64];
  pthread_t th[64];
  int rc = OR_OK;
  const uint64_t per = (n + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; ++t) {
    gbmt_arg *a = &args[t];
    memset(a, 0, sizeof(*a));
    a->k = k;
    a->v = v;
    a->w = w;
    a->lo = per * t < n ? per * t : n;
    a->hi = per * (t + 1) < n ? per * (t + 1) : n;
    a->max_groups = max_groups;
    a->hash_bits = hash_bits;
    a->part.keys = (int32_t *)malloc(cap * 4);
    a->part.cnt = (int64_t *)malloc(cap * 8);
    a->part.sum = (int64_t *)malloc(cap * 8);
    a->part.min = (int64_t *)malloc(cap * 8);
    a->part.max = (int64_t *)malloc(cap * 8);
    a->part.fsum = (double *)malloc(cap * 8);
    if (!a->part.keys || !a->part.cnt || !a->part.sum || !a->part.min || !a->part.max ||
        !a->part.fsum)
      rc = OR_E_OOM;
  }
  if (rc == OR_OK) {
    for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, gbmt_worker, &args[t]);
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    for (int t = 0; t < nthreads; ++t)
      if (args[t].rc != OR_OK) rc = args[t].rc;
  }
  if (rc == OR_OK) {  // merge the partial tables
    const uint64_t hsz = 1ULL << hash_bits, mask = hsz - 1;
    uint32_t *head = (uint32_t *)malloc(hsz * 4);
    uint32_t *next = (uint32_t *)malloc(cap * 4);
    if (!head || !next) rc = OR_E_OOM;
    if (rc == OR_OK) memset(head, 0xff, hsz * 4);
    uint64_t cnt = 0;
    for (int t = 0; t < nthreads && rc == OR_OK; ++t) {
      const or_groups *p = &args[t].part;
      for (uint64_t i = 0; i < args[t].ng; ++i) {
        const int32_t key = p->keys[i];
        const uint64_t h = or_hash_i32(key) & mask;
        uint32_t cur = head[h], last = 0xffffffffU;
        while (cur != 0xffffffffU && out->keys[cur] != key) {
          last = cur;
          cur = next[cur];
        }
        if (cur == 0xffffffffU) {
          if (cnt >= max_groups) {
            rc = OR_E_CAPACITY;
            break;
          }
          cur = (uint32_t)cnt++;
          next[cur] = 0xffffffffU;
          if (last == 0xffffffffU)
            head[h] = cur;
          else
            next[last] = cur;
          out->keys[cur] = key;
          out->cnt[cur] = p->cnt[i];
          out->sum[cur] = p->sum[i];
          out->min[cur] = p->min[i];
          out->max[cur] = p->max[i];
          out->fsum[cur] = p->fsum[i];
        } else {
          out->cnt[cur] += p->cnt[i];
          out->sum[cur] = (int64_t)((uint64_t)out->sum[cur] + (uint64_t)p->sum[i]);
          if (p->min[i] < out->min[cur]) out->min[cur] = p->min[i];
          if (p->max[i] > out->max[cur]) out->max[cur] = p->max[i];
          out->fsum[cur] += p->fsum[i];
        }
      }
    }
    *ngroups = cnt;
    free(head);
    free(next);
  }
  for (int t = 0; t < nthreads; ++t) {
    free(args[t].part.keys);
    free(args[t].part.cnt);
    free(args[t].part.sum);
    free(args[t].part.min);
    free(args[t].part.max);
    free(args[t].part.fsum);
  }
  return rc;
}

/* Dense-array oracle for keys in [0, g): per-key COUNT/SUM/MIN/MAX exact and
 * the float sum accumulated as an exact int64 count of 2^-24 units when every
 * w is a multiple of 2^-24 below 1 (the generator's contract); otherwise a
 * plain FP64 sum.  Absent keys keep cnt = 0. */
typedef struct {
  const int32_t *k;
  const int64_t *v;
  const float *w;
  uint64_t g;
  int64_t *cnt, *sum, *min, *max;
  int64_t *units;
  double *fs;
  pthread_mutex_t mu;
} gd_ctx;
typedef struct {
  gd_ctx *c;
  uint64_t lo, hi;
} gd_arg;
static void *gd_worker(void *p) {
  gd_arg *a = (gd_arg *)p;
  gd_ctx *c = a->c;
  uint64_t g = c->g;
  int64_t *cnt = (int64_t *)calloc(g, 8), *sum = (int64_t *)calloc(g, 8);
  int64_t *mn = (int64_t *)malloc(g * 8), *mx = (int64_t *)malloc(g * 8);
  int64_t *un = (int64_t *)calloc(g, 8);
  double *fs = (double *)calloc(g, 8);
  for (uint64_t x = 0; x < g; ++x) {
    mn[x] = INT64_MAX;
    mx[x] = INT64_MIN;
  }
  for (uint64_t i = a->lo; i < a->hi; ++i) {
    uint64_t key = (uint64_t)(uint32_t)c->k[i];
    int64_t vi = c->v ? c->v[i] : 0;
    cnt[key] += 1;
    sum[key] = (int64_t)((uint64_t)sum[key] + (uint64_t)vi);
    if (vi < mn[key]) mn[key] = vi;
    if (vi > mx[key]) mx[key] = vi;
    if (c->w) {
      double wd = (double)c->w[i];
      un[key] += (int64_t)(wd * 16777216.0);
      fs[key] += wd;
    }
  }
  pthread_mutex_lock(&c->mu);
  for (uint64_t x = 0; x < g; ++x) {
    c->cnt[x] += cnt[x];
    c->sum[x] = (int64_t)((uint64_t)c->sum[x] + (uint64_t)sum[x]);
    if (mn[x] < c->min[x]) c->min[x] = mn[x];
    if (mx[x] > c->max[x]) c->max[x] = mx[x];
    c->units[x] += un[x];
    c->fs[x] += fs[x];
  }
  pthread_mutex_unlock(&c->mu);
  free(cnt);
  free(sum);
  free(mn);
  free(mx);
  free(un);
  free(fs);
  return NULL;
}
int or_groupby_dense(const int32_t *k, const int64_t *v, const float *w,
                     uint64_t n, uint64_t g, int nthreads, or_groups *out) {
  for (uint64_t i = 0; i < n; ++i)
    if (k[i] < 0 || (uint64_t)k[i] >= g) return OR_E_ARG;
  int exact = 1;
  if (w)
    for (uint64_t i = 0; i < n; ++i) {
      double u = (double)w[i] * 16777216.0;
      if (!(w[i] >= 0.0f && w[i] < 1.0f) || u != floor(u)) {
        exact = 0;
        break;
      }
    }
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 64) nthreads = 64;
  gd_ctx c;
  c.k = k;
  c.v = v;
  c.w = w;
  c.g = g;
  c.cnt = out->cnt;
  c.sum = out->sum;
  c.min = out->min;
  c.max = out->max;
  c.units = (int64_t *)calloc(g ? g : 1, 8);
  c.fs = (double *)calloc(g ? g : 1, 8);
  pthread_mutex_init(&c.mu, NULL);
  for (uint64_t x = 0; x < g; ++x) {
    out->keys[x] = (int32_t)x;
    out->cnt[x] = 0;
    out->sum[x] = 0;
    out->min[x] = INT64_MAX;
    out->max[x] = INT64_MIN;
  }
  pthread_t th[64];
  gd_arg args[64];
  uint64_t per = (n + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; ++t) {
    args[t].c = &c;
    args[t].lo = per * t < n ? per * t : n;
    args[t].hi = per * (t + 1) < n ? per * (t + 1) : n;
    pthread_create(&th[t], NULL, gd_worker, &args[t]);
  }
  for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
  for (uint64_t x = 0; x < g; ++x)
    out->fsum[x] = exact ? (double)c.units[x] / 16777216.0 : c.fs[x];
  pthread_mutex_destroy(&c.mu);
  free(c.units);
  free(c.fs);
  return OR_OK;
}

/* ======================================================================= */
/* reduce                                                                   */
/* ======================================================================= */

/* opt::Reduce (reduce-opt.cpp:52-430): identity-initialised accumulators
 * (SUM 0, MIN +max, MAX -max; resetAccumulator :374-430) updated per tuple. */
void or_reduce_i64(const int64_t *v, uint64_t n, int64_t *sum, int64_t *min,
                   int64_t *max) {
  uint64_t s = 0;
  int64_t mn = INT64_MAX, mx = INT64_MIN;
  for (uint64_t i = 0; i < n; ++i) {
    s += (uint64_t)v[i];
    if (v[i] < mn) mn = v[i];
    if (v[i] > mx) mx = v[i];
  }
  *sum = (int64_t)s;
  *min = mn;
  *max = mx;
}

/* ======================================================================= */
/* radix-partition invariants                                               */
/* ======================================================================= */

/* Histogram of partition digits: digit = (fmix32(key) >> shift) & (2^bits-1),
 * fmix32 being the logical-shift 32-bit murmur finaliser of
 * gpu-partitioned-hash-join-chained.cpp:595-610. */
void or_histogram(const int32_t *k, uint64_t n, int shift, int bits,
                  uint64_t *hist) {
  uint32_t mask = (bits >= 32) ? 0xffffffffU : ((1U << bits) - 1U);
  memset(hist, 0, ((uint64_t)mask + 1) * 8);
  for (uint64_t i = 0; i < n; ++i) hist[(fmix32u((uint32_t)k[i]) >> shift) & mask]++;
}

/* ======================================================================= */
/* bloom filters                                                            */
/* ======================================================================= */

/* BloomFilterBuild (bloom-filter-build.cpp:29-58): a zeroed array of `size`
 * one-byte flags (produce_ memsets it), and per input row
 * filter[key & (size - 1)] = true (findInFilter, bloom-filter.cpp:85-107:
 * the key itself, masked; size a power of two).  The mask is applied to the
 * int32 key's bits, so negative keys select high entries. */
void or_bloom_build(uint8_t *filter, uint64_t size, const int32_t *keys,
                    uint64_t n) {
  const uint32_t mask = (uint32_t)(size - 1);
  memset(filter, 0, size);
  for (uint64_t i = 0; i < n; ++i) filter[(uint32_t)keys[i] & mask] = 1;
}

/* BloomFilterProbe (bloom-filter-probe.cpp:43-52): the row passes when its
 * key's flag is set; rows keep their order.  Returns the number passed and
 * writes their row ids. */
uint64_t or_bloom_probe(const uint8_t *filter, uint64_t size,
                        const int32_t *keys, uint64_t n, uint32_t *out_idx) {
  const uint32_t mask = (uint32_t)(size - 1);
  uint64_t m = 0;
  for (uint64_t i = 0; i < n; ++i)
    if (filter[(uint32_t)keys[i] & mask]) out_idx[m++] = (uint32_t)i;
  return m;
}
