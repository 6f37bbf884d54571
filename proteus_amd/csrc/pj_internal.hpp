// Internal declarations shared by the HIP translation units of the
// join / aggregation path (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/proteus_amd.h"

namespace pj {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
void set_error(const char *fmt, ...);
int fail(int code, const char *fmt, ...);

#define PJ_HIP(call)                                                        \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess)                                                   \
      return ::pj::fail(PJ_E_HIP, "%s failed: %s (%s:%d)", #call,           \
                        hipGetErrorString(e_), __FILE__, __LINE__);         \
  } while (0)

#define PJ_TRY(call)               \
  do {                             \
    int rc_ = (call);              \
    if (rc_ != PJ_OK) return rc_;  \
  } while (0)

// ---------------------------------------------------------------------------
// hashing (device + host)
// ---------------------------------------------------------------------------
// hashInt32 of core/olap/lib/expressions/expressions-hasher.cpp:39-58:
// sign-extend, murmur3 fmix with logical shifts, in 64-bit arithmetic.
__host__ __device__ __forceinline__ uint64_t hash_i32(int32_t k) {
  uint64_t h = (uint64_t)(int64_t)k;
  h ^= h >> 16;
  h *= 0x85ebca6bULL;
  h ^= h >> 13;
  h *= 0xc2b2ae35ULL;
  h ^= h >> 16;
  return h;
}

// 32-bit murmur3 finaliser with logical shifts, the in-partition hash of
// gpu-partitioned-hash-join-chained.cpp:595-610.  Radix digits take its top
// bits, LDS buckets its low bits.
__host__ __device__ __forceinline__ uint32_t fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bU;
  h ^= h >> 13;
  h *= 0xc2b2ae35U;
  h ^= h >> 16;
  return h;
}

// Slot hash of the resident dimension tables (build and every probe use
// it): the 32-bit murmur3 finaliser, cheaper than the 64-bit hashInt32.
__host__ __device__ __forceinline__ uint32_t dim_hash(int32_t k) { return fmix32((uint32_t)k); }

// First slot of key k in a resident dimension table.  A direct table
// (range > 0) holds key k only at slot k - lo, so one load decides; a key
// outside [lo, lo + range) is a miss without a load (*in = false, slot 0).
// A hash table starts at dim_hash(k) & mask and is probed linearly.
__host__ __device__ __forceinline__ uint32_t dim_start(int32_t k, uint64_t mask, int32_t lo,
                                                      uint32_t range, bool *in) {
  if (range) {
    const uint32_t o = (uint32_t)k - (uint32_t)lo;
    *in = o < range;
    return *in ? o : 0u;
  }
  *in = true;
  return dim_hash(k) & (uint32_t)mask;
}

// ---------------------------------------------------------------------------
// counter-based RNG (splitmix64 finaliser; Steele, Lea, Flood 2014)
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
constexpr uint64_t kGolden = 0x9e3779b97f4a7c15ULL;
__host__ __device__ __forceinline__ uint64_t rng_key(uint64_t seed,
                                                      uint64_t stream) {
  return mix64(seed + kGolden * (stream + 1));
}
__host__ __device__ __forceinline__ uint64_t rng_at(uint64_t key, uint64_t i) {
  return mix64(key + kGolden * (i + 1));
}
__host__ __device__ __forceinline__ uint64_t mulhi64(uint64_t a, uint64_t b) {
#ifdef __HIP_DEVICE_COMPILE__
  return __umul64hi(a, b);
#else
  return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}
// mulhi64(a, b) for b < 2^32 with two 32 x 32 -> 64 products: with
// a = ah 2^32 + al, (a b) >> 64 = (ah b + ((al b) >> 32)) >> 32 exactly (the
// low 32 bits of al b cannot carry into bit 64), and ah b + 2^32 - 1 < 2^64.
__host__ __device__ __forceinline__ uint64_t mulhi64_32(uint64_t a, uint32_t b) {
  const uint64_t hi = (a >> 32) * (uint64_t)b;
  const uint64_t lo = (a & 0xffffffffULL) * (uint64_t)b;
  return (hi + (lo >> 32)) >> 32;
}

// ---------------------------------------------------------------------------
// device memory: caching pool with an HBM budget
// ---------------------------------------------------------------------------
struct Pool {
  uint64_t budget = 0;      // 0 = unlimited
  uint64_t in_use = 0;      // bytes handed out (user + scratch)
  uint64_t cached = 0;      // bytes held in free lists
  uint64_t clock = 0;       // frees so far: the age stamp of cached blocks
  uint64_t mallocs = 0;     // hipMalloc calls (cache misses)
  uint64_t evictions = 0;   // cached blocks returned with hipFree
  struct Cached {
    void *p;
    uint64_t stamp;  // clock when the block was freed
  };
  std::multimap<uint64_t, Cached> free_blocks;  // size -> block
  // the cached blocks by free stamp, oldest first (the eviction order)
  std::map<uint64_t, std::multimap<uint64_t, Cached>::iterator> lru;
  std::unordered_map<void *, uint64_t> live;    // ptr -> size
};

struct ProfRec {
  std::string name;
  hipEvent_t a, b;
};

struct TuneKey {
  int lb, lp;  // log2 buckets of build / probe sizes
  bool operator<(const TuneKey &o) const {
    return lb < o.lb || (lb == o.lb && lp < o.lp);
  }
};

constexpr uint64_t kBounceBytes = 1 << 20;  // d2h read-backs up to this size go through pinned memory

}  // namespace pj

struct pj_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int num_cus = 0;
  uint64_t hbm_bytes = 0;
  uint64_t lds_per_cu = 0;
  pj::Pool pool;
  std::mutex mu;
  // profiling
  bool prof = false;
  std::vector<pj::ProfRec> prof_pending;
  std::vector<hipEvent_t> ev_free;
  std::map<std::string, std::pair<double, uint64_t>> prof_acc;
  // pinned staging for small results
  void *pinned_small = nullptr;
  // pinned bounce buffer of d2h (pj::kBounceBytes), and its lock
  void *pinned_rb = nullptr;
  std::mutex rb_mu;
  // autotuner
  bool autotune = false;
  std::map<pj::TuneKey, pj_radix_cfg> tuned;
  // batched-query stream slots (owned by batch.hip)
  void *batch = nullptr;
  // bloom filters by id (owned by bloom.hip)
  void *bloom = nullptr;
  // second stream for work that overlaps the main stream's (created on first
  // use), with a fork and a join event
  hipStream_t aux = nullptr;
  hipEvent_t aux_fork = nullptr, aux_join = nullptr;
  // host -> HBM staging workers, pinned buffers and copy stream (stage.hip)
  void *stager = nullptr;
};

namespace pj {

// stops the staging workers after the queued copies and frees their buffers
void stage_release(pj_ctx *ctx);

// scratch allocation from the context pool (rounded, cached, budgeted)
int dalloc(pj_ctx *ctx, uint64_t bytes, void **p);
int dfree(pj_ctx *ctx, void *p);

// RAII scratch buffer
struct Scratch {
  pj_ctx *ctx = nullptr;
  void *p = nullptr;
  Scratch() = default;
  Scratch(const Scratch &) = delete;
  Scratch &operator=(const Scratch &) = delete;
  ~Scratch() {
    if (p) dfree(ctx, p);
  }
  int alloc(pj_ctx *c, uint64_t bytes) {
    ctx = c;
    return dalloc(c, bytes ? bytes : 256, &p);
  }
  template <class T>
  T *as() const {
    return reinterpret_cast<T *>(p);
  }
};

// profiling hooks around launches
void prof_begin(pj_ctx *ctx, const char *name, hipEvent_t *a);
void prof_end(pj_ctx *ctx, const char *name, hipEvent_t a);

#define PJ_LAUNCH(ctx, name, kernel, grid, block, shmem, ...)                 \
  do {                                                                       \
    hipEvent_t ev_a_ = nullptr;                                              \
    ::pj::prof_begin((ctx), (name), &ev_a_);                                 \
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), (shmem),             \
                       (ctx)->stream, __VA_ARGS__);                          \
    hipError_t le_ = hipGetLastError();                                      \
    if (le_ != hipSuccess)                                                   \
      return ::pj::fail(PJ_E_HIP, "launch of %s failed: %s", (name),         \
                        hipGetErrorString(le_));                             \
    ::pj::prof_end((ctx), (name), ev_a_);                                    \
  } while (0)

int h2d(pj_ctx *ctx, void *d, const void *h, uint64_t bytes);
int d2h(pj_ctx *ctx, void *h, const void *d, uint64_t bytes);
int sync(pj_ctx *ctx);
// Returns the auxiliary stream after making it wait for the work queued so
// far on ctx->stream.
int aux_fork(pj_ctx *ctx, hipStream_t *aux);
// Makes ctx->stream wait for the work queued so far on the auxiliary stream.
int aux_join(pj_ctx *ctx);

// Sends launches through ctx to another stream while in scope.
struct StreamSwap {
  pj_ctx *ctx;
  hipStream_t saved;
  StreamSwap(pj_ctx *c, hipStream_t s) : ctx(c), saved(c->stream) { c->stream = s; }
  ~StreamSwap() { ctx->stream = saved; }
  StreamSwap(const StreamSwap &) = delete;
  StreamSwap &operator=(const StreamSwap &) = delete;
};
int zero(pj_ctx *ctx, void *d, uint64_t bytes);
int fill_u32(pj_ctx *ctx, uint32_t *d, uint32_t v, uint64_t n);

__host__ __device__ inline uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }
inline int ilog2_ceil(uint64_t x) {
  int k = 0;
  while ((1ULL << k) < x && k < 63) ++k;
  return k;
}

// streaming 16-byte load (global_load_dwordx4 ... nt)
typedef int v4i_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ int4 ld_nt4(const void *p) {
  v4i_t x = __builtin_nontemporal_load(reinterpret_cast<const v4i_t *>(p));
  return make_int4(x.x, x.y, x.z, x.w);
}

// wave64 helpers --------------------------------------------------------------
__device__ __forceinline__ int64_t wave_sum_i64(int64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ int64_t wave_min_i64(int64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    int64_t x = __shfl_xor(v, o, 64);
    v = x < v ? x : v;
  }
  return v;
}
__device__ __forceinline__ int64_t wave_max_i64(int64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    int64_t x = __shfl_xor(v, o, 64);
    v = x > v ? x : v;
  }
  return v;
}
__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// join aggregate accumulator ---------------------------------------------------
// Build and probe values are int32, so a thread tracks MIN/MAX in 32 bits and
// counts single matches in 32 bits; sums stay 64-bit.  Widened at the flush.
struct JoinAcc {
  int64_t count, sum_b, sum_p;
  uint32_t n1;  // single matches since the last fold into count
  int32_t min_b, max_b, min_p, max_p;
  __device__ __forceinline__ void init() {
    count = 0;
    sum_b = 0;
    sum_p = 0;
    n1 = 0;
    min_b = INT32_MAX;
    max_b = INT32_MIN;
    min_p = INT32_MAX;
    max_p = INT32_MIN;
  }
  // exactly one match
  __device__ __forceinline__ void add1(int32_t b, int32_t pv) {
    n1 += 1;
    sum_b += b;
    sum_p += pv;
    min_b = min(min_b, b);
    max_b = max(max_b, b);
    min_p = min(min_p, pv);
    max_p = max(max_p, pv);
  }
  // up to three matches of one probe row (flags h0..h2 over build values
  // b0..b2), without branches
  __device__ __forceinline__ void add3_if(bool h0, int32_t b0, bool h1, int32_t b1, bool h2,
                                          int32_t b2, int32_t pv) {
    const uint32_t m = (uint32_t)h0 + (uint32_t)h1 + (uint32_t)h2;
    n1 += m;
    sum_b += (int64_t)(h0 ? b0 : 0) + (int64_t)(h1 ? b1 : 0) + (int64_t)(h2 ? b2 : 0);
    sum_p += (int64_t)pv * (int64_t)m;
    int32_t mn = min(h0 ? b0 : INT32_MAX, min(h1 ? b1 : INT32_MAX, h2 ? b2 : INT32_MAX));
    int32_t mx = max(h0 ? b0 : INT32_MIN, max(h1 ? b1 : INT32_MIN, h2 ? b2 : INT32_MIN));
    min_b = min(min_b, mn);
    max_b = max(max_b, mx);
    min_p = m ? min(min_p, pv) : min_p;
    max_p = m ? max(max_p, pv) : max_p;
  }
  // Four probe rows with at most one match each: row c matched iff h[c],
  // with build value b[c] and probe value p[c].
  __device__ __forceinline__ void add4_1(const bool h[4], const int32_t b[4],
                                         const int32_t p[4]) {
    int32_t mnb = INT32_MAX, mxb = INT32_MIN, mnp = INT32_MAX, mxp = INT32_MIN;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int32_t m = -(int32_t)h[c];  // all ones on a match
      n1 += (uint32_t)h[c];
      sum_b += (int64_t)(b[c] & m);
      sum_p += (int64_t)(p[c] & m);
      mnb = min(mnb, h[c] ? b[c] : INT32_MAX);
      mxb = max(mxb, h[c] ? b[c] : INT32_MIN);
      mnp = min(mnp, h[c] ? p[c] : INT32_MAX);
      mxp = max(mxp, h[c] ? p[c] : INT32_MIN);
    }
    min_b = min(min_b, mnb);
    max_b = max(max_b, mxb);
    min_p = min(min_p, mnp);
    max_p = max(max_p, mxp);
  }
  // `m` matches of one probe row whose build values sum to sb and span
  // [mnb, mxb].
  __device__ __forceinline__ void add(int64_t m, int64_t sb, int32_t mnb, int32_t mxb,
                                      int32_t pv) {
    if (m) {
      count += m;
      sum_b += sb;
      sum_p += m * (int64_t)pv;
      min_b = min(min_b, mnb);
      max_b = max(max_b, mxb);
      min_p = min(min_p, pv);
      max_p = max(max_p, pv);
    }
  }
};

// The global join accumulator: kJoinAccCopies copies of 7 slots (count,
// sum_b, sum_p as u64 adds; min_b, max_b, min_p, max_p as signed min/max),
// each copy on its own 128-byte line.  Workgroup b folds into copy
// b % kJoinAccCopies and join_acc_read combines the copies: same-address
// global atomics serialise (about 160 ns apart on MI355X), so a single copy
// made the flush of a large grid cost milliseconds.
constexpr int kJoinAccCopies = 64;
constexpr int kJoinAccStride = 16;  // u64 words per copy
constexpr int kJoinAccWords = kJoinAccCopies * kJoinAccStride;
// Word 8 of the first copy is not an accumulator slot; join_acc_init zeroes
// it, so a join may keep a flag there (the chained build's duplicate-key
// flag) without a clear of its own.
constexpr int kJoinAccFlagWord = 8;

// Reduce a JoinAcc over the workgroup (every thread of the block calls it,
// at most 1024 threads) and fold the result into the block's accumulator
// copy with one set of atomics.  With no match at all the identities are
// INT64_MAX / INT64_MIN.
__device__ __forceinline__ void join_acc_flush(JoinAcc a, unsigned long long *g) {
  __shared__ long long red[16][7];
  const int64_t cnt = a.count + (int64_t)a.n1;
  const bool any = cnt != 0;
  int64_t v[7];
  v[0] = wave_sum_i64(cnt);
  v[1] = wave_sum_i64(a.sum_b);
  v[2] = wave_sum_i64(a.sum_p);
  v[3] = wave_min_i64(any ? (int64_t)a.min_b : INT64_MAX);
  v[4] = wave_max_i64(any ? (int64_t)a.max_b : INT64_MIN);
  v[5] = wave_min_i64(any ? (int64_t)a.min_p : INT64_MAX);
  v[6] = wave_max_i64(any ? (int64_t)a.max_p : INT64_MIN);
  const int wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int i = 0; i < 7; ++i) red[wid][i] = v[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < nw; ++w) {
      v[0] += red[w][0];
      v[1] += red[w][1];
      v[2] += red[w][2];
      v[3] = red[w][3] < v[3] ? red[w][3] : v[3];
      v[4] = red[w][4] > v[4] ? red[w][4] : v[4];
      v[5] = red[w][5] < v[5] ? red[w][5] : v[5];
      v[6] = red[w][6] > v[6] ? red[w][6] : v[6];
    }
    if (v[0] != 0) {
      unsigned long long *c = g + (blockIdx.x % kJoinAccCopies) * kJoinAccStride;
      atomicAdd(&c[0], (unsigned long long)v[0]);
      atomicAdd(&c[1], (unsigned long long)v[1]);
      atomicAdd(&c[2], (unsigned long long)v[2]);
      atomicMin((long long *)&c[3], (long long)v[3]);
      atomicMax((long long *)&c[4], (long long)v[4]);
      atomicMin((long long *)&c[5], (long long)v[5]);
      atomicMax((long long *)&c[6], (long long)v[6]);
    }
  }
}

// d_acc holds kJoinAccWords u64 (kJoinAccWords * 8 bytes).
int join_acc_init(pj_ctx *ctx, unsigned long long *d_acc);
// Resolves the allocation-free entry points' static scratch on `device`
// (which must be current); pj_ctx_create calls it.
int noalloc_scratch_init(int device);
int join_acc_read(pj_ctx *ctx, const unsigned long long *d_acc,
                  pj_join_agg *out);
// Combines a host copy (kJoinAccWords words) of the accumulator copies.
int join_acc_fold(const int64_t *h, pj_join_agg *out);
namespace rel {
// MIN and MAX of n >= 1 int32 rows into d_res[0], d_res[1] (device), queued
// on the context stream with no read-back; part receives the per-workgroup
// partials and must outlive the queued work.
int minmax_i32_device(pj_ctx *ctx, const int32_t *col, uint64_t n, Scratch &part, int64_t *d_res);
}  // namespace rel

// Join -> reduce through a direct table when the build keys are dense (as
// many distinct keys as the span min..max holds, at most kDirectMaxSpan):
// the build values indexed by key - min, 4 bytes a key, and one table read
// per probe row.  *done is false (and nothing is written to out) when the
// keys are not dense or repeat one; the caller then takes another path.
constexpr uint64_t kDirectMaxSpan = 1ULL << 22;  // 16 MiB of values: in the Infinity Cache
int join_direct(pj_ctx *ctx, const int32_t *d_bk, const int32_t *d_bv, uint64_t nb,
                const int32_t *d_pk, const int32_t *d_pv, uint64_t np, pj_join_agg *out,
                bool *done);
// The radix-partitioned join -> reduce with the configuration chosen for
// (nb, np) (the tuned one when a tuning exists), for a caller that holds the
// context lock and has checked the sizes.
int join_partitioned_auto(pj_ctx *ctx, const int32_t *d_bk, const int32_t *d_bv, uint64_t nb,
                          const int32_t *d_pk, const int32_t *d_pv, uint64_t np,
                          pj_join_agg *out);

}  // namespace pj
