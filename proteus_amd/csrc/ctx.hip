// Context, HBM pool, copies, profiling and the seeded device generators.
//
// The reference spreads this over MemoryManager (128 MiB GPU / 1 GiB CPU
// allocation units, core/platform/include/platform/memory/memory-manager.hpp),
// the device buffer pool (core/platform/lib/memory/buffer-manager.cu) and
// gpu_run/stream helpers (core/platform/include/platform/common/gpu/
// gpu-common.hpp:93,383-398).  On one MI355X all of it collapses to one
// context: one stream, one caching HBM pool with a byte budget, pinned staging
// for small results, and status codes instead of aborts.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "pj_internal.hpp"

namespace pj {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

// Sizes are rounded to 256 B below 2 MiB and to 2 MiB above (the reference's
// block granularity, buffer-manager.cuh:62-63), so freed blocks are reused.
static uint64_t round_size(uint64_t b) {
  if (b <= (2ULL << 20)) return (b + 255) & ~255ULL;
  return (b + (2ULL << 20) - 1) & ~((2ULL << 20) - 1);
}

int dalloc(pj_ctx *ctx, uint64_t bytes, void **p) {
  *p = nullptr;
  uint64_t sz = round_size(bytes ? bytes : 1);
  auto &pool = ctx->pool;
  // reuse a cached block of exactly this size class, or up to 2x larger
  auto it = pool.free_blocks.lower_bound(sz);
  if (it != pool.free_blocks.end() && it->first <= sz * 2) {
    *p = it->second.p;
    uint64_t got = it->first;
    pool.lru.erase(it->second.stamp);
    pool.free_blocks.erase(it);
    pool.cached -= got;
    pool.live[*p] = got;
    pool.in_use += got;
    return PJ_OK;
  }
  if (pool.budget && pool.in_use + sz > pool.budget)
    return fail(PJ_E_OOM, "HBM budget exceeded: %llu in use + %llu requested > %llu",
                (unsigned long long)pool.in_use, (unsigned long long)sz,
                (unsigned long long)pool.budget);
  hipError_t e = hipMalloc(p, sz);
  if (e != hipSuccess) {
    // release cached blocks and retry once
    for (auto &kv : pool.free_blocks) (void)hipFree(kv.second.p);
    pool.free_blocks.clear();
    pool.lru.clear();
    pool.cached = 0;
    (void)hipGetLastError();
    e = hipMalloc(p, sz);
    if (e != hipSuccess) {
      *p = nullptr;
      (void)hipGetLastError();
      return fail(PJ_E_OOM, "hipMalloc(%llu) failed: %s", (unsigned long long)sz,
                  hipGetErrorString(e));
    }
  }
  pool.live[*p] = sz;
  pool.in_use += sz;
  ++pool.mallocs;
  return PJ_OK;
}

int dfree(pj_ctx *ctx, void *p) {
  if (!p) return PJ_OK;
  auto &pool = ctx->pool;
  auto it = pool.live.find(p);
  if (it == pool.live.end()) return fail(PJ_E_ARG, "pj_free of unknown pointer %p", p);
  uint64_t sz = it->second;
  pool.live.erase(it);
  pool.in_use -= sz;
  const uint64_t stamp = ++pool.clock;
  auto ins = pool.free_blocks.emplace(sz, Pool::Cached{p, stamp});
  pool.lru.emplace(stamp, ins);
  pool.cached += sz;
  // Keep the cache within a quarter of HBM by dropping the least recently
  // freed blocks: the blocks of an earlier workload go first, and a workload
  // whose own blocks fit stops evicting.  A cached block may still be read
  // by work queued on any of the context's streams (main, auxiliary, batch
  // slots, the out-of-GPU copy stream); hipFree waits for all work on the
  // device before it returns the memory, which is what makes the release
  // safe -- and what makes every eviction a device-wide wait.
  const uint64_t cap = std::max<uint64_t>(8ULL << 30, ctx->hbm_bytes / 4);
  while (pool.cached > cap && !pool.lru.empty()) {
    auto old = pool.lru.begin()->second;
    (void)hipFree(old->second.p);
    pool.cached -= old->first;
    pool.lru.erase(pool.lru.begin());
    pool.free_blocks.erase(old);
    ++pool.evictions;
  }
  return PJ_OK;
}

int h2d(pj_ctx *ctx, void *d, const void *h, uint64_t bytes) {
  if (!bytes) return PJ_OK;
  PJ_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, ctx->stream));
  PJ_HIP(hipStreamSynchronize(ctx->stream));
  return PJ_OK;
}

// Small read-backs (counts, ranges, flags) go through a pinned bounce
// buffer: a copy into pageable memory depends on how the runtime stages it,
// and has been seen to stall a plan for milliseconds.
int d2h(pj_ctx *ctx, void *h, const void *d, uint64_t bytes) {
  if (!bytes) return PJ_OK;
  if (bytes <= kBounceBytes && ctx->pinned_rb) {
    std::lock_guard<std::mutex> g(ctx->rb_mu);
    PJ_HIP(hipMemcpyAsync(ctx->pinned_rb, d, bytes, hipMemcpyDeviceToHost, ctx->stream));
    PJ_HIP(hipStreamSynchronize(ctx->stream));
    memcpy(h, ctx->pinned_rb, bytes);
    return PJ_OK;
  }
  PJ_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, ctx->stream));
  PJ_HIP(hipStreamSynchronize(ctx->stream));
  return PJ_OK;
}

int sync(pj_ctx *ctx) {
  PJ_HIP(hipStreamSynchronize(ctx->stream));
  return PJ_OK;
}

int aux_fork(pj_ctx *ctx, hipStream_t *aux) {
  if (!ctx->aux) {
    PJ_HIP(hipStreamCreateWithFlags(&ctx->aux, hipStreamNonBlocking));
    PJ_HIP(hipEventCreateWithFlags(&ctx->aux_fork, hipEventDisableTiming));
    PJ_HIP(hipEventCreateWithFlags(&ctx->aux_join, hipEventDisableTiming));
  }
  PJ_HIP(hipEventRecord(ctx->aux_fork, ctx->stream));
  PJ_HIP(hipStreamWaitEvent(ctx->aux, ctx->aux_fork, 0));
  *aux = ctx->aux;
  return PJ_OK;
}

int aux_join(pj_ctx *ctx) {
  PJ_HIP(hipEventRecord(ctx->aux_join, ctx->aux));
  PJ_HIP(hipStreamWaitEvent(ctx->stream, ctx->aux_join, 0));
  return PJ_OK;
}

// Clears use a fill kernel, never hipMemsetAsync, so callers may capture the
// work into a hipGraph.
__global__ void k_fill_u32(uint32_t *__restrict__ d, uint32_t v, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t s = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += s) d[i] = v;
}

int fill_u32(pj_ctx *ctx, uint32_t *d, uint32_t v, uint64_t n) {
  if (!n) return PJ_OK;
  uint64_t grid = ceil_div(n, 256);
  if (grid > 4096) grid = 4096;
  PJ_LAUNCH(ctx, "fill_u32", k_fill_u32, grid, 256, 0, d, v, n);
  return PJ_OK;
}

__global__ void k_fill_u8(uint8_t *__restrict__ d, uint8_t v, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t s = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += s) d[i] = v;
}

int zero(pj_ctx *ctx, void *d, uint64_t bytes) {
  if (!bytes) return PJ_OK;
  if ((bytes & 3) == 0 && ((uintptr_t)d & 3) == 0)
    return fill_u32(ctx, (uint32_t *)d, 0, bytes / 4);
  uint64_t grid = ceil_div(bytes, 256);
  if (grid > 4096) grid = 4096;
  PJ_LAUNCH(ctx, "fill_u8", k_fill_u8, grid, 256, 0, (uint8_t *)d, (uint8_t)0, bytes);
  return PJ_OK;
}

void prof_begin(pj_ctx *ctx, const char *name, hipEvent_t *a) {
  *a = nullptr;
  if (!ctx->prof) return;
  hipEvent_t e;
  if (!ctx->ev_free.empty()) {
    e = ctx->ev_free.back();
    ctx->ev_free.pop_back();
  } else if (hipEventCreate(&e) != hipSuccess) {
    return;
  }
  (void)hipEventRecord(e, ctx->stream);
  *a = e;
}

void prof_end(pj_ctx *ctx, const char *name, hipEvent_t a) {
  if (!ctx->prof || !a) return;
  hipEvent_t b;
  if (!ctx->ev_free.empty()) {
    b = ctx->ev_free.back();
    ctx->ev_free.pop_back();
  } else if (hipEventCreate(&b) != hipSuccess) {
    ctx->ev_free.push_back(a);
    return;
  }
  (void)hipEventRecord(b, ctx->stream);
  ctx->prof_pending.push_back(ProfRec{name, a, b});
}

static void prof_collect(pj_ctx *ctx) {
  if (ctx->prof_pending.empty()) return;
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->aux) (void)hipStreamSynchronize(ctx->aux);
  for (auto &r : ctx->prof_pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      auto &acc = ctx->prof_acc[r.name];
      acc.first += ms;
      acc.second += 1;
    }
    ctx->ev_free.push_back(r.a);
    ctx->ev_free.push_back(r.b);
  }
  ctx->prof_pending.clear();
}

// Streaming device copy: 16-byte loads and stores, grid-stride.
__global__ __launch_bounds__(256) void k_copy16(const int4 *__restrict__ src,
                                                int4 *__restrict__ dst, uint64_t n16) {
  uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  const uint64_t s = (uint64_t)gridDim.x * 256;
  for (; i < n16; i += s) dst[i] = ld_nt4(src + i);
}
__global__ void k_copy1(const uint8_t *__restrict__ src, uint8_t *__restrict__ dst, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t s = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += s) dst[i] = src[i];
}

// join accumulator: 7 x u64 in device memory -------------------------------
__global__ void k_join_acc_init(unsigned long long *g) {
  for (int i = threadIdx.x; i < kJoinAccWords; i += blockDim.x) {
    const int s = i % kJoinAccStride;
    unsigned long long v = 0;
    if (s == 3 || s == 5) v = (unsigned long long)INT64_MAX;
    if (s == 4 || s == 6) v = (unsigned long long)INT64_MIN;
    g[i] = v;
  }
}

int join_acc_init(pj_ctx *ctx, unsigned long long *d_acc) {
  PJ_LAUNCH(ctx, "join_acc_init", k_join_acc_init, 1, 256, 0, d_acc);
  return PJ_OK;
}

int join_acc_read(pj_ctx *ctx, const unsigned long long *d_acc,
                  pj_join_agg *out) {
  int64_t h[kJoinAccWords];
  PJ_TRY(d2h(ctx, h, d_acc, sizeof(h)));
  return join_acc_fold(h, out);
}

int join_acc_fold(const int64_t *h, pj_join_agg *out) {
  out->count = 0;
  out->sum_b = 0;
  out->sum_p = 0;
  out->min_b = INT64_MAX;
  out->max_b = INT64_MIN;
  out->min_p = INT64_MAX;
  out->max_p = INT64_MIN;
  for (int c = 0; c < kJoinAccCopies; ++c) {
    const int64_t *x = h + c * kJoinAccStride;
    out->count = (int64_t)((uint64_t)out->count + (uint64_t)x[0]);
    out->sum_b = (int64_t)((uint64_t)out->sum_b + (uint64_t)x[1]);
    out->sum_p = (int64_t)((uint64_t)out->sum_p + (uint64_t)x[2]);
    out->min_b = x[3] < out->min_b ? x[3] : out->min_b;
    out->max_b = x[4] > out->max_b ? x[4] : out->max_b;
    out->min_p = x[5] < out->min_p ? x[5] : out->min_p;
    out->max_p = x[6] > out->max_p ? x[6] : out->max_p;
  }
  return PJ_OK;
}

// ---------------------------------------------------------------------------
// generators
// ---------------------------------------------------------------------------
// Seeded bijection of [0, 2^k): 4 rounds of (x*odd, x ^= x>>s, x += c).
struct PermParams {
  uint64_t m[4], c[4], mask, n;
  int s;
};

static PermParams make_perm(uint64_t seed, uint64_t n) {
  PermParams p;
  int k = 1;
  while (k < 63 && (1ULL << k) < n) ++k;
  p.mask = (1ULL << k) - 1ULL;
  p.s = (k + 1) / 2;
  p.n = n;
  for (int r = 0; r < 4; ++r) {
    uint64_t key = rng_key(seed, 0xFEED0000ULL + (uint64_t)r);
    p.m[r] = rng_at(key, 0) | 1ULL;
    p.c[r] = rng_at(key, 1);
  }
  return p;
}

__device__ __forceinline__ uint64_t perm_apply(const PermParams &p, uint64_t i) {
  if (p.n <= 1) return 0;
  uint64_t x = i;
  do {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      x = (x * p.m[r]) & p.mask;
      x ^= x >> p.s;
      x = (x + p.c[r]) & p.mask;
    }
  } while (x >= p.n);
  return x;
}

__global__ void k_gen_perm(int32_t *out, uint64_t n, PermParams p) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t s = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += s) out[i] = (int32_t)perm_apply(p, i);
}

__global__ void k_gen_uniform_i32(int32_t *out, uint64_t n, uint64_t key,
                                  int64_t lo, uint64_t range) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t s = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += s) out[i] = (int32_t)(lo + (int64_t)mulhi64(rng_at(key, i), range));
}

__global__ void k_gen_uniform_i64(int64_t *out, uint64_t n, uint64_t key,
                                  int64_t lo, uint64_t range) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t s = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += s) out[i] = lo + (int64_t)mulhi64(rng_at(key, i), range);
}

__global__ void k_gen_unit_f32(float *out, uint64_t n, uint64_t key) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t s = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += s)
    out[i] = (float)mulhi64(rng_at(key, i), 1ULL << 24) * (1.0f / 16777216.0f);
}

// Zipf(theta) over ranks [0, n) by rejection-inversion (W. Hormann and
// G. Derflinger, ACM TOMACS 6(3), 1996).  Attempt a draws from stream
// (stream + a * 2^32); at most 64 attempts.
struct ZipfParams {
  double s, hx1, hn, sp;
  uint64_t n;
  uint64_t seed, stream;
};

__host__ __device__ static inline double zh1(double x) {
  return fabs(x) > 1e-8 ? log1p(x) / x : 1.0 - x * (0.5 - x * (1.0 / 3.0 - 0.25 * x));
}
__host__ __device__ static inline double zh2(double x) {
  return fabs(x) > 1e-8 ? expm1(x) / x
                        : 1.0 + x * 0.5 * (1.0 + x * (1.0 / 3.0) * (1.0 + 0.25 * x));
}
__host__ __device__ static inline double zH(double x, double s) {
  double lx = log(x);
  return zh2((1.0 - s) * lx) * lx;
}
__host__ __device__ static inline double zHinv(double x, double s) {
  double t = x * (1.0 - s);
  if (t < -1.0) t = -1.0;
  return exp(zh1(t) * x);
}
__host__ __device__ static inline double zhh(double x, double s) {
  return exp(-s * log(x));
}

static ZipfParams make_zipf(uint64_t seed, uint64_t stream, uint64_t n,
                            double theta) {
  ZipfParams z;
  z.s = theta;
  z.n = n;
  z.seed = seed;
  z.stream = stream;
  z.hx1 = zH(1.5, theta) - 1.0;
  z.hn = zH((double)n + 0.5, theta);
  z.sp = 2.0 - zHinv(zH(2.5, theta) - zhh(2.0, theta), theta);
  return z;
}

__device__ __forceinline__ uint64_t zipf_rank(const ZipfParams &z, uint64_t i) {
  if (z.n <= 1) return 0;
  if (z.s <= 0.0) return mulhi64(rng_at(rng_key(z.seed, z.stream), i), z.n);
  double kd = 1.0;
  for (uint64_t a = 0; a < 64; ++a) {
    uint64_t r = rng_at(rng_key(z.seed, z.stream + (a << 32)), i);
    double u01 = (double)((r >> 11) + 1) * (1.0 / 9007199254740992.0);
    double u = z.hn + u01 * (z.hx1 - z.hn);
    double x = zHinv(u, z.s);
    kd = floor(x + 0.5);
    if (kd < 1.0) kd = 1.0;
    if (kd > (double)z.n) kd = (double)z.n;
    if (kd - x <= z.sp || u >= zH(kd + 0.5, z.s) - zhh(kd, z.s)) break;
  }
  return (uint64_t)kd - 1;
}

__global__ void k_gen_zipf_fk(int32_t *out, uint64_t n, uint64_t key,
                              uint64_t nbuild, uint64_t range, ZipfParams z,
                              PermParams p) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t s = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += s) {
    uint64_t u = mulhi64(rng_at(key, i), range);
    if (u >= nbuild)
      out[i] = (int32_t)u;
    else
      out[i] = (int32_t)perm_apply(p, zipf_rank(z, i));
  }
}

__global__ void k_gen_zipf(int32_t *out, uint64_t n, ZipfParams z, PermParams p) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t s = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += s) out[i] = (int32_t)perm_apply(p, zipf_rank(z, i));
}

static uint64_t gen_grid(uint64_t n) {
  uint64_t g = ceil_div(n, 256);
  return g > 16384 ? 16384 : (g ? g : 1);
}

}  // namespace pj

using namespace pj;

extern "C" {

const char *pj_last_error(void) { return g_err; }

// Lets the host engine (plain C++, C ABI only) report errors through
// pj_last_error.
void pj_set_last_error(const char *msg) {
  std::strncpy(g_err, msg, sizeof(g_err) - 1);
  g_err[sizeof(g_err) - 1] = 0;
}

int pj_ctx_create(int device, uint64_t hbm_budget_bytes, pj_ctx **out) {
  *out = nullptr;
  int ndev = 0;
  PJ_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev)
    return fail(PJ_E_ARG, "device %d out of range (%d devices)", device, ndev);
  PJ_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  PJ_HIP(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(PJ_E_UNSUPPORTED, "device %d is %s; this library is built for gfx950",
                device, prop.gcnArchName);
  // the allocation-free entry points' static scratch, resolved once per
  // device while it is current
  PJ_TRY(noalloc_scratch_init(device));
  pj_ctx *ctx = new pj_ctx();
  ctx->device = device;
  ctx->num_cus = prop.multiProcessorCount;
  ctx->hbm_bytes = prop.totalGlobalMem;
  ctx->lds_per_cu = prop.maxSharedMemoryPerMultiProcessor;
  ctx->pool.budget = hbm_budget_bytes;
  hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete ctx;
    return fail(PJ_E_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
  }
  e = hipHostMalloc(&ctx->pinned_small, 1 << 20, hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc(&ctx->pinned_rb, kBounceBytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    if (ctx->pinned_small) (void)hipHostFree(ctx->pinned_small);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return fail(PJ_E_HIP, "hipHostMalloc: %s", hipGetErrorString(e));
  }
  *out = ctx;
  return PJ_OK;
}

void pj_engine_release_catalog(pj_ctx *ctx);
void pj_batch_release(pj_ctx *ctx);
void pj_bloom_release_all(pj_ctx *ctx);

int pj_ctx_destroy(pj_ctx *ctx) {
  if (!ctx) return PJ_OK;
  (void)hipSetDevice(ctx->device);
  pj::stage_release(ctx);
  (void)hipStreamSynchronize(ctx->stream);
  pj_batch_release(ctx);
  pj_bloom_release_all(ctx);
  pj_engine_release_catalog(ctx);
  for (auto &kv : ctx->pool.live) (void)hipFree(kv.first);
  for (auto &kv : ctx->pool.free_blocks) (void)hipFree(kv.second.p);
  for (auto &r : ctx->prof_pending) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (auto e : ctx->ev_free) (void)hipEventDestroy(e);
  if (ctx->pinned_small) (void)hipHostFree(ctx->pinned_small);
  if (ctx->pinned_rb) (void)hipHostFree(ctx->pinned_rb);
  if (ctx->aux) {
    (void)hipStreamSynchronize(ctx->aux);
    (void)hipEventDestroy(ctx->aux_fork);
    (void)hipEventDestroy(ctx->aux_join);
    (void)hipStreamDestroy(ctx->aux);
  }
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return PJ_OK;
}

int pj_ctx_sync(pj_ctx *ctx) { return sync(ctx); }

int pj_ctx_device_info(pj_ctx *ctx, int *num_cus, uint64_t *hbm_bytes,
                       uint64_t *lds_bytes_per_cu) {
  if (num_cus) *num_cus = ctx->num_cus;
  if (hbm_bytes) *hbm_bytes = ctx->hbm_bytes;
  if (lds_bytes_per_cu) *lds_bytes_per_cu = ctx->lds_per_cu;
  return PJ_OK;
}

int pj_malloc(pj_ctx *ctx, uint64_t bytes, void **d_ptr) {
  std::lock_guard<std::mutex> g(ctx->mu);
  return dalloc(ctx, bytes, d_ptr);
}

int pj_free(pj_ctx *ctx, void *d_ptr) {
  std::lock_guard<std::mutex> g(ctx->mu);
  return dfree(ctx, d_ptr);
}

int pj_malloc_pinned(pj_ctx *ctx, uint64_t bytes, void **h_ptr) {
  (void)ctx;
  PJ_HIP(hipHostMalloc(h_ptr, bytes ? bytes : 1, hipHostMallocDefault));
  return PJ_OK;
}

int pj_free_pinned(pj_ctx *ctx, void *h_ptr) {
  (void)ctx;
  if (h_ptr) PJ_HIP(hipHostFree(h_ptr));
  return PJ_OK;
}

int pj_memcpy_h2d(pj_ctx *ctx, void *d_dst, const void *h_src, uint64_t bytes) {
  return h2d(ctx, d_dst, h_src, bytes);
}

int pj_memcpy_d2h(pj_ctx *ctx, void *h_dst, const void *d_src, uint64_t bytes) {
  return d2h(ctx, h_dst, d_src, bytes);
}

int pj_memset(pj_ctx *ctx, void *d_dst, int byte, uint64_t bytes) {
  if (byte == 0) {
    PJ_TRY(zero(ctx, d_dst, bytes));
  } else {
    uint32_t b = (uint8_t)byte;
    if ((bytes & 3) == 0 && ((uintptr_t)d_dst & 3) == 0) {
      PJ_TRY(fill_u32(ctx, (uint32_t *)d_dst, b * 0x01010101U, bytes / 4));
    } else {
      uint64_t grid = ceil_div(bytes, 256);
      if (grid > 4096) grid = 4096;
      PJ_LAUNCH(ctx, "fill_u8", k_fill_u8, grid, 256, 0, (uint8_t *)d_dst, (uint8_t)b, bytes);
    }
  }
  return sync(ctx);
}

int pj_copy(pj_ctx *ctx, void *d_dst, const void *d_src, uint64_t bytes) {
  if (!bytes) return PJ_OK;
  if ((((uintptr_t)d_dst | (uintptr_t)d_src | bytes) & 15) == 0) {
    const uint64_t n16 = bytes / 16;
    uint64_t grid = std::min<uint64_t>(ceil_div(n16, 256), (uint64_t)ctx->num_cus * 16);
    PJ_LAUNCH(ctx, "copy", k_copy16, grid, 256, 0, (const int4 *)d_src, (int4 *)d_dst, n16);
  } else {
    uint64_t grid = std::min<uint64_t>(ceil_div(bytes, 256), (uint64_t)ctx->num_cus * 16);
    PJ_LAUNCH(ctx, "copy", k_copy1, grid, 256, 0, (const uint8_t *)d_src, (uint8_t *)d_dst,
              bytes);
  }
  return sync(ctx);
}

int pj_mem_in_use(pj_ctx *ctx, uint64_t *bytes) {
  *bytes = ctx->pool.in_use;
  return PJ_OK;
}

int pj_mem_stats(pj_ctx *ctx, uint64_t out[4]) {
  if (!ctx || !out) return fail(PJ_E_ARG, "null argument");
  out[0] = ctx->pool.in_use;
  out[1] = ctx->pool.cached;
  out[2] = ctx->pool.mallocs;
  out[3] = ctx->pool.evictions;
  return PJ_OK;
}

int pj_prof_enable(pj_ctx *ctx, int on) {
  prof_collect(ctx);
  ctx->prof = on != 0;
  return PJ_OK;
}

int pj_prof_reset(pj_ctx *ctx) {
  prof_collect(ctx);
  ctx->prof_acc.clear();
  return PJ_OK;
}

int pj_prof_get(pj_ctx *ctx, const char *prefix, double *total_ms,
                uint64_t *launches) {
  prof_collect(ctx);
  double t = 0;
  uint64_t n = 0;
  size_t pl = std::strlen(prefix);
  for (auto &kv : ctx->prof_acc)
    if (kv.first.compare(0, pl, prefix) == 0) {
      t += kv.second.first;
      n += kv.second.second;
    }
  if (total_ms) *total_ms = t;
  if (launches) *launches = n;
  return PJ_OK;
}

int pj_gen_perm_i32(pj_ctx *ctx, int32_t *d_out, uint64_t n, uint64_t seed) {
  if (!n) return PJ_OK;
  PJ_LAUNCH(ctx, "gen_perm", k_gen_perm, gen_grid(n), 256, 0, d_out, n, make_perm(seed, n));
  return sync(ctx);
}

int pj_gen_uniform_i32(pj_ctx *ctx, int32_t *d_out, uint64_t n, uint64_t seed,
                       uint64_t stream, int64_t lo, uint64_t range) {
  if (!n) return PJ_OK;
  PJ_LAUNCH(ctx, "gen_uniform_i32", k_gen_uniform_i32, gen_grid(n), 256, 0, d_out, n,
            rng_key(seed, stream), lo, range);
  return sync(ctx);
}

int pj_gen_uniform_i64(pj_ctx *ctx, int64_t *d_out, uint64_t n, uint64_t seed,
                       uint64_t stream, int64_t lo, uint64_t range) {
  if (!n) return PJ_OK;
  PJ_LAUNCH(ctx, "gen_uniform_i64", k_gen_uniform_i64, gen_grid(n), 256, 0, d_out, n,
            rng_key(seed, stream), lo, range);
  return sync(ctx);
}

int pj_gen_unit_f32(pj_ctx *ctx, float *d_out, uint64_t n, uint64_t seed,
                    uint64_t stream) {
  if (!n) return PJ_OK;
  PJ_LAUNCH(ctx, "gen_unit_f32", k_gen_unit_f32, gen_grid(n), 256, 0, d_out, n,
            rng_key(seed, stream));
  return sync(ctx);
}

int pj_gen_zipf_fk_i32(pj_ctx *ctx, int32_t *d_out, uint64_t n, uint64_t seed,
                       uint64_t stream, uint64_t nbuild, uint64_t range,
                       double theta, uint64_t perm_seed) {
  if (!n) return PJ_OK;
  if (range < nbuild) return fail(PJ_E_ARG, "range < nbuild");
  PJ_LAUNCH(ctx, "gen_zipf_fk", k_gen_zipf_fk, gen_grid(n), 256, 0, d_out, n,
            rng_key(seed, stream), nbuild, range, make_zipf(seed, stream + 7, nbuild, theta),
            make_perm(perm_seed, nbuild));
  return sync(ctx);
}

int pj_gen_zipf_i32(pj_ctx *ctx, int32_t *d_out, uint64_t n, uint64_t seed,
                    uint64_t stream, uint64_t g, double theta,
                    uint64_t perm_seed) {
  if (!n) return PJ_OK;
  PJ_LAUNCH(ctx, "gen_zipf", k_gen_zipf, gen_grid(n), 256, 0, d_out, n,
            make_zipf(seed, stream, g, theta), make_perm(perm_seed, g));
  return sync(ctx);
}

}  // extern "C"
