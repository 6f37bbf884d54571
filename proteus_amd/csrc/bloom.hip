// Bloom filters of the plan operators bloomfilter-build, bloomfilter-probe
// and bloomfilter-repack: the reference's BloomFilter operators
// (core/olap/lib/operators/bloom-filter/bloom-filter.cpp:36-107,
// bloom-filter-build.cpp:29-58, bloom-filter-probe.cpp:29-52,
// bloom-filter-repack.cpp:43-470) behind RelBuilder::bloomfilter_build /
// _probe / _repack (core/olap/lib/operators/relbuilder.cpp:957-975).
//
// The reference's filter is `filterSize` one-byte flags (filterSize a power
// of two), and key k selects flag (int32)k & (filterSize - 1)
// (BloomFilter::findInFilter, bloom-filter.cpp:85-107): one hash function,
// the identity.  Here the flags are bits, filterSize / 8 bytes (at least one
// 32-bit word): the same set in 1/8 of the bytes, so a filter of 2^20
// entries is 128 KiB and stays in L2 while a fact table streams past it.
//
// Build: a fresh zeroed filter, one atomicOr per key; it replaces the
// context's filter of the same id, as setBloomFilter (bloom-filter.cpp:36-41)
// replaces and frees the registered one.  Probe and repack test bits with
// OP_BLOOM of the expression VM and of the fused probe pipeline (rel.hip);
// pj_bloom_probe_i32 and pj_bloom_repack are the same as C calls.
#include <map>

#include "pj_internal.hpp"
#include "rel_api.hpp"

namespace pj {

struct BloomFilter {
  void *words = nullptr;  // pool block of max(1, size / 32) uint32 words
  int log2 = 0;           // log2 of the filter size, 0..32
};
using BloomMap = std::map<uint64_t, BloomFilter>;

static BloomMap &blooms(pj_ctx *ctx) {
  if (!ctx->bloom) ctx->bloom = new BloomMap();
  return *(BloomMap *)ctx->bloom;
}

__host__ __device__ __forceinline__ uint32_t bloom_mask(int log2) {
  return log2 >= 32 ? 0xffffffffu : (1u << log2) - 1u;
}

static uint64_t bloom_words(int log2) { return log2 <= 5 ? 1 : (1ULL << (log2 - 5)); }

__global__ __launch_bounds__(256) void k_bloom_set(uint32_t *__restrict__ words, uint32_t mask,
                                                   const int32_t *__restrict__ keys, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
    const uint32_t h = (uint32_t)keys[i] & mask;
    atomicOr(&words[h >> 5], 1u << (h & 31));
  }
}

// Filter `id` of the context; the caller holds ctx->mu.
static int lookup_locked(pj_ctx *ctx, uint64_t id, const uint32_t **words, int *log2_size) {
  BloomMap &m = blooms(ctx);
  auto it = m.find(id);
  if (it == m.end())
    return fail(PJ_E_ARG, "bloom filter %llu has not been built", (unsigned long long)id);
  *words = (const uint32_t *)it->second.words;
  *log2_size = it->second.log2;
  return PJ_OK;
}

namespace rel {

int bloom_lookup(pj_ctx *ctx, uint64_t id, const uint32_t **words, int *log2_size) {
  if (!ctx || !words || !log2_size) return fail(PJ_E_ARG, "null argument");
  std::lock_guard<std::mutex> g(ctx->mu);
  return lookup_locked(ctx, id, words, log2_size);
}

}  // namespace rel

// The probe predicate over one int32 column as a VM program.
static rel::Prog bloom_prog(const int32_t *keys, const uint32_t *words, int log2) {
  rel::Prog p;
  p.cols[0] = keys;
  p.ctype[0] = rel::I32;
  p.ncols = 1;
  p.ins[0] = rel::Instr{rel::OP_COL, rel::I32, 0, 0, 0};
  p.ins[1] = rel::Instr{rel::OP_BLOOM, rel::BOOL, log2, 0, (int64_t)(uintptr_t)words};
  p.n = 2;
  p.depth = 1;
  return p;
}

}  // namespace pj

using namespace pj;

extern "C" void pj_bloom_release_all(pj_ctx *ctx) {
  if (!ctx || !ctx->bloom) return;
  BloomMap *m = (BloomMap *)ctx->bloom;
  for (auto &kv : *m) (void)dfree(ctx, kv.second.words);
  delete m;
  ctx->bloom = nullptr;
}

extern "C" int pj_bloom_build_i32(pj_ctx *ctx, uint64_t bloom_id, uint64_t filter_size,
                                  const int32_t *d_keys, uint64_t n) {
  if (!ctx) return fail(PJ_E_ARG, "null context");
  if (!filter_size || (filter_size & (filter_size - 1)) || filter_size > (1ULL << 32))
    return fail(PJ_E_ARG, "bloom filter size %llu is not a power of two in [1, 2^32]",
                (unsigned long long)filter_size);
  if (n && !d_keys) return fail(PJ_E_ARG, "bloom filter build without keys");
  int log2 = 0;
  while ((1ULL << log2) < filter_size) ++log2;
  std::lock_guard<std::mutex> g(ctx->mu);
  (void)hipSetDevice(ctx->device);
  BloomFilter f;
  f.log2 = log2;
  const uint64_t nw = bloom_words(log2);
  PJ_TRY(dalloc(ctx, nw * 4, &f.words));
  // a launch that fails returns from the lambda, and the block is freed below
  auto set_bits = [&]() -> int {
    PJ_TRY(fill_u32(ctx, (uint32_t *)f.words, 0, nw));
    if (n) {
      uint64_t grid = ceil_div(n, (uint64_t)256);
      const uint64_t cap = (uint64_t)ctx->num_cus * 16;
      if (grid > cap) grid = cap;
      PJ_LAUNCH(ctx, "bloom_set", k_bloom_set, grid, 256, 0, (uint32_t *)f.words,
                bloom_mask(log2), d_keys, n);
    }
    return PJ_OK;
  };
  const int rc = set_bits();
  if (rc != PJ_OK) {
    (void)dfree(ctx, f.words);
    return rc;
  }
  BloomMap &m = blooms(ctx);
  auto it = m.find(bloom_id);
  if (it != m.end()) {
    (void)dfree(ctx, it->second.words);  // stream-ordered reuse only
    it->second = f;
  } else {
    m.emplace(bloom_id, f);
  }
  return PJ_OK;
}

extern "C" int pj_bloom_probe_i32(pj_ctx *ctx, uint64_t bloom_id, const int32_t *d_keys,
                                  uint64_t n, uint32_t *d_idx, uint64_t *n_out) {
  if (!ctx || !n_out) return fail(PJ_E_ARG, "null argument");
  *n_out = 0;
  if (n && (!d_keys || !d_idx)) return fail(PJ_E_ARG, "bloom filter probe without keys or output");
  std::lock_guard<std::mutex> g(ctx->mu);
  const uint32_t *words = nullptr;
  int log2 = 0;
  PJ_TRY(lookup_locked(ctx, bloom_id, &words, &log2));
  if (!n) return PJ_OK;
  (void)hipSetDevice(ctx->device);
  PJ_TRY(rel::select(ctx, bloom_prog(d_keys, words, log2), n, d_idx, n_out));
  return sync(ctx);
}

extern "C" int pj_bloom_repack(pj_ctx *ctx, uint64_t bloom_id, const int32_t *d_keys, uint64_t n,
                               int ncols, const void *const *d_cols, const int *esize,
                               void *const *d_out, uint64_t *n_out) {
  if (!ctx || !n_out) return fail(PJ_E_ARG, "null argument");
  *n_out = 0;
  if (ncols < 0 || (ncols && (!d_cols || !esize || !d_out)))
    return fail(PJ_E_ARG, "bloom filter repack with invalid columns");
  for (int c = 0; c < ncols; ++c)
    if (esize[c] != 1 && esize[c] != 4 && esize[c] != 8)
      return fail(PJ_E_ARG, "bloom filter repack of %d-byte elements", esize[c]);
  if (n && !d_keys) return fail(PJ_E_ARG, "bloom filter repack without keys");
  std::lock_guard<std::mutex> g(ctx->mu);
  const uint32_t *words = nullptr;
  int log2 = 0;
  PJ_TRY(lookup_locked(ctx, bloom_id, &words, &log2));
  if (!n) return PJ_OK;
  (void)hipSetDevice(ctx->device);
  Scratch idx;
  PJ_TRY(idx.alloc(ctx, n * 4));
  uint64_t m = 0;
  PJ_TRY(rel::select(ctx, bloom_prog(d_keys, words, log2), n, idx.as<uint32_t>(), &m));
  for (int c = 0; c < ncols; ++c)
    PJ_TRY(rel::gather(ctx, d_cols[c], esize[c], idx.as<const uint32_t>(), m, d_out[c]));
  PJ_TRY(sync(ctx));
  *n_out = m;
  return PJ_OK;
}

extern "C" int pj_bloom_release(pj_ctx *ctx, uint64_t bloom_id) {
  if (!ctx) return fail(PJ_E_ARG, "null context");
  std::lock_guard<std::mutex> g(ctx->mu);
  BloomMap &m = blooms(ctx);
  auto it = m.find(bloom_id);
  if (it == m.end())
    return fail(PJ_E_ARG, "bloom filter %llu has not been built", (unsigned long long)bloom_id);
  (void)hipStreamSynchronize(ctx->stream);
  (void)dfree(ctx, it->second.words);
  m.erase(it);
  return PJ_OK;
}
