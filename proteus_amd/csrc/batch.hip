// Resident dimension tables and batched join -> group-by queries.
//
// A query joins a probe relation S(fk, val, grp) with a resident dimension
// R(key, value) on S.fk = R.key and groups the matches by S.grp:
//     SELECT grp, COUNT(*), SUM(R.value + S.val) GROUP BY grp
// (the plan hashjoin-chained -> groupby of the reference; group keys are
// dense in [0, groups)).
//
// Dimension tables stay in HBM across queries: open addressing over 64-bit
// slots key << 32 | value (empty = key INT32_MIN), load factor <= 1/2.  A
// 2^24-row dimension is a 256 MiB table, the size of MI355X's Infinity
// Cache, so probes of resident dimensions are served mostly from MALL.
// Dimensions of <= kLdsDim rows are copied into LDS once per work tile.
//
// One batch = one persistent launch over a device work queue of
// (query, tile) items: a workgroup takes the next item from a counter,
// binary-searches its query, probes and aggregates the tile (group counts
// and sums in LDS when the query has <= kLdsGroups groups, else with HBM
// atomics) and flushes the LDS partials with one atomic per group.  The
// batch sequence (descriptor upload, output clearing, input generation,
// probe) is captured once per stream slot as a hipGraph and replayed for
// every batch; the plain path issues the same kernels once per query.
#include <algorithm>
#include <vector>

#include "pj_internal.hpp"
#include "rel_api.hpp"

struct pj_dim {
  uint64_t n = 0;
  uint64_t slots = 0;  // table entries: a power of two (hash) or `range` (direct)
  unsigned long long *table = nullptr;  // 64-bit slots key << 32 | value
  uint32_t *ids = nullptr;  // instead, for a direct table of build row ids: the id, or
                            // kNoRow (pipelines of the plan executor)
  int32_t *vals = nullptr;  // instead, for a dense direct table (range = n: every key
                            // of the span present): the value of key lo + i at i
  int32_t lo = 0;      // direct tables: smallest key
  uint32_t range = 0;  // direct tables: max key - lo + 1; 0 = hash table
};

namespace pj {

constexpr unsigned long long kEmptySlot = 0x8000000000000000ULL;  // key INT32_MIN
constexpr uint32_t kLdsDim = 2048;      // dims copied into LDS (<= 4096 slots)
constexpr uint32_t kLdsSlots = 4096;
constexpr uint32_t kLdsGroups = 2048;   // groups aggregated in LDS
constexpr int kBatchBlock = 512;
constexpr uint32_t kMaxBatch = 4096;    // queries per batch

struct QDesc {
  const unsigned long long *table;
  const int32_t *dval; // dense direct table (table null): the value of key dlo + i
  uint64_t mask;       // hash table: slots - 1
  int32_t dlo;         // direct table: smallest key
  uint32_t drange;     // direct table: slots; 0 = hash table
  uint32_t dslots;     // table entries
  const int32_t *fk;
  const int32_t *val;
  const int32_t *grp;
  int32_t *gfk, *gval, *ggrp;  // generator targets (same buffers)
  int64_t *cnt, *sum;
  uint64_t n;
  uint64_t tile0;      // first work item of this query
  uint64_t gen_seed;   // 0: inputs given
  uint64_t fk_range;
  uint32_t dn;         // dimension rows
  uint32_t groups;
  // Staged aggregation of queries with more groups than LDS holds: matches
  // go to per-partition runs of (group & 2047, value) words, partition =
  // group >> 11, and k_batch_agg folds each chunk of a run in LDS.  stage
  // is null for the other queries (and when no scratch was available).
  unsigned long long *stage;  // parts x cap words
  uint32_t *pcur;             // parts cursors
  uint64_t cap;               // words per partition run
  uint64_t agg0;              // first k_batch_agg item of this query
  uint32_t parts;
  uint32_t cpp;               // k_batch_agg chunks per partition
};

struct BatchHdr {
  uint32_t nq;
  uint32_t tile_rows;
  uint64_t total_tiles;
  uint64_t lds_tiles;  // items [0, lds_tiles) belong to LDS-table queries
  uint32_t gen_ctr, probe_ctr_lds, probe_ctr_hbm;
  uint32_t bad_group;  // rows whose group key is outside [0, groups)
  uint64_t total_agg;  // k_batch_agg items
  uint32_t agg_ctr;
};

constexpr uint32_t kStageGroups = 2048;         // groups per staged partition (LDS)
constexpr uint32_t kStageMaxParts = 32;         // staging for <= 65536 groups
constexpr uint32_t kAggChunk = 1u << 18;        // staged words per k_batch_agg item
__device__ __forceinline__ unsigned long long stage_word(uint32_t gl, long long sv) {
  return ((unsigned long long)gl << 53) | ((unsigned long long)sv & ((1ULL << 53) - 1));
}

__global__ void k_dim_fill(unsigned long long *t, uint64_t n) {
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    t[i] = kEmptySlot;
}

// A direct table of build row ids: slot key - lo holds the row (a row id is
// below 2^31, so kNoRow marks an empty slot).  Rows whose mask byte is 0
// (mask may be null) are left out.  flags as k_dim_build.
constexpr uint32_t kNoRow = 0xffffffffu;
__global__ void k_dim_build_ids(const int32_t *__restrict__ k, const uint8_t *__restrict__ mask,
                                uint64_t n, uint32_t *ids, int32_t lo, uint32_t range,
                                unsigned int *flags) {
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    if (mask && !mask[i]) continue;
    const uint32_t off = (uint32_t)k[i] - (uint32_t)lo;
    if (off >= range) {  // cannot happen: lo and range cover every key
      atomicOr(flags, 4u);
      continue;
    }
    if (atomicCAS(&ids[off], kNoRow, (uint32_t)i) != kNoRow) atomicOr(flags, 1u);
  }
}

// A dense direct table (range = n keys, so every key of the span is present
// once they are unique): vals[key - lo] = v[i] (or 0), 4 bytes a key; bits
// (range bits, zeroed) catch a duplicate.  flags as k_dim_build.
__global__ void k_dim_build_dense(const int32_t *__restrict__ k, const int32_t *__restrict__ v,
                                  uint64_t n, int32_t *vals, uint32_t *bits, int32_t lo,
                                  uint32_t range, unsigned int *flags) {
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t off = (uint32_t)k[i] - (uint32_t)lo;
    if (off >= range) {  // cannot happen: lo and range cover every key
      atomicOr(flags, 4u);
      continue;
    }
    const uint32_t bit = 1u << (off & 31);
    if (atomicOr(&bits[off >> 5], bit) & bit) atomicOr(flags, 1u);
    vals[off] = v ? v[i] : 0;
  }
}

// flags: bit 0 duplicate key, bit 1 reserved key INT32_MIN.  The value of
// row i is v[i], or i itself when rows is set (v is then unused), or 0.
// Rows whose keep byte is 0 (keep may be null) are left out.
__global__ void k_dim_build(const int32_t *__restrict__ k, const int32_t *__restrict__ v,
                            const uint8_t *__restrict__ keep, uint64_t n, unsigned long long *t,
                            uint64_t mask, int32_t lo, uint32_t range, unsigned int *flags,
                            bool rows) {
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    if (keep && !keep[i]) continue;
    const int32_t key = k[i];
    if (key == INT32_MIN) {
      atomicOr(flags, 2u);
      continue;
    }
    const uint32_t val = rows ? (uint32_t)i : (uint32_t)(v ? v[i] : 0);
    const unsigned long long w = (unsigned long long)(uint32_t)key << 32 | val;
    bool in;
    uint64_t h = dim_start(key, mask, lo, range, &in);
    if (!in) {  // cannot happen: lo and range cover every key
      atomicOr(flags, 4u);
      continue;
    }
    for (;;) {
      unsigned long long old = atomicCAS(&t[h], kEmptySlot, w);
      if (old == kEmptySlot) break;
      if ((int32_t)(old >> 32) == key || range) {  // a direct slot holds only its key
        atomicOr(flags, 1u);
        break;
      }
      h = (h + 1) & mask;
    }
  }
}

__device__ __forceinline__ uint32_t find_query(const QDesc *d, uint32_t nq, uint64_t item) {
  uint32_t lo = 0, hi = nq;  // last q with tile0 <= item
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (d[mid].tile0 <= item) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Seeded probe inputs: fk uniform [0, fk_range), val uniform [0, 2^31),
// grp uniform [0, groups) -- streams 1, 2, 3 of the query's seed, the same
// formula as pj_gen_uniform_i32.  The workgroup that takes a query's first
// item also clears its outputs (every query owns at least one item); the
// probe kernel starts after this kernel ends.
__global__ __launch_bounds__(256) void k_batch_gen(BatchHdr *hdr, const QDesc *d) {
  __shared__ uint64_t s_item;
  const uint32_t nq = hdr->nq;
  const uint64_t total = hdr->total_tiles;
  const uint32_t T = hdr->tile_rows;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(&hdr->gen_ctr, 1u);
    __syncthreads();
    const uint64_t item = s_item;
    __syncthreads();
    if (item >= total) break;
    const QDesc &q = d[find_query(d, nq, item)];
    if (item == q.tile0) {
      for (uint32_t i = threadIdx.x; i < q.groups; i += blockDim.x) {
        q.cnt[i] = 0;
        q.sum[i] = 0;
      }
      if (q.stage)
        for (uint32_t i = threadIdx.x; i < q.parts; i += blockDim.x) q.pcur[i] = 0;
    }
    if (!q.gen_seed) continue;
    const uint64_t r0 = (item - q.tile0) * T;
    const uint64_t r1 = r0 + T < q.n ? r0 + T : q.n;
    const uint64_t k1 = rng_key(q.gen_seed, 1), k2 = rng_key(q.gen_seed, 2), k3 = rng_key(q.gen_seed, 3);
    // rng_at(k, r) = mix64(k + kGolden (r + 1)): the counter term advances
    // by kGolden * blockDim per step (mod 2^64, as in rng_at); ranges are
    // below 2^32, so the scaling takes two 32-bit products
    const uint32_t fr = (uint32_t)q.fk_range, gr = q.groups;
    uint64_t ctr = kGolden * (r0 + threadIdx.x + 1);
    const uint64_t step = kGolden * blockDim.x;
    for (uint64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x, ctr += step) {
      q.gfk[r] = (int32_t)mulhi64_32(mix64(k1 + ctr), fr);
      q.gval[r] = (int32_t)mulhi64_32(mix64(k2 + ctr), 1u << 31);
      q.ggrp[r] = (int32_t)mulhi64_32(mix64(k3 + ctr), gr);
    }
  }
}

// One probe row: table lookup, then the group update.
template <bool LDS_TAB, bool LDS_GRP>
__device__ __forceinline__ void probe_row(const QDesc &q, const unsigned long long *ltab,
                                          uint32_t *gcnt, unsigned long long *gsum, int32_t k,
                                          int32_t v, int32_t g, uint32_t &bad) {
  int32_t b = 0;
  if (!LDS_TAB && q.dval) {
    const uint32_t off = (uint32_t)k - (uint32_t)q.dlo;
    if (off >= q.drange) return;
    b = q.dval[off];
  } else {
    bool in;
    uint64_t h = dim_start(k, q.mask, q.dlo, q.drange, &in);
    if (!in) return;
    for (;;) {
      const unsigned long long s = LDS_TAB ? ltab[h] : q.table[h];
      if (s == kEmptySlot) return;  // before the key test: a probe key INT32_MIN is a miss
      if ((int32_t)(s >> 32) == k) {
        b = (int32_t)(uint32_t)s;
        break;
      }
      if (q.drange) return;  // a direct slot holds only its own key
      h = (h + 1) & q.mask;
    }
  }
  if ((uint32_t)g >= q.groups) {
    ++bad;
    return;
  }
  const long long sv = (long long)b + (long long)v;
  if (LDS_GRP) {
    atomicAdd(&gcnt[g], 1u);
    atomicAdd(&gsum[g], (unsigned long long)sv);
  } else {
    atomicAdd((unsigned long long *)&q.cnt[g], 1ULL);
    atomicAdd((unsigned long long *)&q.sum[g], (unsigned long long)sv);
  }
}

// Eight rows of one lane: their probe sequences advance together, so the
// table loads of all eight are in flight at once (a row at a time, each
// probe would wait for the previous row's), then the matches aggregate.
template <bool LDS_TAB, bool LDS_GRP>
__device__ __forceinline__ void probe8(const QDesc &q, const unsigned long long *ltab,
                                       uint32_t *gcnt, unsigned long long *gsum,
                                       const int32_t (&k)[8], const int32_t (&v)[8],
                                       const int32_t (&g)[8], uint32_t &bad) {
  const uint32_t mask = (uint32_t)q.mask;
  const bool direct = q.drange != 0;
  int32_t b[8];
  bool hit[8];
  if (!LDS_TAB && q.dval) {  // dense: one 4-byte load, no key test
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const uint32_t off = (uint32_t)k[j] - (uint32_t)q.dlo;
      hit[j] = off < q.drange;
      b[j] = q.dval[hit[j] ? off : 0u];
    }
  } else {
    uint32_t h[8];
    bool act[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      h[j] = dim_start(k[j], q.mask, q.dlo, q.drange, &act[j]);
      hit[j] = false;
      b[j] = 0;
    }
    for (;;) {
      unsigned long long s[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const uint32_t at = act[j] ? h[j] : 0u;
        s[j] = LDS_TAB ? ltab[at] : q.table[at];
      }
      bool any = false;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const bool empty = s[j] == kEmptySlot;
        const bool eq = !empty && (int32_t)(s[j] >> 32) == k[j];
        const bool got = act[j] && eq;
        const bool more = act[j] && !empty && !eq && !direct;
        hit[j] = hit[j] || got;
        b[j] = got ? (int32_t)(uint32_t)s[j] : b[j];
        h[j] = more ? ((h[j] + 1) & mask) : h[j];
        act[j] = more;
        any = any || act[j];
      }
      if (!any) break;
    }
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    if (!hit[j]) continue;
    if ((uint32_t)g[j] >= q.groups) {
      ++bad;
      continue;
    }
    const long long sv = (long long)b[j] + (long long)v[j];
    if (LDS_GRP) {
      atomicAdd(&gcnt[g[j]], 1u);
      atomicAdd(&gsum[g[j]], (unsigned long long)sv);
    } else {
      atomicAdd((unsigned long long *)&q.cnt[g[j]], 1ULL);
      atomicAdd((unsigned long long *)&q.sum[g[j]], (unsigned long long)sv);
    }
  }
}

// The probe of probe8 for a staged query: the eight rows' matches are
// appended to their partitions' runs (ranks from LDS counters, one global
// reservation per partition per step); a match that finds its run full
// falls back to the global atomics.  Every thread of the workgroup calls
// this the same number of times (rows past the end pass valid = false).
__device__ __forceinline__ void stage8(const QDesc &q, bool valid, const int32_t (&k)[8],
                                       const int32_t (&v)[8], const int32_t (&g)[8],
                                       uint32_t &bad, uint32_t *s_pcnt,
                                       unsigned long long *s_pbase) {
  const uint32_t mask = (uint32_t)q.mask;
  const bool direct = q.drange != 0;
  int32_t b[8];
  bool hit[8];
  if (q.dval) {  // dense: one 4-byte load, no key test
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const uint32_t off = (uint32_t)k[j] - (uint32_t)q.dlo;
      hit[j] = valid && off < q.drange;
      b[j] = q.dval[hit[j] ? off : 0u];
    }
  } else {
    uint32_t h[8];
    bool act[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      bool in;
      h[j] = dim_start(k[j], q.mask, q.dlo, q.drange, &in);
      act[j] = valid && in;
      hit[j] = false;
      b[j] = 0;
    }
    for (;;) {
      unsigned long long s[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) s[j] = q.table[act[j] ? h[j] : 0u];
      bool any = false;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const bool empty = s[j] == kEmptySlot;
        const bool eq = !empty && (int32_t)(s[j] >> 32) == k[j];
        const bool got = act[j] && eq;
        const bool more = act[j] && !empty && !eq && !direct;
        hit[j] = hit[j] || got;
        b[j] = got ? (int32_t)(uint32_t)s[j] : b[j];
        h[j] = more ? ((h[j] + 1) & mask) : h[j];
        act[j] = more;
        any = any || act[j];
      }
      if (!any) break;
    }
  }
  uint32_t rank[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    if (hit[j] && (uint32_t)g[j] >= q.groups) {
      ++bad;
      hit[j] = false;
    }
    rank[j] = hit[j] ? atomicAdd(&s_pcnt[(uint32_t)g[j] / kStageGroups], 1u) : 0u;
  }
  __syncthreads();
  if (threadIdx.x < q.parts)
    s_pbase[threadIdx.x] =
        s_pcnt[threadIdx.x] ? atomicAdd(&q.pcur[threadIdx.x], s_pcnt[threadIdx.x]) : 0u;
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    if (!hit[j]) continue;
    const uint32_t p = (uint32_t)g[j] / kStageGroups;
    const unsigned long long pos = s_pbase[p] + rank[j];
    const long long sv = (long long)b[j] + (long long)v[j];
    if (pos < q.cap) {
      q.stage[(uint64_t)p * q.cap + pos] = stage_word((uint32_t)g[j] % kStageGroups, sv);
    } else {
      atomicAdd((unsigned long long *)&q.cnt[g[j]], 1ULL);
      atomicAdd((unsigned long long *)&q.sum[g[j]], (unsigned long long)sv);
    }
  }
  __syncthreads();
  if (threadIdx.x < q.parts) s_pcnt[threadIdx.x] = 0;
  __syncthreads();
}

// Rows [r0, r1) of a staged query's item: steps of 8 rows per lane that
// every thread takes (the steps synchronise the workgroup); the rows after
// the last whole step of the workgroup, or all rows when the columns are
// not 16-byte aligned, go through probe_row's global atomics.
__device__ __forceinline__ void stage_rows(const QDesc &q, uint64_t r0, uint64_t r1, uint32_t &bad,
                                           uint32_t *s_pcnt, unsigned long long *s_pbase) {
  const int tid = threadIdx.x;
  const bool vec = ((((uintptr_t)q.fk | (uintptr_t)q.val | (uintptr_t)q.grp) & 15) == 0) && (r0 % 4 == 0);
  uint64_t r = r0;
  if (vec) {
    const uint64_t full = r0 + (r1 - r0) / 8 * 8;
    for (uint64_t sb = r0; sb < full; sb += (uint64_t)kBatchBlock * 8) {
      const uint64_t base = sb + (uint64_t)tid * 8;
      const bool valid = base < full;
      const uint64_t at = valid ? base : full - 8;
      int4 k[2], v[2], g[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        k[h] = ld_nt4(q.fk + at + 4 * h);
        v[h] = ld_nt4(q.val + at + 4 * h);
        g[h] = ld_nt4(q.grp + at + 4 * h);
      }
      const int32_t kk[8] = {k[0].x, k[0].y, k[0].z, k[0].w, k[1].x, k[1].y, k[1].z, k[1].w};
      const int32_t vv[8] = {v[0].x, v[0].y, v[0].z, v[0].w, v[1].x, v[1].y, v[1].z, v[1].w};
      const int32_t gg[8] = {g[0].x, g[0].y, g[0].z, g[0].w, g[1].x, g[1].y, g[1].z, g[1].w};
      stage8(q, valid, kk, vv, gg, bad, s_pcnt, s_pbase);
    }
    r = full;
  }
  for (uint64_t i = r + tid; i < r1; i += kBatchBlock)
    probe_row<false, false>(q, nullptr, nullptr, nullptr, q.fk[i], q.val[i], q.grp[i], bad);
}

// Rows [r0, r1) of one item: 8 rows per lane per step, 16-byte loads where
// the columns allow.
template <bool LDS_TAB, bool LDS_GRP>
__device__ __forceinline__ void probe_rows(const QDesc &q, uint64_t r0, uint64_t r1,
                                           const unsigned long long *ltab, uint32_t *gcnt,
                                           unsigned long long *gsum, uint32_t &bad) {
  const int tid = threadIdx.x;
  const bool vec = ((((uintptr_t)q.fk | (uintptr_t)q.val | (uintptr_t)q.grp) & 15) == 0) && (r0 % 4 == 0);
  uint64_t r = r0;
  if (vec) {
    const uint64_t full = r0 + (r1 - r0) / 8 * 8;
    for (uint64_t base = r0 + (uint64_t)tid * 8; base < full; base += (uint64_t)kBatchBlock * 8) {
      int4 k[2], v[2], g[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        k[h] = ld_nt4(q.fk + base + 4 * h);
        v[h] = ld_nt4(q.val + base + 4 * h);
        g[h] = ld_nt4(q.grp + base + 4 * h);
      }
      const int32_t kk[8] = {k[0].x, k[0].y, k[0].z, k[0].w, k[1].x, k[1].y, k[1].z, k[1].w};
      const int32_t vv[8] = {v[0].x, v[0].y, v[0].z, v[0].w, v[1].x, v[1].y, v[1].z, v[1].w};
      const int32_t gg[8] = {g[0].x, g[0].y, g[0].z, g[0].w, g[1].x, g[1].y, g[1].z, g[1].w};
      probe8<LDS_TAB, LDS_GRP>(q, ltab, gcnt, gsum, kk, vv, gg, bad);
    }
    r = full;
  }
  for (uint64_t i = r + tid; i < r1; i += kBatchBlock)
    probe_row<LDS_TAB, LDS_GRP>(q, ltab, gcnt, gsum, q.fk[i], q.val[i], q.grp[i], bad);
}

// Persistent probe over items [first, last): LDS_TAB copies the query's
// table into LDS (dimensions of <= kLdsDim rows); otherwise the table is
// read from HBM / Infinity Cache and the LDS holds only group partials.
template <bool LDS_TAB>
__global__ __launch_bounds__(kBatchBlock) void k_batch_probe(BatchHdr *hdr, const QDesc *d) {
  __shared__ unsigned long long ltab[LDS_TAB ? kLdsSlots : 1];
  __shared__ uint32_t gcnt[kLdsGroups];
  __shared__ unsigned long long gsum[kLdsGroups];
  __shared__ uint64_t s_item;
  __shared__ uint32_t s_bad;
  __shared__ uint32_t s_pcnt[kStageMaxParts];
  __shared__ unsigned long long s_pbase[kStageMaxParts];
  if (threadIdx.x < kStageMaxParts) s_pcnt[threadIdx.x] = 0;
  const uint32_t nq = hdr->nq;
  const uint64_t first = LDS_TAB ? 0 : hdr->lds_tiles;
  const uint64_t last = LDS_TAB ? hdr->lds_tiles : hdr->total_tiles;
  uint32_t *ctr = LDS_TAB ? &hdr->probe_ctr_lds : &hdr->probe_ctr_hbm;
  const uint32_t T = hdr->tile_rows;
  const int tid = threadIdx.x;
  uint32_t last_q = 0xffffffffu;
  for (;;) {
    if (tid == 0) {
      s_item = first + atomicAdd(ctr, 1u);
      s_bad = 0;
    }
    __syncthreads();
    const uint64_t item = s_item;
    if (item >= last) break;
    const uint32_t qi = find_query(d, nq, item);
    const QDesc &q = d[qi];
    const bool lds_grp = q.groups <= kLdsGroups;
    if (LDS_TAB && qi != last_q)
      for (uint32_t i = tid; i < q.dslots; i += kBatchBlock) ltab[i] = q.table[i];
    last_q = qi;
    if (lds_grp)
      for (uint32_t i = tid; i < q.groups; i += kBatchBlock) {
        gcnt[i] = 0;
        gsum[i] = 0;
      }
    __syncthreads();
    const uint64_t r0 = (item - q.tile0) * T;
    const uint64_t r1 = r0 + T < q.n ? r0 + T : q.n;
    uint32_t bad = 0;
    if (r0 < r1) {
      if (lds_grp) probe_rows<LDS_TAB, true>(q, r0, r1, ltab, gcnt, gsum, bad);
      else if (!LDS_TAB && q.stage) stage_rows(q, r0, r1, bad, s_pcnt, s_pbase);
      else probe_rows<LDS_TAB, false>(q, r0, r1, ltab, gcnt, gsum, bad);
    }
    if (bad) atomicAdd(&s_bad, bad);
    __syncthreads();
    if (lds_grp)
      for (uint32_t i = tid; i < q.groups; i += kBatchBlock)
        if (gcnt[i]) {
          atomicAdd((unsigned long long *)&q.cnt[i], (unsigned long long)gcnt[i]);
          atomicAdd((unsigned long long *)&q.sum[i], gsum[i]);
        }
    if (tid == 0 && s_bad) atomicAdd(&hdr->bad_group, s_bad);
    __syncthreads();
  }
}

// Folds the staged runs: item = (query, partition, chunk of kAggChunk
// words); the chunk's words are counted and summed per group in LDS and the
// non-zero groups are added to the query's outputs (one atomic pair per
// group per chunk instead of one per match).
__device__ __forceinline__ uint32_t find_agg_query(const QDesc *d, uint32_t nq, uint64_t item) {
  uint32_t lo = 0, hi = nq;  // last q with agg0 <= item
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (d[mid].agg0 <= item) lo = mid;
    else hi = mid;
  }
  return lo;
}
__global__ __launch_bounds__(kBatchBlock) void k_batch_agg(BatchHdr *hdr, const QDesc *d) {
  __shared__ uint32_t acnt[kStageGroups];
  __shared__ unsigned long long asum[kStageGroups];
  __shared__ uint64_t s_item;
  const uint32_t nq = hdr->nq;
  const uint64_t total = hdr->total_agg;
  const int tid = threadIdx.x;
  for (;;) {
    if (tid == 0) s_item = atomicAdd(&hdr->agg_ctr, 1u);
    __syncthreads();
    const uint64_t item = s_item;
    __syncthreads();
    if (item >= total) break;
    const QDesc &q = d[find_agg_query(d, nq, item)];
    const uint64_t local = item - q.agg0;
    const uint32_t p = (uint32_t)(local / q.cpp);
    const uint64_t c0 = (local % q.cpp) * kAggChunk;
    const uint64_t used = min((unsigned long long)q.pcur[p], (unsigned long long)q.cap);
    if (c0 >= used) continue;  // uniform: every thread read the same item
    const uint64_t c1 = min((unsigned long long)(c0 + kAggChunk), (unsigned long long)used);
    for (uint32_t i = tid; i < kStageGroups; i += kBatchBlock) {
      acnt[i] = 0;
      asum[i] = 0;
    }
    __syncthreads();
    const unsigned long long *w = q.stage + (uint64_t)p * q.cap;
    for (uint64_t i = c0 + tid; i < c1; i += kBatchBlock) {
      const unsigned long long x = w[i];
      const uint32_t gl = (uint32_t)(x >> 53);
      const long long sv = (long long)(x << 11) >> 11;
      atomicAdd(&acnt[gl], 1u);
      atomicAdd(&asum[gl], (unsigned long long)sv);
    }
    __syncthreads();
    const uint32_t gbase = p * kStageGroups;
    for (uint32_t i = tid; i < kStageGroups; i += kBatchBlock)
      if (acnt[i]) {
        atomicAdd((unsigned long long *)&q.cnt[gbase + i], (unsigned long long)acnt[i]);
        atomicAdd((unsigned long long *)&q.sum[gbase + i], asum[i]);
      }
    __syncthreads();
  }
}

// One stream slot: pinned + device descriptor buffers, its graph.
struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  void *h_buf = nullptr;  // pinned: BatchHdr then QDesc[kMaxBatch]
  void *d_buf = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint32_t graph_grid = 0;
  bool busy = false;
  std::vector<void *> stage;  // staging of the batch in flight
};

struct Batches {
  std::vector<Slot> slots;
  pj_batch_cfg tuned{};  // defaults for zero fields (pj_batch_autotune)
};

constexpr size_t kBufBytes = 256 + sizeof(QDesc) * kMaxBatch;

static Batches &batches(pj_ctx *ctx) {
  if (!ctx->batch) ctx->batch = new Batches();
  return *(Batches *)ctx->batch;
}

static int fill_desc(const pj_query &q, QDesc *d) {
  if (!q.dim) return fail(PJ_E_ARG, "query without a dimension");
  if (!q.d_cnt || !q.d_sum) return fail(PJ_E_ARG, "query without output arrays");
  if (!q.groups) return fail(PJ_E_ARG, "query with 0 groups");
  if (q.n && (!q.d_fk || !q.d_val || !q.d_grp)) return fail(PJ_E_ARG, "query without probe columns");
  if (q.gen_seed && (!q.fk_range || q.fk_range > (1ULL << 31)))
    return fail(PJ_E_ARG, "generated query needs fk_range in [1, 2^31]");
  if (q.n >= (1ULL << 40)) return fail(PJ_E_ARG, "query probe side too large");
  d->table = q.dim->table;
  d->dval = q.dim->vals;
  d->mask = q.dim->range ? 0 : q.dim->slots - 1;
  d->dlo = q.dim->lo;
  d->drange = q.dim->range;
  d->dslots = (uint32_t)q.dim->slots;
  d->dn = (uint32_t)std::min<uint64_t>(q.dim->n, 0xffffffffULL);
  d->fk = q.d_fk;
  d->val = q.d_val;
  d->grp = q.d_grp;
  d->gfk = const_cast<int32_t *>(q.d_fk);
  d->gval = const_cast<int32_t *>(q.d_val);
  d->ggrp = const_cast<int32_t *>(q.d_grp);
  d->cnt = q.d_cnt;
  d->sum = q.d_sum;
  d->n = q.n;
  d->gen_seed = q.gen_seed;
  d->fk_range = q.fk_range;
  d->groups = q.groups;
  return PJ_OK;
}

static uint32_t tile_rows_of(pj_ctx *ctx, const pj_batch_cfg *cfg) {
  if (cfg && cfg->tile_rows) return cfg->tile_rows;
  const uint32_t t = batches(ctx).tuned.tile_rows;
  return t ? t : 65536u;
}
static uint32_t grid_of(pj_ctx *ctx, const pj_batch_cfg *cfg) {
  int per = cfg && cfg->wg_per_cu > 0 ? cfg->wg_per_cu : batches(ctx).tuned.wg_per_cu;
  if (per <= 0) per = 2;
  return (uint32_t)(ctx->num_cus * per);
}
// Largest dimension probed from LDS (0: none).
static uint32_t lds_dim_of(pj_ctx *ctx, const pj_batch_cfg *cfg) {
  int v = cfg && cfg->lds_dim ? cfg->lds_dim : batches(ctx).tuned.lds_dim;
  if (v == 0) return kLdsDim;
  if (v < 0) return 0;
  return std::min<uint32_t>((uint32_t)v, kLdsDim);
}

// Writes header + descriptors into a host buffer.  Queries whose dimension
// fits LDS come first (their items run in the LDS-table kernel); every query
// owns at least one item (its first item clears its outputs).
// Staging geometry of a query with more groups than LDS holds (parts == 0:
// not staged): partitions of kStageGroups groups, runs sized for uniform
// group keys plus slack (a fuller run spills to atomics).
struct StageGeo {
  uint32_t parts = 0;
  uint64_t cap = 0;
  uint64_t bytes() const { return parts ? parts * cap * 8 + 256 : 0; }
};
static StageGeo stage_geo(const pj_query &q) {
  StageGeo g;
  if (q.groups <= kLdsGroups || q.groups > kStageMaxParts * kStageGroups || !q.n) return g;
  g.parts = (q.groups + kStageGroups - 1) / kStageGroups;
  g.cap = (ceil_div(q.n, g.parts) + 4096 + 3) & ~3ULL;
  return g;
}
// Staging for the queries that want it, from the pool; a query whose
// staging cannot be allocated aggregates with atomics instead.
static void stage_alloc(pj_ctx *ctx, const pj_query *qs, uint32_t nq, std::vector<void *> &ptr,
                        std::vector<void *> &owned) {
  ptr.assign(nq, nullptr);
  for (uint32_t i = 0; i < nq; ++i) {
    const StageGeo g = stage_geo(qs[i]);
    if (!g.parts) continue;
    void *p = nullptr;
    if (dalloc(ctx, g.bytes(), &p) == PJ_OK) {
      ptr[i] = p;
      owned.push_back(p);
    }
  }
}
static void stage_free(pj_ctx *ctx, std::vector<void *> &owned) {
  for (void *p : owned) (void)dfree(ctx, p);
  owned.clear();
}

static int pack(const pj_query *qs, uint32_t nq, uint32_t T, uint32_t lds_max, void *h_buf,
                const std::vector<void *> *stage) {
  BatchHdr *hdr = (BatchHdr *)h_buf;
  QDesc *d = (QDesc *)((char *)h_buf + 256);
  uint64_t tiles = 0, lds_tiles = 0, aggs = 0;
  uint32_t j = 0;
  for (int pass = 0; pass < 2; ++pass)
    for (uint32_t i = 0; i < nq; ++i) {
      if (!qs[i].dim) return fail(PJ_E_ARG, "query without a dimension");
      const bool lds = lds_max && qs[i].dim->table && qs[i].dim->n <= lds_max;
      if (lds != (pass == 0)) continue;
      PJ_TRY(fill_desc(qs[i], &d[j]));
      d[j].tile0 = tiles;
      tiles += std::max<uint64_t>(1, ceil_div(qs[i].n, T));
      if (lds) lds_tiles = tiles;
      // staging only on the HBM-table path (the LDS-table kernel keeps atomics)
      void *sp = (stage && !lds) ? (*stage)[i] : nullptr;
      const StageGeo g = stage_geo(qs[i]);
      d[j].agg0 = aggs;
      if (sp && g.parts) {
        d[j].stage = (unsigned long long *)sp;
        d[j].parts = g.parts;
        d[j].cap = g.cap;
        d[j].pcur = (uint32_t *)((char *)sp + g.parts * g.cap * 8);
        d[j].cpp = (uint32_t)ceil_div(g.cap, kAggChunk);
        aggs += (uint64_t)g.parts * d[j].cpp;
      } else {
        d[j].stage = nullptr;
        d[j].pcur = nullptr;
        d[j].parts = 0;
        d[j].cap = 0;
        d[j].cpp = 0;
      }
      ++j;
    }
  hdr->nq = nq;
  hdr->tile_rows = T;
  hdr->total_tiles = tiles;
  hdr->lds_tiles = lds_tiles;
  hdr->gen_ctr = hdr->probe_ctr_lds = hdr->probe_ctr_hbm = hdr->bad_group = 0;
  hdr->total_agg = aggs;
  hdr->agg_ctr = 0;
  return PJ_OK;
}

// The batch sequence on `s` (graph-capturable: kernels and one copy).
static int enqueue(pj_ctx *ctx, hipStream_t s, void *h_buf, void *d_buf, uint32_t grid,
                   size_t bytes) {
  PJ_HIP(hipMemcpyAsync(d_buf, h_buf, bytes, hipMemcpyHostToDevice, s));
  BatchHdr *hdr = (BatchHdr *)d_buf;
  const QDesc *d = (const QDesc *)((char *)d_buf + 256);
  hipLaunchKernelGGL(k_batch_gen, dim3(grid * 2), dim3(256), 0, s, hdr, d);
  hipLaunchKernelGGL(k_batch_probe<true>, dim3(grid), dim3(kBatchBlock), 0, s, hdr, d);
  hipLaunchKernelGGL(k_batch_probe<false>, dim3(grid * 2), dim3(kBatchBlock), 0, s, hdr, d);
  hipLaunchKernelGGL(k_batch_agg, dim3(grid * 2), dim3(kBatchBlock), 0, s, hdr, d);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(PJ_E_HIP, "batch launch: %s", hipGetErrorString(e));
  return PJ_OK;
}

static int slot_init(pj_ctx *ctx, Slot &sl) {
  PJ_HIP(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
  PJ_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
  PJ_HIP(hipHostMalloc(&sl.h_buf, kBufBytes, hipHostMallocDefault));
  PJ_TRY(dalloc(ctx, kBufBytes, &sl.d_buf));
  return PJ_OK;
}

static void slot_free(pj_ctx *ctx, Slot &sl) {
  if (sl.stream) (void)hipStreamSynchronize(sl.stream);
  if (sl.exec) (void)hipGraphExecDestroy(sl.exec);
  if (sl.graph) (void)hipGraphDestroy(sl.graph);
  if (sl.done) (void)hipEventDestroy(sl.done);
  if (sl.h_buf) (void)hipHostFree(sl.h_buf);
  if (sl.d_buf) (void)dfree(ctx, sl.d_buf);
  stage_free(ctx, sl.stage);
  if (sl.stream) (void)hipStreamDestroy(sl.stream);
  sl = Slot();
}

static int check_bad(const void *h_buf) {
  const BatchHdr *hdr = (const BatchHdr *)h_buf;
  if (hdr->bad_group)
    return fail(PJ_E_ARG, "%u probe rows had a group key outside [0, groups)", hdr->bad_group);
  return PJ_OK;
}

}  // namespace pj

int pj::rel::dim_table(const pj_dim *dim, PipeTable *t) {
  if (!dim) return fail(PJ_E_ARG, "null dimension");
  if (dim->vals) return fail(PJ_E_ARG, "dimension of values, not of build rows");
  t->table = dim->table;
  t->ids = dim->ids;
  t->mask = dim->range ? 0 : dim->slots - 1;
  t->lo = dim->lo;
  t->range = dim->range;
  return PJ_OK;
}

using namespace pj;

extern "C" void pj_batch_release(pj_ctx *ctx) {
  if (!ctx->batch) return;
  Batches *b = (Batches *)ctx->batch;
  for (auto &sl : b->slots) slot_free(ctx, sl);
  delete b;
  ctx->batch = nullptr;
}

// Hash-table slots of a dimension of n keys: open addressing at load <= 1/2;
// tables of mid-sized dimensions (that still fit the L2s) at load <= 1/16,
// which shortens the longest probe sequence a wave waits for.  Dimensions
// probed from LDS keep <= kLdsSlots slots.
static uint64_t dim_hash_slots(uint64_t n) {
  const int lg = ilog2_ceil(n ? n : 1);
  const int spread = (n > kLdsDim && n <= (1u << 20)) ? 4 : 1;
  return std::max<uint64_t>(64, 1ULL << (lg + spread));
}

// Keys whose span is no larger than the hash table they would take (dense
// keys, as primary keys usually are, or a filtered subset of them) -- or, for
// tables of build rows, than twice the build side's rows -- get a direct
// table indexed by key - min: one load per probe, and keys outside the span
// miss without one.  kmin > kmax: no keys.
static void dim_layout(pj_dim *dm, uint64_t keys, uint64_t rows, int64_t kmin, int64_t kmax) {
  const uint64_t hs = dim_hash_slots(keys);
  if (keys && kmin > INT32_MIN) {
    const uint64_t span = (uint64_t)(kmax - kmin) + 1;
    if (span <= std::max<uint64_t>(2 * rows, hs)) {
      dm->lo = (int32_t)kmin;
      dm->range = (uint32_t)span;
    }
  }
  dm->slots = dm->range ? dm->range : hs;
}

static void dim_free(pj_ctx *ctx, pj_dim *dm) {
  if (!dm) return;
  if (dm->table) dfree(ctx, dm->table);
  if (dm->ids) dfree(ctx, dm->ids);
  if (dm->vals) dfree(ctx, dm->vals);
  delete dm;
}

static int dim_flags_error(unsigned int f) {
  if (f & 2) return fail(PJ_E_ARG, "dimension key INT32_MIN is reserved");
  if (f & 1) return fail(PJ_E_ARG, "dimension keys must be unique");
  if (f & 4) return fail(PJ_E_HIP, "dimension key outside its direct table");
  return PJ_OK;
}

// pj_dim_create.  A dense span (as many keys as slots) above the LDS size
// holds just the 4-byte values: every slot is a key, and half the bytes per
// slot put more of a table in each cache level.
static int dim_create(pj_ctx *ctx, const int32_t *d_keys, const int32_t *d_vals, uint64_t n,
                      pj_dim **out) {
  std::lock_guard<std::mutex> g(ctx->mu);
  *out = nullptr;
  if (n && !d_keys) return fail(PJ_E_ARG, "dimension without keys");
  if (n >= (1ULL << 31)) return fail(PJ_E_CAPACITY, "dimension above 2^31 rows");
  pj_dim *dm = new pj_dim();
  dm->n = n;
  int64_t kmin = 1, kmax = 0;
  if (n) {
    int rc = rel::reduce_minmax(ctx, d_keys, rel::I32, n, &kmin, &kmax);
    if (rc != PJ_OK) {
      delete dm;
      return rc;
    }
  }
  dim_layout(dm, n, n, kmin, kmax);
  const bool dense = dm->range && dm->range == n && n > kLdsDim;
  int rc = dense ? dalloc(ctx, dm->slots * 4, (void **)&dm->vals)
                 : dalloc(ctx, dm->slots * 8, (void **)&dm->table);
  if (rc != PJ_OK) {
    delete dm;
    return rc;
  }
  Scratch flags;
  auto fin = [&](int r) {
    if (r != PJ_OK) dim_free(ctx, dm);
    return r;
  };
  if ((rc = flags.alloc(ctx, 4)) != PJ_OK) return fin(rc);
  if ((rc = zero(ctx, flags.p, 4)) != PJ_OK) return fin(rc);
  const dim3 bgrid((uint32_t)std::min<uint64_t>(ceil_div(n, 256), (uint64_t)ctx->num_cus * 16));
  Scratch bits;
  if (dense) {
    const uint64_t words = ceil_div(dm->slots, 32);
    if ((rc = bits.alloc(ctx, words * 4)) != PJ_OK) return fin(rc);
    if ((rc = zero(ctx, bits.p, words * 4)) != PJ_OK) return fin(rc);
    hipLaunchKernelGGL(k_dim_build_dense, bgrid, dim3(256), 0, ctx->stream, d_keys, d_vals, n,
                       dm->vals, bits.as<uint32_t>(), dm->lo, dm->range, flags.as<unsigned int>());
  } else {
    const uint64_t gs = std::min<uint64_t>(ceil_div(dm->slots, 256), (uint64_t)ctx->num_cus * 16);
    hipLaunchKernelGGL(k_dim_fill, dim3(gs), dim3(256), 0, ctx->stream, dm->table, dm->slots);
    if (n)
      hipLaunchKernelGGL(k_dim_build, bgrid, dim3(256), 0, ctx->stream, d_keys, d_vals,
                         (const uint8_t *)nullptr, n, dm->table, dm->range ? 0 : dm->slots - 1,
                         dm->lo, dm->range, flags.as<unsigned int>(), false);
  }
  if (hipGetLastError() != hipSuccess) return fin(fail(PJ_E_HIP, "dimension build launch failed"));
  unsigned int f = 0;
  if ((rc = d2h(ctx, &f, flags.p, 4)) != PJ_OK) return fin(rc);
  if ((rc = dim_flags_error(f)) != PJ_OK) return fin(rc);
  *out = dm;
  return PJ_OK;
}

extern "C" int pj_dim_create(pj_ctx *ctx, const int32_t *d_keys, const int32_t *d_vals, uint64_t n,
                             pj_dim **out) {
  return dim_create(ctx, d_keys, d_vals, n, out);
}

namespace pj {
// Per-workgroup min, max and count of the kept keys (keep may be null):
// part[3 b] = min, part[3 b + 1] = max, part[3 b + 2] = count.
__global__ __launch_bounds__(256) void k_dim_range(const int32_t *__restrict__ k,
                                                   const uint8_t *__restrict__ keep, uint64_t n,
                                                   long long *__restrict__ part) {
  __shared__ int64_t smn[4], smx[4], scn[4];
  int64_t mn = INT64_MAX, mx = INT64_MIN, c = 0;
  // four rows per lane per step, every load issued before any is used
  // (indices clamped to row 0, lanes past the end masked out)
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t b = (uint64_t)blockIdx.x * 256 + threadIdx.x; b < n; b += 4 * stride) {
    int32_t v[4];
    bool on[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint64_t i = b + j * stride;
      on[j] = i < n;
      v[j] = k[on[j] ? i : 0];
      const bool kp = keep ? keep[on[j] ? i : 0] != 0 : true;
      on[j] = on[j] && kp;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (on[j]) {
        mn = v[j] < mn ? v[j] : mn;
        mx = v[j] > mx ? v[j] : mx;
        ++c;
      }
  }
  mn = wave_min_i64(mn);
  mx = wave_max_i64(mx);
  c = wave_sum_i64(c);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    smn[w] = mn;
    smx[w] = mx;
    scn[w] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int x = 1; x < 4; ++x) {
      mn = smn[x] < mn ? smn[x] : mn;
      mx = smx[x] > mx ? smx[x] : mx;
      c += scn[x];
    }
    part[3 * blockIdx.x] = mn;
    part[3 * blockIdx.x + 1] = mx;
    part[3 * blockIdx.x + 2] = c;
  }
}

// Folds the partials of side i (workgroups first[i] .. first[i + 1] - 1)
// into out[3 i .. 3 i + 2], one wave per side: the read-back stays a few
// bytes per side.
__global__ __launch_bounds__(64) void k_dim_range_fold(const long long *__restrict__ part,
                                                       const uint32_t *__restrict__ first,
                                                       long long *__restrict__ out) {
  const int i = blockIdx.x, lane = threadIdx.x;
  int64_t mn = INT64_MAX, mx = INT64_MIN, c = 0;
  for (uint32_t b = first[i] + lane; b < first[i + 1]; b += 64) {
    mn = part[3 * b] < mn ? part[3 * b] : mn;
    mx = part[3 * b + 1] > mx ? part[3 * b + 1] : mx;
    c += part[3 * b + 2];
  }
  mn = wave_min_i64(mn);
  mx = wave_max_i64(mx);
  c = wave_sum_i64(c);
  if (lane == 0) {
    out[3 * i] = mn;
    out[3 * i + 1] = mx;
    out[3 * i + 2] = c;
  }
}
}  // namespace pj

int pj::rel::dims_create_rows(pj_ctx *ctx, int m, const int32_t *const *d_keys,
                              const uint8_t *const *d_keep, const uint64_t *n, pj_dim **out) {
  std::lock_guard<std::mutex> g(ctx->mu);
  for (int i = 0; i < m; ++i) out[i] = nullptr;
  for (int i = 0; i < m; ++i) {
    if (n[i] && !d_keys[i]) return fail(PJ_E_ARG, "dimension without keys");
    if (n[i] >= (1ULL << 31)) return fail(PJ_E_CAPACITY, "dimension above 2^31 rows");
  }
  // 1. min, max and count of every side's kept keys, one read-back for all
  std::vector<uint32_t> first(m + 1, 0);
  for (int i = 0; i < m; ++i)
    first[i + 1] = first[i] + (uint32_t)(n[i] ? std::min<uint64_t>(ceil_div(n[i], 2048), (uint64_t)ctx->num_cus * 4) : 0);
  std::vector<long long> range(3 * (uint64_t)m);
  for (int i = 0; i < m; ++i) {  // sides of no rows
    range[3 * i] = INT64_MAX;
    range[3 * i + 1] = INT64_MIN;
    range[3 * i + 2] = 0;
  }
  if (first[m]) {
    Scratch dpart, dfirst, dout;
    PJ_TRY(dpart.alloc(ctx, (uint64_t)first[m] * 3 * 8));
    PJ_TRY(dfirst.alloc(ctx, (uint64_t)(m + 1) * 4));
    PJ_TRY(dout.alloc(ctx, (uint64_t)m * 3 * 8));
    PJ_HIP(hipMemcpyAsync(dfirst.p, first.data(), (uint64_t)(m + 1) * 4, hipMemcpyHostToDevice,
                          ctx->stream));
    for (int i = 0; i < m; ++i)
      if (n[i])
        PJ_LAUNCH(ctx, "dim_range", k_dim_range, first[i + 1] - first[i], 256, 0, d_keys[i],
                  d_keep[i], n[i], dpart.as<long long>() + 3 * (uint64_t)first[i]);
    PJ_LAUNCH(ctx, "dim_range_fold", k_dim_range_fold, (uint32_t)m, 64, 0,
              dpart.as<const long long>(), dfirst.as<const uint32_t>(), dout.as<long long>());
    PJ_TRY(d2h(ctx, range.data(), dout.p, (uint64_t)m * 3 * 8));
  }
  // 2. every table, then one read-back of every side's flags
  std::vector<pj_dim *> dms(m, nullptr);
  auto drop = [&](int r) {
    for (pj_dim *d : dms) dim_free(ctx, d);
    return r;
  };
  Scratch flags;
  int rc = flags.alloc(ctx, (uint64_t)m * 4);
  if (rc == PJ_OK) rc = zero(ctx, flags.p, (uint64_t)m * 4);
  if (rc != PJ_OK) return rc;
  for (int i = 0; i < m; ++i) {
    const int64_t kmin = range[3 * i], kmax = range[3 * i + 1], cnt = range[3 * i + 2];
    pj_dim *dm = dms[i] = new pj_dim();
    dm->n = n[i];
    dim_layout(dm, (uint64_t)cnt, n[i], kmin, kmax);
    unsigned int *fl = flags.as<unsigned int>() + i;
    const dim3 bgrid((uint32_t)std::min<uint64_t>(ceil_div(n[i], 256), (uint64_t)ctx->num_cus * 16));
    if (dm->range) {  // 4-byte build rows, kNoRow where the span has no key
      if ((rc = dalloc(ctx, dm->slots * 4, (void **)&dm->ids)) != PJ_OK) return drop(rc);
      if ((rc = fill_u32(ctx, dm->ids, kNoRow, dm->slots)) != PJ_OK) return drop(rc);
      if (cnt)
        hipLaunchKernelGGL(k_dim_build_ids, bgrid, dim3(256), 0, ctx->stream, d_keys[i],
                           d_keep[i], n[i], dm->ids, dm->lo, dm->range, fl);
    } else {
      if ((rc = dalloc(ctx, dm->slots * 8, (void **)&dm->table)) != PJ_OK) return drop(rc);
      const uint64_t gs = std::min<uint64_t>(ceil_div(dm->slots, 256), (uint64_t)ctx->num_cus * 16);
      hipLaunchKernelGGL(k_dim_fill, dim3(gs), dim3(256), 0, ctx->stream, dm->table, dm->slots);
      if (cnt)
        hipLaunchKernelGGL(k_dim_build, bgrid, dim3(256), 0, ctx->stream, d_keys[i],
                           (const int32_t *)nullptr, d_keep[i], n[i], dm->table, dm->slots - 1,
                           0, 0u, fl, true);
    }
    if (hipGetLastError() != hipSuccess) return drop(fail(PJ_E_HIP, "dimension build launch failed"));
  }
  std::vector<unsigned int> f(m);
  if ((rc = d2h(ctx, f.data(), flags.p, (uint64_t)m * 4)) != PJ_OK) return drop(rc);
  for (int i = 0; i < m; ++i)
    if ((rc = dim_flags_error(f[i])) != PJ_OK) return drop(rc);
  for (int i = 0; i < m; ++i) out[i] = dms[i];
  return PJ_OK;
}

extern "C" int pj_dim_destroy(pj_ctx *ctx, pj_dim *dim) {
  if (!dim) return PJ_OK;
  std::lock_guard<std::mutex> g(ctx->mu);
  PJ_TRY(sync(ctx));
  dim_free(ctx, dim);
  return PJ_OK;
}

extern "C" int pj_dim_info(pj_ctx *ctx, const pj_dim *dim, uint64_t *rows, uint64_t *table_bytes) {
  (void)ctx;
  if (!dim) return fail(PJ_E_ARG, "null dimension");
  if (rows) *rows = dim->n;
  if (table_bytes) *table_bytes = dim->slots * (dim->table ? 8 : 4);
  return PJ_OK;
}

extern "C" int pj_query_run(pj_ctx *ctx, const pj_query *q, const pj_batch_cfg *cfg) {
  std::lock_guard<std::mutex> g(ctx->mu);
  const uint32_t T = tile_rows_of(ctx, cfg);
  char *h = (char *)ctx->pinned_small;
  static_assert(256 + sizeof(QDesc) <= (1 << 20), "pinned staging too small");
  std::vector<void *> sp, owned;
  stage_alloc(ctx, q, 1, sp, owned);
  int rc = pack(q, 1, T, lds_dim_of(ctx, cfg), h, &sp);
  Scratch dbuf;
  if (rc == PJ_OK) rc = dbuf.alloc(ctx, 256 + sizeof(QDesc));
  if (rc == PJ_OK) rc = enqueue(ctx, ctx->stream, h, dbuf.p, grid_of(ctx, cfg), 256 + sizeof(QDesc));
  if (rc == PJ_OK &&
      hipMemcpyAsync(h, dbuf.p, sizeof(BatchHdr), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
    rc = fail(PJ_E_HIP, "query header copy failed");
  const int rs = sync(ctx);
  stage_free(ctx, owned);
  PJ_TRY(rc);
  PJ_TRY(rs);
  return check_bad(h);
}

extern "C" int pj_batch_slots(pj_ctx *ctx, int n) {
  std::lock_guard<std::mutex> g(ctx->mu);
  if (n < 1 || n > 8) return fail(PJ_E_ARG, "batch slots %d not in [1, 8]", n);
  Batches &b = batches(ctx);
  for (auto &sl : b.slots) slot_free(ctx, sl);
  b.slots.assign(n, Slot());
  for (auto &sl : b.slots) PJ_TRY(slot_init(ctx, sl));
  return PJ_OK;
}

extern "C" int pj_batch_launch(pj_ctx *ctx, int slot, const pj_query *qs, uint32_t nq,
                               const pj_batch_cfg *cfg) {
  std::lock_guard<std::mutex> g(ctx->mu);
  Batches &b = batches(ctx);
  if (slot < 0 || slot >= (int)b.slots.size()) return fail(PJ_E_ARG, "batch slot %d not created", slot);
  if (nq < 1 || nq > kMaxBatch) return fail(PJ_E_ARG, "batch of %u queries not in [1, %u]", nq, kMaxBatch);
  Slot &sl = b.slots[slot];
  if (sl.busy) return fail(PJ_E_ARG, "batch slot %d is busy", slot);
  const uint32_t T = tile_rows_of(ctx, cfg);
  stage_free(ctx, sl.stage);  // left by a launch that failed (the slot is idle)
  std::vector<void *> sp;
  stage_alloc(ctx, qs, nq, sp, sl.stage);
  {
    const int rc = pack(qs, nq, T, lds_dim_of(ctx, cfg), sl.h_buf, &sp);
    if (rc != PJ_OK) {
      stage_free(ctx, sl.stage);
      return rc;
    }
  }
  const uint32_t grid = grid_of(ctx, cfg);
  if (!sl.exec || sl.graph_grid != grid) {
    if (sl.exec) (void)hipGraphExecDestroy(sl.exec);
    if (sl.graph) (void)hipGraphDestroy(sl.graph);
    sl.exec = nullptr;
    sl.graph = nullptr;
    PJ_HIP(hipStreamBeginCapture(sl.stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue(ctx, sl.stream, sl.h_buf, sl.d_buf, grid, kBufBytes);
    hipGraph_t gr = nullptr;
    hipError_t e = hipStreamEndCapture(sl.stream, &gr);
    if (rc != PJ_OK) {
      if (gr) (void)hipGraphDestroy(gr);
      return rc;
    }
    if (e != hipSuccess) return fail(PJ_E_HIP, "batch graph capture: %s", hipGetErrorString(e));
    sl.graph = gr;
    PJ_HIP(hipGraphInstantiate(&sl.exec, sl.graph, nullptr, nullptr, 0));
    sl.graph_grid = grid;
  }
  PJ_HIP(hipGraphLaunch(sl.exec, sl.stream));
  PJ_HIP(hipMemcpyAsync(sl.h_buf, sl.d_buf, sizeof(BatchHdr), hipMemcpyDeviceToHost, sl.stream));
  PJ_HIP(hipEventRecord(sl.done, sl.stream));
  sl.busy = true;
  return PJ_OK;
}

extern "C" int pj_batch_poll(pj_ctx *ctx, int slot, int *done) {
  std::lock_guard<std::mutex> g(ctx->mu);
  Batches &b = batches(ctx);
  if (slot < 0 || slot >= (int)b.slots.size()) return fail(PJ_E_ARG, "batch slot %d not created", slot);
  Slot &sl = b.slots[slot];
  *done = 1;
  if (!sl.busy) return PJ_OK;
  hipError_t e = hipEventQuery(sl.done);
  if (e == hipErrorNotReady) {
    *done = 0;
    return PJ_OK;
  }
  if (e != hipSuccess) return fail(PJ_E_HIP, "batch: %s", hipGetErrorString(e));
  sl.busy = false;
  stage_free(ctx, sl.stage);
  return check_bad(sl.h_buf);
}

extern "C" int pj_batch_wait(pj_ctx *ctx, int slot) {
  std::lock_guard<std::mutex> g(ctx->mu);
  Batches &b = batches(ctx);
  if (slot < 0 || slot >= (int)b.slots.size()) return fail(PJ_E_ARG, "batch slot %d not created", slot);
  Slot &sl = b.slots[slot];
  if (!sl.busy) return PJ_OK;
  PJ_HIP(hipEventSynchronize(sl.done));
  sl.busy = false;
  stage_free(ctx, sl.stage);
  return check_bad(sl.h_buf);
}

// Autotuner: each candidate runs the whole batch on the context stream (no
// graph) twice and keeps the second time; knobs are tuned one after the
// other (tile rows, then workgroups per CU, then the LDS threshold).
extern "C" int pj_batch_autotune(pj_ctx *ctx, const pj_query *qs, uint32_t nq,
                                 pj_batch_cfg *best) {
  std::lock_guard<std::mutex> g(ctx->mu);
  if (nq < 1 || nq > kMaxBatch) return fail(PJ_E_ARG, "batch of %u queries not in [1, %u]", nq, kMaxBatch);
  for (uint32_t i = 0; i < nq; ++i)
    if (qs[i].n && (!qs[i].d_fk || !qs[i].d_val || !qs[i].d_grp))
      return fail(PJ_E_ARG, "autotune query %u without probe columns", i);
  static_assert(kBufBytes <= (1 << 20), "pinned staging too small");
  char *h = (char *)ctx->pinned_small;
  Scratch dbuf;
  PJ_TRY(dbuf.alloc(ctx, kBufBytes));
  hipEvent_t ev[2];
  PJ_HIP(hipEventCreate(&ev[0]));
  PJ_HIP(hipEventCreate(&ev[1]));
  std::vector<void *> sp, owned;
  stage_alloc(ctx, qs, nq, sp, owned);
  auto time_cfg = [&](const pj_batch_cfg &c, float *ms) -> int {
    const size_t bytes = 256 + sizeof(QDesc) * nq;
    for (int rep = 0; rep < 2; ++rep) {
      PJ_TRY(pack(qs, nq, tile_rows_of(ctx, &c), lds_dim_of(ctx, &c), h, &sp));
      PJ_HIP(hipEventRecord(ev[0], ctx->stream));
      PJ_TRY(enqueue(ctx, ctx->stream, h, dbuf.p, grid_of(ctx, &c), bytes));
      PJ_HIP(hipEventRecord(ev[1], ctx->stream));
      PJ_HIP(hipEventSynchronize(ev[1]));
      PJ_HIP(hipEventElapsedTime(ms, ev[0], ev[1]));
    }
    return PJ_OK;
  };
  pj_batch_cfg cur{65536u, 2, 0};
  float best_ms = 0.f;
  int rc = time_cfg(cur, &best_ms);
  auto try_knob = [&](pj_batch_cfg c) {
    float ms = 0.f;
    if (rc == PJ_OK) rc = time_cfg(c, &ms);
    if (rc == PJ_OK && ms < best_ms) {
      best_ms = ms;
      cur = c;
    }
  };
  for (uint32_t t : {16384u, 32768u, 131072u, 262144u}) {
    pj_batch_cfg c = cur;
    c.tile_rows = t;
    try_knob(c);
  }
  const pj_batch_cfg after_t = cur;
  for (int w : {1, 4}) {
    pj_batch_cfg c = after_t;
    c.wg_per_cu = w;
    try_knob(c);
  }
  const pj_batch_cfg after_w = cur;
  for (int l : {-1, 512}) {
    pj_batch_cfg c = after_w;
    c.lds_dim = l;
    try_knob(c);
  }
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  (void)sync(ctx);
  stage_free(ctx, owned);
  PJ_TRY(rc);
  PJ_TRY(d2h(ctx, h, dbuf.p, sizeof(BatchHdr)));  // the last run's group check
  PJ_TRY(check_bad(h));
  batches(ctx).tuned = cur;
  if (best) *best = cur;
  return PJ_OK;
}
