// partitioned-hashjoin-chained and hash-rearrange on one MI355X.
//
// Reference: HashPartitioner + GpuPartitionedHashJoinChained
// (core/olap/lib/operators/gpu/gpu-partitioned-hash-join-chained.cpp:47-404
//  and :612-1586, kernels in gpu-partitioned-hash-join-chained-helpers.cpp),
// GpuHashRearrange (core/olap/lib/operators/gpu/gpu-hash-rearrange.cpp).
//
// Flow of one join (both sides go through the same steps):
//   1. k_chunk_hist + column scans: per-chunk joint histograms of the
//      B = bits1 + bits2 top bits of fmix32(key) -- one read of the key
//      column -- give every (chunk, partition) its exact output offset, and
//      k_pass2_items sizes the pass-2 work items from the digit totals;
//   2. k_radix_pass pass 1 (fan-out 2^bits1, keys replaced by fmix32(key))
//      and pass 2 (2^bits2 inside every pass-1 region): (key, value) pairs,
//      LDS write-combined full lines, no global atomics;
//   3. k_join_units: join work units (<= probe_unit probe tuples each, the
//      reference's decompose_chains idea, helpers.cpp:199-232), largest
//      first;
//   4. k_pjoin_bitmap: per unit, the build partition goes into an LDS
//      presence-bitmap table and the unit's probe tuples are looked up in it
//      (one or two LDS reads, no key compare); units with duplicate build
//      keys or oversized partitions are deferred to k_pjoin, the
//      bucket-sorted LDS table, which takes any keys and processes
//      partitions larger than its capacity in rounds.  Matches fold into the
//      join accumulator (the parent reduce).
// Build sides small enough for one LDS table skip steps 1-2 (B = 0).
#include <algorithm>
#include <cstdio>
#include <vector>

#include "radix.hpp"
#include "rel_api.hpp"

namespace pj {

__global__ __launch_bounds__(256) void k_colsum(const uint32_t *__restrict__ H, int G, int nb,
                                                int slab, uint32_t *__restrict__ P) {
  const int d = blockIdx.x * 256 + threadIdx.x;
  if (d >= nb) return;
  const int c0 = blockIdx.y * slab, c1 = min(c0 + slab, G);
  uint32_t s = 0;
  for (int c = c0; c < c1; c += 8) {  // 8 independent loads in flight
    uint32_t x[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) x[j] = c + j < c1 ? H[(uint64_t)(c + j) * nb + d] : 0u;
#pragma unroll
    for (int j = 0; j < 8; ++j) s += x[j];
  }
  P[(uint64_t)blockIdx.y * nb + d] = s;
}

__global__ __launch_bounds__(256) void k_colscan(const uint32_t *__restrict__ H, int G, int nb,
                                                 int slab, const uint32_t *__restrict__ P,
                                                 uint32_t *__restrict__ E,
                                                 uint32_t *__restrict__ T) {
  const int d = blockIdx.x * 256 + threadIdx.x;
  if (d >= nb) return;
  const int y = blockIdx.y;
  const int c0 = y * slab, c1 = min(c0 + slab, G);
  uint32_t run = 0;
  for (int q = 0; q < y; q += 8) {
    uint32_t x[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) x[j] = q + j < y ? P[(uint64_t)(q + j) * nb + d] : 0u;
#pragma unroll
    for (int j = 0; j < 8; ++j) run += x[j];
  }
  for (int c = c0; c < c1; c += 8) {
    uint32_t x[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) x[j] = c + j < c1 ? H[(uint64_t)(c + j) * nb + d] : 0u;
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (c + j < c1) {
        E[(uint64_t)(c + j) * nb + d] = run;
        run += x[j];
      }
  }
  if (c1 == G) T[d] = run;
}

__global__ __launch_bounds__(kScanBlock) void k_excl_scan(const uint32_t *__restrict__ T, int nb,
                                                          unsigned long long *__restrict__ S,
                                                          uint32_t *__restrict__ ctr) {
  __shared__ unsigned long long tmp[kScanBlock / 64];
  const int tid = threadIdx.x;
  if (tid == 0) *ctr = 0;
  const int per = (nb + kScanBlock - 1) / kScanBlock;
  const int lo = min(tid * per, nb), hi = min(lo + per, nb);
  unsigned long long sum = 0;
  for (int i = lo; i < hi; i += 8) {
    uint32_t x[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) x[j] = i + j < hi ? T[i + j] : 0u;
#pragma unroll
    for (int j = 0; j < 8; ++j) sum += x[j];
  }
  unsigned long long tot;
  unsigned long long run = block_excl_scan<kScanBlock, unsigned long long>(sum, tmp, &tot);
  for (int i = lo; i < hi; i += 8) {
    uint32_t x[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) x[j] = i + j < hi ? T[i + j] : 0u;
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (i + j < hi) {
        S[i + j] = run;
        run += x[j];
      }
  }
  if (tid == 0) S[nb] = tot;
}

__global__ __launch_bounds__(kScanBlock) void k_pass2_items(const uint32_t *__restrict__ T,
                                                            int b1, int b2, int G,
                                                            unsigned long long target,
                                                            uint32_t *__restrict__ items) {
  __shared__ uint32_t tmp[kScanBlock / 64];
  const uint32_t F1 = 1u << b1, F2 = 1u << b2;
  const uint32_t d1 = threadIdx.x;
  unsigned long long t1 = 0;
  if (d1 < F1) {
    const uint32_t *row = T + ((uint64_t)d1 << b2);
    for (uint32_t d2 = 0; d2 < F2; d2 += 8) {
      uint32_t x[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) x[j] = d2 + j < F2 ? row[d2 + j] : 0u;
#pragma unroll
      for (int j = 0; j < 8; ++j) t1 += x[j];
    }
  }
  unsigned long long cpg = (unsigned long long)G;
  if (t1) {
    cpg = target * (unsigned long long)G / t1;
    cpg = cpg < 1 ? 1 : (cpg > (unsigned long long)G ? (unsigned long long)G : cpg);
  }
  const uint32_t ngrp = d1 < F1 ? (uint32_t)(((unsigned long long)G + cpg - 1) / cpg) : 0u;
  uint32_t tot;
  const uint32_t first = block_excl_scan<kScanBlock>(ngrp, tmp, &tot);
  if (d1 < F1) {
    items[d1] = first;
    items[F1 + 1 + d1] = (uint32_t)cpg;
  }
  if (d1 == 0) items[F1] = tot;
}

int colscan_slabs(const pj_ctx *ctx, int G, int nb, int *slab) {
  // about 2048 threads per CU over (bin, slab) pairs, slabs of >= 8 chunks
  const uint64_t want = (uint64_t)ctx->num_cus * 2048;
  uint64_t ns = want / (uint64_t)(nb > 0 ? nb : 1);
  if (ns < 1) ns = 1;
  int sl = (int)ceil_div((uint64_t)(G > 0 ? G : 1), ns);
  if (sl < 8) sl = 8;
  *slab = sl;
  return (int)ceil_div((uint64_t)(G > 0 ? G : 1), (uint64_t)sl);
}

RadixPlan radix_make_plan(pj_ctx *ctx, uint64_t n, int b1, int b2, int nb, int per_cu) {
  RadixPlan p;
  p.b1 = b1;
  p.b2 = b2;
  p.nb = nb;
  uint64_t g = (uint64_t)ctx->num_cus * (per_cu > 0 ? per_cu : 1);
  uint64_t min_chunk = 16384;
  uint64_t gmax = ceil_div(n ? n : 1, min_chunk);
  if (g > gmax) g = gmax;
  if (g < 1) g = 1;
  p.chunk = (ceil_div(n ? n : 1, g) + 3) & ~3ULL;
  p.G = (int)ceil_div(n ? n : 1, p.chunk);
  return p;
}

struct JoinUnit {
  unsigned long long r_begin, s_begin;
  uint32_t r_len, s_len;
};

// Unit list header (device): hdr[0] = units, hdr[1] = claim counter of the
// bitmap kernel, hdr[2] = units it deferred to the CSR kernel, hdr[3] =
// claim counter of the CSR kernel.
constexpr int kUnitHdrWords = 4;

constexpr int ilog2_const(int x) { return x <= 1 ? 0 : 1 + ilog2_const(x >> 1); }

// Probe tuples [s0, s1) of sk / sv: the 16-byte-aligned interior as int4
// vectors (4 tuples per thread per step), the <= 3 tuples before and after
// it one by one.  Three 4-tuple steps rotate through registers A, B, C: a
// step's loads are issued two steps before it is probed.  The loads are
// unconditional (indices clamped to the last vector, results past the end
// unused): a load under a branch, or a register copy of a loaded value,
// makes the compiler wait for outstanding loads there, which would serialise
// the prefetch with the probe.  Without probe values the key column stands
// in for them (probe4 ignores them).
template <int BLOCK, class P4, class P1>
__device__ __forceinline__ void probe_range(const int32_t *__restrict__ sk,
                                            const int32_t *__restrict__ sv,
                                            unsigned long long s0, unsigned long long s1,
                                            P4 &&probe4, P1 &&probe1) {
  const int tid = threadIdx.x;
  const unsigned long long a0 = (s0 + 3) & ~3ULL, a1 = s1 & ~3ULL;
  if (a0 >= a1) {  // short unit: every tuple on its own
    for (unsigned long long x = s0 + tid; x < s1; x += BLOCK) probe1(sk[x], sv ? sv[x] : 0);
    return;
  }
  const int tail_lo = (int)(a0 - s0), tail_hi = (int)(s1 - a1);
  if (tid < tail_lo + tail_hi) {
    const unsigned long long x = tid < tail_lo ? s0 + tid : a1 + (tid - tail_lo);
    probe1(sk[x], sv ? sv[x] : 0);
  }
  const int4 *kq = reinterpret_cast<const int4 *>(sk + a0);
  const int4 *vq = reinterpret_cast<const int4 *>((sv ? sv : sk) + a0);
  const uint32_t nvec = (uint32_t)((a1 - a0) >> 2);  // >= 1 here
  auto cl = [nvec](uint32_t x) { return x < nvec ? x : nvec - 1; };
  uint32_t q = tid;
  int4 kA = ld_nt4(kq + cl(q)), vA = ld_nt4(vq + cl(q));
  int4 kB = ld_nt4(kq + cl(q + BLOCK)), vB = ld_nt4(vq + cl(q + BLOCK));
  // whole triples of steps: a straight-line body, so the compiler's waits
  // stay precise across the back edge
  for (; q + 2 * BLOCK < nvec; q += 3 * BLOCK) {
    const int4 kC = ld_nt4(kq + q + 2 * BLOCK), vC = ld_nt4(vq + q + 2 * BLOCK);
    probe4(kA, vA);
    kA = ld_nt4(kq + cl(q + 3 * BLOCK));
    vA = ld_nt4(vq + cl(q + 3 * BLOCK));
    probe4(kB, vB);
    kB = ld_nt4(kq + cl(q + 4 * BLOCK));
    vB = ld_nt4(vq + cl(q + 4 * BLOCK));
    probe4(kC, vC);
  }
  if (q < nvec) probe4(kA, vA);
  if (q + BLOCK < nvec) probe4(kB, vB);
}

// ---------------------------------------------------------------------------
// Presence-bitmap join (unique build keys, hashed keys).
//
// The radix passes store h = fmix32(key), a bijection, and every key of a
// final partition has the same top B bits.  With LOGH bucket bits taken from
// the bottom of h and the RB = 32 - B - LOGH bits above them as the
// "residue", (bucket, residue) determines the key.  So each bucket holds a
// bitmap of the residues present in it and the index of its first value in
// val[]; values are stored in residue order.  A probe reads the bucket: the
// key is present iff bit r is set, and its value is
// val[start + popcount(bitmap below r)].  No key is stored or compared and
// no bucket is walked.  Two layouts:
//   narrow (RB <= 4): tab[b] = 16-bit bitmap | start << 16, one LDS read;
//   wide   (RB <= 6): bm[b] = 64-bit bitmap, off[b] = 16-bit start.
//
// A unit whose build partition has more than NCAPB rows, or a residue seen
// twice (a duplicate build key), is appended to the deferred list for the
// CSR kernel.  While a unit is probed, the next unit is claimed and its build
// rows are loaded into registers, so a table is built without waiting for
// HBM.  Workgroups claim units from a counter (largest first); a workgroup
// that claims another unit of the partition it has just built reuses the
// table.
template <int BLOCK, int H, int NCAPB, int WAVES, bool WIDE>
__global__ __launch_bounds__(BLOCK, WAVES) void k_pjoin_bitmap(
    const int32_t *__restrict__ rk, const int32_t *__restrict__ rv,
    const int32_t *__restrict__ sk, const int32_t *__restrict__ sv,
    const JoinUnit *__restrict__ units, uint32_t *__restrict__ hdr,
    JoinUnit *__restrict__ deferred, unsigned long long *__restrict__ acc_out) {
  constexpr int LOGH = ilog2_const(H);
  static_assert((1 << LOGH) == H, "H is a power of two");
  static_assert(NCAPB % BLOCK == 0 && H % BLOCK == 0, "whole rows per thread");
  static_assert(NCAPB < 65536, "value starts are 16-bit");
  constexpr uint32_t RMASK = WIDE ? 63u : 15u;
  constexpr int RPT = NCAPB / BLOCK;  // build rows per thread
  constexpr int BPT = H / BLOCK;      // buckets per thread in the scan
  __shared__ uint32_t tab[WIDE ? 1 : H];
  __shared__ unsigned long long bm[WIDE ? H : 1];
  __shared__ uint16_t off[WIDE ? H : 1];
  // a probe of an absent residue may index one past the last value (read,
  // never used)
  __shared__ int32_t val[NCAPB + 4];
  __shared__ uint32_t scan_tmp[BLOCK / 64];
  __shared__ uint32_t s_unit, s_dup;
  const uint32_t nunits = hdr[0];
  const int tid = threadIdx.x;
  auto residue = [](int32_t h) -> uint32_t { return ((uint32_t)h >> LOGH) & RMASK; };
  // (present, value index) of hashed key h in the built table
  auto lookup = [&](int32_t h, uint32_t *idx) -> bool {
    const uint32_t b = (uint32_t)h & (H - 1), r = residue(h);
    if constexpr (WIDE) {
      const unsigned long long m = bm[b];
      *idx = (uint32_t)off[b] + (uint32_t)__popcll(m & ((1ULL << r) - 1ULL));
      return (m >> r) & 1ULL;
    } else {
      const uint32_t w = tab[b];
      *idx = (w >> 16) + (uint32_t)__popc(w & ((1u << r) - 1u));
      return (w >> r) & 1u;
    }
  };
  JoinAcc acc;
  acc.init();
  int32_t bk[RPT], bv[RPT];
  auto fetch = [&](const JoinUnit &X) {
    const int32_t *rkp = rk + X.r_begin;
    const int32_t *rvp = rv ? rv + X.r_begin : nullptr;
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const uint32_t i = (uint32_t)(q * BLOCK + tid);
      bk[q] = i < X.r_len ? rkp[i] : 0;
      bv[q] = (i < X.r_len && rvp) ? rvp[i] : 0;
    }
  };
  if (tid == 0) s_unit = atomicAdd(&hdr[1], 1u);
  __syncthreads();
  uint32_t u = s_unit;
  JoinUnit U{};
  if (u < nunits) {
    U = units[u];
    if (U.r_len <= (uint32_t)NCAPB) fetch(U);
  }
  unsigned long long built_r = ~0ULL;
  while (u < nunits) {
    bool ok = U.r_len <= (uint32_t)NCAPB;
    if (ok && U.r_begin != built_r) {
      __syncthreads();  // the previous table is no longer probed
      for (int i = tid; i < H; i += BLOCK) {
        if constexpr (WIDE)
          bm[i] = 0ULL;
        else
          tab[i] = 0u;
      }
      if (tid == 0) s_dup = 0;
      __syncthreads();
      bool dup = false;
#pragma unroll
      for (int q = 0; q < RPT; ++q)
        if ((uint32_t)(q * BLOCK + tid) < U.r_len) {
          const uint32_t b = (uint32_t)bk[q] & (H - 1), r = residue(bk[q]);
          if constexpr (WIDE) {
            const unsigned long long bit = 1ULL << r;
            dup |= (atomicOr(&bm[b], bit) & bit) != 0ULL;
          } else {
            const uint32_t bit = 1u << r;
            dup |= (atomicOr(&tab[b], bit) & bit) != 0u;
          }
        }
      if (dup) s_dup = 1;
      __syncthreads();
      if (s_dup) {
        ok = false;
        built_r = ~0ULL;
      } else {
        // value starts: exclusive scan of the bitmaps' popcounts
        uint32_t c[BPT], sum = 0;
#pragma unroll
        for (int x = 0; x < BPT; ++x) {
          if constexpr (WIDE)
            c[x] = (uint32_t)__popcll(bm[tid * BPT + x]);
          else
            c[x] = (uint32_t)__popc(tab[tid * BPT + x]);
          sum += c[x];
        }
        uint32_t tot;
        uint32_t run = block_excl_scan<BLOCK>(sum, scan_tmp, &tot);
#pragma unroll
        for (int x = 0; x < BPT; ++x) {
          if constexpr (WIDE)
            off[tid * BPT + x] = (uint16_t)run;
          else
            tab[tid * BPT + x] |= run << 16;
          run += c[x];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < RPT; ++q)
          if ((uint32_t)(q * BLOCK + tid) < U.r_len) {
            uint32_t idx;
            (void)lookup(bk[q], &idx);
            val[idx] = bv[q];
          }
        __syncthreads();
        built_r = U.r_begin;
      }
    }
    if (!ok && tid == 0) deferred[atomicAdd(&hdr[2], 1u)] = U;
    // claim the next unit and start loading its build rows
    if (tid == 0) s_unit = atomicAdd(&hdr[1], 1u);
    __syncthreads();
    const uint32_t un = s_unit;
    __syncthreads();  // s_unit is written again only after every lane read it
    JoinUnit Un{};
    if (un < nunits) {
      Un = units[un];
      if (Un.r_begin != built_r && Un.r_len <= (uint32_t)NCAPB) fetch(Un);
    }
    if (ok) {
      auto probe1 = [&](int32_t key, int32_t pv) {
        uint32_t idx;
        if (lookup(key, &idx)) acc.add1(val[idx], pv);
      };
      auto probe4 = [&](const int4 k4, const int4 v4) {
        const int32_t kk[4] = {k4.x, k4.y, k4.z, k4.w};
        const int32_t vv[4] = {sv ? v4.x : 0, sv ? v4.y : 0, sv ? v4.z : 0, sv ? v4.w : 0};
        bool hit[4];
        uint32_t idx[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) hit[c] = lookup(kk[c], &idx[c]);
        // in bounds also for an absent residue; the value is then unused
        int32_t b[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) b[c] = val[idx[c]];
        acc.add4_1(hit, b, vv);
      };
      probe_range<BLOCK>(sk, sv, U.s_begin, U.s_begin + U.s_len, probe4, probe1);
    }
    u = un;
    U = Un;
  }
  join_acc_flush(acc, acc_out);
}

// ---------------------------------------------------------------------------
// Bucket-sorted ("CSR") join, the general case: any keys, duplicates, build
// partitions of any size.  tab[b] packs the [start, end) range of bucket b in
// ent[] (start in the high, end in the low 16 bits) and entries are (key,
// value) pairs grouped by bucket.  Building it is count -> scan -> scatter
// with LDS atomics (the scatter's cursor is the low half); probing is one
// LDS read for the range plus the bucket's entries, which sit next to each
// other -- the reference's chained LDS table (head / next,
// gpu-partitioned-hash-join-chained.cpp:832-1037) needs one dependent read
// per chain hop instead.  Every entry whose key equals the probe key is a
// match, so duplicate build keys emit every pair.  A build partition larger
// than NCAP is processed in rounds of NCAP rows, each round re-reading the
// unit's probe tuples, so skew never overflows LDS.
//
// HASHED: keys are fmix32(original key) (written by the radix passes); the
// bucket is their low bits and equality of hashes is equality of keys.
// Otherwise keys are raw and hashed here.
// The units are units[0 .. *n_units), claimed with *claim.
// Four waves per SIMD: two 512-thread workgroups per CU, or one of 1024 for
// the largest table.
template <int BLOCK, int H, int NCAP>
constexpr int pjoin_waves_per_simd() {
  return (H * 4 + NCAP * 8 <= 80 * 1024 || BLOCK >= 1024) ? 4 : 2;
}
template <int BLOCK, int H, int NCAP, bool HASHED>
__global__ __launch_bounds__(BLOCK, (pjoin_waves_per_simd<BLOCK, H, NCAP>())) void k_pjoin(
    const int32_t *__restrict__ rk, const int32_t *__restrict__ rv,
    const int32_t *__restrict__ sk, const int32_t *__restrict__ sv,
    const JoinUnit *__restrict__ units, const uint32_t *__restrict__ n_units,
    uint32_t *__restrict__ claim, unsigned long long *__restrict__ acc_out) {
  const uint32_t nunits = *n_units;
  static_assert(NCAP < 65536, "bucket ranges are 16-bit");
  static_assert(NCAP % BLOCK == 0 && H % BLOCK == 0, "whole rows per thread");
  constexpr int RPT = NCAP / BLOCK;  // build rows per thread
  constexpr int BPT = H / BLOCK;     // buckets per thread in the scan
  __shared__ uint32_t tab[H];
  __shared__ int2 ent[NCAP];
  __shared__ uint32_t scan_tmp[BLOCK / 64];
  auto bucket = [](int32_t k) -> uint32_t {
    return (HASHED ? (uint32_t)k : fmix32((uint32_t)k)) & (H - 1);
  };
  const int tid = threadIdx.x;
  JoinAcc acc;
  acc.init();
  __shared__ uint32_t s_unit;
  unsigned long long built_r = ~0ULL;
  for (;;) {
    if (tid == 0) s_unit = atomicAdd(claim, 1u);
    __syncthreads();
    const uint32_t u = s_unit;
    __syncthreads();
    if (u >= nunits) break;
    const JoinUnit U = units[u];
    const uint32_t rounds = (U.r_len + NCAP - 1) / NCAP;
    for (uint32_t round = 0; round < rounds; ++round) {
      const uint32_t rb = round * NCAP;
      const uint32_t rl = (U.r_len - rb) < (uint32_t)NCAP ? (U.r_len - rb) : (uint32_t)NCAP;
      const bool reuse = rounds == 1 && built_r == U.r_begin;
      if (!reuse) {
        __syncthreads();  // the previous table is no longer probed
        for (int i = tid; i < H; i += BLOCK) tab[i] = 0;
        int32_t bk[RPT], bv[RPT];
        uint32_t bb[RPT];
        const int32_t *rkp = rk + U.r_begin + rb;
        const int32_t *rvp = rv ? rv + U.r_begin + rb : nullptr;
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
          const uint32_t i = (uint32_t)(q * BLOCK + tid);
          bk[q] = i < rl ? rkp[i] : 0;
          bv[q] = (i < rl && rvp) ? rvp[i] : 0;
          bb[q] = bucket(bk[q]);
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < RPT; ++q)
          if ((uint32_t)(q * BLOCK + tid) < rl) atomicAdd(&tab[bb[q]], 1u);
        __syncthreads();
        {  // exclusive scan of the bucket counts -> bucket starts
          uint32_t c[BPT], sum = 0;
#pragma unroll
          for (int x = 0; x < BPT; ++x) {
            c[x] = tab[tid * BPT + x];
            sum += c[x];
          }
          uint32_t tot;
          uint32_t run = block_excl_scan<BLOCK>(sum, scan_tmp, &tot);
          // the start in both halves: the scatter's cursor increments the
          // low half, which ends at the bucket's end
#pragma unroll
          for (int x = 0; x < BPT; ++x) {
            tab[tid * BPT + x] = run | (run << 16);
            run += c[x];
          }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < RPT; ++q)
          if ((uint32_t)(q * BLOCK + tid) < rl) {
            const uint32_t pos = atomicAdd(&tab[bb[q]], 1u) & 0xffffu;
            ent[pos] = make_int2(bk[q], bv[q]);
          }
        __syncthreads();
        built_r = rounds == 1 ? U.r_begin : ~0ULL;
      }
      // all entries of a bucket of more than one entry
      auto probe_multi = [&](int32_t key, int32_t pv, uint32_t st, uint32_t en) {
        int64_t m = 0, sb = 0;
        int32_t mnb = INT32_MAX, mxb = INT32_MIN;
        for (uint32_t i = st; i < en; ++i) {
          const int2 kv = ent[i];
          if (kv.x == key) {
            m += 1;
            sb += kv.y;
            mnb = min(mnb, kv.y);
            mxb = max(mxb, kv.y);
          }
        }
        acc.add(m, sb, mnb, mxb, pv);
      };
      auto probe1 = [&](int32_t key, int32_t pv) {
        const uint32_t rg = tab[bucket(key)];
        const uint32_t st = rg >> 16, en = rg & 0xffffu;
        if (st < en) probe_multi(key, pv, st, en);
      };
      auto probe4 = [&](const int4 k4, const int4 v4) {
        const int32_t kk[4] = {k4.x, k4.y, k4.z, k4.w};
        const int32_t vv[4] = {sv ? v4.x : 0, sv ? v4.y : 0, sv ? v4.z : 0, sv ? v4.w : 0};
        uint32_t rg[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) rg[c] = tab[bucket(kk[c])];
        // entries are read only by the lanes whose bucket holds them
        // (exec-masked LDS reads: most buckets hold zero or one entry)
        int2 e0[4], e1[4], e2[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const uint32_t st = rg[c] >> 16, en = rg[c] & 0xffffu;
          e0[c] = e1[c] = e2[c] = make_int2(0, 0);
          if (en > st) e0[c] = ent[st];
          if (en > st + 1) e1[c] = ent[st + 1];
          if (en > st + 2) e2[c] = ent[st + 2];
        }
        // buckets of up to three entries without branches; longer ones
        // finish in probe_multi.  A row with one match among them takes the
        // single-match accumulate; rows matching twice or more the general
        // one.
        bool multi = false, dup = false;
        bool h1m[4];
        int32_t bs[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const uint32_t st = rg[c] >> 16, en = rg[c] & 0xffffu;
          const bool h0 = (en > st) & (e0[c].x == kk[c]);
          const bool h1 = (en > st + 1) & (e1[c].x == kk[c]);
          const bool h2 = (en > st + 2) & (e2[c].x == kk[c]);
          h1m[c] = h0 | h1 | h2;
          // the matching entry's value when there is one match
          bs[c] = (e0[c].y & -(int32_t)h0) | (e1[c].y & -(int32_t)h1) |
                  (e2[c].y & -(int32_t)h2);
          dup |= (h0 & h1) | (h0 & h2) | (h1 & h2);
          multi |= en > st + 3;
        }
        if (!dup) {
          acc.add4_1(h1m, bs, vv);
        } else {
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            const uint32_t st = rg[c] >> 16, en = rg[c] & 0xffffu;
            acc.add3_if((en > st) & (e0[c].x == kk[c]), e0[c].y,
                        (en > st + 1) & (e1[c].x == kk[c]), e1[c].y,
                        (en > st + 2) & (e2[c].x == kk[c]), e2[c].y, vv[c]);
          }
        }
        if (multi) {
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            const uint32_t st = rg[c] >> 16, en = rg[c] & 0xffffu;
            if (en > st + 3) probe_multi(kk[c], vv[c], st + 3, en);
          }
        }
      };
      probe_range<BLOCK>(sk, sv, U.s_begin, U.s_begin + U.s_len, probe4, probe1);
    }
  }
  join_acc_flush(acc, acc_out);
}

// Join work units from the two sides' partition sizes and starts, built on
// the device so the join follows the partitioning without a host round trip.
// Partition q with both sides non-empty gives ceil(|S_q| / probe_unit)
// units.  Units are ordered largest first (|R_q| + probe tuples) by a
// counting sort over size classes of a quarter octave, so the tail of the
// join's dynamic schedule is short; the order within a class is arbitrary
// (the aggregate does not depend on it).  One workgroup.
constexpr int kUnitBlock = 1024;
constexpr int kUnitClasses = 160;  // 4 per octave of a 40-bit weight
__device__ __forceinline__ uint32_t unit_class(uint64_t w) {
  const int lg = 63 - __clzll((long long)(w | 1));
  const uint32_t sub = lg >= 2 ? (uint32_t)(w >> (lg - 2)) & 3u : 0u;
  const uint32_t c = (uint32_t)lg * 4 + sub;
  return c < kUnitClasses ? c : kUnitClasses - 1;
}
__global__ __launch_bounds__(kUnitBlock) void k_join_units(
    const uint32_t *__restrict__ rT, const unsigned long long *__restrict__ rS,
    const uint32_t *__restrict__ sT, const unsigned long long *__restrict__ sS, uint32_t nb,
    uint32_t pu, JoinUnit *__restrict__ units, uint32_t *__restrict__ unit_hdr) {
  __shared__ uint32_t cnt[kUnitClasses];
  const int tid = threadIdx.x;
  for (int c = tid; c < kUnitClasses; c += kUnitBlock) cnt[c] = 0;
  __syncthreads();
  // eight partitions per thread per step, their loads in flight together
  // (out-of-range slots read partition 0 and are skipped)
  constexpr int K = 8;
  for (uint32_t q0 = tid; q0 < nb; q0 += K * kUnitBlock) {
    uint32_t rl[K], sl[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const uint32_t q = q0 + k * kUnitBlock;
      rl[k] = rT[q < nb ? q : 0u];
      sl[k] = sT[q < nb ? q : 0u];
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (q0 + k * kUnitBlock >= nb || !rl[k] || !sl[k]) continue;
      const uint32_t full = sl[k] / pu, tail = sl[k] % pu;
      if (full) atomicAdd(&cnt[unit_class((uint64_t)pu + rl[k])], full);
      if (tail) atomicAdd(&cnt[unit_class((uint64_t)tail + rl[k])], 1u);
    }
  }
  __syncthreads();
  if (tid == 0) {  // class starts, largest class first
    uint32_t run = 0;
    for (int c = kUnitClasses - 1; c >= 0; --c) {
      const uint32_t x = cnt[c];
      cnt[c] = run;
      run += x;
    }
    unit_hdr[0] = run;
    unit_hdr[1] = 0;
    unit_hdr[2] = 0;
    unit_hdr[3] = 0;
  }
  __syncthreads();
  for (uint32_t q0 = tid; q0 < nb; q0 += K * kUnitBlock) {
    uint32_t rl[K], sl[K];
    unsigned long long rs[K], ss[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const uint32_t q = q0 + k * kUnitBlock < nb ? q0 + k * kUnitBlock : 0u;
      rl[k] = rT[q];
      sl[k] = sT[q];
      rs[k] = rS[q];
      ss[k] = sS[q];
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (q0 + k * kUnitBlock >= nb || !rl[k] || !sl[k]) continue;
      const uint32_t full = sl[k] / pu, tail = sl[k] % pu;
      JoinUnit u;
      u.r_begin = rs[k];
      u.r_len = rl[k];
      if (full) {
        const uint32_t at = atomicAdd(&cnt[unit_class((uint64_t)pu + rl[k])], full);
        for (uint32_t f = 0; f < full; ++f) {
          u.s_begin = ss[k] + (unsigned long long)f * pu;
          u.s_len = pu;
          units[at + f] = u;
        }
      }
      if (tail) {
        const uint32_t at = atomicAdd(&cnt[unit_class((uint64_t)tail + rl[k])], 1u);
        u.s_begin = ss[k] + (unsigned long long)full * pu;
        u.s_len = tail;
        units[at] = u;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
template <class T>
static size_t pack_add(std::vector<uint8_t> &buf, const std::vector<T> &v) {
  size_t off = (buf.size() + 255) & ~(size_t)255;
  buf.resize(off + std::max<size_t>(v.size() * sizeof(T), 1));
  if (!v.empty()) std::memcpy(buf.data() + off, v.data(), v.size() * sizeof(T));
  return off;
}

// Tile / block of each pass by fan-out.  Every fan-out writes 32-tuple
// (128-byte) lines: 64-byte lines measured 15-20% slower.  Fan-out 256 needs
// 64 KiB of line buffers, so it runs one 1024-thread workgroup per CU with
// 8192-tuple tiles; smaller ones two 512-thread workgroups with 4096.
template <int F>
static int join_pass(pj_ctx *ctx, const char *name, bool pass2, const int32_t *ik,
                     const int32_t *iv, int32_t *ok, int32_t *ov, uint64_t n,
                     const RadixPlan &p, JointHashed j, RadixHist &h) {
  // pass 1 replaces keys by fmix32(key); pass 2 and the join read digits,
  // buckets and equality off the hashes
  constexpr int T = F >= 256 ? 8192 : 4096;
  constexpr int BLOCK = F >= 256 ? 1024 : 512;
  constexpr int LINE = 32;
  if (pass2)
    return radix_launch_pass<F, T, BLOCK, JointHashed, int32_t, NoVal, false, LINE, false>(
        ctx, name, true, ik, iv, nullptr, ok, ov, nullptr, n, p, j, h);
  if (!iv)
    return radix_launch_pass<F, T, BLOCK, JointHashed, int32_t, NoVal, true, LINE, true>(
        ctx, name, false, ik, iv, nullptr, ok, ov, nullptr, n, p, j, h);
  return radix_launch_pass<F, T, BLOCK, JointHashed, int32_t, NoVal, false, LINE, true>(
      ctx, name, false, ik, iv, nullptr, ok, ov, nullptr, n, p, j, h);
}

static int join_pass_bits(pj_ctx *ctx, const char *name, int bits, bool pass2,
                          const int32_t *ik, const int32_t *iv, int32_t *ok, int32_t *ov,
                          uint64_t n, const RadixPlan &p, JointHashed j, RadixHist &h) {
  switch (bits) {
    case 1: return join_pass<2>(ctx, name, pass2, ik, iv, ok, ov, n, p, j, h);
    case 2: return join_pass<4>(ctx, name, pass2, ik, iv, ok, ov, n, p, j, h);
    case 3: return join_pass<8>(ctx, name, pass2, ik, iv, ok, ov, n, p, j, h);
    case 4: return join_pass<16>(ctx, name, pass2, ik, iv, ok, ov, n, p, j, h);
    case 5: return join_pass<32>(ctx, name, pass2, ik, iv, ok, ov, n, p, j, h);
    case 6: return join_pass<64>(ctx, name, pass2, ik, iv, ok, ov, n, p, j, h);
    case 7: return join_pass<128>(ctx, name, pass2, ik, iv, ok, ov, n, p, j, h);
    case 8: return join_pass<256>(ctx, name, pass2, ik, iv, ok, ov, n, p, j, h);
    default: return fail(PJ_E_ARG, "radix bits %d outside [1, 8]", bits);
  }
}

// One side of the join, partitioned into 2^(b1+b2) dense partitions.
// The four partition arrays start at different odd multiples of 4 KiB
// within their allocations: pass 2 reads and writes the same offsets of
// equally sized buffers, and bases that coincide modulo the HBM channel
// interleave would send its read and write streams to the same channels.
struct Side {
  RadixPlan plan;
  RadixHist hist;
  Scratch k1, v1, k2, v2;
  const int32_t *fk = nullptr, *fv = nullptr;  // final partitioned columns
};
constexpr uint64_t kSkew[4] = {0, 17 * 4096, 33 * 4096, 49 * 4096};
static int32_t *skewed(Scratch &b, int i) { return (int32_t *)((char *)b.p + kSkew[i]); }

// Queues the side's histograms on ctx->stream.
static int side_histograms(pj_ctx *ctx, Side &s, const int32_t *d_k, uint64_t n, int b1,
                           int b2) {
  const int B = b1 + b2;
  s.plan = radix_make_plan(ctx, n, b1, b2, 1 << B, 2);
  PJ_TRY(s.hist.alloc(ctx, s.plan));
  return radix_histograms(ctx, d_k, n, s.plan, JointRadix{32 - B}, s.hist, false);
}

// Queues the side's first radix pass on ctx->stream (after its histograms).
static int side_pass1(pj_ctx *ctx, Side &s, const int32_t *d_k, const int32_t *d_v,
                      uint64_t n, int b1, int b2, const char *tag) {
  JointHashed j{32 - (b1 + b2)};
  PJ_TRY(s.k1.alloc(ctx, n * 4 + 256 + kSkew[0]));
  PJ_TRY(s.v1.alloc(ctx, n * 4 + 256 + kSkew[1]));
  char name[64];
  snprintf(name, sizeof(name), "radix_scatter1_%s", tag);
  PJ_TRY(join_pass_bits(ctx, name, b1, false, d_k, d_v, skewed(s.k1, 0), skewed(s.v1, 1), n,
                        s.plan, j, s.hist));
  s.fk = skewed(s.k1, 0);
  s.fv = skewed(s.v1, 1);
  return PJ_OK;
}

// Queues the side's second radix pass (if any) on ctx->stream.
static int side_pass2(pj_ctx *ctx, Side &s, uint64_t n, int b1, int b2, const char *tag) {
  if (!b2) return PJ_OK;
  JointHashed j{32 - (b1 + b2)};
  PJ_TRY(s.k2.alloc(ctx, n * 4 + 256 + kSkew[2]));
  PJ_TRY(s.v2.alloc(ctx, n * 4 + 256 + kSkew[3]));
  char name[64];
  snprintf(name, sizeof(name), "radix_scatter2_%s", tag);
  PJ_TRY(join_pass_bits(ctx, name, b2, true, s.fk, s.fv, skewed(s.k2, 2), skewed(s.v2, 3), n,
                        s.plan, j, s.hist));
  s.fk = skewed(s.k2, 2);
  s.fv = skewed(s.v2, 3);
  return PJ_OK;
}

// LDS table geometries (bucket count H = 2^lds_log, capacity NCAP rows).
struct Geometry {
  int lds_log;
  int ncap;
  int per_cu;         // workgroups per CU the LDS allows
  int block = 512;    // threads per workgroup
};
static Geometry geometry_for(int lds_log) {
  switch (lds_log) {
    case 11: return {11, 3072, 4};
    case 12: return {12, 6144, 2};
    case 14: return {14, 11264, 1, 1024};
    default: return {13, 5632, 2};
  }
}

// CSR join over units[0 .. *n_units) claimed with *claim.
template <int BLOCK, int H, int NCAP>
static int launch_pjoin_csr_g(pj_ctx *ctx, bool hashed, const int32_t *rk, const int32_t *rv,
                              const int32_t *sk, const int32_t *sv, const JoinUnit *units,
                              const uint32_t *n_units, uint32_t *claim, uint32_t grid,
                              unsigned long long *acc) {
  if (hashed)
    PJ_LAUNCH(ctx, "pjoin_probe", (k_pjoin<BLOCK, H, NCAP, true>), grid, BLOCK, 0,
              rk, rv, sk, sv, units, n_units, claim, acc);
  else
    PJ_LAUNCH(ctx, "pjoin_probe", (k_pjoin<BLOCK, H, NCAP, false>), grid, BLOCK, 0,
              rk, rv, sk, sv, units, n_units, claim, acc);
  return PJ_OK;
}

static int launch_pjoin_csr(pj_ctx *ctx, int lds_log, bool hashed, const int32_t *rk,
                            const int32_t *rv, const int32_t *sk, const int32_t *sv,
                            const JoinUnit *units, const uint32_t *n_units, uint32_t *claim,
                            uint32_t grid, unsigned long long *acc) {
  switch (lds_log) {
    case 11:
      return launch_pjoin_csr_g<512, 2048, 3072>(ctx, hashed, rk, rv, sk, sv, units, n_units,
                                                 claim, grid, acc);
    case 12:
      return launch_pjoin_csr_g<512, 4096, 6144>(ctx, hashed, rk, rv, sk, sv, units, n_units,
                                                 claim, grid, acc);
    case 13:
      return launch_pjoin_csr_g<512, 8192, 5632>(ctx, hashed, rk, rv, sk, sv, units, n_units,
                                                 claim, grid, acc);
    case 14:
      return launch_pjoin_csr_g<1024, 16384, 11264>(ctx, hashed, rk, rv, sk, sv, units,
                                                    n_units, claim, grid, acc);
    default:
      return fail(PJ_E_ARG, "lds_log %d not in [11, 14]", lds_log);
  }
}

// The presence-bitmap join needs hashed keys (radix_bits B > 0) whose bits
// between the bucket bits and the partition bits fit the bitmap:
//   lds_log 13: narrow table, 8192 buckets, 32 - B - 13 <= 4 (B >= 15), up
//               to 4608 rows in 50 KiB;
//   lds_log 14: wide table, 4096 buckets, 32 - B - 12 <= 6 (B >= 14), up to
//               9216 rows in 76 KiB.
// Both run two 512-thread workgroups per CU.
static bool bitmap_join_ok(int radix_bits, int lds_log) {
  if (radix_bits <= 0) return false;
  if (lds_log == 13) return 32 - radix_bits - 13 <= 4;
  if (lds_log == 14) return 32 - radix_bits - 12 <= 6;
  return false;
}
constexpr int kBitmapPerCu = 2;

// Joins the units of the list at `du` (header, then the units at +256):
// the bitmap kernel first when it applies, which defers the units it cannot
// take to `deferred`, then the CSR kernel over what is left.
static int launch_join(pj_ctx *ctx, const pj_radix_cfg &cfg, int radix_bits, const int32_t *rk,
                       const int32_t *rv, const int32_t *sk, const int32_t *sv, uint8_t *du,
                       JoinUnit *deferred, uint64_t max_units, unsigned long long *acc) {
  const Geometry geo = geometry_for(cfg.lds_log);
  const uint64_t mul = cfg.grid_mul > 0 ? cfg.grid_mul : 1;
  uint32_t *hdr = reinterpret_cast<uint32_t *>(du);
  const JoinUnit *units = reinterpret_cast<const JoinUnit *>(du + 256);
  // workgroups beyond the number of units exit at their first claim
  const uint32_t csr_grid =
      (uint32_t)std::min<uint64_t>((uint64_t)ctx->num_cus * geo.per_cu * mul, max_units);
  if (!bitmap_join_ok(radix_bits, cfg.lds_log))
    return launch_pjoin_csr(ctx, cfg.lds_log, radix_bits > 0, rk, rv, sk, sv, units, hdr,
                            hdr + 1, csr_grid, acc);
  const uint32_t grid =
      (uint32_t)std::min<uint64_t>((uint64_t)ctx->num_cus * kBitmapPerCu * mul, max_units);
  if (cfg.lds_log == 13)
    PJ_LAUNCH(ctx, "pjoin_probe", (k_pjoin_bitmap<512, 8192, 4608, 4, false>), grid, 512, 0,
              rk, rv, sk, sv, units, hdr, deferred, acc);
  else
    PJ_LAUNCH(ctx, "pjoin_probe", (k_pjoin_bitmap<512, 4096, 9216, 4, true>), grid, 512, 0,
              rk, rv, sk, sv, units, hdr, deferred, acc);
  // the deferred units (duplicate build keys, oversized partitions); with
  // none, every workgroup exits at its first claim
  return launch_pjoin_csr(ctx, cfg.lds_log, true, rk, rv, sk, sv, deferred, hdr + 2, hdr + 3,
                          csr_grid, acc);
}

static int join_partitioned(pj_ctx *ctx, const int32_t *d_bk, const int32_t *d_bv,
                            uint64_t nb, const int32_t *d_pk, const int32_t *d_pv,
                            uint64_t np, pj_radix_cfg cfg, pj_join_agg *out) {
  const int B = cfg.bits1 + cfg.bits2;
  Scratch acc;
  PJ_TRY(acc.alloc(ctx, kJoinAccWords * 8));
  PJ_TRY(join_acc_init(ctx, acc.as<unsigned long long>()));
  Side R, S;
  Scratch dunits;  // unit header (kUnitHdrWords x u32, 256 B), then the units
  Scratch ddeferred;  // units the bitmap join left to the CSR join
  uint64_t max_units = 0;
  const int32_t *rk = d_bk, *rv = d_bv, *sk = d_pk, *sv = d_pv;
  if (nb && np) {
    if (B == 0) {
      std::vector<JoinUnit> units;
      for (uint64_t s0 = 0; s0 < np; s0 += cfg.probe_unit) {
        JoinUnit u;
        u.r_begin = 0;
        u.r_len = (uint32_t)nb;
        u.s_begin = s0;
        u.s_len = (uint32_t)std::min<uint64_t>(cfg.probe_unit, np - s0);
        units.push_back(u);
      }
      std::vector<uint32_t> hdr(kUnitHdrWords, 0u);
      hdr[0] = (uint32_t)units.size();
      std::vector<uint8_t> buf;
      pack_add(buf, hdr);
      pack_add(buf, units);  // at offset 256
      PJ_TRY(dunits.alloc(ctx, buf.size()));
      PJ_TRY(h2d(ctx, dunits.p, buf.data(), buf.size()));
      max_units = units.size();
    } else {
      // Schedule: the build side's histograms and pass 1 run on the
      // auxiliary stream beside the probe side's histograms and scans on the
      // main one (the smaller side's kernels alone leave HBM bandwidth
      // unused); the probe side's pass 1 waits for the build side's pass 1
      // and runs alone; the build side's pass 2 waits for the probe side's
      // pass 1 and runs beside the probe side's pass 2.  Left free, the two
      // sides' first passes shared the device in an order that changed from
      // run to run, and the probe side's pass 1 took 4.1 to 6.6 ms.
      hipStream_t aux;
      PJ_TRY(aux_fork(ctx, &aux));
      const int b1 = cfg.bits1, b2 = cfg.bits2;
      int rc;
      {
        StreamSwap sw(ctx, aux);
        rc = side_histograms(ctx, R, d_bk, nb, b1, b2);
        if (rc == PJ_OK) rc = side_pass1(ctx, R, d_bk, d_bv, nb, b1, b2, "build");
      }
      if (rc == PJ_OK) rc = side_histograms(ctx, S, d_pk, np, b1, b2);
      if (rc == PJ_OK) rc = aux_join(ctx);
      if (rc == PJ_OK) rc = side_pass1(ctx, S, d_pk, d_pv, np, b1, b2, "probe");
      if (rc == PJ_OK) rc = aux_fork(ctx, &aux);
      if (rc == PJ_OK) {
        StreamSwap sw(ctx, aux);
        rc = side_pass2(ctx, R, nb, b1, b2, "build");
      }
      if (rc == PJ_OK) rc = side_pass2(ctx, S, np, b1, b2, "probe");
      // join even on failure, so no scratch is released under running work
      const int rj = aux_join(ctx);
      if (rc != PJ_OK) {
        (void)hipStreamSynchronize(aux);
        return rc;
      }
      PJ_TRY(rj);
      rk = R.fk;
      rv = d_bv ? R.fv : nullptr;
      sk = S.fk;
      sv = d_pv ? S.fv : nullptr;
      const uint32_t NB = 1u << B;
      max_units = NB + np / cfg.probe_unit;
      PJ_TRY(dunits.alloc(ctx, 256 + max_units * sizeof(JoinUnit)));
      uint8_t *du = dunits.as<uint8_t>();
      PJ_LAUNCH(ctx, "pjoin_units", k_join_units, 1, kUnitBlock, 0,
                R.hist.T.as<const uint32_t>(), R.hist.S.as<const unsigned long long>(),
                S.hist.T.as<const uint32_t>(), S.hist.S.as<const unsigned long long>(), NB,
                cfg.probe_unit, reinterpret_cast<JoinUnit *>(du + 256),
                reinterpret_cast<uint32_t *>(du));
    }
  }
  if (max_units) {
    if (bitmap_join_ok(B, cfg.lds_log))
      PJ_TRY(ddeferred.alloc(ctx, max_units * sizeof(JoinUnit)));
    PJ_TRY(launch_join(ctx, cfg, B, rk, rv, sk, sv, dunits.as<uint8_t>(),
                       ddeferred.as<JoinUnit>(), max_units, acc.as<unsigned long long>()));
  }
  return join_acc_read(ctx, acc.as<unsigned long long>(), out);
}

static void derive_bits(pj_radix_cfg &c, uint64_t nb) {
  const Geometry g = geometry_for(c.lds_log);
  const uint64_t target = (uint64_t)g.ncap * 3 / 4;
  int B = 0;
  while (B < 15 && (nb >> B) > target) ++B;
  if (B <= 8) {
    c.bits1 = B;
    c.bits2 = 0;
  } else {
    // the odd bit goes to pass 2: a fan-out of 256 runs closer to copy speed
    // over pass 1's output (runs already grouped by the first digit) than
    // over the original input
    c.bits1 = B / 2;
    c.bits2 = B - c.bits1;
  }
}

int validate_cfg(pj_radix_cfg &c, uint64_t nb, uint64_t np) {
  // Large builds take the 8192-row tables (fan-out 2^7 in both passes and the
  // wide bitmap join); smaller ones the 4096-row tables.
  if (c.lds_log == 0) c.lds_log = nb > (1ULL << 26) ? 14 : 13;
  // about 1024 join units (four per CU of a 256-CU device), 4096..131072
  // probe tuples each: a small probe side split into 65536-tuple units leaves
  // most workgroups without one (2^24 probe tuples over 256 partitions made
  // 256); at 2^30 probe tuples 131072-tuple units join 1.6% faster than
  // 65536-tuple ones
  if (c.probe_unit == 0) {
    uint64_t pu = 4096;
    while (pu < 131072 && pu * 2 * 1024 <= np) pu *= 2;
    c.probe_unit = (uint32_t)pu;
  }
  if (c.lds_log < 11 || c.lds_log > 14) return fail(PJ_E_ARG, "lds_log %d not in [11, 14]", c.lds_log);
  if (c.bits1 == 0 && c.bits2 == 0) derive_bits(c, nb);
  if (c.bits1 < 0 || c.bits1 > 8 || c.bits2 < 0 || c.bits2 > 8 || (c.bits1 == 0 && c.bits2 > 0))
    return fail(PJ_E_ARG, "radix bits (%d, %d) invalid: each in [0, 8], bits2 needs bits1",
                c.bits1, c.bits2);
  if (c.bits1 + c.bits2 > 15)
    return fail(PJ_E_ARG, "bits1 + bits2 = %d exceeds 15", c.bits1 + c.bits2);
  if (c.probe_unit < 1024 || (c.probe_unit & 3))
    return fail(PJ_E_ARG, "probe_unit must be a multiple of 4 and >= 1024");
  if (c.grid_mul < 0 || c.grid_mul > 8) return fail(PJ_E_ARG, "grid_mul %d not in [0, 8]", c.grid_mul);
  return PJ_OK;
}

// ---------------------------------------------------------------------------
// hash-rearrange: one pass with bucket = hashInt32(key) URem buckets
// ---------------------------------------------------------------------------
template <int F>
static int rearrange_f(pj_ctx *ctx, const int32_t *d_k, const int32_t *d_v, uint64_t n,
                       uint32_t buckets, int32_t *d_ok, int32_t *d_ov, uint64_t *h_off) {
  RadixPlan p = radix_make_plan(ctx, n, 0, 0, (int)buckets, 2);
  RadixHist h;
  PJ_TRY(h.alloc(ctx, p));
  JointMod j{buckets};
  PJ_TRY(radix_histograms(ctx, d_k, n, p, j, h, true));
  Scratch tmpv;
  int32_t *ov = d_ov;
  if (!ov) {  // the pass always moves a value column
    PJ_TRY(tmpv.alloc(ctx, n * 4 + 256));
    ov = tmpv.as<int32_t>();
  }
  if (d_v)
    PJ_TRY((radix_launch_pass<F, 4096, 256, JointMod, int32_t, NoVal, false>(
        ctx, "rearrange_scatter", false, d_k, d_v, nullptr, d_ok, ov, nullptr, n, p, j, h)));
  else
    PJ_TRY((radix_launch_pass<F, 4096, 256, JointMod, int32_t, NoVal, true>(
        ctx, "rearrange_scatter", false, d_k, d_v, nullptr, d_ok, ov, nullptr, n, p, j, h)));
  PJ_TRY(sync(ctx));
  for (uint32_t b = 0; b <= buckets; ++b) h_off[b] = h.hS[b];
  return PJ_OK;
}

}  // namespace pj

using namespace pj;

extern "C" int pj_join_partitioned_config(pj_ctx *ctx, uint64_t nb, uint64_t np,
                                          pj_radix_cfg *cfg) {
  pj_radix_cfg c{};
  auto it = ctx->tuned.find(TuneKey{ilog2_ceil(nb ? nb : 1), ilog2_ceil(np ? np : 1)});
  if (it != ctx->tuned.end()) c = it->second;
  PJ_TRY(validate_cfg(c, nb, np));
  *cfg = c;
  return PJ_OK;
}

namespace pj {
// The direct table of dense build keys: lo = range[0], span = range[1] -
// range[0] + 1 (the keys' MIN and MAX, computed on the device).  When span is
// the row count nb, vals[key - lo] = value (0 without values) and
// seen[key - lo] = 1 (seen zeroed); plain stores: a repeated key leaves some
// slot of seen at 0, which the probe kernel finds.  Otherwise the keys are
// not dense: flag bit 1, and nothing else happens.
__device__ __forceinline__ bool direct_span(const int64_t *range, uint64_t nb, int32_t *lo,
                                            uint32_t *span) {
  const int64_t mn = range[0], mx = range[1];
  *lo = (int32_t)mn;
  *span = (uint32_t)(mx - mn + 1);
  return (uint64_t)(mx - mn + 1) == nb;
}
__global__ __launch_bounds__(256) void k_direct_build(const int32_t *__restrict__ bk,
                                                      const int32_t *__restrict__ bv, uint64_t nb,
                                                      const int64_t *__restrict__ range,
                                                      int32_t *__restrict__ vals,
                                                      uint8_t *__restrict__ seen,
                                                      uint32_t *__restrict__ flag) {
  int32_t lo;
  uint32_t span;
  if (!direct_span(range, nb, &lo, &span)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(flag, 2u);
    return;
  }
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < nb; i += stride) {
    const uint32_t off = (uint32_t)bk[i] - (uint32_t)lo;  // < span: lo and span cover every key
    vals[off] = bv ? bv[i] : 0;
    seen[off] = 1;
  }
}

// Probe of the direct table with the fused reduce: four probe rows per lane
// per step (16-byte loads, three steps in flight), one table read each.  Each
// workgroup first checks its slice of seen (a slot at 0: flag bit 0).
__global__ __launch_bounds__(256) void k_direct_probe(const int32_t *__restrict__ pk,
                                                      const int32_t *__restrict__ pv, uint64_t np,
                                                      uint64_t nb, const int64_t *__restrict__ range,
                                                      const int32_t *__restrict__ vals,
                                                      const uint8_t *__restrict__ seen,
                                                      uint32_t *__restrict__ flag,
                                                      unsigned long long *__restrict__ acc_out) {
  int32_t lo;
  uint32_t span;
  if (!direct_span(range, nb, &lo, &span)) return;  // uniform: no flush, no result
  {
    const uint64_t per = ceil_div((uint64_t)span, (uint64_t)gridDim.x);
    const uint64_t a0 = (uint64_t)blockIdx.x * per, a1 = a0 + per < span ? a0 + per : span;
    bool hole = false;
    for (uint64_t i = a0 + threadIdx.x; i < a1; i += 256) hole |= seen[i] == 0;
    if (hole) atomicOr(flag, 1u);
  }
  JoinAcc acc;
  acc.init();
  auto probe1 = [&](int32_t k, int32_t p) {
    const uint32_t off = (uint32_t)k - (uint32_t)lo;
    if (off < span) acc.add1(vals[off], p);
  };
  auto probe4 = [&](const int4 k4, const int4 v4) {
    const int32_t kk[4] = {k4.x, k4.y, k4.z, k4.w};
    const int32_t vv[4] = {pv ? v4.x : 0, pv ? v4.y : 0, pv ? v4.z : 0, pv ? v4.w : 0};
    bool hit[4];
    int32_t b[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const uint32_t off = (uint32_t)kk[c] - (uint32_t)lo;
      hit[c] = off < span;
      b[c] = vals[hit[c] ? off : 0u];
    }
    acc.add4_1(hit, b, vv);
  };
  // every workgroup takes one contiguous slice of the probe rows
  const uint64_t per = (ceil_div(np, (uint64_t)gridDim.x) + 3) & ~3ULL;
  const uint64_t s0 = (uint64_t)blockIdx.x * per;
  const uint64_t s1 = s0 + per < np ? s0 + per : np;
  if ((((uintptr_t)pk | (uintptr_t)pv) & 15) == 0) {
    if (s0 < s1) probe_range<256>(pk, pv, s0, s1, probe4, probe1);
  } else {  // columns not 16-byte aligned: one row at a time
    for (uint64_t x = s0 + threadIdx.x; x < s1; x += 256) probe1(pk[x], pv ? pv[x] : 0);
  }
  join_acc_flush(acc, acc_out);
}
}  // namespace pj

int pj::join_direct(pj_ctx *ctx, const int32_t *d_bk, const int32_t *d_bv, uint64_t nb,
                    const int32_t *d_pk, const int32_t *d_pv, uint64_t np, pj_join_agg *out,
                    bool *done) {
  *done = false;
  if (nb < 2 || nb > kDirectMaxSpan || !np) return PJ_OK;
  // the key range stays on the device: the table is sized for nb keys, and
  // the kernels do nothing when the span is not nb; one read-back at the end
  // brings the result and the flags together
  Scratch part, range, vals, seen, acc;
  PJ_TRY(range.alloc(ctx, 16));
  PJ_TRY(rel::minmax_i32_device(ctx, d_bk, nb, part, range.as<int64_t>()));
  PJ_TRY(vals.alloc(ctx, nb * 4));
  PJ_TRY(seen.alloc(ctx, nb));
  PJ_TRY(zero(ctx, seen.p, nb));
  PJ_TRY(acc.alloc(ctx, kJoinAccWords * 8));
  PJ_TRY(join_acc_init(ctx, acc.as<unsigned long long>()));  // clears the flag word too
  uint32_t *flag = reinterpret_cast<uint32_t *>(acc.as<unsigned long long>() + kJoinAccFlagWord);
  const uint64_t bgrid = std::min<uint64_t>(ceil_div(nb, 256), (uint64_t)ctx->num_cus * 16);
  PJ_LAUNCH(ctx, "join_direct_build", k_direct_build, bgrid, 256, 0, d_bk, d_bv, nb,
            range.as<const int64_t>(), vals.as<int32_t>(), seen.as<uint8_t>(), flag);
  // slices of >= 16 Ki rows, at most 8 workgroups per CU
  const uint64_t pgrid = std::max<uint64_t>(
      1, std::min<uint64_t>(ceil_div(np, 16384), (uint64_t)ctx->num_cus * 8));
  PJ_LAUNCH(ctx, "join_direct_probe", k_direct_probe, pgrid, 256, 0, d_pk, d_pv, np, nb,
            range.as<const int64_t>(), vals.as<const int32_t>(), seen.as<const uint8_t>(), flag,
            acc.as<unsigned long long>());
  // the flag word sits in the accumulator's first copy
  int64_t h[kJoinAccWords];
  PJ_TRY(d2h(ctx, h, acc.p, sizeof(h)));
  if (((const uint32_t *)h)[2 * kJoinAccFlagWord]) return PJ_OK;  // not dense, or a repeat
  PJ_TRY(join_acc_fold(h, out));
  *done = true;
  return PJ_OK;
}

int pj::join_partitioned_auto(pj_ctx *ctx, const int32_t *d_bk, const int32_t *d_bv,
                              uint64_t nb, const int32_t *d_pk, const int32_t *d_pv, uint64_t np,
                              pj_join_agg *out) {
  pj_radix_cfg c{};
  auto it = ctx->tuned.find(TuneKey{ilog2_ceil(nb ? nb : 1), ilog2_ceil(np ? np : 1)});
  if (it != ctx->tuned.end()) c = it->second;
  PJ_TRY(validate_cfg(c, nb, np));
  return join_partitioned(ctx, d_bk, d_bv, nb, d_pk, d_pv, np, c, out);
}

extern "C" int pj_autotune_enable(pj_ctx *ctx, int on) {
  ctx->autotune = on != 0;
  return PJ_OK;
}

int pj_autotune_join(pj_ctx *ctx, const int32_t *d_bk, const int32_t *d_bv, uint64_t nb,
                     const int32_t *d_pk, const int32_t *d_pv, uint64_t np);

extern "C" int pj_join_partitioned_reduce(pj_ctx *ctx, const int32_t *d_bk,
                                          const int32_t *d_bv, uint64_t nb,
                                          const int32_t *d_pk, const int32_t *d_pv,
                                          uint64_t np, uint64_t max_build,
                                          uint64_t max_probe, const pj_radix_cfg *cfg,
                                          pj_join_agg *out) {
  std::lock_guard<std::mutex> g(ctx->mu);
  if (nb > max_build)
    return fail(PJ_E_CAPACITY, "build input %llu exceeds maxBuildInputSize %llu",
                (unsigned long long)nb, (unsigned long long)max_build);
  if (np > max_probe)
    return fail(PJ_E_CAPACITY, "probe input %llu exceeds maxProbeInputSize %llu",
                (unsigned long long)np, (unsigned long long)max_probe);
  if (nb >= (1ULL << 32) || np >= (1ULL << 40))
    return fail(PJ_E_CAPACITY, "input above the supported row counts");
  pj_radix_cfg c{};
  if (cfg) {
    c = *cfg;
  } else {
    TuneKey key{ilog2_ceil(nb ? nb : 1), ilog2_ceil(np ? np : 1)};
    auto it = ctx->tuned.find(key);
    if (it == ctx->tuned.end() && ctx->autotune && nb && np) {
      PJ_TRY(pj_autotune_join(ctx, d_bk, d_bv, nb, d_pk, d_pv, np));
      it = ctx->tuned.find(key);
    }
    if (it != ctx->tuned.end()) c = it->second;
  }
  PJ_TRY(validate_cfg(c, nb, np));
  return join_partitioned(ctx, d_bk, d_bv, nb, d_pk, d_pv, np, c, out);
}

// Sweep a small grid of configurations on the given inputs and remember the
// fastest per (log2 build, log2 probe) size class.
int pj_autotune_join(pj_ctx *ctx, const int32_t *d_bk, const int32_t *d_bv, uint64_t nb,
                     const int32_t *d_pk, const int32_t *d_pv, uint64_t np) {
  std::vector<pj_radix_cfg> cands;
  for (int lds : {12, 13, 14}) {
    pj_radix_cfg base{};
    base.lds_log = lds;
    if (validate_cfg(base, nb, np) != PJ_OK) continue;
    int B = base.bits1 + base.bits2;
    for (int b1 = std::max(1, B - 8); b1 <= std::min(8, B); ++b1) {
      pj_radix_cfg c = base;
      c.bits1 = b1;
      c.bits2 = B - b1;
      if (B == 0) c.bits1 = 0;
      for (uint32_t pu : {32768u, 65536u}) {
        c.probe_unit = pu;
        cands.push_back(c);
      }
      if (B == 0) break;
    }
  }
  hipEvent_t a, b;
  PJ_HIP(hipEventCreate(&a));
  PJ_HIP(hipEventCreate(&b));
  float best = 1e30f;
  pj_radix_cfg bestc = cands.empty() ? pj_radix_cfg{} : cands[0];
  bool saved_prof = ctx->prof;
  ctx->prof = false;
  for (auto &c : cands) {
    pj_join_agg o;
    float tsum = 0;
    bool ok = true;
    for (int rep = 0; rep < 3 && ok; ++rep) {
      (void)hipEventRecord(a, ctx->stream);
      ok = join_partitioned(ctx, d_bk, d_bv, nb, d_pk, d_pv, np, c, &o) == PJ_OK;
      (void)hipEventRecord(b, ctx->stream);
      (void)hipEventSynchronize(b);
      float ms = 0;
      (void)hipEventElapsedTime(&ms, a, b);
      if (rep) tsum += ms;
    }
    if (ok && tsum < best) {
      best = tsum;
      bestc = c;
    }
  }
  ctx->prof = saved_prof;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  ctx->tuned[TuneKey{ilog2_ceil(nb), ilog2_ceil(np)}] = bestc;
  return PJ_OK;
}

extern "C" int pj_hash_rearrange_i32(pj_ctx *ctx, const int32_t *d_k,
                                     const int32_t *d_v, uint64_t n, uint32_t buckets,
                                     int32_t *d_out_k, int32_t *d_out_v,
                                     uint64_t *h_offsets) {
  std::lock_guard<std::mutex> g(ctx->mu);
  if (buckets < 1 || buckets > 256)
    return fail(PJ_E_ARG, "buckets %u outside [1, 256]", buckets);
  if (n == 0) {
    for (uint32_t b = 0; b <= buckets; ++b) h_offsets[b] = 0;
    return PJ_OK;
  }
  // the scatter always moves a value column; without one, move the row id
  if (buckets <= 32)
    return rearrange_f<32>(ctx, d_k, d_v, n, buckets, d_out_k, d_out_v, h_offsets);
  return rearrange_f<256>(ctx, d_k, d_v, n, buckets, d_out_k, d_out_v, h_offsets);
}
