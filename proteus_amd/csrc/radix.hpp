// Radix partitioning kernels shared by the partitioned join, hash-rearrange
// and the partitioned group-by.
//
// Reference: HashPartitioner::close and build_partitions
// (core/olap/lib/operators/gpu/gpu-partitioned-hash-join-chained.cpp:249-404,
//  gpu-partitioned-hash-join-chained-helpers.cpp:243-481).  The reference
// chains 4096-tuple buckets per partition, reserving space with 64-bit
// atomics plus a retry loop, after materialising the input with one global
// atomicAdd per tuple.  Here:
//   * one histogram kernel counts every (pass-1, pass-2) digit pair at once,
//     so both passes write into exactly sized, 32-element-aligned regions;
//   * each workgroup keeps, per output digit, a 32-element (128 B) LDS
//     write-combining buffer; a tile is shuffled through LDS by digit and
//     only FULL aligned 128 B lines are written, their space reserved with
//     one atomic per (tile, digit);
//   * the < 32-element remainders left when a workgroup finishes a region are
//     written from the BACK of the region ("tails"), so full lines and tails
//     never interleave and every region ends up dense.
#pragma once

#include "pj_internal.hpp"

namespace pj {

constexpr int kLine = 32;  // elements per 128 B line of int32

// Digit functors -----------------------------------------------------------
struct DigitRadix {  // (fmix32(key) >> shift) & mask
  int shift;
  uint32_t mask;
  __device__ __forceinline__ uint32_t operator()(int32_t k) const {
    return (fmix32((uint32_t)k) >> shift) & mask;
  }
};
struct DigitMod {  // hashInt32(key) URem buckets (gpu-hash-rearrange.cpp:338-348)
  uint32_t buckets;
  __device__ __forceinline__ uint32_t operator()(int32_t k) const {
    return (uint32_t)(hash_i32(k) % (uint64_t)buckets);
  }
};

// Block-wide exclusive scan of one value per thread (threads >= count pass 0).
template <int BLOCK>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *tmp,
                                                    uint32_t *total) {
  constexpr int NW = BLOCK / 64;
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t incl = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t x = __shfl_up(incl, o, 64);
    if (lane >= o) incl += x;
  }
  if (lane == 63) tmp[wid] = incl;
  __syncthreads();
  uint32_t pre = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    uint32_t t = tmp[w];
    pre += (w < wid) ? t : 0;
    tot += t;
  }
  *total = tot;
  __syncthreads();
  return pre + incl - v;
}

// Histogram of (fmix32(key) >> shift) over 2^BITS bins: one LDS histogram per
// workgroup, written out as a row of `partial` (no global atomics), summed by
// k_hist_reduce.
template <int BITS, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_hist(const int32_t *__restrict__ keys,
                                                uint64_t n, int shift,
                                                uint32_t *__restrict__ partial) {
  constexpr int NB = 1 << BITS;
  __shared__ uint32_t h[NB];
  for (int i = threadIdx.x; i < NB; i += BLOCK) h[i] = 0;
  __syncthreads();
  uint64_t nvec = n / 4;
  const int4 *kv = reinterpret_cast<const int4 *>(keys);
  bool vec = ((uintptr_t)keys & 15) == 0;
  uint64_t tid = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  uint64_t nthr = (uint64_t)gridDim.x * BLOCK;
  if (vec) {
    for (uint64_t i = tid; i < nvec; i += nthr) {
      int4 k = ld_nt4(&kv[i]);
      atomicAdd(&h[fmix32((uint32_t)k.x) >> shift], 1u);
      atomicAdd(&h[fmix32((uint32_t)k.y) >> shift], 1u);
      atomicAdd(&h[fmix32((uint32_t)k.z) >> shift], 1u);
      atomicAdd(&h[fmix32((uint32_t)k.w) >> shift], 1u);
    }
    for (uint64_t i = nvec * 4 + tid; i < n; i += nthr)
      atomicAdd(&h[fmix32((uint32_t)keys[i]) >> shift], 1u);
  } else {
    for (uint64_t i = tid; i < n; i += nthr)
      atomicAdd(&h[fmix32((uint32_t)keys[i]) >> shift], 1u);
  }
  __syncthreads();
  uint32_t *row = partial + (uint64_t)blockIdx.x * NB;
  for (int i = threadIdx.x; i < NB; i += BLOCK) row[i] = h[i];
}

// Joint histogram of the top `bits` bits of fmix32(key) (bits <= 15), read
// back to the host.
int radix_histogram(pj_ctx *ctx, const int32_t *d_k, uint64_t n, int bits,
                    std::vector<uint32_t> &h);

__global__ void k_hist_reduce(const uint32_t *__restrict__ partial, int rows,
                              int nb, uint32_t *__restrict__ hist);

// Histogram of an arbitrary digit functor over `F` bins (hash-rearrange).
template <int F, int BLOCK, class Digit>
__global__ __launch_bounds__(BLOCK) void k_hist_digit(
    const int32_t *__restrict__ keys, uint64_t n, Digit dig,
    uint32_t *__restrict__ partial) {
  __shared__ uint32_t h[F];
  for (int i = threadIdx.x; i < F; i += BLOCK) h[i] = 0;
  __syncthreads();
  uint64_t tid = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  uint64_t nthr = (uint64_t)gridDim.x * BLOCK;
  for (uint64_t i = tid; i < n; i += nthr) atomicAdd(&h[dig(keys[i])], 1u);
  __syncthreads();
  for (int i = threadIdx.x; i < F; i += BLOCK) partial[(uint64_t)blockIdx.x * F + i] = h[i];
}

// Reserve c elements at the back of a region: returns the first slot.
__device__ __forceinline__ unsigned long long reserve_tail(unsigned long long *ct,
                                                           uint32_t c) {
  unsigned long long old = atomicAdd(ct, (unsigned long long)(-(long long)c));
  return old - c;
}

// Scatter plan: input regions (one per parent digit), the flat tile list over
// them, and the output cursors of every (region, digit) pair.
struct ScatterPlan {
  const uint64_t *reg_start;   // [nreg] first input element of each region
  const uint64_t *reg_count;   // [nreg]
  const uint32_t *tile_begin;  // [nreg + 1] prefix of tiles per region
  uint32_t nreg;
  uint32_t ntiles;
  uint32_t tiles_per_wg;
  unsigned long long *cfull;  // [nreg * F] next full-line slot (front)
  unsigned long long *ctail;  // [nreg * F] end of the unreserved space (back)
};

// Value-column traits.  NoVal marks an absent second value column.
struct NoVal {};
template <class V>
struct ValIO {
  static constexpr bool present = true;
  // Load elements e..e+3 of a tile (vector loads when aligned and in range).
  __device__ static __forceinline__ void load4(const V *__restrict__ p, uint64_t base,
                                               uint32_t e, uint32_t len, bool vec,
                                               V out[4]) {
    if (vec && e + 3 < len) {
      const int4 *q = reinterpret_cast<const int4 *>(p + base + e);
      if constexpr (sizeof(V) == 4) {
        int4 a = ld_nt4(q);
        out[0] = __builtin_bit_cast(V, a.x);
        out[1] = __builtin_bit_cast(V, a.y);
        out[2] = __builtin_bit_cast(V, a.z);
        out[3] = __builtin_bit_cast(V, a.w);
      } else {
        static_assert(sizeof(V) == 8, "4- or 8-byte values");
        int4 a = ld_nt4(q);
        int4 b = ld_nt4(q + 1);
        out[0] = __builtin_bit_cast(V, make_int2(a.x, a.y));
        out[1] = __builtin_bit_cast(V, make_int2(a.z, a.w));
        out[2] = __builtin_bit_cast(V, make_int2(b.x, b.y));
        out[3] = __builtin_bit_cast(V, make_int2(b.z, b.w));
      }
    } else {
#pragma unroll
      for (int c = 0; c < 4; ++c) out[c] = (e + c < len) ? p[base + e + c] : V{};
    }
  }
  __device__ static __forceinline__ bool aligned(const V *p, uint64_t base) {
    return ((uintptr_t)(p + base) & 15) == 0;
  }
};
template <>
struct ValIO<NoVal> {
  static constexpr bool present = false;
  __device__ static __forceinline__ void load4(const NoVal *, uint64_t, uint32_t, uint32_t,
                                               bool, NoVal out[4]) {}
  __device__ static __forceinline__ bool aligned(const NoVal *, uint64_t) { return true; }
};

// One radix scatter pass over (key, V1[, V2]) rows.  Per tile of T rows:
//   load -> digit + LDS rank -> scan -> reserve full lines (1 atomic/digit)
//   -> shuffle into LDS by digit -> write full lines (32 rows; a half-wave
//   per line) -> keep remainders in the per-digit LDS line buffers.
// Buffers are flushed as tails when the workgroup leaves a region.
// With ROWID the V1 (int32) column is the input row index instead of a load
// (region 0 must start at row 0).
template <int F, int T, int BLOCK, class Digit, class V1, class V2, bool ROWID>
__global__ __launch_bounds__(BLOCK) void k_scatter(
    const int32_t *__restrict__ in_k, const V1 *__restrict__ in_v,
    const V2 *__restrict__ in_w, int32_t *__restrict__ out_k,
    V1 *__restrict__ out_v, V2 *__restrict__ out_w, ScatterPlan plan, Digit dig) {
  constexpr int IPT = T / BLOCK;
  static_assert(IPT % 4 == 0, "tile must give each thread a multiple of 4 items");
  static_assert(F <= BLOCK, "one thread per digit");
  static_assert(!ROWID || sizeof(V1) == 4, "row ids are int32");
  constexpr bool HW = ValIO<V2>::present;
  constexpr int MAXLINES = T / kLine + F;
  constexpr int TW = HW ? T : 1;
  constexpr int FW = HW ? F : 1;

  __shared__ int32_t st_k[T];
  __shared__ V1 st_v[T];
  __shared__ V2 st_w[TW];
  __shared__ int32_t lf_k[F][kLine];
  __shared__ V1 lf_v[F][kLine];
  __shared__ V2 lf_w[FW][kLine];
  __shared__ uint32_t hist[F], off[F], lcnt[F], nfull[F], linepre[F];
  __shared__ unsigned long long gbase[F];
  __shared__ uint16_t line_digit[MAXLINES];
  __shared__ uint32_t scan_tmp[BLOCK / 64];
  __shared__ uint32_t s_total_lines;

  const int tid = threadIdx.x;
  uint32_t t0 = blockIdx.x * plan.tiles_per_wg;
  uint32_t t1 = t0 + plan.tiles_per_wg;
  if (t1 > plan.ntiles) t1 = plan.ntiles;
  if (t0 >= t1) return;

  uint32_t r = 0;
  {
    uint32_t lo = 0, hi = plan.nreg;  // last r with tile_begin[r] <= t0
    while (hi - lo > 1) {
      uint32_t mid = (lo + hi) >> 1;
      if (plan.tile_begin[mid] <= t0) lo = mid; else hi = mid;
    }
    r = lo;
  }
  if (tid < F) lcnt[tid] = 0;

  auto flush_tails = [&](uint32_t reg) {
    __syncthreads();
    if (tid < F) {
      uint32_t c = lcnt[tid];
      gbase[tid] = c ? reserve_tail(&plan.ctail[(uint64_t)reg * F + tid], c) : 0ULL;
    }
    __syncthreads();
    for (int x = tid; x < F * kLine; x += BLOCK) {
      int d = x / kLine, s = x % kLine;
      if ((uint32_t)s < lcnt[d]) {
        unsigned long long g = gbase[d] + s;
        out_k[g] = lf_k[d][s];
        out_v[g] = lf_v[d][s];
        if constexpr (HW) out_w[g] = lf_w[d][s];
      }
    }
    __syncthreads();
    if (tid < F) lcnt[tid] = 0;
  };

  for (uint32_t t = t0; t < t1; ++t) {
    while (t >= plan.tile_begin[r + 1]) {
      flush_tails(r);
      ++r;
    }
    const uint64_t rstart = plan.reg_start[r];
    const uint64_t rcount = plan.reg_count[r];
    const uint64_t toff = (uint64_t)(t - plan.tile_begin[r]) * T;
    const uint64_t base = rstart + toff;
    const uint32_t len = (uint32_t)((rcount - toff) < (uint64_t)T ? (rcount - toff) : T);

    // ---- load ----------------------------------------------------------
    int32_t k[IPT];
    V1 v[IPT];
    V2 w[IPT];
    uint32_t dg[IPT], rk[IPT];
    const bool vec = (((uintptr_t)(in_k + base) & 15) == 0) &&
                     (ROWID || ValIO<V1>::aligned(in_v, base)) &&
                     ValIO<V2>::aligned(in_w, base);
#pragma unroll
    for (int q = 0; q < IPT / 4; ++q) {
      uint32_t e = (uint32_t)(q * BLOCK + tid) * 4;
      ValIO<int32_t>::load4(in_k, base, e, len, vec, &k[4 * q]);
      if constexpr (ROWID) {
#pragma unroll
        for (int c = 0; c < 4; ++c) v[4 * q + c] = __builtin_bit_cast(V1, (int32_t)(base + e + c));
      } else {
        ValIO<V1>::load4(in_v, base, e, len, vec, &v[4 * q]);
      }
      ValIO<V2>::load4(in_w, base, e, len, vec, &w[4 * q]);
    }
    if (tid < F) hist[tid] = 0;
    __syncthreads();
    // ---- digits and LDS ranks ------------------------------------------
#pragma unroll
    for (int q = 0; q < IPT / 4; ++q) {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        uint32_t e = (uint32_t)(q * BLOCK + tid) * 4 + c;
        int i = 4 * q + c;
        if (e < len) {
          dg[i] = dig(k[i]);
          rk[i] = atomicAdd(&hist[dg[i]], 1u);
        } else {
          dg[i] = 0xffffffffu;
          rk[i] = 0;
        }
      }
    }
    __syncthreads();
    // ---- scan, full-line reservation ------------------------------------
    {
      uint32_t hv = tid < F ? hist[tid] : 0;
      uint32_t tot;
      uint32_t o = block_excl_scan<BLOCK>(hv, scan_tmp, &tot);
      uint32_t n = tid < F ? lcnt[tid] + hv : 0;
      uint32_t full = n & ~(uint32_t)(kLine - 1);
      uint32_t ltot;
      uint32_t lp = block_excl_scan<BLOCK>(full / kLine, scan_tmp, &ltot);
      if (tid < F) {
        off[tid] = o;
        nfull[tid] = full;
        linepre[tid] = lp;
        gbase[tid] = full ? atomicAdd(&plan.cfull[(uint64_t)r * F + tid],
                                      (unsigned long long)full)
                          : 0ULL;
      }
      if (tid == 0) s_total_lines = ltot;
    }
    __syncthreads();
    // ---- shuffle into LDS by digit; line -> digit map --------------------
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
      if (dg[i] != 0xffffffffu) {
        uint32_t p = off[dg[i]] + rk[i];
        st_k[p] = k[i];
        st_v[p] = v[i];
        if constexpr (HW) st_w[p] = w[i];
      }
    }
    if (tid < F) {
      uint32_t nl = nfull[tid] / kLine, lp = linepre[tid];
      for (uint32_t l = 0; l < nl; ++l) line_digit[lp + l] = (uint16_t)tid;
    }
    __syncthreads();
    // ---- write full lines: one half-wave per 32-row line ----------------
    {
      const uint32_t nlines = s_total_lines;
      const int hw = tid >> 5, l32 = tid & 31;
      for (uint32_t L = hw; L < nlines; L += BLOCK / 32) {
        uint32_t d = line_digit[L];
        uint32_t j = (L - linepre[d]) * kLine + l32;
        uint32_t lc = lcnt[d];
        int32_t ok;
        V1 ov;
        V2 ow;
        if (j < lc) {
          ok = lf_k[d][j];
          ov = lf_v[d][j];
          if constexpr (HW) ow = lf_w[d][j];
        } else {
          uint32_t src = off[d] + j - lc;
          ok = st_k[src];
          ov = st_v[src];
          if constexpr (HW) ow = st_w[src];
        }
        unsigned long long g = gbase[d] + j;
        __builtin_nontemporal_store(ok, out_k + g);
        __builtin_nontemporal_store(ov, out_v + g);
        if constexpr (HW) __builtin_nontemporal_store(ow, out_w + g);
      }
    }
    __syncthreads();
    // ---- keep remainders ---------------------------------------------------
    for (int x = tid; x < F * kLine; x += BLOCK) {
      int d = x / kLine, s = x % kLine;
      uint32_t lc = lcnt[d], h = hist[d], full = nfull[d], n = lc + h;
      int64_t src = -1;
      if (full) {
        if ((uint32_t)s < n - full) src = (int64_t)off[d] + full + s - lc;
      } else if ((uint32_t)s >= lc && (uint32_t)s < n) {
        src = (int64_t)off[d] + s - lc;
      }
      if (src >= 0) {
        lf_k[d][s] = st_k[src];
        lf_v[d][s] = st_v[src];
        if constexpr (HW) lf_w[d][s] = st_w[src];
      }
    }
    __syncthreads();
    if (tid < F) lcnt[tid] = lcnt[tid] + hist[tid] - nfull[tid];
  }
  flush_tails(r);
}

}  // namespace pj
