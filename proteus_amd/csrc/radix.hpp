// Radix partitioning shared by the partitioned join, hash-rearrange and the
// partitioned group-by.
//
// Reference: HashPartitioner::close and build_partitions
// (core/olap/lib/operators/gpu/gpu-partitioned-hash-join-chained.cpp:249-404,
//  gpu-partitioned-hash-join-chained-helpers.cpp:243-481).  The reference
// materialises its input with one global atomicAdd per tuple, then chains
// 4096-tuple buckets per partition, reserving space with 64-bit atomics and a
// retry loop.  Here partitioning is deterministic and atomic-free:
//   1. k_chunk_hist: the input is cut into G chunks (one per pass-1
//      workgroup); every chunk's joint histogram over all NB = 2^(b1+b2)
//      (pass-1 digit, pass-2 digit) bins is one row of H[G][NB];
//   2. k_colscan + k_excl_scan: E[c][d] = sum of H[c'][d] over c' < c, the
//      bin totals T[d] and the dense partition starts S[d] = sum T[d' < d];
//   3. pass 1: chunk c writes its digit-d1 tuples to
//         S[d1 << b2] + sum over d2 of E[c][d1][d2]     (exact, no atomics);
//   4. pass 2: the tuples chunk c wrote for digit d1 form one contiguous
//      segment whose pass-2 histogram is H[c][d1][*], so digit d2 goes to
//         S[d1][d2] + E[c][d1][d2].
// Every workgroup shuffles a tile through LDS by digit and keeps one 32-tuple
// (128 B per column) write-combining buffer per digit; it only writes whole
// aligned lines except the first and last line of each of its runs.  The next
// tile is loaded into registers while the current one is shuffled.
// The result: both sides laid out densely, final partition d occupying
// [S[d], S[d] + T[d]).
#pragma once

#include <algorithm>
#include <vector>

#include "pj_internal.hpp"

namespace pj {

constexpr int kLine = 32;  // tuples per write-combining line (128 B of int32)

// Joint digit functors: bin in [0, NB) of a key.  d1 = bin >> b2,
// d2 = bin & (2^b2 - 1).
// A Joint with kRange = true also has k_chunk_hist record the keys' min and
// max in range[0] / range[1] (initialised to INT32_MAX / INT32_MIN).
struct JointRadix {  // top B bits of fmix32(key)
  static constexpr bool kRange = false;
  int shift;         // 32 - B, B >= 1
  __device__ __forceinline__ uint32_t operator()(int32_t k) const {
    return fmix32((uint32_t)k) >> shift;
  }
};
struct JointHashed {  // top B bits of a key that already is fmix32(original key)
  static constexpr bool kRange = false;
  int shift;
  __device__ __forceinline__ uint32_t operator()(int32_t h) const { return (uint32_t)h >> shift; }
};
struct JointMod {  // hashInt32(key) URem buckets (gpu-hash-rearrange.cpp:338-348)
  static constexpr bool kRange = false;
  uint32_t buckets;
  __device__ __forceinline__ uint32_t operator()(int32_t k) const {
    return (uint32_t)(hash_i32(k) % (uint64_t)buckets);
  }
};
struct JointLow {  // key bits [shift, shift + B): B = bits above the low `shift`
  static constexpr bool kRange = true;
  int shift;
  uint32_t mask;  // 2^B - 1
  int *range;     // min / max of the keys, recorded by k_chunk_hist
  __device__ __forceinline__ uint32_t operator()(int32_t k) const {
    return ((uint32_t)k >> shift) & mask;
  }
};

// Block-wide exclusive scan of one value per thread.
template <int BLOCK, class U>
__device__ __forceinline__ U block_excl_scan(U v, U *tmp, U *total) {
  constexpr int NW = BLOCK / 64;
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  U incl = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    U x = __shfl_up(incl, o, 64);
    if (lane >= o) incl += x;
  }
  if (lane == 63) tmp[wid] = incl;
  __syncthreads();
  U pre = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    U t = tmp[w];
    pre += (w < wid) ? t : 0;
    tot += t;
  }
  *total = tot;
  __syncthreads();
  return pre + incl - v;
}

// Value-column traits.  NoVal marks an absent column.
struct NoVal {};
template <class V>
struct ValIO {
  static constexpr bool present = true;
  // Elements e..e+3 of a tile whose valid elements are [lo, len): one or two
  // 16-byte loads when all four are valid and the address is aligned.
  __device__ static __forceinline__ void load4(const V *__restrict__ p, uint64_t base,
                                               uint32_t e, uint32_t lo, uint32_t len,
                                               bool vec, V out[4]) {
    if (vec && e >= lo && e + 3 < len) {
      const char *q = reinterpret_cast<const char *>(p + base + e);
      if constexpr (sizeof(V) == 4) {
        int4 a = ld_nt4(q);
        out[0] = __builtin_bit_cast(V, a.x);
        out[1] = __builtin_bit_cast(V, a.y);
        out[2] = __builtin_bit_cast(V, a.z);
        out[3] = __builtin_bit_cast(V, a.w);
      } else {
        static_assert(sizeof(V) == 8, "4- or 8-byte values");
        int4 a = ld_nt4(q);
        int4 b = ld_nt4(q + 16);
        out[0] = __builtin_bit_cast(V, make_int2(a.x, a.y));
        out[1] = __builtin_bit_cast(V, make_int2(a.z, a.w));
        out[2] = __builtin_bit_cast(V, make_int2(b.x, b.y));
        out[3] = __builtin_bit_cast(V, make_int2(b.z, b.w));
      }
    } else {
#pragma unroll
      for (int c = 0; c < 4; ++c)
        out[c] = (e + c >= lo && e + c < len) ? p[base + e + c] : V{};
    }
  }
  __device__ static __forceinline__ bool aligned(const V *p, uint64_t base) {
    return ((uintptr_t)(p + base) & 15) == 0;
  }
  // Elements 4*vi .. 4*vi+3 of a 16-byte-aligned column, unconditionally.
  __device__ static __forceinline__ void load4_vec(const V *__restrict__ p, uint64_t vi,
                                                   V out[4]) {
    const char *q = reinterpret_cast<const char *>(p + vi * 4);
    if constexpr (sizeof(V) == 4) {
      int4 a = ld_nt4(q);
      out[0] = __builtin_bit_cast(V, a.x);
      out[1] = __builtin_bit_cast(V, a.y);
      out[2] = __builtin_bit_cast(V, a.z);
      out[3] = __builtin_bit_cast(V, a.w);
    } else {
      int4 a = ld_nt4(q);
      int4 b = ld_nt4(q + 16);
      out[0] = __builtin_bit_cast(V, make_int2(a.x, a.y));
      out[1] = __builtin_bit_cast(V, make_int2(a.z, a.w));
      out[2] = __builtin_bit_cast(V, make_int2(b.x, b.y));
      out[3] = __builtin_bit_cast(V, make_int2(b.z, b.w));
    }
  }
};
template <>
struct ValIO<NoVal> {
  static constexpr bool present = false;
  __device__ static __forceinline__ void load4(const NoVal *, uint64_t, uint32_t, uint32_t,
                                               uint32_t, bool, NoVal *) {}
  __device__ static __forceinline__ bool aligned(const NoVal *, uint64_t) { return true; }
  __device__ static __forceinline__ void load4_vec(const NoVal *, uint64_t, NoVal *) {}
};

// Per-chunk joint histogram: chunk c = [c*chunk, min(n, (c+1)*chunk)) into
// row c of H (nb u32 bins, dynamic LDS).  chunk is a multiple of 4.
template <class Joint, int HB>
__global__ __launch_bounds__(HB) void k_chunk_hist(const int32_t *__restrict__ keys,
                                                     uint64_t n, uint64_t chunk, int nb,
                                                     Joint joint, uint32_t *__restrict__ H) {
  extern __shared__ uint32_t hsh[];
  for (int i = threadIdx.x; i < nb; i += HB) hsh[i] = 0;
  __syncthreads();
  int kmin = INT32_MAX, kmax = INT32_MIN;
  auto note = [&](int32_t k) {
    kmin = min(kmin, k);
    kmax = max(kmax, k);
  };
  const uint64_t lo = (uint64_t)blockIdx.x * chunk;
  const uint64_t hi = lo + chunk < n ? lo + chunk : n;
  const uint64_t nvec = lo < hi ? (hi - lo) >> 2 : 0;
  if (nvec && ((uintptr_t)(keys + lo) & 15) == 0) {
    // Sets of U vectors per thread, two sets in flight: the loads of the next
    // set are issued before the LDS atomics of the current one.  Loads are
    // unconditional (clamped to the last vector) so the compiler keeps them
    // in flight; a clamped duplicate adds 0.
    constexpr int U = 4;
    const int4 *kv = reinterpret_cast<const int4 *>(keys + lo);
    const uint64_t last = nvec - 1;
    const uint64_t nsets = (nvec + U * HB - 1) / (U * HB);
    auto vix = [&](uint64_t set, int u) {
      return (set * U + u) * HB + threadIdx.x;
    };
    auto load = [&](uint64_t set, int4 (&k)[U]) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint64_t v = vix(set, u);
        k[u] = ld_nt4(kv + (v < last ? v : last));
      }
    };
    auto count = [&](uint64_t set, const int4 (&k)[U]) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint32_t inc = vix(set, u) < nvec ? 1u : 0u;
        atomicAdd(&hsh[joint(k[u].x)], inc);
        atomicAdd(&hsh[joint(k[u].y)], inc);
        atomicAdd(&hsh[joint(k[u].z)], inc);
        atomicAdd(&hsh[joint(k[u].w)], inc);
        if (Joint::kRange) {  // clamped duplicates are keys of the chunk too
          note(k[u].x);
          note(k[u].y);
          note(k[u].z);
          note(k[u].w);
        }
      }
    };
    int4 A[U], B[U];
    load(0, A);
    uint64_t s = 0;
    for (; s + 1 < nsets; s += 2) {
      load(s + 1, B);
      count(s, A);
      load(s + 2 < nsets ? s + 2 : nsets - 1, A);
      count(s + 1, B);
    }
    if (s < nsets) count(s, A);
    const uint64_t j = lo + nvec * 4 + threadIdx.x;  // < 4 tail keys
    if (j < hi) {
      atomicAdd(&hsh[joint(keys[j])], 1u);
      if (Joint::kRange) note(keys[j]);
    }
  } else if (lo < hi) {
    for (uint64_t j = lo + threadIdx.x; j < hi; j += HB) {
      atomicAdd(&hsh[joint(keys[j])], 1u);
      if (Joint::kRange) note(keys[j]);
    }
  }
  if constexpr (Joint::kRange) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      kmin = min(kmin, __shfl_xor(kmin, o, 64));
      kmax = max(kmax, __shfl_xor(kmax, o, 64));
    }
    if ((threadIdx.x & 63) == 0 && kmin <= kmax) {
      atomicMin(&joint.range[0], kmin);
      atomicMax(&joint.range[1], kmax);
    }
  }
  __syncthreads();
  uint32_t *row = H + (uint64_t)blockIdx.x * nb;
  for (int i = threadIdx.x; i < nb; i += HB) row[i] = hsh[i];
}

// Column-wise exclusive scan over chunks, E[c][d] = sum_{c'<c} H[c'][d] and
// T[d] = sum over all chunks, as two launches over (bin, slab of `slab`
// chunks) threads: k_colsum writes each slab's per-bin sum to P[slab][d];
// k_colscan adds the sums of the earlier slabs and walks its own.
__global__ void k_colsum(const uint32_t *__restrict__ H, int G, int nb, int slab,
                         uint32_t *__restrict__ P);
__global__ void k_colscan(const uint32_t *__restrict__ H, int G, int nb, int slab,
                          const uint32_t *__restrict__ P, uint32_t *__restrict__ E,
                          uint32_t *__restrict__ T);
// Number of slabs (return) and chunks per slab for G chunks and nb bins.
int colscan_slabs(const pj_ctx *ctx, int G, int nb, int *slab);
// Exclusive scan of T[0..nb) into S[0..nb], and *ctr = 0 (the pass-2 item
// counter), in one workgroup of kScanBlock threads: a small workgroup finds a
// free slot sooner while the other side's passes hold the CUs.
constexpr int kScanBlock = 256;
__global__ void k_excl_scan(const uint32_t *__restrict__ T, int nb,
                            unsigned long long *__restrict__ S, uint32_t *__restrict__ ctr);

// Pass-2 work items sized by the pass-1 digits' totals (one workgroup of
// kScanBlock threads, 2^b1 <= kScanBlock): digit d1 of T1 tuples is cut into
// ceil(G / cpg) items of cpg consecutive chunks, cpg = clamp(target * G /
// T1, 1, G), so an item holds about `target` tuples however skewed the
// digits are (a hot key's digit is cut into single chunks).
// items[0 .. F1) = first item of each digit, items[F1] = the item count,
// items[F1 + 1 + d1] = cpg of digit d1.
__global__ void k_pass2_items(const uint32_t *__restrict__ T, int b1, int b2, int G,
                              unsigned long long target, uint32_t *__restrict__ items);

// One radix pass.  PASS2 = false: workgroup c partitions input chunk c on the
// pass-1 digit.  PASS2 = true: workgroups take (d1, group of cpg chunks) items
// of the pass-1 output from a counter and partition them on the pass-2 digit.
// ROWID: the V1 (int32) column is the input row index.
template <int F, int T, int BLOCK, class Joint, class V1, class V2, bool ROWID, bool PASS2,
          int LINE = kLine, bool HASHKEYS = false, bool VEC = false>
__global__ __launch_bounds__(BLOCK) void k_radix_pass(
    const int32_t *__restrict__ in_k, const V1 *__restrict__ in_v,
    const V2 *__restrict__ in_w, int32_t *__restrict__ out_k, V1 *__restrict__ out_v,
    V2 *__restrict__ out_w, uint64_t n, uint64_t chunk, int G, int nb, int b2, uint32_t cpg,
    Joint joint,
    const uint32_t *__restrict__ H, const uint32_t *__restrict__ E,
    const unsigned long long *__restrict__ S, uint32_t *__restrict__ seg_counter,
    const uint32_t *__restrict__ items) {
  constexpr int IPT = T / BLOCK;
  static_assert(IPT % 4 == 0, "tile must give each thread a multiple of 4 items");
  static_assert(F <= BLOCK, "one thread per digit");
  static_assert(!ROWID || sizeof(V1) == 4, "row ids are int32");
  static_assert(!(ROWID && PASS2), "row ids only in pass 1");
  constexpr bool HW = ValIO<V2>::present;
  static_assert(LINE == 16 || LINE == 32, "16- or 32-tuple lines");
  constexpr int LOGL = LINE == 32 ? 5 : 4;
  constexpr int MAXLINES = T / LINE + 2 * F;
  constexpr int TW = HW ? T : 1;
  constexpr int FW = HW ? F : 1;

  __shared__ int32_t st_k[T];
  __shared__ V1 st_v[T];
  __shared__ V2 st_w[TW];
  __shared__ int32_t lf_k[F][LINE];
  __shared__ V1 lf_v[F][LINE];
  __shared__ V2 lf_w[FW][LINE];
  // per-digit state: hist / off are dense (rank atomics and per-tuple reads
  // spread over all banks), the rest is read per line with one wide load
  struct Meta {
    unsigned long long cursor;  // next output position of the digit
    uint32_t lcnt;              // tuples waiting in the digit's line buffer
    uint32_t nwrite;            // tuples written this tile
    uint32_t linepre;           // first line of the digit in this tile
    uint32_t pad;
  };
  __shared__ Meta meta[F];
  __shared__ uint32_t hist[F];  // tuples of the digit in this tile
  __shared__ uint32_t off[F];   // start of the digit's run in the stage
  __shared__ uint16_t line_digit[MAXLINES];
  __shared__ uint32_t scan_tmp[BLOCK / 64];
  __shared__ uint32_t s_total_lines;
  __shared__ unsigned long long s_seg_lo, s_seg_hi;
  __shared__ uint32_t s_d1, s_c0, s_c1, s_live;
  // pass 2: the item table, read once (a claim then searches LDS, not HBM)
  __shared__ uint32_t s_items[PASS2 ? 2 * 256 + 2 : 1];

  const int tid = threadIdx.x;
  const uint32_t F2 = 1u << b2;
  const uint32_t F1 = (uint32_t)nb >> b2;  // pass-1 digits
  // number of digits of this pass (<= F)
  const uint32_t ND = PASS2 ? F2 : F1;
  if (PASS2 && items) {
    for (uint32_t i = tid; i < 2 * F1 + 1; i += BLOCK) s_items[i] = items[i];
    __syncthreads();
  }

  for (;;) {
    // ---- pick the segment and set its cursors -------------------------------
    if (!PASS2) {
      const uint64_t c = blockIdx.x;
      if (tid == 0) {
        uint64_t a = c * chunk, b = a + chunk;
        s_seg_lo = a < n ? a : n;
        s_seg_hi = b < n ? b : n;
      }
      // cursor[d1] = S[d1 << b2] + sum_{d2} E[c][d1][d2]
      const uint64_t row = c * (uint64_t)nb;
      for (int d = tid; d < F; d += BLOCK) {
        unsigned long long cur = 0;
        if ((uint32_t)d < F1) {
          cur = S[(uint64_t)d << b2];
          for (uint32_t d2 = 0; d2 < F2; ++d2) cur += E[row + ((uint64_t)d << b2) + d2];
        }
        meta[d].cursor = cur;
        meta[d].lcnt = 0;
        hist[d] = 0;
      }
    } else {
      // work item = (d1, group of consecutive chunks): the runs those chunks
      // wrote for d1 are adjacent, so the item is one contiguous range.  With
      // `items` (k_pass2_items) digit d1 has its own group size, so items
      // hold about the same number of tuples whatever the skew; without, every
      // digit is cut into groups of `cpg` chunks.
      if (tid == 0) {
        const uint32_t seg = atomicAdd(seg_counter, 1u);
        uint32_t d1 = 0, c0 = 0, c1 = 0, live = 0;
        if (items) {
          if (seg < s_items[F1]) {
            uint32_t lo = 0, hi = F1 - 1;  // the last digit whose first item <= seg
            while (lo < hi) {
              const uint32_t mid = (lo + hi + 1) >> 1;
              if (s_items[mid] <= seg) lo = mid; else hi = mid - 1;
            }
            const uint32_t gs = s_items[F1 + 1 + lo];
            d1 = lo;
            c0 = (seg - s_items[lo]) * gs;
            c1 = c0 + gs < (uint32_t)G ? c0 + gs : (uint32_t)G;
            live = 1;
          }
        } else {
          const uint32_t ngrp = ((uint32_t)G + cpg - 1) / cpg;
          if (seg < ngrp * F1) {
            d1 = seg / ngrp;
            c0 = (seg % ngrp) * cpg;
            c1 = c0 + cpg < (uint32_t)G ? c0 + cpg : (uint32_t)G;
            live = 1;
          }
        }
        s_d1 = d1;
        s_c0 = c0;
        s_c1 = c1;
        s_live = live;
      }
      __syncthreads();
      if (!s_live) break;
      const uint32_t d1 = s_d1, c0 = s_c0, c1 = s_c1;
      const uint64_t row0 = (uint64_t)c0 * (uint64_t)nb + ((uint64_t)d1 << b2);
      const uint64_t row1 = (uint64_t)c1 * (uint64_t)nb + ((uint64_t)d1 << b2);
      if (tid < 64) {
        unsigned long long pre0 = 0, pre1 = 0;
        for (uint32_t d2 = tid; d2 < F2; d2 += 64) {
          pre0 += E[row0 + d2];
          if (c1 < (uint32_t)G) pre1 += E[row1 + d2];
        }
        for (int o = 32; o > 0; o >>= 1) {
          pre0 += __shfl_xor(pre0, o, 64);
          pre1 += __shfl_xor(pre1, o, 64);
        }
        if (tid == 0) {
          const unsigned long long r0 = S[(uint64_t)d1 << b2];
          s_seg_lo = r0 + pre0;
          s_seg_hi = c1 < (uint32_t)G ? r0 + pre1 : S[(uint64_t)(d1 + 1) << b2];
        }
      }
      for (uint32_t d = tid; d < (uint32_t)F; d += BLOCK) {
        meta[d].cursor = d < F2 ? S[((uint64_t)d1 << b2) + d] + E[row0 + d] : 0ULL;
        meta[d].lcnt = 0;
        hist[d] = 0;
      }
    }
    __syncthreads();
    const uint64_t seg_lo = s_seg_lo, seg_hi = s_seg_hi;
    // tiles cover [base0, seg_hi) with base0 = seg_lo rounded down to 4
    const uint64_t base0 = seg_lo & ~3ULL;
    const uint32_t lo0 = (uint32_t)(seg_lo - base0);
    const uint64_t span = seg_hi > base0 ? seg_hi - base0 : 0;
    const uint32_t ntiles = (uint32_t)((span + T - 1) / T);

    // Tiles alternate between register sets A and B: tile t + 1 streams in
    // while tile t is shuffled and written.  Alternating (instead of copying
    // the next tile's registers over the current ones) and, with VEC,
    // unconditional loads keep the compiler from waiting for the prefetch
    // before the current tile is processed.
    // VEC (host-checked: every column 16-byte aligned, n a multiple of 4):
    // whole aligned vectors, indices clamped to the column's last vector;
    // elements outside the tile's [lo, len) are loaded but never used.
    auto load_tile = [&](uint32_t t, int32_t *kk, V1 *vv, V2 *ww) {
      const uint64_t base = base0 + (uint64_t)t * T;
      if constexpr (VEC) {
        const uint64_t vlast = n / 4 - 1;
#pragma unroll
        for (int q = 0; q < IPT / 4; ++q) {
          const uint32_t e = (uint32_t)(q * BLOCK + tid) * 4;
          uint64_t vi = (base + e) / 4;
          vi = vi < vlast ? vi : vlast;
          ValIO<int32_t>::load4_vec(in_k, vi, &kk[4 * q]);
          if constexpr (ROWID) {
#pragma unroll
            for (int c = 0; c < 4; ++c)
              vv[4 * q + c] = __builtin_bit_cast(V1, (int32_t)(base + e + c));
          } else {
            ValIO<V1>::load4_vec(in_v, vi, &vv[4 * q]);
          }
          ValIO<V2>::load4_vec(in_w, vi, &ww[4 * q]);
        }
      } else {
        const uint32_t len = (uint32_t)((seg_hi - base) < (uint64_t)T ? (seg_hi - base) : T);
        const uint32_t lo = t == 0 ? lo0 : 0;
        const bool vec = (((uintptr_t)(in_k + base) & 15) == 0) &&
                         (ROWID || ValIO<V1>::aligned(in_v, base)) &&
                         ValIO<V2>::aligned(in_w, base);
#pragma unroll
        for (int q = 0; q < IPT / 4; ++q) {
          uint32_t e = (uint32_t)(q * BLOCK + tid) * 4;
          ValIO<int32_t>::load4(in_k, base, e, lo, len, vec, &kk[4 * q]);
          if constexpr (ROWID) {
#pragma unroll
            for (int c = 0; c < 4; ++c)
              vv[4 * q + c] = __builtin_bit_cast(V1, (int32_t)(base + e + c));
          } else {
            ValIO<V1>::load4(in_v, base, e, lo, len, vec, &vv[4 * q]);
          }
          ValIO<V2>::load4(in_w, base, e, lo, len, vec, &ww[4 * q]);
        }
      }
    };
    auto process_tile = [&](uint32_t t, int32_t *k, V1 *v, V2 *w) {
      const uint64_t base = base0 + (uint64_t)t * T;
      const uint32_t len = (uint32_t)((seg_hi - base) < (uint64_t)T ? (seg_hi - base) : T);
      const uint32_t lo = t == 0 ? lo0 : 0;
      // ---- digits and LDS ranks (hist was cleared after the last tile) ------
      // HASHKEYS: the output carries fmix32(key), a bijection, so later
      // passes and the join take digits / buckets / equality from it directly
      if constexpr (HASHKEYS) {
#pragma unroll
        for (int i = 0; i < IPT; ++i) k[i] = (int32_t)fmix32((uint32_t)k[i]);
      }
      uint32_t dg[IPT], rk[IPT];
#pragma unroll
      for (int q = 0; q < IPT / 4; ++q) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          uint32_t e = (uint32_t)(q * BLOCK + tid) * 4 + c;
          int i = 4 * q + c;
          if (e >= lo && e < len) {
            uint32_t j = joint(k[i]);
            dg[i] = PASS2 ? (j & (F2 - 1)) : (j >> b2);
            rk[i] = atomicAdd(&hist[dg[i]], 1u);
          } else {
            dg[i] = 0xffffffffu;
            rk[i] = 0;
          }
        }
      }
      __syncthreads();
      // ---- one scan of (run length, line count) packed as 16 + 16 bits -----
      {
        uint32_t hv = 0, nl = 0, nw = 0;
        if ((uint32_t)tid < ND) {
          const Meta m = meta[tid];
          hv = hist[tid];
          const unsigned long long end = m.cursor + m.lcnt + hv;
          const unsigned long long aend = end & ~(unsigned long long)(LINE - 1);
          if (aend > m.cursor) {
            nw = (uint32_t)(aend - m.cursor);
            nl = (uint32_t)(((aend - 1) >> LOGL) - (m.cursor >> LOGL) + 1);
          }
        }
        uint32_t tot;
        const uint32_t pre = block_excl_scan<BLOCK>(hv | (nl << 16), scan_tmp, &tot);
        if (tid < F) {
          off[tid] = pre & 0xffffu;
          meta[tid].nwrite = nw;
          meta[tid].linepre = pre >> 16;
        }
        if (tid == 0) s_total_lines = tot >> 16;
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
        const Meta m = meta[tid];
        if (m.nwrite) {
          const uint32_t nl =
              (uint32_t)(((m.cursor + m.nwrite - 1) >> LOGL) - (m.cursor >> LOGL) + 1);
          for (uint32_t l = 0; l < nl; ++l) line_digit[m.linepre + l] = (uint16_t)tid;
        }
      }
      __syncthreads();
      // ---- write lines: LINE lanes per line ----------------------------------
      {
        const uint32_t nlines = s_total_lines;
        const int hw = tid / LINE, l32 = tid % LINE;
        for (uint32_t L = hw; L < nlines; L += BLOCK / LINE) {
          const uint32_t d = line_digit[L];
          const Meta m = meta[d];
          const unsigned long long pos =
              (((m.cursor >> LOGL) + (L - m.linepre)) << LOGL) + (unsigned long long)l32;
          if (pos >= m.cursor && pos < m.cursor + m.nwrite) {
            const uint32_t j = (uint32_t)(pos - m.cursor);
            int32_t ok;
            V1 ov;
            V2 ow;
            if (j < m.lcnt) {
              ok = lf_k[d][j];
              ov = lf_v[d][j];
              if constexpr (HW) ow = lf_w[d][j];
            } else {
              const uint32_t src = off[d] + j - m.lcnt;
              ok = st_k[src];
              ov = st_v[src];
              if constexpr (HW) ow = st_w[src];
            }
            __builtin_nontemporal_store(ok, out_k + pos);
            __builtin_nontemporal_store(ov, out_v + pos);
            if constexpr (HW) __builtin_nontemporal_store(ow, out_w + pos);
          }
        }
      }
      __syncthreads();
      // ---- keep the remainders (< LINE per digit), LINE lanes per digit -----
      // A digit's first write completes its first (partial) line and later
      // writes are whole lines, so nwrite > lcnt whenever nwrite > 0.
      for (int x = tid; x < F * LINE; x += BLOCK) {
        const int d = x / LINE, s = x % LINE;
        const Meta m = meta[d];
        const uint32_t nn = m.lcnt + hist[d], o = off[d];
        int64_t src = -1;
        if (m.nwrite) {
          if ((uint32_t)s < nn - m.nwrite) src = (int64_t)o + m.nwrite + s - m.lcnt;
        } else if ((uint32_t)s >= m.lcnt && (uint32_t)s < nn) {
          src = (int64_t)o + s - m.lcnt;
        }
        if (src >= 0) {
          lf_k[d][s] = st_k[src];
          lf_v[d][s] = st_v[src];
          if constexpr (HW) lf_w[d][s] = st_w[src];
        }
      }
      __syncthreads();
      if (tid < F) {
        Meta &m = meta[tid];
        m.cursor += m.nwrite;
        m.lcnt = m.lcnt + hist[tid] - m.nwrite;
        hist[tid] = 0;
      }
      __syncthreads();
    };
    int32_t kA[IPT], kB[IPT];
    V1 vA[IPT], vB[IPT];
    V2 wA[IPT], wB[IPT];
    if (ntiles) load_tile(0, kA, vA, wA);
    uint32_t t = 0;
    for (; t + 1 < ntiles; t += 2) {
      load_tile(t + 1, kB, vB, wB);
      process_tile(t, kA, vA, wA);
      // past the last tile: the last tile again (cached), never processed
      load_tile(t + 2 < ntiles ? t + 2 : ntiles - 1, kA, vA, wA);
      process_tile(t + 1, kB, vB, wB);
    }
    if (t < ntiles) process_tile(t, kA, vA, wA);
    // ---- flush the remainders of this segment --------------------------------
    for (int x = tid; x < F * LINE; x += BLOCK) {
      const int d = x / LINE, s = x % LINE;
      if ((uint32_t)s < meta[d].lcnt) {
        const unsigned long long pos = meta[d].cursor + s;
        out_k[pos] = lf_k[d][s];
        out_v[pos] = lf_v[d][s];
        if constexpr (HW) out_w[pos] = lf_w[d][s];
      }
    }
    __syncthreads();
    if (!PASS2) break;
  }
}

// Host side ------------------------------------------------------------------
struct RadixPlan {
  int b1 = 0, b2 = 0;  // fan-out bits of pass 1 / pass 2 (b2 = 0: one pass)
  int nb = 0;          // joint bins (2^(b1+b2), or the bucket count)
  int G = 0;           // pass-1 chunks (workgroups)
  uint64_t chunk = 0;  // tuples per chunk, a multiple of 4
};

// Chunking of n tuples for pass 1: about `per_cu` chunks per CU.
RadixPlan radix_make_plan(pj_ctx *ctx, uint64_t n, int b1, int b2, int nb, int per_cu);

// Device scratch of one side's histograms.
struct RadixHist {
  Scratch H, E, T, S, P, ctr, items;
  std::vector<uint32_t> hT;            // partition sizes (host)
  std::vector<unsigned long long> hS;  // partition starts (host), nb + 1
  int slab = 0, nslab = 0;             // column-scan geometry
  int alloc(pj_ctx *ctx, const RadixPlan &p) {
    nslab = colscan_slabs(ctx, p.G, p.nb, &slab);
    PJ_TRY(P.alloc(ctx, (uint64_t)nslab * p.nb * 4));
    PJ_TRY(H.alloc(ctx, (uint64_t)p.G * p.nb * 4));
    PJ_TRY(E.alloc(ctx, (uint64_t)p.G * p.nb * 4));
    PJ_TRY(T.alloc(ctx, (uint64_t)p.nb * 4));
    PJ_TRY(S.alloc(ctx, (uint64_t)(p.nb + 1) * 8));
    PJ_TRY(ctr.alloc(ctx, 256));
    PJ_TRY(items.alloc(ctx, (2 * 256 + 2) * 4));
    return PJ_OK;
  }
};

// Partition sizes and starts of h to the host (waits for ctx->stream).
inline int radix_readback(pj_ctx *ctx, const RadixPlan &p, RadixHist &h) {
  h.hT.resize(p.nb);
  h.hS.resize(p.nb + 1);
  PJ_TRY(d2h(ctx, h.hT.data(), h.T.p, (uint64_t)p.nb * 4));
  PJ_TRY(d2h(ctx, h.hS.data(), h.S.p, (uint64_t)(p.nb + 1) * 8));
  return PJ_OK;
}

// Histogram + scans for one side into h (device), optionally reading T and S
// back to the host.
template <class Joint>
int radix_histograms(pj_ctx *ctx, const int32_t *d_k, uint64_t n, const RadixPlan &p,
                     Joint joint, RadixHist &h, bool readback) {
  static bool attr_set = false;
  if (!attr_set) {
    (void)hipFuncSetAttribute((const void *)k_chunk_hist<Joint, 1024>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, 131072);
    attr_set = true;
  }
  // (512-thread workgroups, which would leave wave slots for the other
  // side's scans, measured no faster)
  PJ_LAUNCH(ctx, "radix_hist", (k_chunk_hist<Joint, 1024>), p.G, 1024, (size_t)p.nb * 4, d_k, n,
            p.chunk, p.nb, joint, h.H.as<uint32_t>());
  const dim3 cgrid((uint32_t)ceil_div(p.nb, 256), (uint32_t)h.nslab);
  PJ_LAUNCH(ctx, "radix_colsum", k_colsum, cgrid, 256, 0, h.H.as<const uint32_t>(), p.G, p.nb,
            h.slab, h.P.as<uint32_t>());
  PJ_LAUNCH(ctx, "radix_colscan", k_colscan, cgrid, 256, 0, h.H.as<const uint32_t>(), p.G, p.nb,
            h.slab, h.P.as<const uint32_t>(), h.E.as<uint32_t>(), h.T.as<uint32_t>());
  PJ_LAUNCH(ctx, "radix_scan", k_excl_scan, 1, kScanBlock, 0, h.T.as<const uint32_t>(), p.nb,
            h.S.as<unsigned long long>(), h.ctr.as<uint32_t>());
  if (p.b2 > 0) {
    // about 8 pass-2 items per CU (finer items measured slower: the per-item
    // setup outweighs the shorter tail)
    const unsigned long long target =
        std::max<unsigned long long>(1, n / ((uint64_t)ctx->num_cus * 8));
    PJ_LAUNCH(ctx, "radix_items", k_pass2_items, 1, kScanBlock, 0, h.T.as<const uint32_t>(),
              p.b1, p.b2, p.G, target, h.items.as<uint32_t>());
  }
  if (readback) PJ_TRY(radix_readback(ctx, p, h));
  return PJ_OK;
}

// One pass launch (pass2 selects the segment-walking variant).
template <int F, int T, int BLOCK, class Joint, class V1, class V2, bool ROWID, int LINE,
          bool HASHKEYS, bool VEC>
int radix_launch_pass_v(pj_ctx *ctx, const char *name, bool pass2, const int32_t *ik,
                        const V1 *iv, const V2 *iw, int32_t *ok, V1 *ov, V2 *ow, uint64_t n,
                        const RadixPlan &p, Joint joint, RadixHist &h) {
  if (!pass2) {
    PJ_LAUNCH(ctx, name, (k_radix_pass<F, T, BLOCK, Joint, V1, V2, ROWID, false, LINE, HASHKEYS, VEC>),
              p.G, BLOCK, 0, ik, iv, iw, ok, ov, ow, n, p.chunk, p.G, p.nb, p.b2, 1u, joint, h.H.as<const uint32_t>(), h.E.as<const uint32_t>(),
              h.S.as<const unsigned long long>(), h.ctr.as<uint32_t>(), nullptr);
  } else {
    if constexpr (!ROWID) {
      const uint32_t F1 = 1u << p.b1;
      constexpr uint64_t lds = (uint64_t)T * (4 + sizeof(V1) + (ValIO<V2>::present ? sizeof(V2) : 0)) +
                               (uint64_t)F * LINE * (4 + sizeof(V1) + (ValIO<V2>::present ? sizeof(V2) : 0));
      (void)F1;
      const uint64_t per_cu = lds > 80000 ? 1 : 2;
      // items sized by k_pass2_items (radix_histograms); workgroups past the
      // last item exit at their first claim
      const uint64_t grid = (uint64_t)ctx->num_cus * per_cu;
      PJ_LAUNCH(ctx, name, (k_radix_pass<F, T, BLOCK, Joint, V1, V2, false, true, LINE, HASHKEYS, VEC>),
                grid, BLOCK, 0, ik, iv, iw, ok, ov, ow, n, p.chunk, p.G, p.nb, p.b2, 1u, joint, h.H.as<const uint32_t>(), h.E.as<const uint32_t>(),
                h.S.as<const unsigned long long>(), h.ctr.as<uint32_t>(), h.items.as<const uint32_t>());
    } else {
      return fail(PJ_E_ARG, "row-id values only in pass 1");
    }
  }
  return PJ_OK;
}

// Picks the whole-vector load variant when every input column is 16-byte
// aligned and n is a multiple of 4.
template <int F, int T, int BLOCK, class Joint, class V1, class V2, bool ROWID, int LINE = kLine,
          bool HASHKEYS = false>
int radix_launch_pass(pj_ctx *ctx, const char *name, bool pass2, const int32_t *ik,
                      const V1 *iv, const V2 *iw, int32_t *ok, V1 *ov, V2 *ow, uint64_t n,
                      const RadixPlan &p, Joint joint, RadixHist &h) {
  auto a16 = [](const void *q) { return ((uintptr_t)q & 15) == 0; };
  const bool vec = n >= 4 && n % 4 == 0 && a16(ik) && (ROWID || a16(iv)) &&
                   (!ValIO<V2>::present || a16(iw));
  if (vec)
    return radix_launch_pass_v<F, T, BLOCK, Joint, V1, V2, ROWID, LINE, HASHKEYS, true>(
        ctx, name, pass2, ik, iv, iw, ok, ov, ow, n, p, joint, h);
  return radix_launch_pass_v<F, T, BLOCK, Joint, V1, V2, ROWID, LINE, HASHKEYS, false>(
      ctx, name, pass2, ik, iv, iw, ok, ov, ow, n, p, joint, h);
}

}  // namespace pj
