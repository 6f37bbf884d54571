// Device sort of the plan operator `sort` (the reference's GPU sort,
// core/olap/lib/operators/gpu/gpu-sort.cpp with
// core/platform/lib/util/sort/gpu/sort-util.cu, which sorts packed records
// with thrust): a stable LSD radix sort of a row permutation.  For each sort
// expression, from the last to the first, every row's value (read through the
// current permutation) is turned into an unsigned key whose order is the
// column's order, and the (key, row) pairs are stably radix-sorted by it.
// Stability makes the result lexicographic over the expressions, with ties
// kept in input order.
//
// Keys: 32-bit types (int32, bool, dictionary ranks) flip the sign bit into
// a 32-bit key; int64 flips it into a 64-bit key; FP64 (and float32, widened)
// takes its bit pattern, flipped whole when negative and with the sign bit
// set otherwise, after mapping -0.0 to +0.0 so that the two compare equal.
// DESC inverts the key; NONE directions are skipped.
//
// The radix sort (8-bit digits, least significant first):
//   * one read of the keys ORs them and their complements together; a digit
//     position where both ORs leave a byte with no bit set in common is one
//     every key shares, and is skipped (narrow key ranges sort in fewer
//     passes);
//   * per pass, the input is cut into G chunks, one workgroup each (as many
//     as fit the CUs at once); k_rs_chunk_hist counts each chunk's digits,
//     column scans give every (chunk, digit) its exact output offset, and
//     k_rs_scatter walks its chunk tile by tile in order, ranking each tile
//     stably and writing it out digit by digit from LDS.
//   * stable ranking of an 8192-key tile: each of the 8 waves owns a
//     contiguous eighth of it, 16 rows of 64 keys.  Row by row, lanes holding the same digit
//     find each other with eight ballots; a key's rank is its wave's running
//     count of the digit plus the matching lanes before it, and the lowest
//     such lane advances the count.  Wave offsets per digit and the tile's
//     digit starts follow from the counts.  No atomics: the order is fixed by
//     the input order alone.
#include <algorithm>
#include <vector>

#include "pj_internal.hpp"
#include "radix.hpp"
#include "rel_api.hpp"

namespace pj {
namespace rel {

// Scatter geometry per key type: BLOCK threads with IPT keys each make a
// tile; PER_CU workgroups fit a CU's LDS at once.
template <class K>
struct RsCfg;
template <>
struct RsCfg<uint32_t> {
  static constexpr int BLOCK = 512, IPT = 16, PER_CU = 2;
};
template <>
struct RsCfg<unsigned long long> {
  static constexpr int BLOCK = 512, IPT = 16, PER_CU = 1;
};
constexpr int kRsHistWaves = 16;  // k_rs_chunk_hist: 1024 threads

template <class K>
__device__ __forceinline__ uint32_t rs_digit(K k, int shift) {
  return (uint32_t)(k >> shift) & 255u;
}

// Workgroup b writes the OR of the complements of its keys to ors[2 b] and
// the OR of its keys to ors[2 b + 1]: bit i varies among the keys iff it is
// set in both ORs over all workgroups (folded on the host: same-address
// atomics from every wave serialise in L2).
template <class K>
__global__ __launch_bounds__(256) void k_rs_digits(const K *__restrict__ keys, uint64_t n,
                                                   unsigned long long *__restrict__ ors) {
  __shared__ unsigned long long s0[4], s1[4];
  constexpr int R = 8;  // loads in flight per thread
  unsigned long long o0 = 0, o1 = 0;
  const uint64_t stride = (uint64_t)gridDim.x * 256 * R;
  for (uint64_t b = (uint64_t)blockIdx.x * 256 * R + threadIdx.x; b < n; b += stride) {
    K k[R];
#pragma unroll
    for (int j = 0; j < R; ++j) k[j] = b + j * 256 < n ? keys[b + j * 256] : keys[b];
#pragma unroll
    for (int j = 0; j < R; ++j) {
      o0 |= (unsigned long long)(K)~k[j];
      o1 |= (unsigned long long)k[j];
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    o0 |= __shfl_xor(o0, off);
    o1 |= __shfl_xor(o1, off);
  }
  if ((threadIdx.x & 63) == 0) {
    s0[threadIdx.x >> 6] = o0;
    s1[threadIdx.x >> 6] = o1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    ors[2 * blockIdx.x] = s0[0] | s0[1] | s0[2] | s0[3];
    ors[2 * blockIdx.x + 1] = s1[0] | s1[1] | s1[2] | s1[3];
  }
}

// Digit counts of chunk c = [c * chunk, min(n, (c + 1) * chunk)) into row c
// of H[G][256].  One LDS copy per wave spreads same-digit atomics.
template <class K>
__global__ __launch_bounds__(kRsHistWaves * 64) void k_rs_chunk_hist(
    const K *__restrict__ keys, uint64_t n, uint64_t chunk, int shift, uint32_t *__restrict__ H) {
  constexpr int B = kRsHistWaves * 64;
  __shared__ uint32_t h[kRsHistWaves][256];
  const int tid = threadIdx.x, w = tid >> 6;
  for (int i = tid; i < kRsHistWaves * 256; i += B) (&h[0][0])[i] = 0;
  __syncthreads();
  const uint64_t lo = (uint64_t)blockIdx.x * chunk;
  const uint64_t hi = lo + chunk < n ? lo + chunk : n;
  // eight keys per thread in flight before their atomics; a wave whose 64
  // keys share one digit (the skewed high bytes of floating-point keys)
  // adds 64 once instead of 64 atomics on one LDS address
  uint64_t j = lo + tid;
  for (; j + 7 * B < hi; j += 8 * B) {
    K x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) x[u] = keys[j + u * B];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const uint32_t d = rs_digit(x[u], shift);
      const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
      if (__ballot(d == d0) == ~0ULL) {
        if ((tid & 63) == 0) atomicAdd(&h[w][d0], 64u);
      } else {
        atomicAdd(&h[w][d], 1u);
      }
    }
  }
  for (; j < hi; j += B) atomicAdd(&h[w][rs_digit(keys[j], shift)], 1u);
  __syncthreads();
  if (tid < 256) {
    uint32_t s = 0;
#pragma unroll
    for (int x = 0; x < kRsHistWaves; ++x) s += h[x][tid];
    H[(uint64_t)blockIdx.x * 256 + tid] = s;
  }
}

// One stable pass: chunk c of (ik, iv) to its (digit, chunk) ranges of
// (ok, ov): digit d of chunk c starts at S[d] + E[c][d].
template <class K, int BLOCK = RsCfg<K>::BLOCK, int IPT = RsCfg<K>::IPT>
__global__ __launch_bounds__(BLOCK) void k_rs_scatter(
    const K *__restrict__ ik, const uint32_t *__restrict__ iv, K *__restrict__ ok,
    uint32_t *__restrict__ ov, uint64_t n, uint64_t chunk, int shift,
    const uint32_t *__restrict__ E, const unsigned long long *__restrict__ S) {
  constexpr int kRsTile = BLOCK * IPT, kRsWaves = BLOCK / 64, kRsIpt = IPT;
  static_assert(BLOCK >= 256, "one thread per digit");
  __shared__ __align__(16) K st_k[kRsTile];
  __shared__ uint32_t st_v[kRsTile];
  __shared__ uint32_t cntw[kRsWaves][256];  // per wave: running digit counts
  __shared__ unsigned long long cursor[256];
  __shared__ uint32_t start[256];
  __shared__ uint32_t scan_tmp[kRsWaves];

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const bool dthread = tid < 256;  // threads 0..255 own one digit each
  const unsigned long long lt = (1ull << lane) - 1ull;
  const uint64_t c = blockIdx.x;
  const uint64_t lo = c * chunk;
  const uint64_t hi = lo + chunk < n ? lo + chunk : n;
  if (dthread) cursor[tid] = S[tid] + E[c * 256 + tid];
  const uint64_t span = hi > lo ? hi - lo : 0;
  const uint32_t ntiles = (uint32_t)((span + kRsTile - 1) / kRsTile);
  // item i of a thread is tile element w * 64 * kRsIpt + i * 64 + lane: each
  // wave owns a contiguous quarter of the tile, so input order is (wave,
  // item, lane) and a wave's running counts rank its keys
  auto elem = [&](int i) { return (uint32_t)(w * 64 * kRsIpt + i * 64 + lane); };

  K kA[kRsIpt], kB[kRsIpt];
  uint32_t vA[kRsIpt], vB[kRsIpt];
  auto load = [&](uint32_t t, K *k, uint32_t *v) {
    const uint64_t base = lo + (uint64_t)t * kRsTile;
    const uint32_t len = (uint32_t)(hi - base < (uint64_t)kRsTile ? hi - base : kRsTile);
#pragma unroll
    for (int i = 0; i < kRsIpt; ++i) {
      const uint32_t e = elem(i);
      k[i] = e < len ? ik[base + e] : K(0);
      v[i] = e < len ? iv[base + e] : 0u;
    }
  };
  auto process = [&](uint32_t t, const K *k, const uint32_t *v) {
    const uint64_t base = lo + (uint64_t)t * kRsTile;
    const uint32_t len = (uint32_t)(hi - base < (uint64_t)kRsTile ? hi - base : kRsTile);
    for (int x = tid; x < kRsWaves * 256; x += BLOCK) (&cntw[0][0])[x] = 0;
    __syncthreads();
    uint32_t dg[kRsIpt], rk[kRsIpt];
#pragma unroll
    for (int i = 0; i < kRsIpt; ++i) {
      const bool valid = elem(i) < len;
      const uint32_t d = valid ? rs_digit(k[i], shift) : 0u;
      unsigned long long peers = __ballot(valid);
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const unsigned long long m = __ballot(bit);
        peers &= bit ? m : ~m;
      }
      const uint32_t before = (uint32_t)__popcll(peers & lt);
      const uint32_t run = cntw[w][d];  // every lane reads before the leader writes
      dg[i] = d;
      rk[i] = run + before;
      if (valid && before == 0) cntw[w][d] = run + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    // per digit (one per thread): wave counts -> wave offsets, tile total
    uint32_t tot_d = 0;
    if (dthread) {
#pragma unroll
      for (int x = 0; x < kRsWaves; ++x) {
        const uint32_t cx = cntw[x][tid];
        cntw[x][tid] = tot_d;
        tot_d += cx;
      }
    }
    uint32_t tot;
    const uint32_t st = block_excl_scan<BLOCK, uint32_t>(tot_d, scan_tmp, &tot);
    if (dthread) start[tid] = st;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kRsIpt; ++i) {
      if (elem(i) < len) {
        const uint32_t p = start[dg[i]] + cntw[w][dg[i]] + rk[i];
        st_k[p] = k[i];
        st_v[p] = v[i];
      }
    }
    __syncthreads();
    for (uint32_t p = tid; p < len; p += BLOCK) {
      const K key = st_k[p];
      const uint32_t d = rs_digit(key, shift);
      const unsigned long long o = cursor[d] + (p - start[d]);
      ok[o] = key;
      ov[o] = st_v[p];
    }
    __syncthreads();
    if (dthread) cursor[tid] += tot_d;
    // the next tile's first barrier orders this update before its reads
  };
  __syncthreads();
  if (ntiles) load(0, kA, vA);
  uint32_t t = 0;
  for (; t + 1 < ntiles; t += 2) {
    load(t + 1, kB, vB);
    process(t, kA, vA);
    if (t + 2 < ntiles) load(t + 2, kA, vA);
    process(t + 1, kB, vB);
  }
  if (t < ntiles) process(t, kA, vA);
}

// Sorts (keys, vals) by keys, stably.  On return *keys / *vals point at the
// sorted arrays, which are either the inputs or the alternates.
template <class K>
int radix_sort_pairs(pj_ctx *ctx, K **keys, K **keys_alt, uint32_t **vals, uint32_t **vals_alt,
                     uint64_t n) {
  constexpr int P = sizeof(K);
  if (n < 2) return PJ_OK;
  const uint64_t dgrid = std::min<uint64_t>(ceil_div(n, (uint64_t)256 * 8), (uint64_t)ctx->num_cus * 4);
  Scratch dh;
  PJ_TRY(dh.alloc(ctx, dgrid * 16));
  PJ_LAUNCH(ctx, "sort_digits", k_rs_digits<K>, dgrid, 256, 0, *keys, n,
            dh.as<unsigned long long>());
  std::vector<unsigned long long> ors(dgrid * 2);
  PJ_TRY(d2h(ctx, ors.data(), dh.p, dgrid * 16));
  unsigned long long o0 = 0, o1 = 0;
  for (uint64_t b = 0; b < dgrid; ++b) {
    o0 |= ors[2 * b];
    o1 |= ors[2 * b + 1];
  }
  const unsigned long long varying = o0 & o1;
  std::vector<int> passes;
  for (int p = 0; p < P; ++p)
    if ((varying >> (8 * p)) & 255u) passes.push_back(p);
  if (passes.empty()) return PJ_OK;
  // workgroups that fit the CUs at once (LDS-bound), chunks whole tiles
  constexpr uint64_t kTile = (uint64_t)RsCfg<K>::BLOCK * RsCfg<K>::IPT;
  uint64_t G = std::min<uint64_t>((uint64_t)ctx->num_cus * RsCfg<K>::PER_CU, ceil_div(n, kTile));
  const uint64_t chunk = ceil_div(ceil_div(n, G), kTile) * kTile;
  G = ceil_div(n, chunk);
  int slab = 0;
  const int nslab = colscan_slabs(ctx, (int)G, 256, &slab);
  Scratch H, E, T, S, Pp, ctr;
  PJ_TRY(H.alloc(ctx, G * 256 * 4));
  PJ_TRY(E.alloc(ctx, G * 256 * 4));
  PJ_TRY(T.alloc(ctx, 256 * 4));
  PJ_TRY(S.alloc(ctx, 257 * 8));
  PJ_TRY(Pp.alloc(ctx, (uint64_t)nslab * 256 * 4));
  PJ_TRY(ctr.alloc(ctx, 256));
  for (int p : passes) {
    const int shift = 8 * p;
    PJ_LAUNCH(ctx, "sort_hist", k_rs_chunk_hist<K>, G, kRsHistWaves * 64, 0, *keys, n, chunk, shift,
              H.as<uint32_t>());
    const dim3 cgrid(1, (uint32_t)nslab);
    PJ_LAUNCH(ctx, "sort_colsum", k_colsum, cgrid, 256, 0, H.as<const uint32_t>(), (int)G, 256,
              slab, Pp.as<uint32_t>());
    PJ_LAUNCH(ctx, "sort_colscan", k_colscan, cgrid, 256, 0, H.as<const uint32_t>(), (int)G,
              256, slab, Pp.as<const uint32_t>(), E.as<uint32_t>(), T.as<uint32_t>());
    PJ_LAUNCH(ctx, "sort_scan", k_excl_scan, 1, kScanBlock, 0, T.as<const uint32_t>(), 256,
              S.as<unsigned long long>(), ctr.as<uint32_t>());
    PJ_LAUNCH(ctx, "sort_scatter", k_rs_scatter<K>, G, RsCfg<K>::BLOCK, 0, *keys, *vals, *keys_alt,
              *vals_alt, n, chunk, shift, E.as<const uint32_t>(), S.as<const unsigned long long>());
    std::swap(*keys, *keys_alt);
    std::swap(*vals, *vals_alt);
  }
  return PJ_OK;
}

// The sort key of row perm[i] of a column whose rows are `stride` bytes
// apart (K = uint32_t for 32-bit types, uint64_t otherwise).
template <class K>
__global__ __launch_bounds__(256) void k_sort_key(const char *__restrict__ col, uint64_t stride,
                                                  int t, int desc,
                                                  const int32_t *__restrict__ rank, uint32_t nrank,
                                                  const uint32_t *__restrict__ perm, uint64_t n,
                                                  K *__restrict__ key) {
  // four rows per thread: their row ids, then their values, are loaded
  // together so the gathers' latencies overlap
  constexpr int R = 4;
  const uint64_t gs = (uint64_t)gridDim.x * 256 * R;
  const bool wide = t == I64 || t == F64;
  for (uint64_t b = (uint64_t)blockIdx.x * 256 * R + threadIdx.x; b < n; b += gs) {
    bool ok[R];
    uint32_t pr[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
      ok[j] = b + j * 256 < n;
      pr[j] = perm[ok[j] ? b + j * 256 : b];
    }
    unsigned long long raw[R];
    if (wide) {
#pragma unroll
      for (int j = 0; j < R; ++j) raw[j] = *(const unsigned long long *)(col + (uint64_t)pr[j] * stride);
    } else if (t == BOOL) {
#pragma unroll
      for (int j = 0; j < R; ++j) raw[j] = *(const uint8_t *)(col + (uint64_t)pr[j] * stride);
    } else {  // I32, F32
#pragma unroll
      for (int j = 0; j < R; ++j) raw[j] = *(const uint32_t *)(col + (uint64_t)pr[j] * stride);
    }
#pragma unroll
    for (int j = 0; j < R; ++j) {
      if (!ok[j]) continue;
      unsigned long long u;
      switch (t) {
        case I32: {
          int32_t v = (int32_t)(uint32_t)raw[j];
          if (rank) v = (uint32_t)v < nrank ? rank[v] : (int32_t)nrank;  // unknown codes last
          u = (uint32_t)v ^ 0x80000000u;
          if (desc) u = ~u & 0xffffffffULL;
          break;
        }
        case BOOL: {
          u = raw[j] != 0;
          if (desc) u = ~u & 0xffffffffULL;
          break;
        }
        case I64: {
          u = raw[j] ^ 0x8000000000000000ULL;
          if (desc) u = ~u;
          break;
        }
        default: {  // F64, F32
          double d = t == F32 ? (double)__builtin_bit_cast(float, (uint32_t)raw[j])
                              : __builtin_bit_cast(double, raw[j]);
          if (d == 0.0) d = 0.0;
          const unsigned long long bb = __builtin_bit_cast(unsigned long long, d);
          u = (bb >> 63) ? ~bb : (bb | 0x8000000000000000ULL);
          if (desc) u = ~u;
          break;
        }
      }
      key[b + j * 256] = (K)u;
    }
  }
}

__global__ __launch_bounds__(256) void k_sort_iota(uint32_t *__restrict__ p, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) p[i] = (uint32_t)i;
}

static int type_bytes(int t) {
  return t == I64 || t == F64 ? 8 : t == BOOL ? 1 : 4;
}

// Small sorts (n <= kRankSortMax rows): every row is ranked by comparing it
// with every other row on the order-preserving keys of all sort expressions
// at once (key 0 most significant), ties by row index, so the result is the
// stable lexicographic order; perm[rank] = row.  Workgroup b ranks rows
// [64 b, 64 b + 64), one per lane; its 16 waves split the rows compared
// against into sixteenths and add their partial ranks in LDS, so a lane's
// chain of dependent LDS compares is n / 16 long.  The keys (nk x n u64)
// sit in LDS of every workgroup and every lane of a wave reads the same row
// j at a time (broadcast).  A handful of launches replace the radix passes'
// ten per key, which at hundreds of rows are launch latency.
constexpr uint64_t kRankSortMax = 2048;
constexpr uint64_t kRankSortLds = 65536;  // bytes of keys in LDS
constexpr int kRankWaves = 16;
__global__ __launch_bounds__(1024) void k_rank_sort(const unsigned long long *__restrict__ keys,
                                                    int nk, uint32_t n,
                                                    uint32_t *__restrict__ perm) {
  extern __shared__ unsigned long long sk[];
  __shared__ uint32_t part[kRankWaves][64];
  for (uint32_t x = threadIdx.x; x < (uint32_t)nk * n; x += blockDim.x) sk[x] = keys[x];
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t i = blockIdx.x * 64 + lane;
  const uint32_t per = (n + kRankWaves - 1) / kRankWaves;
  const uint32_t j0 = w * per, j1 = min(n, j0 + per);
  uint32_t rank = 0;
  if (i < n)
    for (uint32_t j = j0; j < j1; ++j) {
      // row j before row i: first differing key smaller, or all equal and j < i
      int c = 0;
      for (int k = 0; k < nk && c == 0; ++k) {
        const unsigned long long a = sk[(uint64_t)k * n + j], b = sk[(uint64_t)k * n + i];
        c = a < b ? -1 : (a > b ? 1 : 0);
      }
      rank += (c < 0 || (c == 0 && j < i)) ? 1u : 0u;
    }
  part[w][lane] = rank;
  __syncthreads();
  if (w == 0 && i < n) {
    uint32_t r = 0;
#pragma unroll
    for (int q = 0; q < kRankWaves; ++q) r += part[q][lane];
    perm[r] = i;
  }
}

// int32 values back from their ascending (desc = 0) or descending 32-bit
// sort keys: the inverse of k_sort_key's I32 case without ranks.
__global__ __launch_bounds__(256) void k_sort_unkey_i32(const uint32_t *__restrict__ key,
                                                        uint64_t n, int desc,
                                                        int32_t *__restrict__ out) {
  const uint64_t gs = (uint64_t)gridDim.x * 256;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += gs) {
    const uint32_t u = desc ? ~key[i] : key[i];
    out[i] = (int32_t)(u ^ 0x80000000u);
  }
}

// Sorts the row permutation perm (n entries, initialised here) by the keys;
// strides[k] (bytes) defaults to the column's element size.  key0_out (may
// be null): when key 0 is a plain int32 key with a direction and the radix
// passes ran, its values in sorted order (n entries) and *key0_done = true.
static int sort_perm_strided(pj_ctx *ctx, const SortKey *keys, const uint64_t *strides, int nk,
                             uint64_t n, uint32_t *perm, int32_t *key0_out = nullptr,
                             bool *key0_done = nullptr) {
  if (nk < 0 || (nk && !keys) || (n && !perm)) return fail(PJ_E_ARG, "sort with invalid arguments");
  if (!n) return PJ_OK;
  if (n > (uint64_t)INT32_MAX) return fail(PJ_E_CAPACITY, "sort of more than 2^31-1 rows");
  bool wide = false;
  for (int k = 0; k < nk; ++k) {
    if (!keys[k].col) return fail(PJ_E_ARG, "sort key %d without a column", k);
    if (keys[k].t != I32 && keys[k].t != I64 && keys[k].t != F64 && keys[k].t != F32 &&
        keys[k].t != BOOL)
      return fail(PJ_E_ARG, "sort key %d of unknown type %d", k, keys[k].t);
    if (keys[k].rank && (keys[k].t != I32 || keys[k].nrank > (uint32_t)INT32_MAX))
      return fail(PJ_E_ARG, "sort key %d: ranks need int32 codes", k);
    wide |= keys[k].dir && !(keys[k].t == I32 || keys[k].t == BOOL);
  }
  uint64_t grid = ceil_div(n, (uint64_t)256);
  if (grid > (uint64_t)ctx->num_cus * 16) grid = (uint64_t)ctx->num_cus * 16;
  int nka = 0;  // sort expressions with a direction
  for (int k = 0; k < nk; ++k) nka += keys[k].dir != 0;
  if (n <= kRankSortMax && (uint64_t)nka * n * 8 <= kRankSortLds) {
    if (!nka) {
      PJ_LAUNCH(ctx, "sort_iota", k_sort_iota, grid, 256, 0, perm, n);
      return sync(ctx);
    }
    Scratch iota, kk;
    PJ_TRY(iota.alloc(ctx, n * 4));
    PJ_TRY(kk.alloc(ctx, (uint64_t)nka * n * 8));
    PJ_LAUNCH(ctx, "sort_iota", k_sort_iota, grid, 256, 0, iota.as<uint32_t>(), n);
    int at = 0;
    for (int k = 0; k < nk; ++k) {
      if (!keys[k].dir) continue;
      const uint64_t stride = strides ? strides[k] : (uint64_t)type_bytes(keys[k].t);
      PJ_LAUNCH(ctx, "sort_key", k_sort_key<unsigned long long>, grid, 256, 0,
                (const char *)keys[k].col, stride, keys[k].t, keys[k].dir < 0 ? 1 : 0,
                keys[k].rank, keys[k].nrank, iota.as<const uint32_t>(), n,
                kk.as<unsigned long long>() + (uint64_t)at * n);
      ++at;
    }
    PJ_LAUNCH(ctx, "sort_rank", k_rank_sort, (uint32_t)ceil_div(n, (uint64_t)64), 1024,
              (size_t)nka * n * 8,
              kk.as<const unsigned long long>(), nka, (uint32_t)n, perm);
    return sync(ctx);
  }
  Scratch p1, k0, k1;
  const uint32_t *k0_sorted = nullptr;  // key 0's 32-bit sort keys after its passes
  PJ_TRY(p1.alloc(ctx, n * 4));
  PJ_TRY(k0.alloc(ctx, n * (wide ? 8 : 4)));
  PJ_TRY(k1.alloc(ctx, n * (wide ? 8 : 4)));
  PJ_LAUNCH(ctx, "sort_iota", k_sort_iota, grid, 256, 0, perm, n);
  uint32_t *va = perm, *vb = p1.as<uint32_t>();
  for (int k = nk - 1; k >= 0; --k) {
    if (!keys[k].dir) continue;
    const uint64_t stride = strides ? strides[k] : (uint64_t)type_bytes(keys[k].t);
    const int desc = keys[k].dir < 0 ? 1 : 0;
    if (keys[k].t == I32 || keys[k].t == BOOL) {
      uint32_t *ka = k0.as<uint32_t>(), *kb = k1.as<uint32_t>();
      PJ_LAUNCH(ctx, "sort_key", k_sort_key<uint32_t>, grid, 256, 0, (const char *)keys[k].col,
                stride, keys[k].t, desc, keys[k].rank, keys[k].nrank, va, n, ka);
      PJ_TRY(radix_sort_pairs<uint32_t>(ctx, &ka, &kb, &va, &vb, n));
      if (k == 0) k0_sorted = ka;
    } else {
      unsigned long long *ka = k0.as<unsigned long long>(), *kb = k1.as<unsigned long long>();
      PJ_LAUNCH(ctx, "sort_key", k_sort_key<unsigned long long>, grid, 256, 0,
                (const char *)keys[k].col, stride, keys[k].t, desc, keys[k].rank, keys[k].nrank,
                va, n, ka);
      PJ_TRY(radix_sort_pairs<unsigned long long>(ctx, &ka, &kb, &va, &vb, n));
    }
  }
  if (va != perm) PJ_HIP(hipMemcpyAsync(perm, va, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
  // key 0 was sorted last: its sort keys are in the final order
  if (key0_out && key0_done && nk >= 1 && keys[0].dir && keys[0].t == I32 && !keys[0].rank) {
    PJ_LAUNCH(ctx, "sort_unkey", k_sort_unkey_i32, grid, 256, 0, k0_sorted, n,
              keys[0].dir < 0 ? 1 : 0, key0_out);
    *key0_done = true;
  }
  return sync(ctx);
}

int sort_perm(pj_ctx *ctx, const SortKey *keys, int nk, uint64_t n, uint32_t *perm,
              int32_t *key0_out, bool *key0_done) {
  if (key0_done) *key0_done = false;
  return sort_perm_strided(ctx, keys, nullptr, nk, n, perm, key0_out, key0_done);
}

// out row i = in row perm[i], rows of `words` 4-byte words
__global__ __launch_bounds__(256) void k_gather_rows(const uint32_t *__restrict__ in,
                                                     const uint32_t *__restrict__ perm,
                                                     uint64_t n, uint32_t words,
                                                     uint32_t *__restrict__ out) {
  const uint64_t total = n * words;
  const uint64_t gs = (uint64_t)gridDim.x * 256;
  for (uint64_t x = (uint64_t)blockIdx.x * 256 + threadIdx.x; x < total; x += gs) {
    const uint64_t r = x / words, w = x - r * words;
    out[x] = in[(uint64_t)perm[r] * words + w];
  }
}

}  // namespace rel
}  // namespace pj

using namespace pj;
using namespace pj::rel;

// qsort_<fields> of core/platform/lib/util/sort/gpu/sort-util.hpp:27-58.
extern "C" int pj_sort_records(pj_ctx *ctx, void *d_rows, uint64_t n, const char *fields) {
  if (!ctx || !fields || (n && !d_rows)) return fail(PJ_E_ARG, "pj_sort_records: null argument");
  const int nf = (int)strlen(fields);
  if (nf < 1 || nf > 8) return fail(PJ_E_ARG, "pj_sort_records: 1 to 8 fields, got '%s'", fields);
  // record layout: fields in order at their natural alignment, the record
  // padded to its widest field
  std::vector<SortKey> keys(nf);
  std::vector<uint64_t> offs(nf), strides(nf);
  uint64_t off = 0, align = 4;
  for (int i = 0; i < nf; ++i) {
    const char f = fields[i];
    if (f != 'i' && f != 'l')
      return fail(PJ_E_ARG, "pj_sort_records: field type '%c' is neither 'i' (int32) nor 'l' (int64)", f);
    const uint64_t sz = f == 'i' ? 4 : 8;
    off = (off + sz - 1) / sz * sz;
    offs[i] = off;
    off += sz;
    align = std::max(align, sz);
  }
  const uint64_t rec = (off + align - 1) / align * align;
  if (!n) return PJ_OK;
  if (((uintptr_t)d_rows & (align - 1)) != 0)
    return fail(PJ_E_ARG, "pj_sort_records: rows must be %llu-byte aligned", (unsigned long long)align);
  for (int i = 0; i < nf; ++i) {
    keys[i] = SortKey{(const char *)d_rows + offs[i], fields[i] == 'i' ? I32 : I64, 1, nullptr, 0};
    strides[i] = rec;
  }
  Scratch perm, tmp;
  PJ_TRY(perm.alloc(ctx, n * 4));
  PJ_TRY(tmp.alloc(ctx, n * rec));
  PJ_TRY(sort_perm_strided(ctx, keys.data(), strides.data(), nf, n, perm.as<uint32_t>()));
  uint64_t grid = std::min<uint64_t>(ceil_div(n * (rec / 4), (uint64_t)256), (uint64_t)ctx->num_cus * 16);
  PJ_LAUNCH(ctx, "sort_gather", k_gather_rows, grid, 256, 0, (const uint32_t *)d_rows,
            perm.as<const uint32_t>(), n, (uint32_t)(rec / 4), tmp.as<uint32_t>());
  PJ_HIP(hipMemcpyAsync(d_rows, tmp.p, n * rec, hipMemcpyDeviceToDevice, ctx->stream));
  return sync(ctx);
}
