// groupby / hashgroupby-chained and reduce on one MI355X.
//
// Reference: GpuHashGroupByChained (core/olap/lib/operators/gpu/
// gpu-hash-group-by-chained.cpp:77-556) and HashGroupByChained::generate_scan
// (core/olap/lib/operators/hash-group-by-chained.cpp:638-853): one global
// chained table in HBM, CAS insert + chain walk + one global atomic per
// aggregate per row, then a scan kernel that emits the live entries.
//
// Global atomics on MI355X sustain only ~2e10/s (~6e9/s on 1024 hot
// addresses), so one atomic per aggregate per row would take >100 ms for
// 2^29 rows.  Here every row is aggregated in LDS:
//   * a workgroup owns an LDS open-addressing table of SLOTS slots:
//       slot = 1 << 32 | key (u64, 0 = empty) claimed with one LDS CAS,
//       COUNT/SUM/MIN/MAX (int64) and FSUM (double) updated with LDS atomics;
//   * a "unit" (a row range of one partition) is folded into the table, and
//     the table's live entries are written out compactly (the scan);
//   * when maxInputSize exceeds what one table holds, the rows are first
//     radix-partitioned on the top bits of fmix32(key) (the join's scatter
//     kernel, 16-byte rows), so each partition's groups fit one table;
//   * units of the same partition are merged by the same kernel (entries in,
//     entries out) until one unit per partition remains -- no global atomics;
//   * dense keys skip the hashing.  Partition p holds the keys whose bits
//     [11, 11 + pd) equal p, and its table is direct: slot = key & 2047.  Two
//     keys of one partition with the same slot agree on their low 11 + pd
//     bits, so they differ by a multiple of 2^(11+pd); when the keys' span
//     (max - min + 1) is at most 2^(11+pd) that cannot happen, and the key of
//     slot s is kmin + ((p << 11 | s) - kmin) mod 2^(11+pd).  The histogram
//     pass records min and max on the way (pd > 0); with no partitions the
//     aggregation kernel checks each unit's span itself and abandons the
//     attempt when one exceeds 2048.  A direct table has no probe loop and
//     no CAS, and a slot is live when its count is non-zero.  The probe loop
//     of the hashed table, not the LDS atomics, bounds the hashed kernel: at
//     2^10 groups a direct table aggregates 2^29 rows in 1.5 ms against 2.8.
// Exact int64 arithmetic for COUNT/SUM/MIN/MAX; FSUM accumulates in FP64
// (Proteus float is FP64, core/olap/lib/values/expressionTypes.cpp:87-89).
#include <algorithm>
#include <memory>
#include <type_traits>
#include <vector>

#include "radix.hpp"

namespace pj {

constexpr int kGbSlots = 2048;   // LDS table slots (96 KiB)
constexpr int kGbCap = 1024;     // groups per table the planner aims for
constexpr int kGbBlock = 1024;  // one workgroup per CU (96 KiB of LDS), 16 waves
constexpr int kGbMergeFanIn = 64;     // merge fan-in with many partitions
constexpr int kGbMergeFanInFew = 8;   // with at most 64 partitions
constexpr int kGbDirectLog = 11;      // a direct table covers 2^11 = kGbSlots keys
static_assert((1 << kGbDirectLog) == kGbSlots, "direct range = table size");

// Table modes of k_gb_agg.
enum GbMode : int {
  kGbHashed = 0,  // fmix32 slots, CAS insert
  kGbDirect = 1,  // slot = key & 2047, key = kbase + ((part << 11 | s) - kbase) & wmask
  kGbProbe = 2,   // level 0 without partitions: direct, each unit checks its own span
};

struct GbEntries {  // SoA entry arrays of one level
  int32_t *k;
  int64_t *cnt, *sum, *mn, *mx;
  double *fs;
};

struct GbUnit {
  unsigned long long begin;  // first row (level 0) or first source unit (merge)
  uint32_t len;              // rows (level 0) or source units (merge)
  uint32_t part;
};

__device__ __forceinline__ bool gb_insert(unsigned long long *tk, int32_t key,
                                          uint32_t &slot) {
  const unsigned long long mine = (1ULL << 32) | (uint32_t)key;
  uint32_t h = fmix32((uint32_t)key) & (kGbSlots - 1);
  for (int probe = 0; probe < kGbSlots; ++probe) {
    unsigned long long w = tk[h];
    if (w == mine) {
      slot = h;
      return true;
    }
    if (w == 0ULL) {
      unsigned long long old = atomicCAS(&tk[h], 0ULL, mine);
      if (old == 0ULL || old == mine) {
        slot = h;
        return true;
      }
    }
    h = (h + 1) & (kGbSlots - 1);
  }
  return false;
}

// The slot of `key` in the unit's table: direct modes take its low 11
// bits; hashed finds or claims it with gb_insert (false: the table is full).
template <int MODE>
__device__ __forceinline__ bool gb_slot(unsigned long long *tk, int32_t key, uint32_t &s) {
  if (MODE != kGbHashed) {
    s = (uint32_t)key & (kGbSlots - 1);
    return true;
  }
  return gb_insert(tk, key, s);
}

// MIN / MAX of slot s from v: both bounds are read with one 16-byte load
// (tmm[2s] = min, tmm[2s + 1] = max) and an atomic runs only where v would
// change a bound.  A bound only moves outwards, so a stale read can only
// cause an atomic that changes nothing, never skip one that is needed; on
// hot slots (skewed keys) most rows then need no MIN / MAX atomic at all.
__device__ __forceinline__ void gb_minmax(long long *tmm, uint32_t s, long long v) {
  const longlong2 mm = *reinterpret_cast<const longlong2 *>(&tmm[2 * s]);
  if (v < mm.x) atomicMin(&tmm[2 * s], v);
  if (v > mm.y) atomicMax(&tmm[2 * s + 1], v);
}

// Adds row (key, v, w) to the LDS table; false when the table is full.
template <bool HV, bool HW, int MODE>
__device__ __forceinline__ bool gb_add_row(unsigned long long *tk, unsigned long long *tcnt,
                                           long long *tsum, long long *tmm, double *tfs,
                                           int32_t key, int64_t v, float w) {
  uint32_t s;
  if (!gb_slot<MODE>(tk, key, s)) return false;
  atomicAdd(&tcnt[s], 1ULL);
  if (HV) {
    atomicAdd((unsigned long long *)&tsum[s], (unsigned long long)v);
    gb_minmax(tmm, s, (long long)v);
  }
  if (HW) atomicAdd(&tfs[s], (double)w);
  return true;
}

// One unit per workgroup iteration.  MERGE = inputs are entries of the
// previous level's units; otherwise raw rows (k, v if HV, w if HW).
// MODE kGbDirect: the key of slot s is kbase + (((part << 11) | s) - kbase)
// & wmask.  MODE kGbProbe (level 0, no partitions): each unit's key span is
// checked; a span above 2048 sets *abort and every workgroup stops at its
// next unit, otherwise the unit's own minimum is kbase with wmask 2047; the
// keys' min / max also go to range[0] / range[1].
template <bool MERGE, bool HV, bool HW, int MODE>
__global__ __launch_bounds__(kGbBlock) void k_gb_agg(
    const GbUnit *__restrict__ units, uint32_t nunits,
    const int32_t *__restrict__ rk, const int64_t *__restrict__ rv,
    const float *__restrict__ rw, GbEntries in, const uint32_t *__restrict__ in_count,
    GbEntries out, uint32_t *__restrict__ out_count, uint32_t *__restrict__ overflow,
    bool vec, uint32_t kbase, uint32_t wmask, int *__restrict__ range,
    uint32_t *__restrict__ abort) {
  static_assert(MODE != kGbProbe || !MERGE, "the span check is for level 0");
  __shared__ unsigned long long tk[MODE == kGbHashed ? kGbSlots : 1];
  __shared__ unsigned long long tcnt[kGbSlots];
  __shared__ long long tsum[kGbSlots];
  __shared__ __align__(16) long long tmm[2 * kGbSlots];  // MIN, MAX interleaved
  __shared__ double tfs[kGbSlots];
  __shared__ uint32_t wave_off[kGbBlock / 64 + 1];
  __shared__ int s_lo, s_hi;
  __shared__ uint32_t s_abort;
  const int tid = threadIdx.x;
  // units round-robin over the grid (the host makes many similar units)
  for (uint32_t u = blockIdx.x; u < nunits; u += gridDim.x) {
    __syncthreads();
    for (int i = tid; i < kGbSlots; i += kGbBlock) {
      if (MODE == kGbHashed) tk[i] = 0ULL;
      tcnt[i] = 0ULL;
      tsum[i] = 0;
      tmm[2 * i] = INT64_MAX;
      tmm[2 * i + 1] = INT64_MIN;
      tfs[i] = 0.0;
    }
    if (MODE == kGbProbe && tid == 0) {
      s_lo = INT32_MAX;
      s_hi = INT32_MIN;
      s_abort = __hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    if (MODE == kGbProbe && s_abort) break;
    const GbUnit U = units[u];
    int lo = INT32_MAX, hi = INT32_MIN;  // kGbProbe: this thread's keys
    auto note = [&](int32_t k) {
      if (MODE == kGbProbe) {
        lo = min(lo, k);
        hi = max(hi, k);
      }
    };
    bool ovf = false;
    if (!MERGE) {
      // Rows [s0, s1): with `vec` (every column 16-byte aligned) the
      // 4-aligned interior is read as vectors of 4 rows
      // (key int4, v 2 x int4, w float4), three register sets rotating so
      // that a step's loads are issued two steps before it is aggregated;
      // the loads are unconditional (indices clamped to the last vector)
      // so the compiler keeps them in flight.  The <= 3 rows before and
      // after the interior are added one by one.
      const uint64_t s0 = U.begin, s1 = U.begin + U.len;
      const uint64_t a0 = (s0 + 3) & ~3ULL, a1 = s1 & ~3ULL;
      auto row = [&](uint64_t x) {
        const int32_t k = rk[x];
        note(k);
        ovf |= !gb_add_row<HV, HW, MODE>(tk, tcnt, tsum, tmm, tfs, k, HV ? rv[x] : 0,
                                         HW ? rw[x] : 0.f);
      };
      if (!vec || a0 >= a1) {
        for (uint64_t x = s0 + tid; x < s1; x += kGbBlock) row(x);
      } else {
        const int head = (int)(a0 - s0), tail = (int)(s1 - a1);
        if (tid < head + tail) row(tid < head ? s0 + tid : a1 + (tid - head));
        struct Vec {
          int4 k;
          int4 v0, v1;
          float4 w;
        };
        const uint32_t nvec = (uint32_t)((a1 - a0) >> 2);  // >= 1 here
        const int4 *kq = reinterpret_cast<const int4 *>(rk + a0);
        const int4 *vq = HV ? reinterpret_cast<const int4 *>(rv + a0) : nullptr;
        const float4 *wq = HW ? reinterpret_cast<const float4 *>(rw + a0) : nullptr;
        auto load = [&](uint32_t q) {
          q = q < nvec ? q : nvec - 1;
          Vec r;
          r.k = ld_nt4(kq + q);
          if (HV) {
            r.v0 = ld_nt4(vq + 2 * q);
            r.v1 = ld_nt4(vq + 2 * q + 1);
          }
          if (HW) {
            const int4 t = ld_nt4(wq + q);
            r.w = make_float4(__int_as_float(t.x), __int_as_float(t.y), __int_as_float(t.z),
                              __int_as_float(t.w));
          }
          return r;
        };
        auto agg4 = [&](const Vec &r) {
          const int32_t kk[4] = {r.k.x, r.k.y, r.k.z, r.k.w};
          const int64_t vv[4] = {
              HV ? (int64_t)(((uint64_t)(uint32_t)r.v0.y << 32) | (uint32_t)r.v0.x) : 0,
              HV ? (int64_t)(((uint64_t)(uint32_t)r.v0.w << 32) | (uint32_t)r.v0.z) : 0,
              HV ? (int64_t)(((uint64_t)(uint32_t)r.v1.y << 32) | (uint32_t)r.v1.x) : 0,
              HV ? (int64_t)(((uint64_t)(uint32_t)r.v1.w << 32) | (uint32_t)r.v1.z) : 0};
          const float ww[4] = {HW ? r.w.x : 0.f, HW ? r.w.y : 0.f, HW ? r.w.z : 0.f,
                               HW ? r.w.w : 0.f};
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            note(kk[c]);
            ovf |= !gb_add_row<HV, HW, MODE>(tk, tcnt, tsum, tmm, tfs, kk[c], vv[c], ww[c]);
          }
        };
        uint32_t q = tid;
        Vec A = load(q), B = load(q + kGbBlock);
        for (; q + 2 * kGbBlock < nvec; q += 3 * kGbBlock) {
          const Vec C = load(q + 2 * kGbBlock);
          agg4(A);
          A = load(q + 3 * kGbBlock);
          agg4(B);
          B = load(q + 4 * kGbBlock);
          agg4(C);
        }
        if (q < nvec) agg4(A);
        if (q + kGbBlock < nvec) agg4(B);
      }
    } else {
      for (uint32_t q = 0; q < U.len; ++q) {
        const uint64_t src = U.begin + q;
        const uint32_t n = in_count[src];
        const uint64_t b = src * kGbSlots;
        for (uint32_t j = tid; j < n; j += kGbBlock) {
          uint32_t s;
          if (!gb_slot<MODE>(tk, in.k[b + j], s)) {
            ovf = true;
            continue;
          }
          atomicAdd(&tcnt[s], (unsigned long long)in.cnt[b + j]);
          atomicAdd((unsigned long long *)&tsum[s], (unsigned long long)in.sum[b + j]);
          const long long mn = in.mn[b + j], mx = in.mx[b + j];
          const longlong2 mm = *reinterpret_cast<const longlong2 *>(&tmm[2 * s]);
          if (mn < mm.x) atomicMin(&tmm[2 * s], mn);
          if (mx > mm.y) atomicMax(&tmm[2 * s + 1], mx);
          atomicAdd(&tfs[s], in.fs[b + j]);
        }
      }
    }
    if (ovf) atomicOr(overflow, 1u);
    uint32_t kb = kbase, wm = wmask, part = U.part;
    if (MODE == kGbProbe) {
#pragma unroll
      for (int o = 32; o >= 1; o >>= 1) {
        lo = min(lo, __shfl_xor(lo, o, 64));
        hi = max(hi, __shfl_xor(hi, o, 64));
      }
      if ((tid & 63) == 0 && lo <= hi) {
        atomicMin(&s_lo, lo);
        atomicMax(&s_hi, hi);
      }
    }
    __syncthreads();
    if (MODE == kGbProbe) {
      if (s_lo <= s_hi) {
        if ((int64_t)s_hi - (int64_t)s_lo >= kGbSlots) {
          if (tid == 0) __hip_atomic_store(abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;  // uniform: every thread read the same s_lo / s_hi
        }
        if (tid == 0) {
          atomicMin(&range[0], s_lo);
          atomicMax(&range[1], s_hi);
        }
      }
      kb = (uint32_t)s_lo;
      wm = kGbSlots - 1;
      part = 0;
    }
    // scan: compact live slots into this unit's output area (wave ballots)
    const int lane = tid & 63, wid = tid >> 6;
    uint32_t running = 0;
    for (int base = 0; base < kGbSlots; base += kGbBlock) {
      int i = base + tid;
      bool live = MODE == kGbHashed ? tk[i] != 0ULL : tcnt[i] != 0ULL;
      unsigned long long bal = __ballot(live);
      uint32_t before = __popcll(bal & ((1ULL << lane) - 1ULL));
      if (lane == 0) wave_off[wid] = (uint32_t)__popcll(bal);
      __syncthreads();
      uint32_t woff = 0, tot = 0;
      for (int x = 0; x < kGbBlock / 64; ++x) {
        woff += x < wid ? wave_off[x] : 0;
        tot += wave_off[x];
      }
      if (live) {
        uint64_t o = (uint64_t)u * kGbSlots + running + woff + before;
        out.k[o] = MODE == kGbHashed
                       ? (int32_t)(uint32_t)tk[i]
                       : (int32_t)(kb + ((((part << kGbDirectLog) | (uint32_t)i) - kb) & wm));
        out.cnt[o] = (int64_t)tcnt[i];
        out.sum[o] = tsum[i];
        out.mn[o] = tmm[2 * i];
        out.mx[o] = tmm[2 * i + 1];
        out.fs[o] = tfs[i];
      }
      running += tot;
      __syncthreads();
    }
    if (tid == 0) out_count[u] = running;
  }
}

// Final compaction: unit u's entries go to [dst[u], dst[u] + count[u]).
__global__ void k_gb_compact(GbEntries in, const uint32_t *__restrict__ count,
                             const unsigned long long *__restrict__ dst, uint32_t nunits,
                             int32_t *ok, int64_t *ocnt, int64_t *osum, int64_t *omin,
                             int64_t *omax, double *ofs) {
  for (uint32_t u = blockIdx.x; u < nunits; u += gridDim.x) {
    uint32_t n = count[u];
    uint64_t s = (uint64_t)u * kGbSlots, d = dst[u];
    for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) {
      ok[d + j] = in.k[s + j];
      if (ocnt) ocnt[d + j] = in.cnt[s + j];
      if (osum) osum[d + j] = in.sum[s + j];
      if (omin) omin[d + j] = in.mn[s + j];
      if (omax) omax[d + j] = in.mx[s + j];
      if (ofs) ofs[d + j] = in.fs[s + j];
    }
  }
}

struct GbLevel {
  Scratch k, cnt, sum, mn, mx, fs, count;
  GbEntries e{};
  int alloc(pj_ctx *ctx, uint64_t nunits) {
    uint64_t n = nunits * (uint64_t)kGbSlots;
    PJ_TRY(k.alloc(ctx, n * 4));
    PJ_TRY(cnt.alloc(ctx, n * 8));
    PJ_TRY(sum.alloc(ctx, n * 8));
    PJ_TRY(mn.alloc(ctx, n * 8));
    PJ_TRY(mx.alloc(ctx, n * 8));
    PJ_TRY(fs.alloc(ctx, n * 8));
    PJ_TRY(count.alloc(ctx, nunits * 4));
    e = GbEntries{k.as<int32_t>(), cnt.as<int64_t>(), sum.as<int64_t>(), mn.as<int64_t>(),
                  mx.as<int64_t>(), fs.as<double>()};
    return PJ_OK;
  }
};

template <int F, class Joint>
static int gb_pass(pj_ctx *ctx, const char *name, bool pass2, const int32_t *ik,
                   const int64_t *iv, const float *iw, int32_t *ok, int64_t *ov, float *ow,
                   uint64_t n, const RadixPlan &p, Joint j, RadixHist &h) {
  return radix_launch_pass<F, 2048, 256, Joint, int64_t, float, false>(
      ctx, name, pass2, ik, iv, iw, ok, ov, ow, n, p, j, h);
}

template <class Joint>
static int gb_pass_bits(pj_ctx *ctx, const char *name, int bits, bool pass2,
                        const int32_t *ik, const int64_t *iv, const float *iw, int32_t *ok,
                        int64_t *ov, float *ow, uint64_t n, const RadixPlan &p, Joint j,
                        RadixHist &h) {
  switch (bits) {
    case 1: return gb_pass<2>(ctx, name, pass2, ik, iv, iw, ok, ov, ow, n, p, j, h);
    case 2: return gb_pass<4>(ctx, name, pass2, ik, iv, iw, ok, ov, ow, n, p, j, h);
    case 3: return gb_pass<8>(ctx, name, pass2, ik, iv, iw, ok, ov, ow, n, p, j, h);
    case 4: return gb_pass<16>(ctx, name, pass2, ik, iv, iw, ok, ov, ow, n, p, j, h);
    case 5: return gb_pass<32>(ctx, name, pass2, ik, iv, iw, ok, ov, ow, n, p, j, h);
    case 6: return gb_pass<64>(ctx, name, pass2, ik, iv, iw, ok, ov, ow, n, p, j, h);
    case 7: return gb_pass<128>(ctx, name, pass2, ik, iv, iw, ok, ov, ow, n, p, j, h);
    default: return fail(PJ_E_ARG, "group-by radix bits %d outside [1, 7]", bits);
  }
}

// How the group-by partitions its rows and which tables it uses.
struct GbPlan {
  int pbits = 0;         // partition bits (0: one partition)
  int mode = kGbHashed;  // kGbHashed: fmix32 digits; kGbDirect: JointLow digits
  uint32_t kbase = 0;    // kGbDirect: the key of slot s of partition p is
  uint32_t wmask = 0;    //   kbase + (((p << 11) | s) - kbase) & wmask
};

// Histogram of the partition digits (read back to the host).  Direct plans
// partition on key bits [11, 11 + pbits) and record min / max in `range`.
static int gb_hist(pj_ctx *ctx, const int32_t *d_k, uint64_t n, int pbits, bool direct,
                   int *range, RadixPlan &plan, RadixHist &rh) {
  const int b1 = pbits <= 7 ? pbits : (pbits + 1) / 2;
  plan = radix_make_plan(ctx, n, b1, pbits - b1, 1 << pbits, 2);
  PJ_TRY(rh.alloc(ctx, plan));
  if (direct)
    return radix_histograms(ctx, d_k, n, plan,
                            JointLow{kGbDirectLog, (1u << pbits) - 1u, range}, rh, true);
  return radix_histograms(ctx, d_k, n, plan, JointRadix{32 - pbits}, rh, true);
}

// The one or two radix passes after gb_hist, with its digit function `j`.
template <class Joint>
static int gb_scatter(pj_ctx *ctx, const RadixPlan &plan, Joint j, RadixHist &rh, uint64_t n,
                      const int32_t *&pk, const int64_t *&pv, const float *&pw,
                      Scratch (&bufs)[6]) {
  // The six partition arrays start at different odd multiples of 4 KiB
  // inside their allocations: pass 2 reads and writes the same offsets of
  // equally sized arrays, and bases that coincide modulo the HBM channel
  // interleave send its read and write streams to the same channels.
  constexpr uint64_t kSkew[6] = {0, 17 * 4096, 33 * 4096, 49 * 4096, 65 * 4096, 81 * 4096};
  auto at = [&](int i) { return (char *)bufs[i].p + kSkew[i]; };
  PJ_TRY(bufs[0].alloc(ctx, n * 4 + 256 + kSkew[0]));
  PJ_TRY(bufs[1].alloc(ctx, n * 8 + 256 + kSkew[1]));
  PJ_TRY(bufs[2].alloc(ctx, n * 4 + 256 + kSkew[2]));
  int32_t *ok1 = (int32_t *)at(0);
  int64_t *ov1 = (int64_t *)at(1);
  float *ow1 = (float *)at(2);
  PJ_TRY(gb_pass_bits(ctx, "gb_scatter1", plan.b1, false, pk, pv, pw, ok1, ov1, ow1, n, plan, j,
                      rh));
  pk = ok1;
  pv = ov1;
  pw = ow1;
  if (plan.b2) {
    PJ_TRY(bufs[3].alloc(ctx, n * 4 + 256 + kSkew[3]));
    PJ_TRY(bufs[4].alloc(ctx, n * 8 + 256 + kSkew[4]));
    PJ_TRY(bufs[5].alloc(ctx, n * 4 + 256 + kSkew[5]));
    int32_t *ok2 = (int32_t *)at(3);
    int64_t *ov2 = (int64_t *)at(4);
    float *ow2 = (float *)at(5);
    PJ_TRY(gb_pass_bits(ctx, "gb_scatter2", plan.b2, true, pk, pv, pw, ok2, ov2, ow2, n, plan,
                        j, rh));
    pk = ok2;
    pv = ov2;
    pw = ow2;
  }
  return PJ_OK;
}

// Launches k_gb_agg in mode `mode` (a runtime value) with HV / HW fixed.
template <bool MERGE, bool HV, bool HW, class... A>
static int gb_launch(pj_ctx *ctx, const char *name, int mode, uint32_t grid, A... args) {
  switch (mode) {
    case kGbHashed:
      PJ_LAUNCH(ctx, name, (k_gb_agg<MERGE, HV, HW, kGbHashed>), grid, kGbBlock, 0, args...);
      return PJ_OK;
    case kGbDirect:
      PJ_LAUNCH(ctx, name, (k_gb_agg<MERGE, HV, HW, kGbDirect>), grid, kGbBlock, 0, args...);
      return PJ_OK;
    default:
      if constexpr (!MERGE) {
        PJ_LAUNCH(ctx, name, (k_gb_agg<false, HV, HW, kGbProbe>), grid, kGbBlock, 0, args...);
        return PJ_OK;
      }
      return fail(PJ_E_ARG, "group-by merge in probe mode");
  }
}

// One group-by under plan `pl`.  pl.pbits > 0 needs gb_hist's plan and
// histogram for the same digits (rh != nullptr) or, for hashed plans, makes
// them.  *overflowed: a hashed table filled up (more partitions needed);
// *aborted: a kGbProbe unit spanned more than 2048 keys (no result).
static int groupby_run(pj_ctx *ctx, const int32_t *d_k, const int64_t *d_v, const float *d_w,
                       uint64_t n, const GbPlan &pl, RadixPlan *plan_in, RadixHist *rh_in,
                       uint64_t max_groups, int32_t *d_ok, int64_t *d_cnt, int64_t *d_sum,
                       int64_t *d_min, int64_t *d_max, double *d_fs, uint64_t *n_groups,
                       bool *overflowed, bool *aborted) {
  *overflowed = false;
  *aborted = false;
  *n_groups = 0;
  // ---- partition (optional) ---------------------------------------------
  const int32_t *pk = d_k;
  const int64_t *pv = d_v;
  const float *pw = d_w;
  std::vector<uint64_t> part_start{0}, part_count{n};
  Scratch bufs[6];
  RadixPlan own_plan;
  RadixHist own_rh;
  Scratch zv, zw;  // zero columns when v / w are absent but partitioning runs
  if (pl.pbits > 0 && n) {
    RadixPlan *plan = plan_in;
    RadixHist *rh = rh_in;
    if (!rh) {
      if (pl.mode != kGbHashed) return fail(PJ_E_ARG, "direct group-by without its histogram");
      PJ_TRY(gb_hist(ctx, d_k, n, pl.pbits, false, nullptr, own_plan, own_rh));
      plan = &own_plan;
      rh = &own_rh;
    }
    if (!pv) {
      PJ_TRY(zv.alloc(ctx, n * 8));
      PJ_TRY(zero(ctx, zv.p, n * 8));
      pv = zv.as<int64_t>();
    }
    if (!pw) {
      PJ_TRY(zw.alloc(ctx, n * 4));
      PJ_TRY(zero(ctx, zw.p, n * 4));
      pw = zw.as<float>();
    }
    if (pl.mode == kGbDirect)
      PJ_TRY(gb_scatter(ctx, *plan, JointLow{kGbDirectLog, (1u << pl.pbits) - 1u, nullptr}, *rh,
                        n, pk, pv, pw, bufs));
    else
      PJ_TRY(gb_scatter(ctx, *plan, JointRadix{32 - pl.pbits}, *rh, n, pk, pv, pw, bufs));
    const int NB = 1 << pl.pbits;
    part_start.assign(rh->hS.begin(), rh->hS.begin() + NB);
    part_count.assign(rh->hT.begin(), rh->hT.end());
    if (!d_v) pv = nullptr;
    if (!d_w) pw = nullptr;
  }
  // ---- level 0 units -------------------------------------------------------
  // about eight units per workgroup (one per CU): with units taken
  // round-robin, the last round is at most one unit of eight
  const uint64_t target_units = (uint64_t)ctx->num_cus * 8;
  uint64_t unit_rows = std::max<uint64_t>(ceil_div(n ? n : 1, target_units), 16384);
  std::vector<GbUnit> units;
  for (size_t p = 0; p < part_start.size(); ++p) {
    for (uint64_t r = 0; r < part_count[p]; r += unit_rows) {
      GbUnit u;
      u.begin = part_start[p] + r;
      u.len = (uint32_t)std::min<uint64_t>(unit_rows, part_count[p] - r);
      u.part = (uint32_t)p;
      units.push_back(u);
    }
  }
  if (units.empty()) return PJ_OK;
  // flags: [0] table overflow, [1] probe abort, [2..3] probe min / max
  Scratch flags, dunits;
  PJ_TRY(flags.alloc(ctx, 16));
  const int32_t init[4] = {0, 0, INT32_MAX, INT32_MIN};
  PJ_TRY(h2d(ctx, flags.p, init, 16));
  uint32_t *ovf = flags.as<uint32_t>();
  uint32_t *abort = ovf + 1;
  int *range = flags.as<int>() + 2;
  auto level = std::make_unique<GbLevel>();
  PJ_TRY(level->alloc(ctx, units.size()));
  PJ_TRY(dunits.alloc(ctx, units.size() * sizeof(GbUnit)));
  PJ_TRY(h2d(ctx, dunits.p, units.data(), units.size() * sizeof(GbUnit)));
  {
    // one 1024-thread workgroup per CU (96 KiB of LDS each)
    uint32_t grid = (uint32_t)std::min<uint64_t>(units.size(), (uint64_t)ctx->num_cus);
    GbEntries none{};
    // The vector loads need 16-byte-aligned columns; otherwise every row is
    // added one by one.
    auto a16 = [](const void *q) { return ((uintptr_t)q & 15) == 0; };
    const bool vec = a16(pk) && (!pv || a16(pv)) && (!pw || a16(pw));
    auto launch = [&](auto hv, auto hw) -> int {
      return gb_launch<false, decltype(hv)::value, decltype(hw)::value>(
          ctx, "gb_agg", pl.mode, grid, dunits.as<const GbUnit>(), (uint32_t)units.size(), pk,
          pv, pw, none, (const uint32_t *)nullptr, level->e, level->count.as<uint32_t>(), ovf,
          vec, pl.kbase, pl.wmask, range, abort);
    };
    using T_ = std::true_type;
    using F_ = std::false_type;
    if (pv && pw) PJ_TRY(launch(T_{}, T_{}));
    else if (pv) PJ_TRY(launch(T_{}, F_{}));
    else if (pw) PJ_TRY(launch(F_{}, T_{}));
    else PJ_TRY(launch(F_{}, F_{}));
  }
  // The merges' tables: a probe run whose units all fit their spans merges
  // directly when the whole key span fits one table (kbase = min), hashed
  // otherwise (each unit's entries are correct either way).
  GbPlan mp = pl;
  if (pl.mode == kGbProbe) {
    int32_t f[4];
    PJ_TRY(d2h(ctx, f, flags.p, 16));
    if (f[1]) {
      *aborted = true;
      return PJ_OK;
    }
    const bool narrow = f[2] <= f[3] && (int64_t)f[3] - (int64_t)f[2] < kGbSlots;
    mp.mode = narrow ? kGbDirect : kGbHashed;
    mp.kbase = (uint32_t)f[2];
    mp.wmask = kGbSlots - 1;
  }
  // ---- merge levels: fold units of the same partition ------------------
  // Units are ordered by partition.  Each level merges up to `fan` units of
  // one partition into one, until every partition has one unit.  With few
  // partitions (many units each) a small fan-in keeps the first levels wide
  // enough for every CU (2^10 groups: merges 0.32 -> 0.12 ms); with many
  // partitions most need no merge, and a large fan-in saves levels.
  const size_t fan = part_start.size() <= 64 ? (size_t)kGbMergeFanInFew : (size_t)kGbMergeFanIn;
  while (true) {
    std::vector<GbUnit> next;
    bool single = true;
    for (size_t i = 0; i < units.size();) {
      size_t j = i;
      while (j < units.size() && units[j].part == units[i].part) ++j;
      if (j - i > 1) single = false;
      for (size_t a = i; a < j; a += fan) {
        GbUnit m;
        m.begin = a;
        m.len = (uint32_t)std::min<size_t>(fan, j - a);
        m.part = units[i].part;
        next.push_back(m);
      }
      i = j;
    }
    if (single) break;
    auto nl = std::make_unique<GbLevel>();
    PJ_TRY(nl->alloc(ctx, next.size()));
    Scratch dn;
    PJ_TRY(dn.alloc(ctx, next.size() * sizeof(GbUnit)));
    PJ_TRY(h2d(ctx, dn.p, next.data(), next.size() * sizeof(GbUnit)));
    uint32_t grid = (uint32_t)std::min<uint64_t>(next.size(), (uint64_t)ctx->num_cus);
    PJ_TRY((gb_launch<true, true, true>(
        ctx, "gb_merge", mp.mode, grid, dn.as<const GbUnit>(), (uint32_t)next.size(),
        (const int32_t *)nullptr, (const int64_t *)nullptr, (const float *)nullptr, level->e,
        level->count.as<const uint32_t>(), nl->e, nl->count.as<uint32_t>(), ovf, false,
        mp.kbase, mp.wmask, range, abort)));
    PJ_TRY(sync(ctx));
    units.swap(next);
    level.swap(nl);
  }
  uint32_t of = 0;
  PJ_TRY(d2h(ctx, &of, ovf, 4));
  if (of) {
    *overflowed = true;
    return PJ_OK;
  }
  std::vector<uint32_t> cnt(units.size());
  PJ_TRY(d2h(ctx, cnt.data(), level->count.p, units.size() * 4));
  std::vector<unsigned long long> dst(units.size());
  uint64_t tot = 0;
  for (size_t i = 0; i < units.size(); ++i) {
    dst[i] = tot;
    tot += cnt[i];
  }
  *n_groups = tot;
  if (tot > max_groups)
    return fail(PJ_E_CAPACITY, "group-by produced %llu groups > maxInputSize %llu",
                (unsigned long long)tot, (unsigned long long)max_groups);
  Scratch ddst;
  PJ_TRY(ddst.alloc(ctx, dst.size() * 8));
  PJ_TRY(h2d(ctx, ddst.p, dst.data(), dst.size() * 8));
  uint32_t grid = (uint32_t)std::min<uint64_t>(units.size(), 4096);
  PJ_LAUNCH(ctx, "gb_compact", k_gb_compact, grid, 256, 0, level->e,
            level->count.as<const uint32_t>(), ddst.as<const unsigned long long>(),
            (uint32_t)units.size(), d_ok, d_cnt, d_sum, d_min, d_max, d_fs);
  return sync(ctx);
}

// ---------------------------------------------------------------------------
// reduce
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_reduce_i64(const int64_t *__restrict__ v, uint64_t n,
                                                    unsigned long long *__restrict__ acc) {
  int64_t s = 0, mn = INT64_MAX, mx = INT64_MIN;
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t st = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) {
    int64_t x = v[i];
    s += x;
    mn = x < mn ? x : mn;
    mx = x > mx ? x : mx;
  }
  s = wave_sum_i64(s);
  mn = wave_min_i64(mn);
  mx = wave_max_i64(mx);
  // one atomic triple per workgroup: same-address atomics serialise in L2
  __shared__ int64_t rs[4], rmn[4], rmx[4];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    rs[w] = s;
    rmn[w] = mn;
    rmx[w] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int x = 1; x < 4; ++x) {
      s += rs[x];
      mn = rmn[x] < mn ? rmn[x] : mn;
      mx = rmx[x] > mx ? rmx[x] : mx;
    }
    atomicAdd(&acc[0], (unsigned long long)s);
    atomicMin((long long *)&acc[1], (long long)mn);
    atomicMax((long long *)&acc[2], (long long)mx);
  }
}

__global__ void k_reduce_init(unsigned long long *acc) {
  if (threadIdx.x == 0) {
    acc[0] = 0;
    acc[1] = (unsigned long long)INT64_MAX;
    acc[2] = (unsigned long long)INT64_MIN;
  }
}

}  // namespace pj

using namespace pj;


extern "C" int pj_groupby_i32(pj_ctx *ctx, const int32_t *d_k, const int64_t *d_v,
                              const float *d_w, uint64_t n, int hash_bits,
                              uint64_t max_groups, int32_t *d_out_k, int64_t *d_cnt,
                              int64_t *d_sum, int64_t *d_min, int64_t *d_max, double *d_fsum,
                              uint64_t *n_groups) {
  std::lock_guard<std::mutex> g(ctx->mu);
  (void)hash_bits;
  *n_groups = 0;
  if (!d_out_k) return fail(PJ_E_ARG, "output key column is required");
  if (n >= (1ULL << 40)) return fail(PJ_E_CAPACITY, "input above the supported row count");
  int pbits = 0;
  while (pbits < 14 && ((uint64_t)kGbCap << pbits) < max_groups) ++pbits;
  if (!n) return PJ_OK;
  bool ovf = false, aborted = false;
  // Direct tables first: one table covers 2048 keys, so the direct plan
  // needs one partition bit fewer than the hashed one for the same bound.
  int pd = pbits > 0 ? pbits - 1 : 0;
  if (pd > 0) {
    Scratch range;
    PJ_TRY(range.alloc(ctx, 8));
    const int32_t init[2] = {INT32_MAX, INT32_MIN};
    PJ_TRY(h2d(ctx, range.p, init, 8));
    RadixPlan plan;
    RadixHist rh;
    PJ_TRY(gb_hist(ctx, d_k, n, pd, true, range.as<int>(), plan, rh));
    int32_t mm[2];
    PJ_TRY(d2h(ctx, mm, range.p, 8));
    const uint64_t span = (uint64_t)((int64_t)mm[1] - (int64_t)mm[0]) + 1;
    if (span <= (uint64_t)kGbSlots) {
      pd = 0;  // the bound overstates the groups: one direct table, no partitions
    } else {
      const int dn = ilog2_ceil(ceil_div(span, (uint64_t)kGbSlots));
      if (dn <= pd) {
        if (dn <= 7 && pd > 7) {
          // fewer digits that still cover the span save the second pass
          pd = dn;
          PJ_TRY(gb_hist(ctx, d_k, n, pd, true, range.as<int>(), plan, rh));
        }
        GbPlan pl{pd, kGbDirect, (uint32_t)mm[0], (uint32_t)((1ULL << (kGbDirectLog + pd)) - 1)};
        return groupby_run(ctx, d_k, d_v, d_w, n, pl, &plan, &rh, max_groups, d_out_k, d_cnt,
                           d_sum, d_min, d_max, d_fsum, n_groups, &ovf, &aborted);
      }
      pd = -1;  // sparse keys: hashed tables
    }
  }
  if (pd == 0) {
    GbPlan pl{0, kGbProbe, 0u, 0u};
    PJ_TRY(groupby_run(ctx, d_k, d_v, d_w, n, pl, nullptr, nullptr, max_groups, d_out_k, d_cnt,
                       d_sum, d_min, d_max, d_fsum, n_groups, &ovf, &aborted));
    if (!aborted) return PJ_OK;
  }
  while (true) {
    GbPlan pl{pbits, kGbHashed, 0u, 0u};
    PJ_TRY(groupby_run(ctx, d_k, d_v, d_w, n, pl, nullptr, nullptr, max_groups, d_out_k, d_cnt,
                       d_sum, d_min, d_max, d_fsum, n_groups, &ovf, &aborted));
    if (!ovf) return PJ_OK;
    if (pbits >= 14)
      return fail(PJ_E_CAPACITY,
                  "group-by input has more distinct keys than maxInputSize %llu allows",
                  (unsigned long long)max_groups);
    ++pbits;  // a table overflowed: more partitions
  }
}

extern "C" int pj_reduce_i64(pj_ctx *ctx, const int64_t *d_v, uint64_t n, int64_t *sum,
                             int64_t *min, int64_t *max) {
  std::lock_guard<std::mutex> g(ctx->mu);
  Scratch acc;
  PJ_TRY(acc.alloc(ctx, 64));
  PJ_LAUNCH(ctx, "reduce_init", k_reduce_init, 1, 64, 0, acc.as<unsigned long long>());
  if (n) {
    uint64_t grid = std::min<uint64_t>(ceil_div(n, 256 * 8), (uint64_t)ctx->num_cus * 8);
    PJ_LAUNCH(ctx, "reduce_i64", k_reduce_i64, grid ? grid : 1, 256, 0, d_v, n,
              acc.as<unsigned long long>());
  }
  int64_t h[3];
  PJ_TRY(d2h(ctx, h, acc.p, sizeof(h)));
  if (sum) *sum = h[0];
  if (min) *min = h[1];
  if (max) *max = h[2];
  return PJ_OK;
}
