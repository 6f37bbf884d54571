// hashjoin-chained on one MI355X: a global open-addressing table in HBM.
//
// Reference: HashJoinChained / GpuHashJoinChained
// (core/olap/lib/operators/hash-join-chained.cpp:233-566,
//  core/olap/lib/operators/gpu/gpu-hash-join-chained.cpp:31-73).
// The reference keeps a chained table: head[2^hash_bits] plus packets
// {next, key, payloads}, one atomicAdd on a single counter and one atomicXchg
// per build row, and a dependent head -> packet walk per probe.  Here the
// build rows go into a linear-probing table of 64-bit slots
//     slot = key << 32 | (build_row + 1),   0 = empty,
// claimed with one 64-bit CAS (no shared counter, no sentinel key: every
// occupied slot has a non-zero low word).  A probe reads consecutive slots of
// one or two cache lines until an empty slot, emitting every key match, so
// duplicates on the build side behave as the reference's chain walk.  The
// slot index is hashInt32(key) mod 2^bits exactly as the reference's bucket
// (hash-join-chained.cpp:233-239); the table has max(2^hash_bits,
// 2*next_pow2(n_build)) slots so the load factor stays <= 1/2.
#include "pj_internal.hpp"

namespace pj {

// Inserts the build rows.  A row whose CAS fails on a slot holding its own
// key has a duplicate on the build side; that is reported in *dup.  The test
// is exact: two rows with one key walk the same slot sequence from the same
// start, so whichever claims the later slot failed its CAS on the other's.
__global__ __launch_bounds__(256) void k_chained_build(
    const int32_t *__restrict__ bk, uint64_t nb,
    unsigned long long *__restrict__ slots, uint64_t mask, uint32_t *__restrict__ dup) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  bool seen = false;
  for (; i < nb; i += stride) {
    int32_t k = bk[i];
    unsigned long long w =
        ((unsigned long long)(uint32_t)k << 32) | (unsigned long long)(i + 1);
    uint64_t h = hash_i32(k) & mask;
    while (true) {
      unsigned long long old = atomicCAS(&slots[h], 0ULL, w);
      if (old == 0ULL) break;
      seen |= (uint32_t)(old >> 32) == (uint32_t)k;
      h = (h + 1) & mask;
    }
  }
  if (seen) atomicOr(dup, 1u);
}

// Probe with the fused reduce.  Each thread handles 4 consecutive probe rows
// per step.  With unique build keys (*dup == 0, a uniform branch) a row ends
// at its key or at an empty slot, and the four rows walk in lockstep: every
// round issues the four slot loads together, then the four build-value
// gathers of the matches.  With duplicate keys every row walks its run of
// occupied slots to the empty slot that ends it, summing all its matches
// (the reference's chain walk, hash-join-chained.cpp:321-566).
__global__ __launch_bounds__(256) void k_chained_probe_reduce(
    const int32_t *__restrict__ pk, const int32_t *__restrict__ pv, uint64_t np,
    const int32_t *__restrict__ bv,
    const unsigned long long *__restrict__ slots, uint64_t mask,
    const uint32_t *__restrict__ dup, unsigned long long *__restrict__ acc_out) {
  JoinAcc acc;
  acc.init();
  const bool uniq = *dup == 0u;
  uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t nthr = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t base = tid * 4; base < np; base += nthr * 4) {
    int32_t k[4];
    int32_t v[4];
    uint64_t h[4];
    bool live[4];
    if (base + 4 <= np && ((uintptr_t)(pk + base) & 15) == 0 &&
        (!pv || ((uintptr_t)(pv + base) & 15) == 0)) {
      const int4 kq = *reinterpret_cast<const int4 *>(pk + base);
      const int4 vq = pv ? *reinterpret_cast<const int4 *>(pv + base) : make_int4(0, 0, 0, 0);
      k[0] = kq.x, k[1] = kq.y, k[2] = kq.z, k[3] = kq.w;
      v[0] = vq.x, v[1] = vq.y, v[2] = vq.z, v[3] = vq.w;
#pragma unroll
      for (int u = 0; u < 4; ++u) live[u] = true;
    } else {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        uint64_t j = base + u;
        live[u] = j < np;
        k[u] = live[u] ? pk[j] : 0;
        v[u] = (live[u] && pv) ? pv[j] : 0;
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) h[u] = hash_i32(k[u]) & mask;
    unsigned long long w[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) w[u] = slots[h[u]];
    if (uniq) {
      bool act[4], hit[4];
      uint32_t row[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        act[u] = live[u];
        hit[u] = false;
        row[u] = 0;
      }
      while (true) {
        bool more = false;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const bool empty = w[u] == 0ULL;
          const bool match = !empty && (int32_t)(uint32_t)(w[u] >> 32) == k[u];
          if (act[u] && match) {
            hit[u] = true;
            row[u] = (uint32_t)(w[u] & 0xffffffffULL) - 1u;
          }
          act[u] = act[u] && !empty && !match;
          h[u] = act[u] ? ((h[u] + 1) & mask) : h[u];
          more |= act[u];
        }
        if (!more) break;
        // rows that are done reload their own slot (a cache hit), so the
        // four loads are issued without branches
#pragma unroll
        for (int u = 0; u < 4; ++u) w[u] = slots[h[u]];
      }
      int32_t b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) b[u] = bv ? bv[row[u]] : 0;
      acc.add4_1(hit, b, v);
    } else {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (!live[u]) continue;
        int64_t m = 0, sb = 0;
        int32_t mnb = INT32_MAX, mxb = INT32_MIN;
        uint64_t p = h[u];
        unsigned long long x = w[u];
        while (x != 0ULL) {
          if ((int32_t)(uint32_t)(x >> 32) == k[u]) {
            uint64_t r = (x & 0xffffffffULL) - 1;
            int32_t bb = bv ? bv[r] : 0;
            m += 1;
            sb += bb;
            mnb = min(mnb, bb);
            mxb = max(mxb, bb);
          }
          p = (p + 1) & mask;
          x = slots[p];
        }
        acc.add(m, sb, mnb, mxb, v[u]);
      }
    }
  }
  join_acc_flush(acc, acc_out);
}

// Materialising probe: count a lane's matches, reserve them with one atomic
// per wave (wave64 prefix sum), then walk again and write the pairs.
__global__ __launch_bounds__(256) void k_chained_probe_pairs(
    const int32_t *__restrict__ pk, uint64_t np,
    const unsigned long long *__restrict__ slots, uint64_t mask,
    uint32_t *__restrict__ out_b, uint32_t *__restrict__ out_p, uint64_t cap,
    unsigned long long *__restrict__ cursor) {
  uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t nthr = (uint64_t)gridDim.x * blockDim.x;
  uint64_t rounds = ceil_div(np, nthr);
  for (uint64_t r = 0; r < rounds; ++r) {
    uint64_t j = r * nthr + tid;
    bool live = j < np;
    int32_t k = live ? pk[j] : 0;
    uint64_t h0 = hash_i32(k) & mask;
    uint32_t m = 0;
    if (live) {
      for (uint64_t p = h0;; p = (p + 1) & mask) {
        unsigned long long w = slots[p];
        if (w == 0ULL) break;
        m += ((int32_t)(uint32_t)(w >> 32) == k);
      }
    }
    // inclusive scan of m over the wave
    uint32_t incl = m;
    int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      uint32_t x = __shfl_up(incl, o, 64);
      if (lane >= o) incl += x;
    }
    uint32_t total = __shfl(incl, 63, 64);
    unsigned long long base = 0;
    if (lane == 63 && total) base = atomicAdd(cursor, (unsigned long long)total);
    base = __shfl(base, 63, 64);
    uint64_t pos = base + incl - m;
    if (live && m) {
      for (uint64_t p = h0;; p = (p + 1) & mask) {
        unsigned long long w = slots[p];
        if (w == 0ULL) break;
        if ((int32_t)(uint32_t)(w >> 32) == k) {
          if (pos < cap) {
            out_b[pos] = (uint32_t)((w & 0xffffffffULL) - 1);
            out_p[pos] = (uint32_t)j;
          }
          ++pos;
        }
      }
    }
  }
}

static int chained_check(int hash_bits, uint64_t nb, uint64_t max_build) {
  if (hash_bits < 0 || hash_bits > 32)
    return fail(PJ_E_ARG, "hash_bits %d out of range [0, 32]", hash_bits);
  if (nb > max_build)
    return fail(PJ_E_CAPACITY,
                "Build input size exceeded given parameter: %llu rows > "
                "maxBuildInputSize %llu",
                (unsigned long long)nb, (unsigned long long)max_build);
  if (nb >= 0xffffffffULL) return fail(PJ_E_CAPACITY, "build side above 2^32-1 rows");
  return PJ_OK;
}

// Builds the table; *dup (zeroed by the caller) reports duplicate build keys.
static int chained_table(pj_ctx *ctx, const int32_t *d_bk, uint64_t nb,
                         int hash_bits, uint64_t max_build, Scratch &tab,
                         uint64_t *mask, uint32_t *dup) {
  PJ_TRY(chained_check(hash_bits, nb, max_build));
  uint64_t slots = 1ULL << hash_bits;
  uint64_t need = 2ULL << ilog2_ceil(nb ? nb : 1);
  if (slots < need) slots = need;
  if (slots < 64) slots = 64;
  PJ_TRY(tab.alloc(ctx, slots * 8));
  PJ_TRY(zero(ctx, tab.p, slots * 8));
  *mask = slots - 1;
  if (nb) {
    uint64_t grid = ceil_div(nb, 256);
    if (grid > 8192) grid = 8192;
    PJ_LAUNCH(ctx, "join_chained_build", k_chained_build, grid, 256, 0, d_bk, nb,
              tab.as<unsigned long long>(), *mask, dup);
  }
  return PJ_OK;
}

// hashjoin-chained -> reduce runs as the radix-partitioned join when the
// probe side is large enough for partitioning to pay: the partitioned join
// costs a fixed ~0.1-0.2 ms of small kernels, while every chained probe is
// a dependent random read of a table that is beyond the L2 once the build
// side passes 2^18 rows.  Results are the same either way (same counts and
// sums); only the device algorithm differs.
static bool chained_goes_partitioned(uint64_t nb, uint64_t np) {
  if (!nb || !np) return false;
  return np >= (1ULL << 24) || (np >= (1ULL << 22) && nb >= (1ULL << 19));
}

static uint64_t probe_grid(pj_ctx *ctx, uint64_t rows_per_thread, uint64_t np) {
  uint64_t g = ceil_div(np, 256 * rows_per_thread);
  uint64_t cap = (uint64_t)ctx->num_cus * 16;
  if (g > cap) g = cap;
  return g ? g : 1;
}

// ---------------------------------------------------------------------------
// Allocation-free variant: the table and the accumulator live in a
// file-scope device array instead of the context pool.  Every device holds
// its own copy of the array (at its own address) from module load until the
// process ends; host code reaches it only through the address
// hipGetSymbolAddress returns on that device, never through the array's name.
// ---------------------------------------------------------------------------
constexpr uint64_t kNoAllocSlots = 1ULL << 22;  // 32 MiB of 64-bit slots
constexpr uint64_t kNoAllocAccWords = kJoinAccWords;  // 8 KiB: the accumulator copies
static __device__ __align__(256) unsigned long long scratch[kNoAllocSlots + kNoAllocAccWords];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned long long *g_scratch_ptr[kMaxDevices];

// Resolves the scratch array's address on `device`, which must be current.
// Called once per device by pj_ctx_create.
int noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices)
    return fail(PJ_E_ARG, "device %d outside [0, %d)", device, kMaxDevices);
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (g_scratch_ptr[device]) return PJ_OK;
  void *p = nullptr;
  PJ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) return fail(PJ_E_HIP, "hipGetSymbolAddress(scratch) returned NULL on device %d", device);
  g_scratch_ptr[device] = static_cast<unsigned long long *>(p);
  return PJ_OK;
}

static unsigned long long *noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  return g_scratch_ptr[device];
}

}  // namespace pj

using namespace pj;

extern "C" int pj_join_chained_reduce_noalloc(pj_ctx *ctx, const int32_t *d_bk,
                                              const int32_t *d_bv, uint64_t nb,
                                              const int32_t *d_pk, const int32_t *d_pv,
                                              uint64_t np, int hash_bits,
                                              uint64_t max_build, pj_join_agg *out) {
  std::lock_guard<std::mutex> g(ctx->mu);
  PJ_TRY(chained_check(hash_bits, nb, max_build));
  // the same geometry as chained_table()
  uint64_t slots = 1ULL << hash_bits;
  uint64_t need = 2ULL << ilog2_ceil(nb ? nb : 1);
  if (slots < need) slots = need;
  if (slots < 64) slots = 64;
  if (slots > kNoAllocSlots)
    return fail(PJ_E_CAPACITY,
                "pj_join_chained_reduce_noalloc: a table of %llu slots (hash_bits %d, %llu "
                "build rows) exceeds the static scratch of %llu slots (at most 2^21 build rows "
                "and hash_bits <= 22)",
                (unsigned long long)slots, hash_bits, (unsigned long long)nb,
                (unsigned long long)kNoAllocSlots);
  unsigned long long *base = noalloc_scratch(ctx->device);
  if (!base)
    return fail(PJ_E_HIP, "static scratch of device %d was not resolved", ctx->device);
  unsigned long long *tab = base;
  unsigned long long *acc = base + kNoAllocSlots;
  uint32_t *dup = reinterpret_cast<uint32_t *>(acc + kJoinAccFlagWord);
  const uint64_t mask = slots - 1;
  // the previous call's contents stay behind: clear what this call reads
  // (join_acc_init also clears the duplicate-key flag)
  PJ_TRY(fill_u32(ctx, reinterpret_cast<uint32_t *>(tab), 0u, slots * 2));
  PJ_TRY(join_acc_init(ctx, acc));
  if (nb) {
    uint64_t grid = ceil_div(nb, 256);
    if (grid > 8192) grid = 8192;
    PJ_LAUNCH(ctx, "join_chained_build", k_chained_build, grid, 256, 0, d_bk, nb, tab, mask,
              dup);
  }
  if (np && nb)
    PJ_LAUNCH(ctx, "join_chained_probe", k_chained_probe_reduce, probe_grid(ctx, 4, np), 256,
              0, d_pk, d_pv, np, d_bv, (const unsigned long long *)tab, mask,
              (const uint32_t *)dup, acc);
  return join_acc_read(ctx, acc, out);
}

extern "C" int pj_join_chained_reduce(pj_ctx *ctx, const int32_t *d_bk,
                                      const int32_t *d_bv, uint64_t nb,
                                      const int32_t *d_pk, const int32_t *d_pv,
                                      uint64_t np, int hash_bits,
                                      uint64_t max_build, pj_join_agg *out) {
  std::lock_guard<std::mutex> g(ctx->mu);
  PJ_TRY(chained_check(hash_bits, nb, max_build));
  // dense build keys (a primary key 0..n-1): a direct table of the values
  bool done = false;
  PJ_TRY(join_direct(ctx, d_bk, d_bv, nb, d_pk, d_pv, np, out, &done));
  if (done) return PJ_OK;
  if (chained_goes_partitioned(nb, np))
    return join_partitioned_auto(ctx, d_bk, d_bv, nb, d_pk, d_pv, np, out);
  Scratch tab, acc;
  uint64_t mask = 0;
  PJ_TRY(acc.alloc(ctx, kJoinAccWords * 8));
  PJ_TRY(join_acc_init(ctx, acc.as<unsigned long long>()));
  uint32_t *dup = reinterpret_cast<uint32_t *>(acc.as<unsigned long long>() + kJoinAccFlagWord);
  PJ_TRY(chained_table(ctx, d_bk, nb, hash_bits, max_build, tab, &mask, dup));
  if (np && nb)
    PJ_LAUNCH(ctx, "join_chained_probe", k_chained_probe_reduce, probe_grid(ctx, 4, np), 256,
              0, d_pk, d_pv, np, d_bv, tab.as<const unsigned long long>(), mask,
              (const uint32_t *)dup, acc.as<unsigned long long>());
  return join_acc_read(ctx, acc.as<unsigned long long>(), out);
}

extern "C" int pj_join_chained_pairs(pj_ctx *ctx, const int32_t *d_bk,
                                     uint64_t nb, const int32_t *d_pk,
                                     uint64_t np, int hash_bits,
                                     uint64_t max_build, uint32_t *d_out_b,
                                     uint32_t *d_out_p, uint64_t cap,
                                     uint64_t *n_out) {
  std::lock_guard<std::mutex> g(ctx->mu);
  *n_out = 0;
  Scratch tab, cur;
  uint64_t mask = 0;
  PJ_TRY(cur.alloc(ctx, 16));  // the pair cursor, then the duplicate-key flag
  PJ_TRY(zero(ctx, cur.p, 16));
  PJ_TRY(chained_table(ctx, d_bk, nb, hash_bits, max_build, tab, &mask,
                       reinterpret_cast<uint32_t *>(cur.as<unsigned long long>() + 1)));
  if (np && nb)
    PJ_LAUNCH(ctx, "join_chained_pairs", k_chained_probe_pairs, probe_grid(ctx, 1, np), 256, 0,
              d_pk, np, tab.as<const unsigned long long>(), mask, d_out_b, d_out_p, cap,
              cur.as<unsigned long long>());
  uint64_t total = 0;
  PJ_TRY(d2h(ctx, &total, cur.p, 8));
  *n_out = total;
  if (total > cap)
    return fail(PJ_E_CAPACITY, "join produced %llu pairs, output capacity %llu",
                (unsigned long long)total, (unsigned long long)cap);
  return PJ_OK;
}
