// partitioned-hashjoin-chained and hash-rearrange on one MI355X.
//
// Reference: HashPartitioner + GpuPartitionedHashJoinChained
// (core/olap/lib/operators/gpu/gpu-partitioned-hash-join-chained.cpp:47-404
//  and :612-1586, kernels in gpu-partitioned-hash-join-chained-helpers.cpp),
// GpuHashRearrange (core/olap/lib/operators/gpu/gpu-hash-rearrange.cpp).
//
// Flow of one join (both sides go through the same steps):
//   1. k_hist: joint histogram of the B = bits1 + bits2 top bits of
//      fmix32(key) -- one read of the key column;
//   2. host plan: 32-element-aligned region per pass-1 digit and per final
//      partition, tile lists, join work units (<= probe_unit probe tuples each,
//      the reference's decompose_chains idea, helpers.cpp:199-232);
//   3. k_scatter pass 1 (fan-out 2^bits1) and pass 2 (2^bits2 inside every
//      pass-1 region), (key, value) pairs, LDS write-combined full lines;
//   4. k_pjoin: per work unit, the build partition goes into an LDS chained
//      table (head[H] u32, entries (key, value) int2, next u16 -- the
//      reference's LDS layout, :657-692, sized for up to NCAP rows instead of
//      4096), then the unit's probe tuples walk it; matches fold into the
//      join accumulator (the parent reduce).  A build partition larger than
//      NCAP is processed in rounds of NCAP rows (each round re-reads the
//      unit's probe tuples), so skew never overflows LDS.
// Build partitions small enough for one LDS table skip steps 1-3 (B = 0).
#include <algorithm>
#include <cstdio>
#include <vector>

#include "radix.hpp"

namespace pj {

__global__ void k_hist_reduce(const uint32_t *__restrict__ partial, int rows,
                              int nb, uint32_t *__restrict__ hist) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  uint32_t s = 0;
  for (int r = 0; r < rows; ++r) s += partial[(uint64_t)r * nb + i];
  hist[i] = s;
}

// Dynamic-LDS histogram over 2^bits bins (bits <= 15).
__global__ __launch_bounds__(1024) void k_hist_dyn(const int32_t *__restrict__ keys,
                                                   uint64_t n, int bits,
                                                   uint32_t *__restrict__ partial) {
  extern __shared__ uint32_t hsh[];
  const int nb = 1 << bits;
  const int shift = 32 - bits;
  for (int i = threadIdx.x; i < nb; i += 1024) hsh[i] = 0;
  __syncthreads();
  uint64_t tid = (uint64_t)blockIdx.x * 1024 + threadIdx.x;
  uint64_t nthr = (uint64_t)gridDim.x * 1024;
  if (((uintptr_t)keys & 15) == 0) {
    const int4 *kv = reinterpret_cast<const int4 *>(keys);
    uint64_t nvec = n / 4;
    for (uint64_t i = tid; i < nvec; i += nthr) {
      int4 k = ld_nt4(&kv[i]);
      atomicAdd(&hsh[fmix32((uint32_t)k.x) >> shift], 1u);
      atomicAdd(&hsh[fmix32((uint32_t)k.y) >> shift], 1u);
      atomicAdd(&hsh[fmix32((uint32_t)k.z) >> shift], 1u);
      atomicAdd(&hsh[fmix32((uint32_t)k.w) >> shift], 1u);
    }
    for (uint64_t i = nvec * 4 + tid; i < n; i += nthr)
      atomicAdd(&hsh[fmix32((uint32_t)keys[i]) >> shift], 1u);
  } else {
    for (uint64_t i = tid; i < n; i += nthr)
      atomicAdd(&hsh[fmix32((uint32_t)keys[i]) >> shift], 1u);
  }
  __syncthreads();
  uint32_t *row = partial + (uint64_t)blockIdx.x * nb;
  for (int i = threadIdx.x; i < nb; i += 1024) row[i] = hsh[i];
}

struct JoinUnit {
  unsigned long long r_begin, s_begin;
  uint32_t r_len, s_len;
};

// Per-unit LDS chained table + probe.  Units [wg_begin[b], wg_begin[b+1])
// belong to workgroup b; consecutive units of one partition reuse the table.
template <int BLOCK, int H, int NCAP>
__global__ __launch_bounds__(BLOCK) void k_pjoin(
    const int32_t *__restrict__ rk, const int32_t *__restrict__ rv,
    const int32_t *__restrict__ sk, const int32_t *__restrict__ sv,
    const JoinUnit *__restrict__ units, const uint32_t *__restrict__ wg_begin,
    unsigned long long *__restrict__ acc_out) {
  static_assert(NCAP < 65535, "next indices are 16-bit");
  __shared__ uint32_t head[H];
  __shared__ int2 ent[NCAP];
  __shared__ uint16_t nxt[NCAP];
  const int tid = threadIdx.x;
  JoinAcc acc;
  acc.init();
  const uint32_t u0 = wg_begin[blockIdx.x], u1 = wg_begin[blockIdx.x + 1];
  unsigned long long built_r = ~0ULL;
  for (uint32_t u = u0; u < u1; ++u) {
    const JoinUnit U = units[u];
    const uint32_t rounds = (U.r_len + NCAP - 1) / NCAP;
    for (uint32_t round = 0; round < rounds; ++round) {
      const uint32_t rb = round * NCAP;
      const uint32_t rl = (U.r_len - rb) < (uint32_t)NCAP ? (U.r_len - rb) : (uint32_t)NCAP;
      const bool reuse = rounds == 1 && built_r == U.r_begin;
      if (!reuse) {
        __syncthreads();
        for (int i = tid; i < H; i += BLOCK) head[i] = 0xffffffffu;
        __syncthreads();
        const int32_t *rkp = rk + U.r_begin + rb;
        const int32_t *rvp = rv ? rv + U.r_begin + rb : nullptr;
        for (uint32_t i = tid; i < rl; i += BLOCK) {
          int32_t k = rkp[i];
          int32_t v = rvp ? rvp[i] : 0;
          ent[i] = make_int2(k, v);
          uint32_t old = atomicExch(&head[fmix32((uint32_t)k) & (H - 1)], i);
          nxt[i] = (uint16_t)old;
        }
        __syncthreads();
        built_r = rounds == 1 ? U.r_begin : ~0ULL;
      }
      const int32_t *skp = sk + U.s_begin;
      const int32_t *svp = sv ? sv + U.s_begin : nullptr;
      const uint32_t sl = U.s_len;
      const bool vec = (((uintptr_t)skp & 15) == 0) && (!svp || (((uintptr_t)svp & 15) == 0));
      for (uint32_t j = tid * 4; j < sl; j += BLOCK * 4) {
        int32_t k[4], v[4];
        if (vec && j + 3 < sl) {
          int4 kk = ld_nt4(skp + j);
          k[0] = kk.x; k[1] = kk.y; k[2] = kk.z; k[3] = kk.w;
          if (svp) {
            int4 vv = ld_nt4(svp + j);
            v[0] = vv.x; v[1] = vv.y; v[2] = vv.z; v[3] = vv.w;
          } else {
            v[0] = v[1] = v[2] = v[3] = 0;
          }
        } else {
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            bool ok = j + c < sl;
            k[c] = ok ? skp[j + c] : 0;
            v[c] = (ok && svp) ? svp[j + c] : 0;
          }
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          if (j + c >= sl) break;
          int64_t m = 0, sb = 0, mnb = INT64_MAX, mxb = INT64_MIN;
          uint32_t e = head[fmix32((uint32_t)k[c]) & (H - 1)];
          while (e != 0xffffffffu) {
            int2 kv = ent[e];
            if (kv.x == k[c]) {
              int64_t b = kv.y;
              m += 1;
              sb += b;
              mnb = b < mnb ? b : mnb;
              mxb = b > mxb ? b : mxb;
            }
            uint16_t nx = nxt[e];
            e = nx == 0xffff ? 0xffffffffu : (uint32_t)nx;
          }
          acc.add(m, sb, mnb, mxb, (int64_t)v[c]);
        }
      }
    }
  }
  join_acc_flush(acc, acc_out);
}

// ---------------------------------------------------------------------------
// host planning
// ---------------------------------------------------------------------------
struct Pack {  // host staging of plan arrays, uploaded with one copy
  std::vector<uint8_t> buf;
  template <class T>
  size_t add(const std::vector<T> &v) {
    size_t off = (buf.size() + 255) & ~(size_t)255;
    buf.resize(off + std::max<size_t>(v.size() * sizeof(T), 1));
    if (!v.empty()) std::memcpy(buf.data() + off, v.data(), v.size() * sizeof(T));
    return off;
  }
};

struct PassPlan {
  std::vector<uint64_t> reg_start, reg_count;  // input regions
  std::vector<uint32_t> tile_begin;
  std::vector<unsigned long long> cfull, ctail;  // output cursors
  uint32_t ntiles = 0, tiles_per_wg = 0, grid = 0;
  size_t o_rs = 0, o_rc = 0, o_tb = 0, o_cf = 0, o_ct = 0;
};

constexpr int kTile = 4096;
constexpr int kScatterBlock = 256;

static uint64_t align32(uint64_t x) { return (x + 31) & ~31ULL; }

static int scatter_occupancy(int fbits) { return fbits <= 7 ? 2 : 1; }

static void plan_pass(PassPlan &p, const std::vector<uint64_t> &rs,
                      const std::vector<uint64_t> &rc,
                      const std::vector<uint64_t> &out_start,
                      const std::vector<uint64_t> &out_count, int num_cus,
                      int fbits) {
  p.reg_start = rs;
  p.reg_count = rc;
  p.tile_begin.assign(rs.size() + 1, 0);
  for (size_t r = 0; r < rs.size(); ++r)
    p.tile_begin[r + 1] = p.tile_begin[r] + (uint32_t)ceil_div(rc[r], kTile);
  p.ntiles = p.tile_begin.back();
  p.cfull.resize(out_start.size());
  p.ctail.resize(out_start.size());
  for (size_t i = 0; i < out_start.size(); ++i) {
    p.cfull[i] = out_start[i];
    p.ctail[i] = out_start[i] + out_count[i];
  }
  uint64_t g = (uint64_t)num_cus * scatter_occupancy(fbits);
  if (g > p.ntiles) g = p.ntiles ? p.ntiles : 1;
  p.tiles_per_wg = (uint32_t)ceil_div(p.ntiles ? p.ntiles : 1, g);
  p.grid = (uint32_t)ceil_div(p.ntiles ? p.ntiles : 1, p.tiles_per_wg);
}

static void pack_pass(Pack &pk, PassPlan &p) {
  p.o_rs = pk.add(p.reg_start);
  p.o_rc = pk.add(p.reg_count);
  p.o_tb = pk.add(p.tile_begin);
  p.o_cf = pk.add(p.cfull);
  p.o_ct = pk.add(p.ctail);
}

static ScatterPlan device_plan(const PassPlan &p, uint8_t *dbase) {
  ScatterPlan s;
  s.reg_start = reinterpret_cast<const uint64_t *>(dbase + p.o_rs);
  s.reg_count = reinterpret_cast<const uint64_t *>(dbase + p.o_rc);
  s.tile_begin = reinterpret_cast<const uint32_t *>(dbase + p.o_tb);
  s.nreg = (uint32_t)p.reg_start.size();
  s.ntiles = p.ntiles;
  s.tiles_per_wg = p.tiles_per_wg;
  s.cfull = reinterpret_cast<unsigned long long *>(dbase + p.o_cf);
  s.ctail = reinterpret_cast<unsigned long long *>(dbase + p.o_ct);
  return s;
}

template <int F>
static int launch_scatter_f(pj_ctx *ctx, const char *name, const int32_t *ik,
                            const int32_t *iv, int32_t *ok, int32_t *ov,
                            const ScatterPlan &sp, uint32_t grid, int shift) {
  DigitRadix dig{shift, (uint32_t)(F - 1)};
  if (iv)
    PJ_LAUNCH(ctx, name,
              (k_scatter<F, kTile, kScatterBlock, DigitRadix, int32_t, NoVal, false>), grid,
              kScatterBlock, 0, ik, iv, (const NoVal *)nullptr, ok, ov, (NoVal *)nullptr, sp,
              dig);
  else
    PJ_LAUNCH(ctx, name,
              (k_scatter<F, kTile, kScatterBlock, DigitRadix, int32_t, NoVal, true>), grid,
              kScatterBlock, 0, ik, iv, (const NoVal *)nullptr, ok, ov, (NoVal *)nullptr, sp,
              dig);
  return PJ_OK;
}

static int launch_scatter(pj_ctx *ctx, const char *name, int fbits,
                          const int32_t *ik, const int32_t *iv, int32_t *ok,
                          int32_t *ov, const ScatterPlan &sp, uint32_t grid,
                          int shift) {
  if (sp.ntiles == 0) return PJ_OK;
  switch (fbits) {
    case 1: return launch_scatter_f<2>(ctx, name, ik, iv, ok, ov, sp, grid, shift);
    case 2: return launch_scatter_f<4>(ctx, name, ik, iv, ok, ov, sp, grid, shift);
    case 3: return launch_scatter_f<8>(ctx, name, ik, iv, ok, ov, sp, grid, shift);
    case 4: return launch_scatter_f<16>(ctx, name, ik, iv, ok, ov, sp, grid, shift);
    case 5: return launch_scatter_f<32>(ctx, name, ik, iv, ok, ov, sp, grid, shift);
    case 6: return launch_scatter_f<64>(ctx, name, ik, iv, ok, ov, sp, grid, shift);
    case 7: return launch_scatter_f<128>(ctx, name, ik, iv, ok, ov, sp, grid, shift);
    case 8: return launch_scatter_f<256>(ctx, name, ik, iv, ok, ov, sp, grid, shift);
    default: return fail(PJ_E_ARG, "radix bits %d outside [1, 8]", fbits);
  }
}

int radix_histogram(pj_ctx *ctx, const int32_t *d_k, uint64_t n, int bits,
                    std::vector<uint32_t> &h) {
  const int nb = 1 << bits;
  h.assign(nb, 0);
  if (!n) return PJ_OK;
  uint64_t grid = std::min<uint64_t>((uint64_t)ctx->num_cus, ceil_div(n, 1024 * 16));
  if (grid < 1) grid = 1;
  Scratch part, hist;
  PJ_TRY(part.alloc(ctx, grid * nb * 4));
  PJ_TRY(hist.alloc(ctx, (uint64_t)nb * 4));
  static bool attr_set = false;
  if (!attr_set) {
    (void)hipFuncSetAttribute((const void *)k_hist_dyn,
                              hipFuncAttributeMaxDynamicSharedMemorySize, 131072);
    attr_set = true;
  }
  PJ_LAUNCH(ctx, "radix_hist", k_hist_dyn, grid, 1024, (size_t)nb * 4, d_k, n, bits,
            part.as<uint32_t>());
  PJ_LAUNCH(ctx, "radix_hist_reduce", k_hist_reduce, ceil_div(nb, 256), 256, 0,
            part.as<const uint32_t>(), (int)grid, nb, hist.as<uint32_t>());
  return d2h(ctx, h.data(), hist.p, (uint64_t)nb * 4);
}

// Partitioned side: pass-1 regions, final regions, and both pass plans.
struct Side {
  uint64_t n = 0;
  std::vector<uint32_t> hist;               // [2^B]
  std::vector<uint64_t> r1_start, r1_count; // [2^bits1]
  std::vector<uint64_t> f_start, f_count;   // [2^B] final partitions
  uint64_t total1 = 0, total2 = 0;
  PassPlan p1, p2;
  Scratch k1, v1, k2, v2;
};

static void plan_side(Side &s, int bits1, int bits2, int num_cus) {
  const int F1 = 1 << bits1, F2 = 1 << bits2;
  s.r1_start.assign(F1, 0);
  s.r1_count.assign(F1, 0);
  s.f_start.assign((size_t)F1 * F2, 0);
  s.f_count.assign((size_t)F1 * F2, 0);
  uint64_t pos = 0;
  for (int d1 = 0; d1 < F1; ++d1) {
    uint64_t c = 0;
    for (int d2 = 0; d2 < F2; ++d2) c += s.hist[(size_t)d1 * F2 + d2];
    pos = align32(pos);
    s.r1_start[d1] = pos;
    s.r1_count[d1] = c;
    pos += c;
  }
  s.total1 = pos;
  if (bits2 == 0) {
    s.f_start = s.r1_start;
    s.f_count = s.r1_count;
    s.total2 = 0;
  } else {
    pos = 0;
    for (size_t i = 0; i < (size_t)F1 * F2; ++i) {
      pos = align32(pos);
      s.f_start[i] = pos;
      s.f_count[i] = s.hist[i];
      pos += s.hist[i];
    }
    s.total2 = pos;
  }
  plan_pass(s.p1, std::vector<uint64_t>{0}, std::vector<uint64_t>{s.n}, s.r1_start,
            s.r1_count, num_cus, bits1);
  if (bits2)
    plan_pass(s.p2, s.r1_start, s.r1_count, s.f_start, s.f_count, num_cus, bits2);
}

static int run_side(pj_ctx *ctx, Side &s, const int32_t *d_k, const int32_t *d_v,
                    int bits1, int bits2, uint8_t *dplan, const int32_t **fk,
                    const int32_t **fv, const char *tag) {
  const int B = bits1 + bits2;
  PJ_TRY(s.k1.alloc(ctx, s.total1 * 4 + 256));
  PJ_TRY(s.v1.alloc(ctx, s.total1 * 4 + 256));
  char name[64];
  snprintf(name, sizeof(name), "radix_scatter1_%s", tag);
  PJ_TRY(launch_scatter(ctx, name, bits1, d_k, d_v, s.k1.as<int32_t>(), s.v1.as<int32_t>(),
                        device_plan(s.p1, dplan), s.p1.grid, 32 - bits1));
  if (bits2) {
    PJ_TRY(s.k2.alloc(ctx, s.total2 * 4 + 256));
    PJ_TRY(s.v2.alloc(ctx, s.total2 * 4 + 256));
    snprintf(name, sizeof(name), "radix_scatter2_%s", tag);
    PJ_TRY(launch_scatter(ctx, name, bits2, s.k1.as<int32_t>(), s.v1.as<int32_t>(),
                          s.k2.as<int32_t>(), s.v2.as<int32_t>(), device_plan(s.p2, dplan),
                          s.p2.grid, 32 - B));
    *fk = s.k2.as<int32_t>();
    *fv = s.v2.as<int32_t>();
  } else {
    *fk = s.k1.as<int32_t>();
    *fv = s.v1.as<int32_t>();
  }
  return PJ_OK;
}

// LDS table geometries.  NCAP rows, H heads; occupancy from the LDS bytes.
struct Geometry {
  int lds_log;
  int ncap;
};
static Geometry geometry_for(int lds_log) {
  switch (lds_log) {
    case 11: return {11, 3072};
    case 12: return {12, 6144};
    default: return {13, 12288};
  }
}

constexpr int kJoinBlock = 512;

static int launch_pjoin(pj_ctx *ctx, int lds_log, const int32_t *rk, const int32_t *rv,
                        const int32_t *sk, const int32_t *sv, const JoinUnit *units,
                        const uint32_t *wgb, uint32_t grid, unsigned long long *acc) {
  switch (lds_log) {
    case 11:
      PJ_LAUNCH(ctx, "pjoin_probe", (k_pjoin<kJoinBlock, 2048, 3072>), grid, kJoinBlock, 0, rk,
                rv, sk, sv, units, wgb, acc);
      return PJ_OK;
    case 12:
      PJ_LAUNCH(ctx, "pjoin_probe", (k_pjoin<kJoinBlock, 4096, 6144>), grid, kJoinBlock, 0, rk,
                rv, sk, sv, units, wgb, acc);
      return PJ_OK;
    case 13:
      PJ_LAUNCH(ctx, "pjoin_probe", (k_pjoin<kJoinBlock, 8192, 12288>), grid, kJoinBlock, 0,
                rk, rv, sk, sv, units, wgb, acc);
      return PJ_OK;
    default:
      return fail(PJ_E_ARG, "lds_log %d not in {11, 12, 13}", lds_log);
  }
}
static int join_occupancy(int lds_log) { return lds_log >= 13 ? 1 : (lds_log == 12 ? 2 : 4); }

// Default configuration: final partitions of ~2/3 of the LDS table capacity.
static pj_radix_cfg default_cfg(uint64_t nb) {
  pj_radix_cfg c{};
  c.lds_log = 12;
  c.probe_unit = 65536;
  c.grid_mul = 0;
  Geometry g = geometry_for(c.lds_log);
  uint64_t target = (uint64_t)g.ncap * 2 / 3;
  int B = 0;
  while (B < 15 && (nb >> B) > target) ++B;
  if (B <= 8) {
    c.bits1 = B;
    c.bits2 = 0;
  } else {
    c.bits1 = B / 2 + (B & 1);
    c.bits2 = B - c.bits1;
  }
  return c;
}

static void balance_units(const std::vector<JoinUnit> &units, uint32_t grid,
                          std::vector<uint32_t> &wgb) {
  wgb.assign(grid + 1, (uint32_t)units.size());
  uint64_t total = 0;
  for (auto &u : units) total += (uint64_t)u.s_len + u.r_len;
  uint64_t per = ceil_div(total ? total : 1, grid);
  uint64_t acc = 0;
  uint32_t w = 0;
  wgb[0] = 0;
  for (uint32_t i = 0; i < units.size(); ++i) {
    while (w + 1 <= grid && acc >= per * (w + 1)) wgb[++w] = i;
    acc += (uint64_t)units[i].s_len + units[i].r_len;
  }
  while (w + 1 <= grid) wgb[++w] = (uint32_t)units.size();
}

static int join_partitioned(pj_ctx *ctx, const int32_t *d_bk, const int32_t *d_bv,
                            uint64_t nb, const int32_t *d_pk, const int32_t *d_pv,
                            uint64_t np, pj_radix_cfg cfg, pj_join_agg *out) {
  const int B = cfg.bits1 + cfg.bits2;
  Scratch acc;
  PJ_TRY(acc.alloc(ctx, 64));
  PJ_TRY(join_acc_init(ctx, acc.as<unsigned long long>()));
  Geometry geo = geometry_for(cfg.lds_log);
  std::vector<JoinUnit> units;
  Side R, S;
  R.n = nb;
  S.n = np;
  Pack pk;
  const int32_t *rk = d_bk, *rv = d_bv, *sk = d_pk, *sv = d_pv;
  if (nb && np) {
    if (B == 0) {
      for (uint64_t s0 = 0; s0 < np; s0 += cfg.probe_unit) {
        JoinUnit u;
        u.r_begin = 0;
        u.r_len = (uint32_t)nb;
        u.s_begin = s0;
        u.s_len = (uint32_t)std::min<uint64_t>(cfg.probe_unit, np - s0);
        units.push_back(u);
      }
    } else {
      PJ_TRY(radix_histogram(ctx, d_bk, nb, B, R.hist));
      PJ_TRY(radix_histogram(ctx, d_pk, np, B, S.hist));
      plan_side(R, cfg.bits1, cfg.bits2, ctx->num_cus);
      plan_side(S, cfg.bits1, cfg.bits2, ctx->num_cus);
      for (size_t q = 0; q < R.f_start.size(); ++q) {
        uint64_t rl = R.f_count[q], sl = S.f_count[q];
        if (!rl || !sl) continue;
        for (uint64_t s0 = 0; s0 < sl; s0 += cfg.probe_unit) {
          JoinUnit u;
          u.r_begin = R.f_start[q];
          u.r_len = (uint32_t)rl;
          u.s_begin = S.f_start[q] + s0;
          u.s_len = (uint32_t)std::min<uint64_t>(cfg.probe_unit, sl - s0);
          units.push_back(u);
        }
      }
      pack_pass(pk, R.p1);
      pack_pass(pk, S.p1);
      if (cfg.bits2) {
        pack_pass(pk, R.p2);
        pack_pass(pk, S.p2);
      }
    }
  }
  uint32_t grid = 1;
  std::vector<uint32_t> wgb;
  size_t o_units = 0, o_wgb = 0;
  if (!units.empty()) {
    uint64_t g = (uint64_t)ctx->num_cus * join_occupancy(geo.lds_log) *
                 (cfg.grid_mul > 0 ? cfg.grid_mul : 1);
    grid = (uint32_t)std::min<uint64_t>(g, units.size());
    balance_units(units, grid, wgb);
    o_units = pk.add(units);
    o_wgb = pk.add(wgb);
  }
  Scratch dplan;
  if (!pk.buf.empty()) {
    PJ_TRY(dplan.alloc(ctx, pk.buf.size()));
    PJ_TRY(h2d(ctx, dplan.p, pk.buf.data(), pk.buf.size()));
  }
  uint8_t *dp = dplan.as<uint8_t>();
  if (nb && np && B > 0) {
    PJ_TRY(run_side(ctx, R, d_bk, d_bv, cfg.bits1, cfg.bits2, dp, &rk, &rv, "build"));
    PJ_TRY(run_side(ctx, S, d_pk, d_pv, cfg.bits1, cfg.bits2, dp, &sk, &sv, "probe"));
    if (!d_bv) rv = nullptr;
    if (!d_pv) sv = nullptr;
  }
  if (!units.empty()) {
    PJ_TRY(launch_pjoin(ctx, geo.lds_log, rk, rv, sk, sv,
                        reinterpret_cast<const JoinUnit *>(dp + o_units),
                        reinterpret_cast<const uint32_t *>(dp + o_wgb), grid,
                        acc.as<unsigned long long>()));
  }
  return join_acc_read(ctx, acc.as<unsigned long long>(), out);
}

int validate_cfg(pj_radix_cfg &c, uint64_t nb) {
  pj_radix_cfg d = default_cfg(nb);
  if (c.lds_log == 0) c.lds_log = d.lds_log;
  if (c.probe_unit == 0) c.probe_unit = d.probe_unit;
  if (c.bits1 == 0 && c.bits2 == 0) {
    // derive the bits for the chosen table size
    Geometry g = geometry_for(c.lds_log);
    uint64_t target = (uint64_t)g.ncap * 2 / 3;
    int B = 0;
    while (B < 15 && (nb >> B) > target) ++B;
    if (B <= 8) {
      c.bits1 = B;
      c.bits2 = 0;
    } else {
      c.bits1 = B / 2 + (B & 1);
      c.bits2 = B - c.bits1;
    }
  }
  if (c.bits1 < 0 || c.bits1 > 8 || c.bits2 < 0 || c.bits2 > 8 || (c.bits1 == 0 && c.bits2 > 0))
    return fail(PJ_E_ARG, "radix bits (%d, %d) invalid: each in [0, 8], bits2 needs bits1",
                c.bits1, c.bits2);
  if (c.bits1 + c.bits2 > 15)
    return fail(PJ_E_ARG, "bits1 + bits2 = %d exceeds 15", c.bits1 + c.bits2);
  if (c.lds_log < 11 || c.lds_log > 13) return fail(PJ_E_ARG, "lds_log %d not in [11, 13]", c.lds_log);
  if (c.probe_unit < 1024 || (c.probe_unit & 3))
    return fail(PJ_E_ARG, "probe_unit must be a multiple of 4 and >= 1024");
  if (c.grid_mul < 0 || c.grid_mul > 8) return fail(PJ_E_ARG, "grid_mul %d not in [0, 8]", c.grid_mul);
  return PJ_OK;
}

// ---------------------------------------------------------------------------
// hash-rearrange
// ---------------------------------------------------------------------------
template <int F>
static int rearrange_f(pj_ctx *ctx, const int32_t *d_k, const int32_t *d_v, uint64_t n,
                       uint32_t buckets, int32_t *d_ok, int32_t *d_ov, uint64_t *h_off) {
  DigitMod dig{buckets};
  uint64_t grid = std::min<uint64_t>((uint64_t)ctx->num_cus * 2, ceil_div(n, 4096));
  if (grid < 1) grid = 1;
  Scratch part, hist;
  PJ_TRY(part.alloc(ctx, grid * F * 4));
  PJ_TRY(hist.alloc(ctx, F * 4));
  PJ_LAUNCH(ctx, "rearrange_hist", (k_hist_digit<F, 256, DigitMod>), grid, 256, 0, d_k, n, dig,
            part.as<uint32_t>());
  PJ_LAUNCH(ctx, "radix_hist_reduce", k_hist_reduce, ceil_div(F, 256), 256, 0,
            part.as<const uint32_t>(), (int)grid, F, hist.as<uint32_t>());
  std::vector<uint32_t> hc(F);
  PJ_TRY(d2h(ctx, hc.data(), hist.p, F * 4));
  // Output regions must be 32-element aligned for full-line writes, so the
  // scatter goes to an aligned staging buffer that is then compacted.
  std::vector<uint64_t> st(F), cnt(F), dense(F + 1, 0);
  uint64_t pos = 0;
  for (int b = 0; b < F; ++b) {
    pos = align32(pos);
    st[b] = pos;
    cnt[b] = hc[b];
    pos += hc[b];
    dense[b + 1] = dense[b] + hc[b];
  }
  PassPlan p;
  plan_pass(p, std::vector<uint64_t>{0}, std::vector<uint64_t>{n}, st, cnt, ctx->num_cus, 5);
  Pack pk;
  pack_pass(pk, p);
  Scratch dplan, sk, sv;
  PJ_TRY(dplan.alloc(ctx, pk.buf.size()));
  PJ_TRY(h2d(ctx, dplan.p, pk.buf.data(), pk.buf.size()));
  PJ_TRY(sk.alloc(ctx, pos * 4 + 256));
  PJ_TRY(sv.alloc(ctx, pos * 4 + 256));
  if (p.ntiles) {
    if (d_v)
      PJ_LAUNCH(ctx, "rearrange_scatter",
                (k_scatter<F, kTile, kScatterBlock, DigitMod, int32_t, NoVal, false>), p.grid,
                kScatterBlock, 0, d_k, d_v, (const NoVal *)nullptr, sk.as<int32_t>(),
                sv.as<int32_t>(), (NoVal *)nullptr, device_plan(p, dplan.as<uint8_t>()), dig);
    else
      PJ_LAUNCH(ctx, "rearrange_scatter",
                (k_scatter<F, kTile, kScatterBlock, DigitMod, int32_t, NoVal, true>), p.grid,
                kScatterBlock, 0, d_k, d_v, (const NoVal *)nullptr, sk.as<int32_t>(),
                sv.as<int32_t>(), (NoVal *)nullptr, device_plan(p, dplan.as<uint8_t>()), dig);
  }
  for (uint32_t b = 0; b < buckets; ++b) {
    if (!cnt[b]) continue;
    PJ_HIP(hipMemcpyAsync(d_ok + dense[b], sk.as<int32_t>() + st[b], cnt[b] * 4,
                          hipMemcpyDeviceToDevice, ctx->stream));
    if (d_ov)
      PJ_HIP(hipMemcpyAsync(d_ov + dense[b], sv.as<int32_t>() + st[b], cnt[b] * 4,
                            hipMemcpyDeviceToDevice, ctx->stream));
  }
  PJ_TRY(sync(ctx));
  for (uint32_t b = 0; b <= buckets; ++b) h_off[b] = dense[b];
  return PJ_OK;
}

}  // namespace pj

using namespace pj;

extern "C" int pj_join_partitioned_config(pj_ctx *ctx, uint64_t nb, uint64_t np,
                                          pj_radix_cfg *cfg) {
  (void)np;
  pj_radix_cfg c{};
  auto it = ctx->tuned.find(TuneKey{ilog2_ceil(nb ? nb : 1), ilog2_ceil(np ? np : 1)});
  if (it != ctx->tuned.end()) c = it->second;
  PJ_TRY(validate_cfg(c, nb));
  *cfg = c;
  return PJ_OK;
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
  PJ_TRY(validate_cfg(c, nb));
  return join_partitioned(ctx, d_bk, d_bv, nb, d_pk, d_pv, np, c, out);
}

// Sweep a small grid of configurations on the given inputs and remember the
// fastest per (log2 build, log2 probe) size class.
int pj_autotune_join(pj_ctx *ctx, const int32_t *d_bk, const int32_t *d_bv, uint64_t nb,
                     const int32_t *d_pk, const int32_t *d_pv, uint64_t np) {
  std::vector<pj_radix_cfg> cands;
  for (int lds : {12, 13}) {
    pj_radix_cfg base{};
    base.lds_log = lds;
    if (validate_cfg(base, nb) != PJ_OK) continue;
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
