// Out-of-GPU co-partitioned join -> reduce: the reference's "out-of-gpu-join"
// (core/olap/lib/plan/plan-parser.cpp:531-850, core/olap/lib/operators/
// packet-zip.cpp): both inputs live in host memory and may exceed the HBM
// budget.  The reference hash-rearranges them on CPU threads into
// numOfBuckets buckets (bucket = hashInt32(key) URem buckets, HashRearrange)
// and ZipCollect ships each bucket pair to a GpuPartitionedHashJoinChained.
//
// Here the GPU does the bucketing too, streaming: each side goes to HBM in
// chunks that fit the budget, is rearranged by bucket on the device
// (pj_hash_rearrange_i32, the same bucket function), and comes back to a
// pinned host arena, where bucket b of the side is a list of pieces (one per
// chunk).  Then each bucket's pieces of both sides are copied to HBM and
// joined with the radix-partitioned join; the per-bucket aggregates combine
// exactly (int64 counts and sums add, minima and maxima combine), since
// equal keys fall in the same bucket.  PCIe carries 24 B per tuple (8 in,
// 8 back, 8 in again).  The caller's (pageable) chunks come in through the
// context's pinned staging ring, the next chunk while the current one is
// rearranged; a chunk's copy back to the arena overlaps the next chunk's
// upload; a bucket's pieces go up from the pinned arena as one queue of
// asynchronous copies that the bucket's join follows on the same stream.
#include <algorithm>
#include <vector>

#include "pj_internal.hpp"

namespace pj {

namespace {

struct Piece {
  uint64_t off, len;  // in the side's host arena
};

struct DevBuf {
  pj_ctx *ctx;
  void *p = nullptr;
  explicit DevBuf(pj_ctx *c) : ctx(c) {}
  int alloc(uint64_t bytes) { return pj_malloc(ctx, bytes ? bytes : 4, &p); }
  ~DevBuf() {
    if (p) (void)pj_free(ctx, p);
  }
};

struct HostArena {
  void *p = nullptr;
  int alloc(uint64_t bytes) {
    PJ_HIP(hipHostMalloc(&p, bytes ? bytes : 8, hipHostMallocDefault));
    return PJ_OK;
  }
  ~HostArena() {
    if (p) (void)hipHostFree(p);
  }
};

// One side: host (k, v) -> arena keys / values grouped by bucket per chunk.
// Each chunk's rearranged output goes back to the pinned arena on a copy
// stream while the next chunk comes in (the link is full duplex); two output
// buffer sets alternate, each reused only after its copy has finished.
struct CopyStream {
  hipStream_t s = nullptr;
  hipEvent_t done[2] = {nullptr, nullptr};
  int init() {
    PJ_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    for (auto &e : done) PJ_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return PJ_OK;
  }
  ~CopyStream() {
    if (s) (void)hipStreamSynchronize(s);
    for (auto &e : done)
      if (e) (void)hipEventDestroy(e);
    if (s) (void)hipStreamDestroy(s);
  }
};

// Staging tickets not yet waited for; the destructor waits for them and for
// the context stream, so no copy still writes into buffers that are freed
// after an error.  Declared after the buffers it protects.
struct Tickets {
  pj_ctx *ctx;
  std::vector<uint64_t> open;
  explicit Tickets(pj_ctx *c) : ctx(c) {}
  int begin(const void *h, uint64_t bytes, void *d, uint64_t *t) {
    PJ_TRY(pj_stage_begin(ctx, nullptr, h, 0, bytes, d, t));
    open.push_back(*t);
    return PJ_OK;
  }
  int wait(uint64_t t) {
    open.erase(std::find(open.begin(), open.end(), t));
    return pj_stage_wait(ctx, t);
  }
  ~Tickets() {
    for (uint64_t t : open) (void)pj_stage_wait(ctx, t);
    (void)hipStreamSynchronize(ctx->stream);
  }
};

int bucket_side(pj_ctx *ctx, const int32_t *hk, const int32_t *hv, uint64_t n, uint32_t buckets,
                uint64_t chunk, int32_t *ak, int32_t *av, std::vector<std::vector<Piece>> &pieces) {
  pieces.assign(buckets, {});
  if (!n) return PJ_OK;
  const uint64_t c = std::min(chunk, n);
  const uint64_t nchunks = ceil_div(n, c);
  // two input and two output buffer sets: chunk i + 1 comes in while chunk
  // i is rearranged, chunk i goes back while chunk i + 1 is rearranged
  DevBuf ik0(ctx), iv0(ctx), ik1(ctx), iv1(ctx), ok0(ctx), ov0(ctx), ok1(ctx), ov1(ctx);
  PJ_TRY(ik0.alloc(c * 4));
  PJ_TRY(ok0.alloc(c * 4));
  PJ_TRY(ok1.alloc(c * 4));
  if (nchunks > 1) PJ_TRY(ik1.alloc(c * 4));
  if (hv) {
    PJ_TRY(iv0.alloc(c * 4));
    PJ_TRY(ov0.alloc(c * 4));
    PJ_TRY(ov1.alloc(c * 4));
    if (nchunks > 1) PJ_TRY(iv1.alloc(c * 4));
  }
  DevBuf *ik[2] = {&ik0, &ik1}, *iv[2] = {&iv0, &iv1};
  DevBuf *ok[2] = {&ok0, &ok1}, *ov[2] = {&ov0, &ov1};
  CopyStream cs;
  PJ_TRY(cs.init());
  Tickets tk(ctx);
  uint64_t tkk[2] = {0, 0}, tkv[2] = {0, 0};
  auto stage = [&](uint64_t i) -> int {
    const uint64_t base = i * c, len = std::min(c, n - base);
    const int s = (int)(i & 1);
    PJ_TRY(tk.begin(hk + base, len * 4, ik[s]->p, &tkk[s]));
    if (hv) PJ_TRY(tk.begin(hv + base, len * 4, iv[s]->p, &tkv[s]));
    return PJ_OK;
  };
  std::vector<uint64_t> off(buckets + 1);
  PJ_TRY(stage(0));
  for (uint64_t i = 0; i < nchunks; ++i) {
    const int b = (int)(i & 1);
    const uint64_t base = i * c, len = std::min(c, n - base);
    // input set b ^ 1 was last read by chunk i - 1's rearrange, which has
    // completed (pj_hash_rearrange_i32 synchronises)
    if (i + 1 < nchunks) PJ_TRY(stage(i + 1));
    PJ_TRY(tk.wait(tkk[b]));  // the context stream now waits for chunk i's DMA
    if (hv) PJ_TRY(tk.wait(tkv[b]));
    if (i >= 2) PJ_HIP(hipEventSynchronize(cs.done[b]));  // output set b is free
    PJ_TRY(pj_hash_rearrange_i32(ctx, (const int32_t *)ik[b]->p,
                                 hv ? (const int32_t *)iv[b]->p : nullptr, len, buckets,
                                 (int32_t *)ok[b]->p, hv ? (int32_t *)ov[b]->p : nullptr,
                                 off.data()));
    // the rearrange has synchronised: its output is complete
    PJ_HIP(hipMemcpyAsync(ak + base, ok[b]->p, len * 4, hipMemcpyDeviceToHost, cs.s));
    if (hv) PJ_HIP(hipMemcpyAsync(av + base, ov[b]->p, len * 4, hipMemcpyDeviceToHost, cs.s));
    PJ_HIP(hipEventRecord(cs.done[b], cs.s));
    for (uint32_t q = 0; q < buckets; ++q)
      if (off[q + 1] > off[q]) pieces[q].push_back(Piece{base + off[q], off[q + 1] - off[q]});
  }
  PJ_HIP(hipStreamSynchronize(cs.s));
  return PJ_OK;
}

uint64_t total(const std::vector<Piece> &ps) {
  uint64_t t = 0;
  for (const Piece &p : ps) t += p.len;
  return t;
}

// The pieces of one bucket, concatenated into device memory: asynchronous
// copies from the pinned arena on the context stream, which the bucket's
// join then follows.
int gather_pieces(pj_ctx *ctx, const int32_t *arena, const std::vector<Piece> &ps, int32_t *dst) {
  uint64_t at = 0;
  for (const Piece &p : ps) {
    PJ_HIP(hipMemcpyAsync(dst + at, arena + p.off, p.len * 4, hipMemcpyHostToDevice,
                          ctx->stream));
    at += p.len;
  }
  return PJ_OK;
}

}  // namespace
}  // namespace pj

using namespace pj;

extern "C" int pj_join_out_of_gpu_reduce(pj_ctx *ctx, const int32_t *h_bk, const int32_t *h_bv,
                                         uint64_t nb, const int32_t *h_pk, const int32_t *h_pv,
                                         uint64_t np, uint32_t num_buckets, uint64_t hbm_budget,
                                         pj_join_agg *out) {
  if (!ctx || !out) return fail(PJ_E_ARG, "null argument");
  *out = pj_join_agg{0, 0, 0, INT64_MAX, INT64_MIN, INT64_MAX, INT64_MIN};
  if ((nb && !h_bk) || (np && !h_pk)) return fail(PJ_E_ARG, "out-of-GPU join without keys");
  if (!nb || !np) return PJ_OK;
  if (!hbm_budget) hbm_budget = ctx->hbm_bytes / 2;
  // a bucket pair in HBM takes its 8 B per tuple plus the partitioned join's
  // scratch, about 3x in all; buckets sized for half that, for skew
  constexpr uint64_t kJoinBytes = 24;
  // The device bucketing is one radix pass (pj_hash_rearrange_i32), which
  // takes at most kMaxBuckets buckets.
  constexpr uint32_t kMaxBuckets = 256;
  if (num_buckets > kMaxBuckets)
    return fail(PJ_E_ARG, "out-of-GPU join with %u buckets: at most %u", num_buckets, kMaxBuckets);
  uint32_t buckets = num_buckets;
  if (!buckets) {
    const uint64_t need = (nb + np) * kJoinBytes * 2;
    const uint64_t want = std::max<uint64_t>(1, ceil_div(need, hbm_budget));
    if (want > kMaxBuckets)
      return fail(PJ_E_CAPACITY,
                  "out-of-GPU join: %llu + %llu tuples need %llu buckets under an HBM budget of "
                  "%llu bytes, more than the %u one bucketing pass makes (raise the budget)",
                  (unsigned long long)nb, (unsigned long long)np, (unsigned long long)want,
                  (unsigned long long)hbm_budget, kMaxBuckets);
    buckets = (uint32_t)want;
  }
  // chunks of the bucketing phase: 24 B per tuple in HBM (input and two
  // output sets), several chunks per side so that copies overlap
  const uint64_t chunk = std::max<uint64_t>(1 << 16, hbm_budget / 64);
  HostArena bk, bv, pk, pv;
  PJ_TRY(bk.alloc(nb * 4));
  PJ_TRY(pk.alloc(np * 4));
  if (h_bv) PJ_TRY(bv.alloc(nb * 4));
  if (h_pv) PJ_TRY(pv.alloc(np * 4));
  std::vector<std::vector<Piece>> bpieces, ppieces;
  PJ_TRY(bucket_side(ctx, h_bk, h_bv, nb, buckets, chunk, (int32_t *)bk.p, (int32_t *)bv.p, bpieces));
  PJ_TRY(bucket_side(ctx, h_pk, h_pv, np, buckets, chunk, (int32_t *)pk.p, (int32_t *)pv.p, ppieces));
  uint64_t max_nb = 0, max_np = 0;
  for (uint32_t b = 0; b < buckets; ++b) {
    const uint64_t rb = total(bpieces[b]), sb = total(ppieces[b]);
    if ((rb + sb) * kJoinBytes > hbm_budget)
      return fail(PJ_E_CAPACITY,
                  "out-of-GPU join: bucket %u holds %llu + %llu tuples, beyond the HBM budget "
                  "of %llu bytes (use more buckets)",
                  b, (unsigned long long)rb, (unsigned long long)sb, (unsigned long long)hbm_budget);
    max_nb = std::max(max_nb, rb);
    max_np = std::max(max_np, sb);
  }
  DevBuf drk(ctx), drv(ctx), dsk(ctx), dsv(ctx);
  // after the buffers: no queued piece copy outlives them on an error return
  struct StreamDrain {
    pj_ctx *ctx;
    ~StreamDrain() { (void)hipStreamSynchronize(ctx->stream); }
  } drain{ctx};
  PJ_TRY(drk.alloc(max_nb * 4));
  PJ_TRY(dsk.alloc(max_np * 4));
  if (h_bv) PJ_TRY(drv.alloc(max_nb * 4));
  if (h_pv) PJ_TRY(dsv.alloc(max_np * 4));
  for (uint32_t b = 0; b < buckets; ++b) {
    const uint64_t rb = total(bpieces[b]), sb = total(ppieces[b]);
    if (!rb || !sb) continue;
    PJ_TRY(gather_pieces(ctx, (const int32_t *)bk.p, bpieces[b], (int32_t *)drk.p));
    PJ_TRY(gather_pieces(ctx, (const int32_t *)pk.p, ppieces[b], (int32_t *)dsk.p));
    if (h_bv) PJ_TRY(gather_pieces(ctx, (const int32_t *)bv.p, bpieces[b], (int32_t *)drv.p));
    if (h_pv) PJ_TRY(gather_pieces(ctx, (const int32_t *)pv.p, ppieces[b], (int32_t *)dsv.p));
    pj_join_agg a;
    PJ_TRY(pj_join_partitioned_reduce(ctx, (const int32_t *)drk.p,
                                      h_bv ? (const int32_t *)drv.p : nullptr, rb,
                                      (const int32_t *)dsk.p,
                                      h_pv ? (const int32_t *)dsv.p : nullptr, sb, rb, sb,
                                      nullptr, &a));
    out->count += a.count;
    out->sum_b += a.sum_b;
    out->sum_p += a.sum_p;
    out->min_b = std::min(out->min_b, a.min_b);
    out->max_b = std::max(out->max_b, a.max_b);
    out->min_p = std::min(out->min_p, a.min_p);
    out->max_p = std::max(out->max_p, a.max_p);
  }
  return PJ_OK;
}
