// Serving loop over the batched-query entry points: Poisson (or any)
// arrivals, admission by an HBM budget for the queries' probe inputs,
// continuous batching across stream slots.  Uses only the C ABI.
#include <algorithm>
#include <chrono>
#include <deque>
#include <thread>
#include <vector>

#include "engine.hpp"

namespace {

using Clock = std::chrono::steady_clock;

struct InFlight {
  std::vector<uint32_t> qs;
  std::vector<void *> bufs;
  uint64_t bytes = 0;
  bool busy = false;
};

uint64_t input_bytes(const pj_query &q) {
  return (q.gen_seed && !q.d_fk) ? q.n * 12 : 0;
}

}  // namespace

extern "C" int pj_serve_run(pj_ctx *ctx, pj_query *qs, const double *arrival_s, uint32_t nq,
                            const pj_serve_cfg *cfg, double *done_s, pj_serve_stats *stats) {
  if (!ctx || (!qs && nq) || !arrival_s || !done_s || !cfg)
    return pj_fail(PJ_E_ARG, "pj_serve_run: null argument");
  for (uint32_t i = 1; i < nq; ++i)
    if (arrival_s[i] < arrival_s[i - 1]) return pj_fail(PJ_E_ARG, "arrivals must be sorted");
  const uint64_t budget = cfg->hbm_budget ? cfg->hbm_budget : ~0ULL;
  for (uint32_t i = 0; i < nq; ++i)
    if (input_bytes(qs[i]) > budget)
      return pj_fail(PJ_E_CAPACITY, "query %u needs %llu bytes of inputs, budget %llu", i,
                     (unsigned long long)input_bytes(qs[i]), (unsigned long long)budget);
  pj_serve_stats st{};
  const auto t0 = Clock::now();
  auto now = [&]() { return std::chrono::duration<double>(Clock::now() - t0).count(); };
  auto alloc_inputs = [&](pj_query &q, std::vector<void *> &bufs) -> int {
    if (!input_bytes(q)) return PJ_OK;
    void *p = nullptr;
    int rc = pj_malloc(ctx, q.n * 12, &p);
    if (rc != PJ_OK) return rc;
    bufs.push_back(p);
    q.d_fk = (const int32_t *)p;
    q.d_val = (const int32_t *)((char *)p + q.n * 4);
    q.d_grp = (const int32_t *)((char *)p + q.n * 8);
    return PJ_OK;
  };
  auto release = [&](std::vector<void *> &bufs, const std::vector<uint32_t> &ids) {
    for (void *p : bufs) pj_free(ctx, p);
    bufs.clear();
    for (uint32_t i : ids)
      if (qs[i].gen_seed) qs[i].d_fk = qs[i].d_val = qs[i].d_grp = nullptr;
  };

  if (cfg->mode == 1) {
    for (uint32_t i = 0; i < nq; ++i) {
      double wait = arrival_s[i] - now();
      if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
      std::vector<void *> bufs;
      int rc = alloc_inputs(qs[i], bufs);
      if (rc == PJ_OK) rc = pj_query_run(ctx, &qs[i], &cfg->batch);
      release(bufs, {i});
      if (rc != PJ_OK) return rc;
      done_s[i] = now();
      st.batches++;
      st.max_in_flight_bytes = std::max(st.max_in_flight_bytes, input_bytes(qs[i]));
    }
    st.wall_s = now();
    if (stats) *stats = st;
    return PJ_OK;
  }

  const int nslots = cfg->slots > 0 ? cfg->slots : 2;
  const uint32_t max_batch = cfg->max_batch ? std::min<uint32_t>(cfg->max_batch, 4096) : 4096;
  int rc = pj_batch_slots(ctx, nslots);
  if (rc != PJ_OK) return rc;
  std::vector<InFlight> fl(nslots);
  std::deque<uint32_t> waiting;
  uint32_t next = 0, completed = 0;
  uint64_t in_flight = 0;
  std::vector<pj_query> batch;
  int err = PJ_OK;
  while (completed < nq && err == PJ_OK) {
    bool progress = false;
    const double t = now();
    while (next < nq && arrival_s[next] <= t) waiting.push_back(next++);
    for (int s = 0; s < nslots; ++s) {
      if (!fl[s].busy) continue;
      int done = 0;
      if ((err = pj_batch_poll(ctx, s, &done)) != PJ_OK) break;
      if (!done) continue;
      const double tc = now();
      for (uint32_t i : fl[s].qs) done_s[i] = tc;
      completed += (uint32_t)fl[s].qs.size();
      in_flight -= fl[s].bytes;
      release(fl[s].bufs, fl[s].qs);
      fl[s] = InFlight();
      progress = true;
    }
    if (err != PJ_OK) break;
    for (int s = 0; s < nslots && !waiting.empty(); ++s) {
      if (fl[s].busy) continue;
      InFlight &f = fl[s];
      batch.clear();
      while (!waiting.empty() && batch.size() < max_batch) {
        pj_query &q = qs[waiting.front()];
        const uint64_t b = input_bytes(q);
        if (in_flight + f.bytes + b > budget) break;
        if ((err = alloc_inputs(q, f.bufs)) != PJ_OK) break;
        f.bytes += b;
        f.qs.push_back(waiting.front());
        batch.push_back(q);
        waiting.pop_front();
      }
      if (err != PJ_OK) break;
      if (batch.empty()) break;  // budget full: wait for a completion
      if ((err = pj_batch_launch(ctx, s, batch.data(), (uint32_t)batch.size(), &cfg->batch)) != PJ_OK)
        break;
      f.busy = true;
      in_flight += f.bytes;
      st.max_in_flight_bytes = std::max(st.max_in_flight_bytes, in_flight);
      st.batches++;
      progress = true;
    }
    if (err != PJ_OK) break;
    if (!progress) {
      bool any_busy = false;
      for (auto &f : fl) any_busy |= f.busy;
      if (!any_busy && waiting.empty() && next < nq) {
        double wait = arrival_s[next] - now();
        if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.001)));
      } else {
        std::this_thread::sleep_for(std::chrono::microseconds(20));
      }
    }
  }
  // drain on error
  for (int s = 0; s < nslots; ++s)
    if (fl[s].busy) {
      pj_batch_wait(ctx, s);
      release(fl[s].bufs, fl[s].qs);
    }
  if (err != PJ_OK) return err;
  st.wall_s = now();
  if (stats) *stats = st;
  return PJ_OK;
}
