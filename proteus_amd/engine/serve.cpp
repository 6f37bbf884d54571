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
  uint64_t bytes = 0;  // used bytes of the slot's arena
  uint64_t rows = 0;   // probe rows of the batch
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
  pj_serve_stats st{};
  const auto t0 = Clock::now();
  auto now = [&]() { return std::chrono::duration<double>(Clock::now() - t0).count(); };
  // Probe inputs are carved from preallocated arenas (one per slot), so
  // admission never calls the allocator: 16-byte aligned columns.
  // Queries whose probe columns this call placed in an arena: only their
  // pointers are reset afterwards; caller-supplied columns are left alone.
  std::vector<char> placed(nq, 0);
  auto place = [&](uint32_t i, char *base) {
    pj_query &q = qs[i];
    const uint64_t c = (q.n * 4 + 15) & ~15ULL;
    q.d_fk = (const int32_t *)base;
    q.d_val = (const int32_t *)(base + c);
    q.d_grp = (const int32_t *)(base + 2 * c);
    placed[i] = 1;
  };
  // The host may set *cfg->cancel at any time: no further query is admitted,
  // batches on the device finish, the queries not run get done_s = -1 and
  // the call returns PJ_E_CANCELLED.
  auto cancelled = [&]() { return cfg->cancel && *cfg->cancel != 0; };
  for (uint32_t i = 0; i < nq; ++i) done_s[i] = -1.0;
  auto padded = [&](const pj_query &q) -> uint64_t {
    return input_bytes(q) ? 3 * ((q.n * 4 + 15) & ~15ULL) : 0;
  };
  auto release = [&](const std::vector<uint32_t> &ids) {
    for (uint32_t i : ids)
      if (placed[i]) {
        qs[i].d_fk = qs[i].d_val = qs[i].d_grp = nullptr;
        placed[i] = 0;
      }
  };
  uint64_t arena_bytes = 0;
  for (uint32_t i = 0; i < nq; ++i) arena_bytes = std::max(arena_bytes, padded(qs[i]));
  if (arena_bytes > budget)
    return pj_fail(PJ_E_CAPACITY, "a query needs %llu bytes of inputs, budget %llu",
                   (unsigned long long)arena_bytes, (unsigned long long)budget);

  if (cfg->mode == 1) {
    void *arena = nullptr;
    if (arena_bytes) {
      int rc = pj_malloc(ctx, arena_bytes, &arena);
      if (rc != PJ_OK) return rc;
    }
    for (uint32_t i = 0; i < nq; ++i) {
      double wait = arrival_s[i] - now();
      while (wait > 0 && !cancelled()) {
        std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.001)));
        wait = arrival_s[i] - now();
      }
      if (cancelled()) {
        if (arena) pj_free(ctx, arena);
        st.wall_s = now();
        if (stats) *stats = st;
        return pj_fail(PJ_E_CANCELLED, "serving loop cancelled after %u of %u queries", i, nq);
      }
      if (input_bytes(qs[i])) place(i, (char *)arena);
      int rc = pj_query_run(ctx, &qs[i], &cfg->batch);
      release({i});
      if (rc != PJ_OK) {
        if (arena) pj_free(ctx, arena);
        return rc;
      }
      done_s[i] = now();
      st.batches++;
      st.max_in_flight_bytes = std::max(st.max_in_flight_bytes, input_bytes(qs[i]));
    }
    st.wall_s = now();
    if (arena) pj_free(ctx, arena);
    if (stats) *stats = st;
    return PJ_OK;
  }

  // Size-class lanes: slot 0 serves only light queries (probe rows <
  // heavy_rows), so they never queue behind a heavy batch and their kernels
  // run beside it on the device; the other slots take queries in arrival
  // order.  Every slot's arena must hold the largest query it may take.
  int nslots = cfg->slots > 0 ? std::min(cfg->slots, 8) : 2;
  const uint64_t heavy_rows = cfg->heavy_rows ? cfg->heavy_rows : (1ULL << 22);
  uint64_t light_max = 0, any_max = arena_bytes;
  for (uint32_t i = 0; i < nq; ++i)
    if (qs[i].n < heavy_rows) light_max = std::max(light_max, padded(qs[i]));
  std::vector<uint64_t> slot_bytes;
  if (nslots >= 2) {
    uint64_t light_arena = cfg->hbm_budget ? std::max(light_max, budget / (4 * (uint64_t)nslots)) : light_max;
    uint64_t rest = cfg->hbm_budget ? budget - std::min(budget, light_arena) : 0;
    int heavy = nslots - 1;
    if (cfg->hbm_budget && any_max)
      heavy = (int)std::min<uint64_t>(heavy, rest / any_max);
    if (heavy < 1) {  // budget too small for lanes: one general slot
      nslots = 1;
      slot_bytes.push_back(cfg->hbm_budget ? budget : any_max);
    } else {
      nslots = heavy + 1;
      slot_bytes.push_back(light_arena);
      for (int i = 0; i < heavy; ++i) slot_bytes.push_back(cfg->hbm_budget ? rest / heavy : any_max);
    }
  } else {
    slot_bytes.push_back(cfg->hbm_budget ? budget : any_max);
  }
  const bool lanes = nslots >= 2;
  const uint32_t max_batch = cfg->max_batch ? std::min<uint32_t>(cfg->max_batch, 4096) : 4096;
  // A batch completes as a whole, so a large query batched with others
  // waits for them: batches stop taking queries at max_rows probe rows.
  // (Served mix: p99 54 ms uncapped, 44 ms at 2^26 rows, same queries/s.)
  const uint64_t max_rows = cfg->max_batch_rows ? cfg->max_batch_rows : (1ULL << 26);
  int rc = pj_batch_slots(ctx, nslots);
  if (rc != PJ_OK) return rc;
  std::vector<void *> arenas(nslots, nullptr);
  auto free_arenas = [&]() {
    for (void *a : arenas)
      if (a) pj_free(ctx, a);
  };
  if (arena_bytes)
    for (int s = 0; s < nslots; ++s)
      if (slot_bytes[s] && (rc = pj_malloc(ctx, slot_bytes[s], &arenas[s])) != PJ_OK) {
        free_arenas();
        return rc;
      }
  std::vector<InFlight> fl(nslots);
  std::deque<uint32_t> light, heavy;  // waiting, arrival order each
  uint32_t next = 0, completed = 0;
  uint64_t in_flight = 0;
  std::vector<pj_query> batch;
  int err = PJ_OK;
  auto is_light = [&](uint32_t i) { return lanes && qs[i].n < heavy_rows; };
  bool stop = false;  // cancelled: admit nothing more, drain
  while (completed < nq && err == PJ_OK) {
    bool progress = false;
    if (!stop && cancelled()) stop = true;
    if (stop) {
      bool any_busy = false;
      for (auto &f : fl) any_busy |= f.busy;
      if (!any_busy) break;
    }
    const double t = now();
    while (next < nq && arrival_s[next] <= t) {
      (is_light(next) ? light : heavy).push_back(next);
      ++next;
    }
    for (int s = 0; s < nslots; ++s) {
      if (!fl[s].busy) continue;
      int done = 0;
      if ((err = pj_batch_poll(ctx, s, &done)) != PJ_OK) break;
      if (!done) continue;
      const double tc = now();
      for (uint32_t i : fl[s].qs) done_s[i] = tc;
      completed += (uint32_t)fl[s].qs.size();
      in_flight -= fl[s].bytes;
      release(fl[s].qs);
      fl[s] = InFlight();
      progress = true;
    }
    if (err != PJ_OK) break;
    for (int s = 0; s < nslots && !stop; ++s) {
      if (fl[s].busy) continue;
      InFlight &f = fl[s];
      batch.clear();
      while (batch.size() < max_batch) {
        // slot 0 (with lanes) takes light queries only; the others take the
        // earliest waiting query of either class
        std::deque<uint32_t> *src = nullptr;
        if (lanes && s == 0) {
          src = light.empty() ? nullptr : &light;
        } else if (!light.empty() && (heavy.empty() || light.front() < heavy.front())) {
          src = &light;
        } else if (!heavy.empty()) {
          src = &heavy;
        }
        if (!src) break;
        pj_query &q = qs[src->front()];
        const uint64_t b = padded(q);
        if (f.bytes + b > slot_bytes[s]) break;
        if (!batch.empty() && f.rows + q.n > max_rows) break;
        if (b) place(src->front(), (char *)arenas[s] + f.bytes);
        f.bytes += b;
        f.rows += q.n;
        f.qs.push_back(src->front());
        batch.push_back(q);
        src->pop_front();
      }
      if (batch.empty()) continue;
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
      if (!any_busy && light.empty() && heavy.empty() && next < nq) {
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
      release(fl[s].qs);
    }
  free_arenas();
  if (err != PJ_OK) return err;
  st.wall_s = now();
  if (stats) *stats = st;
  if (completed < nq)
    return pj_fail(PJ_E_CANCELLED, "serving loop cancelled after %u of %u queries", completed, nq);
  return PJ_OK;
}
