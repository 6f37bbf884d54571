// Host -> HBM staging through a ring of pinned buffers.
//
// The reference moves host-resident blocks of a relation to the device in
// its cpu-to-gpu operator (core/olap/lib/operators/cpu-to-gpu.cpp:69-239),
// from files that the storage manager loads (core/storage/lib/
// storage-manager.cpp:300-330; a column of a prepared relation is the
// headerless file <rel>.<attr>, binary-block-plugin.cpp:67-100).  Here a copy
// is cut into chunks that worker threads fill into pinned buffers, from a
// file (pread) or from pageable host memory (memcpy).  Each filled chunk is
// DMA'd to HBM on a copy stream of the context at once, so filling, DMA and
// the kernels of the context's stream all overlap; pj_stage_wait orders the
// context's stream after a copy without waiting on the host for the DMA.
//
// Each worker owns two pinned chunk buffers and alternates between them; a
// buffer is refilled only after the event recorded behind its last DMA has
// completed.  The last chunk of a copy to be issued records the copy's done
// event on the copy stream.
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <atomic>
#include <condition_variable>
#include <deque>
#include <thread>

#include "pj_internal.hpp"

namespace pj {
namespace {

constexpr uint64_t kChunk = 8ull << 20;  // bytes per pinned buffer
constexpr int kMaxWorkers = 8;

struct Job {
  int fd = -1;
  const char *host = nullptr;
  uint64_t offset = 0, bytes = 0;
  char *dst = nullptr;
  std::atomic<uint64_t> left{0};  // chunks not yet issued
  hipEvent_t done = nullptr;
  bool issued = false;
  int err = PJ_OK;
  std::string msg;
  std::string path;
};

struct Chunk {
  Job *job;
  uint64_t off, len;
};

struct Stager {
  int device = 0;
  hipStream_t copy = nullptr;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Chunk> q;
  bool stop = false;
  std::vector<std::thread> threads;
  std::map<uint64_t, Job *> jobs;
  uint64_t next_ticket = 1;
  std::vector<hipEvent_t> ev_free;
  int worker_err = PJ_OK;
  std::string worker_msg;
  int ready = 0;  // workers that allocated their buffers
};

// A worker: fills chunks into its two pinned buffers and issues their DMA.
void worker(Stager *st) {
  void *buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool used[2] = {false, false};
  bool ok = hipSetDevice(st->device) == hipSuccess;
  for (int i = 0; ok && i < 2; ++i)
    ok = hipHostMalloc(&buf[i], kChunk, hipHostMallocDefault) == hipSuccess &&
         hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) == hipSuccess;
  {
    std::lock_guard<std::mutex> g(st->mu);
    if (!ok && st->worker_err == PJ_OK) {
      st->worker_err = PJ_E_HIP;
      st->worker_msg = "staging worker: pinned buffer allocation failed";
    }
    ++st->ready;
  }
  st->cv_done.notify_all();
  int turn = 0;
  for (;;) {
    Chunk c;
    {
      std::unique_lock<std::mutex> lk(st->mu);
      st->cv_work.wait(lk, [&] { return st->stop || !st->q.empty(); });
      if (st->q.empty()) break;  // stop, and no work left
      c = st->q.front();
      st->q.pop_front();
    }
    Job *j = c.job;
    const int s = turn;
    turn ^= 1;
    int err = ok ? PJ_OK : PJ_E_HIP;
    std::string msg;
    if (err == PJ_OK && used[s] && hipEventSynchronize(ev[s]) != hipSuccess) {
      err = PJ_E_HIP;
      msg = "staging: buffer event wait failed";
    }
    if (err == PJ_OK) {
      char *b = (char *)buf[s];
      if (j->host) {
        memcpy(b, j->host + c.off, c.len);
      } else {
        uint64_t got = 0;
        while (got < c.len) {
          ssize_t r = pread(j->fd, b + got, c.len - got, (off_t)(j->offset + c.off + got));
          if (r <= 0) {
            err = PJ_E_ARG;
            msg = "staging: short read of " + j->path;
            break;
          }
          got += (uint64_t)r;
        }
      }
    }
    if (err == PJ_OK) {
      if (hipMemcpyAsync(j->dst + c.off, buf[s], c.len, hipMemcpyHostToDevice, st->copy) != hipSuccess ||
          hipEventRecord(ev[s], st->copy) != hipSuccess) {
        err = PJ_E_HIP;
        msg = "staging: hipMemcpyAsync failed";
      }
      used[s] = true;
    }
    if (err != PJ_OK) {
      std::lock_guard<std::mutex> g(st->mu);
      if (j->err == PJ_OK) {
        j->err = err;
        j->msg = msg.empty() ? "staging worker unavailable" : msg;
      }
    }
    if (j->left.fetch_sub(1) == 1) {
      // every chunk of the copy is issued: its done event follows them
      hipError_t e = hipEventRecord(j->done, st->copy);
      std::lock_guard<std::mutex> g(st->mu);
      if (e != hipSuccess && j->err == PJ_OK) {
        j->err = PJ_E_HIP;
        j->msg = "staging: hipEventRecord failed";
      }
      if (j->fd >= 0) close(j->fd);
      j->fd = -1;
      j->issued = true;
      st->cv_done.notify_all();
    }
  }
  for (int i = 0; i < 2; ++i) {
    if (used[i]) (void)hipEventSynchronize(ev[i]);
    if (ev[i]) (void)hipEventDestroy(ev[i]);
    if (buf[i]) (void)hipHostFree(buf[i]);
  }
}

int stager_of(pj_ctx *ctx, Stager **out) {
  if (!ctx->stager) {
    auto *st = new Stager();
    st->device = ctx->device;
    if (hipStreamCreateWithFlags(&st->copy, hipStreamNonBlocking) != hipSuccess) {
      delete st;
      return fail(PJ_E_HIP, "staging: copy stream creation failed");
    }
    unsigned hw = std::thread::hardware_concurrency();
    int n = (int)std::min<unsigned>(kMaxWorkers, std::max(2u, hw / 2));
    for (int i = 0; i < n; ++i) st->threads.emplace_back(worker, st);
    std::unique_lock<std::mutex> lk(st->mu);
    st->cv_done.wait(lk, [&] { return st->ready == n; });
    ctx->stager = st;
    if (st->worker_err != PJ_OK) return fail(st->worker_err, "%s", st->worker_msg.c_str());
  }
  *out = (Stager *)ctx->stager;
  return PJ_OK;
}

}  // namespace

void stage_release(pj_ctx *ctx) {
  auto *st = (Stager *)ctx->stager;
  if (!st) return;
  {
    std::lock_guard<std::mutex> g(st->mu);
    st->stop = true;
  }
  st->cv_work.notify_all();
  for (auto &t : st->threads) t.join();  // the queue drains first
  (void)hipStreamSynchronize(st->copy);
  for (auto &kv : st->jobs) {
    if (kv.second->done) (void)hipEventDestroy(kv.second->done);
    if (kv.second->fd >= 0) close(kv.second->fd);
    delete kv.second;
  }
  for (auto e : st->ev_free) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(st->copy);
  delete st;
  ctx->stager = nullptr;
}

}  // namespace pj

using namespace pj;

extern "C" int pj_stage_begin(pj_ctx *ctx, const char *path, const void *h_src, uint64_t offset,
                              uint64_t bytes, void *d_dst, uint64_t *ticket) {
  if (!ctx || !ticket || (!path && !h_src) || (path && h_src) || (bytes && !d_dst))
    return fail(PJ_E_ARG, "pj_stage_begin: give exactly one of path and h_src, a ticket and a destination");
  *ticket = 0;
  int fd = -1;
  if (path) {
    fd = open(path, O_RDONLY | O_CLOEXEC);
    if (fd < 0) return fail(PJ_E_ARG, "pj_stage_begin: cannot open %s", path);
    struct stat sb;
    if (fstat(fd, &sb) != 0 || (uint64_t)sb.st_size < offset + bytes) {
      unsigned long long have = fstat(fd, &sb) == 0 ? (unsigned long long)sb.st_size : 0;
      close(fd);
      return fail(PJ_E_ARG, "pj_stage_begin: %s holds %llu bytes, the copy needs %llu", path, have,
                  (unsigned long long)(offset + bytes));
    }
  }
  Stager *st = nullptr;
  int rc = stager_of(ctx, &st);
  if (rc != PJ_OK) {
    if (fd >= 0) close(fd);
    return rc;
  }
  auto *j = new Job();
  j->fd = fd;
  j->host = (const char *)h_src;
  j->offset = path ? offset : 0;
  if (h_src) j->host += offset;
  j->bytes = bytes;
  j->dst = (char *)d_dst;
  if (path) j->path = path;
  std::unique_lock<std::mutex> lk(st->mu);
  if (!st->ev_free.empty()) {
    j->done = st->ev_free.back();
    st->ev_free.pop_back();
  } else if (hipEventCreateWithFlags(&j->done, hipEventDisableTiming) != hipSuccess) {
    lk.unlock();
    if (fd >= 0) close(fd);
    delete j;
    return fail(PJ_E_HIP, "pj_stage_begin: hipEventCreate failed");
  }
  const uint64_t id = st->next_ticket++;
  st->jobs[id] = j;
  *ticket = id;
  if (bytes == 0) {
    // nothing to move: done once the copy stream reaches this point
    hipError_t e = hipEventRecord(j->done, st->copy);
    if (fd >= 0) close(fd);
    j->fd = -1;
    j->issued = true;
    if (e != hipSuccess) {
      j->err = PJ_E_HIP;
      j->msg = "pj_stage_begin: hipEventRecord failed";
    }
    return PJ_OK;
  }
  const uint64_t n = (bytes + kChunk - 1) / kChunk;
  j->left.store(n);
  for (uint64_t i = 0; i < n; ++i) {
    const uint64_t off = i * kChunk;
    st->q.push_back(Chunk{j, off, std::min(kChunk, bytes - off)});
  }
  lk.unlock();
  st->cv_work.notify_all();
  return PJ_OK;
}

extern "C" int pj_stage_wait(pj_ctx *ctx, uint64_t ticket) {
  if (!ctx || !ctx->stager) return fail(PJ_E_ARG, "pj_stage_wait: no copy with ticket %llu", (unsigned long long)ticket);
  auto *st = (Stager *)ctx->stager;
  std::unique_lock<std::mutex> lk(st->mu);
  auto it = st->jobs.find(ticket);
  if (it == st->jobs.end())
    return fail(PJ_E_ARG, "pj_stage_wait: no copy with ticket %llu", (unsigned long long)ticket);
  Job *j = it->second;
  st->cv_done.wait(lk, [&] { return j->issued; });
  st->jobs.erase(it);
  const int err = j->err;
  const std::string msg = j->msg;
  hipError_t e = err == PJ_OK ? hipStreamWaitEvent(ctx->stream, j->done, 0) : hipSuccess;
  st->ev_free.push_back(j->done);
  delete j;
  lk.unlock();
  if (err != PJ_OK) {
    // the copy stream may still write to the destination: drain it before
    // the caller reuses or frees that memory
    (void)hipStreamSynchronize(st->copy);
    return fail(err, "%s", msg.c_str());
  }
  if (e != hipSuccess) return fail(PJ_E_HIP, "pj_stage_wait: hipStreamWaitEvent: %s", hipGetErrorString(e));
  return PJ_OK;
}

extern "C" int pj_stage_file(pj_ctx *ctx, const char *path, uint64_t offset, uint64_t bytes, void *d_dst) {
  if (!path) return fail(PJ_E_ARG, "pj_stage_file: null path");
  uint64_t t = 0;
  PJ_TRY(pj_stage_begin(ctx, path, nullptr, offset, bytes, d_dst, &t));
  PJ_TRY(pj_stage_wait(ctx, t));
  PJ_HIP(hipStreamSynchronize(ctx->stream));
  return PJ_OK;
}
