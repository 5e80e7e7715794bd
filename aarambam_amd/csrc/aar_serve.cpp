// Serving loop: continuous batching of IC jobs on one device.
//
// The reference runs one initial-condition job per process (main.c:53-102).
// A server instead keeps one context (twiddle tables, the plan cache of
// device-resident tables, streams) and one HBM arena resident and accepts a
// stream of jobs of mixed Nmesh, table and seed:
//   * admission is FIFO by footprint: the job at the head of the queue is
//     admitted as soon as a stream slot is free and the arena has a contiguous
//     piece for the job's own buffers (field slots, seed table, particles;
//     first fit, neighbours merged on free);
//   * the tables of a job (k-space tables, modal schedule, transfer table,
//     glass) come from the context's plan cache, so jobs that share them with
//     an earlier job upload nothing but their seed table;
//   * admitted jobs are prepared by a small pool of worker threads (the host
//     seed table, the uploads) and enqueued on their slot's stream, so the
//     admission thread never waits on host work and the kernels of up to
//     max_active jobs overlap on the device;
//   * completion is signalled by a host function enqueued after the job's last
//     kernel, which timestamps it and wakes the scheduler to admit more work.
// Latency of a job = completion time - submission time; queueing = admission
// time - submission time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/aarambam.h"
#include "aar_common.h"
#include "aar_internal.h"

namespace {

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Req {
  long long id = 0;
  aar_params p{};
  std::vector<double> alpha, x0, tk, tt;
  std::vector<float> glass;
  std::vector<uint8_t> header;
  aar_tables t{};
  int prec = AAR_FP32, sched = AAR_SCHED_BATCHED;
  bool keep = false;
  size_t bytes = 0, off = 0;
  int slot = -1;
  aar_job* job = nullptr;
  int status = AAR_OK;
  enum State { PENDING, RUNNING, DONE } state = PENDING;
  bool finished = false;  // set by the completion host function
  bool prepared = false;  // the worker has handed over the job (or its failure)
  double t_submit = 0, t_start = 0, t_done = 0;
  aar_server* srv = nullptr;
};

}  // namespace

struct aar_server {
  aar_ctx* ctx = nullptr;
  int device = 0;
  char* arena = nullptr;
  size_t budget = 0;
  std::map<size_t, size_t> free_;  // offset -> bytes of each free piece
  int max_active = 1;
  std::vector<hipStream_t> streams;
  std::vector<bool> busy;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Req*> pending;
  std::deque<Req*> admitted;  // admitted, waiting for a preparing worker
  std::vector<Req*> running;  // admitted and not yet retired
  std::unordered_map<long long, Req*> reqs;
  std::thread th;
  std::vector<std::thread> workers;
  bool stop = false, stop_workers = false;
  long long next_id = 1, submitted = 0, completed = 0, peak_bytes = 0, in_use = 0;
  int peak_active = 0;

  bool arena_take(size_t bytes, size_t* off) {
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second >= bytes) {
        *off = it->first;
        const size_t rest = it->second - bytes;
        free_.erase(it);
        if (rest) free_[*off + bytes] = rest;
        return true;
      }
    return false;
  }

  void arena_give(size_t off, size_t bytes) {
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }

  void loop();
  void worker();
};

namespace {

void on_done(void* arg) {  // host function on the job's stream: no HIP calls here
  Req* r = (Req*)arg;
  aar_server* s = r->srv;
  {
    std::lock_guard<std::mutex> g(s->mu);
    r->t_done = now_s();
    r->finished = true;
  }
  s->cv.notify_all();
}

size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

void aar_server::loop() {
  hipSetDevice(device);
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    // retire finished jobs
    for (size_t q = 0; q < running.size();) {
      Req* r = running[q];
      if (!r->finished || !r->prepared) {
        ++q;
        continue;
      }
      running.erase(running.begin() + q);
      busy[r->slot] = false;
      if (!r->keep || r->status) {
        aar_job* jb = r->job;
        r->job = nullptr;
        lk.unlock();
        if (jb) aar_job_destroy(jb);
        lk.lock();
        arena_give(r->off, r->bytes);
        in_use -= (long long)r->bytes;
      }
      r->state = Req::DONE;
      ++completed;
      cv.notify_all();
    }
    // admit in arrival order while a slot and a piece of the arena are free
    while (!stop && !pending.empty()) {
      Req* r = pending.front();
      int slot = -1;
      for (int q = 0; q < max_active; ++q)
        if (!busy[q]) {
          slot = q;
          break;
        }
      size_t off = 0;
      if (slot < 0 || !arena_take(r->bytes, &off)) break;
      pending.pop_front();
      busy[slot] = true;
      r->slot = slot;
      r->off = off;
      r->state = Req::RUNNING;
      r->t_start = now_s();
      running.push_back(r);
      in_use += (long long)r->bytes;
      if (in_use > peak_bytes) peak_bytes = in_use;
      if ((int)running.size() > peak_active) peak_active = (int)running.size();
      admitted.push_back(r);
      cv.notify_all();
    }
    if (stop && running.empty()) break;
    bool work = false;
    for (Req* r : running) work |= r->finished && r->prepared;
    if (!work) cv.wait(lk);
  }
}

// Prepares admitted jobs inside their arena pieces and enqueues them on their slots' streams.
void aar_server::worker() {
  hipSetDevice(device);
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    cv.wait(lk, [&] { return stop_workers || !admitted.empty(); });
    if (admitted.empty()) break;  // stop_workers, and nothing left to prepare
    Req* r = admitted.front();
    admitted.pop_front();
    lk.unlock();
    hipStream_t st = streams[r->slot];
    aar_job* jb = nullptr;
    int rc = aar_job_prepare_ws(ctx, &r->p, &r->t, r->prec, r->sched, arena + r->off, r->bytes, st, 0, &jb, nullptr,
                                nullptr);
    if (!rc) rc = aar_job_run(jb, st);
    if (!rc && hipLaunchHostFunc(st, on_done, r) != hipSuccess) rc = AAR_ERR_HIP;
    if (rc) hipStreamSynchronize(st);  // nothing of a failed job may still use its piece
    lk.lock();
    r->job = jb;
    r->prepared = true;
    if (rc) {
      r->status = rc;
      r->finished = true;
      r->t_done = now_s();
    }
    cv.notify_all();
  }
}

extern "C" int aar_server_create(aar_ctx* ctx, long long hbm_budget, int max_active, aar_server** out) {
  if (!ctx || !out || hbm_budget <= 0 || max_active < 1 || max_active > 16) return AAR_ERR_BAD_ARG;
  const int device = aar_ctx_device(ctx);
  AAR_HIP(hipSetDevice(device));
  aar_server* s = new aar_server();
  s->ctx = ctx;
  s->device = device;
  s->budget = round256((size_t)hbm_budget);
  if (hipMalloc((void**)&s->arena, s->budget) != hipSuccess) {
    delete s;
    return AAR_ERR_ALLOC;
  }
  s->free_[0] = s->budget;
  s->max_active = max_active;
  s->streams.resize(max_active);
  s->busy.assign(max_active, false);
  for (auto& st : s->streams)
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
      st = nullptr;
      aar_server_destroy(s);
      return AAR_ERR_HIP;
    }
  s->th = std::thread([s] { s->loop(); });
  for (int q = 0; q < std::min(max_active, 4); ++q) s->workers.emplace_back([s] { s->worker(); });
  *out = s;
  return AAR_OK;
}

extern "C" int aar_server_submit(aar_server* s, const aar_params* p, const aar_tables* t, int precision,
                                 int schedule, int keep, long long* ticket) {
  if (!s || !p || !t || !ticket || !t->alpha || !t->x0 || !t->glass || t->nglass <= 0) return AAR_ERR_BAD_ARG;
  if (precision != AAR_FP32 && precision != AAR_FP64) return AAR_ERR_BAD_ARG;
  if (schedule != AAR_SCHED_PLAIN && schedule != AAR_SCHED_BATCHED) return AAR_ERR_BAD_ARG;
  const long long b = aar_job_ws_bytes(p, t, precision, schedule, 0);  // tables live in the plan cache
  if (b < 0) return AAR_ERR_BAD_ARG;
  if ((size_t)b > s->budget) return AAR_ERR_HBM_BUDGET;
  Req* r = new Req();
  r->p = *p;
  const size_t nm = (size_t)p->N_modes;
  r->alpha.assign(t->alpha, t->alpha + nm * nm * nm);
  r->x0.assign(t->x0, t->x0 + 2 * nm);
  if (t->ntransfer > 0 && t->transfer_k && t->transfer_t) {
    r->tk.assign(t->transfer_k, t->transfer_k + t->ntransfer);
    r->tt.assign(t->transfer_t, t->transfer_t + t->ntransfer);
  }
  r->glass.assign(t->glass, t->glass + 3 * (size_t)t->nglass);
  if (t->glass_header) r->header.assign(t->glass_header, t->glass_header + 256);
  r->t = *t;
  r->t.alpha = r->alpha.data();
  r->t.x0 = r->x0.data();
  r->t.transfer_k = r->tk.empty() ? nullptr : r->tk.data();
  r->t.transfer_t = r->tt.empty() ? nullptr : r->tt.data();
  r->t.glass = r->glass.data();
  r->t.glass_header = r->header.empty() ? nullptr : r->header.data();
  r->prec = precision;
  r->sched = schedule;
  r->keep = keep != 0;
  r->bytes = round256((size_t)b);
  r->srv = s;
  {
    std::lock_guard<std::mutex> g(s->mu);
    if (s->stop) {
      delete r;
      return AAR_ERR_BAD_ARG;
    }
    r->id = s->next_id++;
    r->t_submit = now_s();
    s->reqs[r->id] = r;
    s->pending.push_back(r);
    ++s->submitted;
    *ticket = r->id;
  }
  s->cv.notify_all();
  return AAR_OK;
}

extern "C" int aar_server_wait(aar_server* s, long long ticket, double* latency_s, double* queue_s) {
  if (!s) return AAR_ERR_BAD_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return AAR_ERR_BAD_ARG;
  Req* r = it->second;
  s->cv.wait(lk, [&] { return r->state == Req::DONE; });
  if (latency_s) *latency_s = r->t_done - r->t_submit;
  if (queue_s) *queue_s = r->t_start - r->t_submit;
  return r->status;
}

static Req* done_req(aar_server* s, long long ticket) {
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end() || it->second->state != Req::DONE) return nullptr;
  return it->second;
}

extern "C" int aar_server_fetch(aar_server* s, long long ticket, float* pos, float* vel) {
  if (!s) return AAR_ERR_BAD_ARG;
  Req* r = done_req(s, ticket);
  if (!r || !r->job) return r && r->status ? r->status : AAR_ERR_BAD_ARG;
  return aar_job_fetch(r->job, pos, vel);
}

extern "C" int aar_server_device_output(aar_server* s, long long ticket, float** pos, float** vel,
                                        long long* npart) {
  if (!s) return AAR_ERR_BAD_ARG;
  Req* r = done_req(s, ticket);
  if (!r || !r->job) return r && r->status ? r->status : AAR_ERR_BAD_ARG;
  return aar_job_device_output(r->job, pos, vel, npart);
}

extern "C" int aar_server_release(aar_server* s, long long ticket) {
  if (!s) return AAR_ERR_BAD_ARG;
  Req* r = done_req(s, ticket);
  if (!r) return AAR_ERR_BAD_ARG;
  aar_job* jb = nullptr;
  {
    std::lock_guard<std::mutex> g(s->mu);
    jb = r->job;
    r->job = nullptr;
    s->reqs.erase(ticket);
  }
  if (jb) {
    hipSetDevice(s->device);
    aar_job_destroy(jb);
    std::lock_guard<std::mutex> g(s->mu);
    s->arena_give(r->off, r->bytes);
    s->in_use -= (long long)r->bytes;
  }
  delete r;
  s->cv.notify_all();
  return AAR_OK;
}

extern "C" int aar_server_stats(aar_server* s, long long* submitted, long long* completed, int* peak_active,
                                long long* peak_bytes) {
  if (!s) return AAR_ERR_BAD_ARG;
  std::lock_guard<std::mutex> g(s->mu);
  if (submitted) *submitted = s->submitted;
  if (completed) *completed = s->completed;
  if (peak_active) *peak_active = s->peak_active;
  if (peak_bytes) *peak_bytes = s->peak_bytes;
  return AAR_OK;
}

extern "C" int aar_server_destroy(aar_server* s) {
  if (!s) return AAR_OK;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
    for (Req* r : s->pending) {  // never admitted: cancelled
      r->status = AAR_ERR_BAD_ARG;
      r->state = Req::DONE;
    }
    s->pending.clear();
  }
  s->cv.notify_all();
  if (s->th.joinable()) s->th.join();  // the loop ends once every running job has finished
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop_workers = true;
  }
  s->cv.notify_all();
  for (auto& w : s->workers)
    if (w.joinable()) w.join();
  hipSetDevice(s->device);
  for (auto& kv : s->reqs) {
    if (kv.second->job) aar_job_destroy(kv.second->job);
    delete kv.second;
  }
  s->reqs.clear();
  for (hipStream_t st : s->streams)
    if (st) hipStreamDestroy(st);
  if (s->arena) hipFree(s->arena);
  delete s;
  return AAR_OK;
}
