// Serving runtime: a resident scene answers requests of (body_i, body_j)
// pairs under fresh per-pair poses with GJK+EPA, continuously batched.
//
//  * Slots: each slot owns a HIP stream, device buffers for up to `capacity`
//    pairs (pairs, poses, results, workspace) and pinned host staging.  The
//    number of slots follows the device-memory budget.
//  * Dispatcher thread: when a slot is free and requests are pending, it
//    packs pending requests in arrival order into the slot while they fit
//    (admission by the slot's memory), uploads them, launches GJK+EPA and the
//    result download on the slot's stream, and records an event.  Requests
//    that arrive while other slots run are admitted as soon as a slot frees,
//    so their work overlaps on the device.
//  * Completion thread: polls the events of running slots, copies each
//    request's results to the caller's buffers, stamps its latency (submit to
//    results on the host) and wakes its waiter.
// Each pair is computed by the same kernels and lane layout whatever batch it
// lands in, so a request's results do not depend on the mix.
//
// Shutdown (ogjk_server_destroy): new submissions are refused, requests not
// yet launched are cancelled (their wait returns an error), launched batches
// complete and are answered as usual, and the server is freed only after
// every thread blocked in ogjk_server_wait has returned.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ogjk_internal.hpp"
#include "opengjk/ogjk.h"

// error reporting shared with ogjk_api.hip (printf-style message)
ogjk_status ogjk_fail(ogjk_status code, const char* fmt, ...);

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  int64_t ticket = 0;
  const int32_t* pairs = nullptr;
  const float* poses = nullptr;
  int32_t n = 0;
  gkSimplex* simplices = nullptr;
  float* distances = nullptr;
  float* normals = nullptr;
  double* f64 = nullptr;
  Clock::time_point t_submit;
  double latency_s = 0.0;
  int64_t batch = -1;  // launch sequence number that carried it
  bool done = false;
  bool cancelled = false;  // never launched: the server was destroyed first
  ogjk_status status = OGJK_OK;
};

// Per batch of `total` pairs, inputs and results are packed so that each
// direction takes one copy: in = pairs (8 B) then poses (56 B) per pair; out =
// simplices (108 B), then distances (4 B), then normals (12 B) per pair.
constexpr size_t kInBytes = 64, kOutBytes = sizeof(gkSimplex) + 16;

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  char* d_in = nullptr;
  char* d_out = nullptr;
  double* d_f64 = nullptr;
  void* d_ws = nullptr;
  char* h_in = nullptr;
  char* h_out = nullptr;
  double* h_f64 = nullptr;
  int32_t total = 0;
  std::vector<std::shared_ptr<Request>> reqs;
  std::vector<int32_t> offs;
  bool busy = false;      // taken by the dispatcher
  bool launched = false;  // its work and event are enqueued
  bool with_f64 = false;
};

}  // namespace

struct ogjk_server_s {
  ogjk::SceneArgs sa{};
  int32_t capacity = 0;
  std::vector<Slot> slots;
  std::mutex mu;
  std::condition_variable cv_dispatch, cv_complete, cv_done;
  std::deque<std::shared_ptr<Request>> pending;
  std::unordered_map<int64_t, std::shared_ptr<Request>> tickets;
  std::thread dispatcher, completer;
  bool stop = false;     // threads exit
  bool closing = false;  // destroy has begun: no new submissions
  int waiters = 0;       // threads inside ogjk_server_wait
  int64_t next_ticket = 1;
  int64_t launches = 0;
  int64_t pairs_done = 0;
  int device = 0;
  hipError_t async_error = hipSuccess;

  // Cancel what is not launched, let launched batches complete, wait for
  // every waiter to leave, then stop the threads.
  void shutdown() {
    {
      std::unique_lock<std::mutex> lock(mu);
      closing = true;
      for (auto& r : pending) {
        r->status = OGJK_ERR_INVALID;
        r->cancelled = true;
        r->done = true;
      }
      pending.clear();
      cv_done.notify_all();
      cv_done.wait(lock, [&] {
        for (const Slot& sl : slots)
          if (sl.busy) return false;
        return waiters == 0;
      });
      stop = true;
    }
    cv_dispatch.notify_all();
    cv_complete.notify_all();
    if (dispatcher.joinable()) dispatcher.join();
    if (completer.joinable()) completer.join();
  }

  ~ogjk_server_s() {
    {
      std::lock_guard<std::mutex> lock(mu);
      stop = true;
    }
    cv_dispatch.notify_all();
    cv_complete.notify_all();
    if (dispatcher.joinable()) dispatcher.join();
    if (completer.joinable()) completer.join();
    for (Slot& s : slots) {
      if (s.stream) (void)hipStreamSynchronize(s.stream);
      (void)hipFree(s.d_in);
      (void)hipFree(s.d_out);
      (void)hipFree(s.d_f64);
      (void)hipFree(s.d_ws);
      (void)hipHostFree(s.h_in);
      (void)hipHostFree(s.h_out);
      (void)hipHostFree(s.h_f64);
      if (s.done) (void)hipEventDestroy(s.done);
      if (s.stream) (void)hipStreamDestroy(s.stream);
    }
  }

  static size_t slot_device_bytes(int32_t cap) {
    const size_t n = (size_t)cap;
    return n * (kInBytes + kOutBytes + 8 * OGJK_F64_RECORD) + ogjk::workspace_bytes(cap);
  }

  hipError_t init_slot(Slot& s) {
    const size_t n = (size_t)capacity;
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking)) != hipSuccess) return e;
    if ((e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming)) != hipSuccess) return e;
    if ((e = hipMalloc(&s.d_in, kInBytes * n)) != hipSuccess) return e;
    if ((e = hipMalloc(&s.d_out, kOutBytes * n)) != hipSuccess) return e;
    if ((e = hipMalloc(&s.d_f64, 8 * OGJK_F64_RECORD * n)) != hipSuccess) return e;
    if ((e = hipMalloc(&s.d_ws, ogjk::workspace_bytes(capacity))) != hipSuccess) return e;
    if ((e = hipHostMalloc(&s.h_in, kInBytes * n)) != hipSuccess) return e;
    if ((e = hipHostMalloc(&s.h_out, kOutBytes * n)) != hipSuccess) return e;
    if ((e = hipHostMalloc(&s.h_f64, 8 * OGJK_F64_RECORD * n)) != hipSuccess) return e;
    return hipSuccess;
  }

  // Stage, upload, launch and download one batch on slot s (no lock held).
  hipError_t launch(Slot& s) {
    int32_t total = 0;
    s.with_f64 = false;
    for (const auto& r : s.reqs) {
      total += r->n;
      s.with_f64 |= r->f64 != nullptr;
    }
    s.total = total;
    const size_t t = (size_t)total;
    int32_t* h_pairs = (int32_t*)s.h_in;
    float* h_poses = (float*)(s.h_in + 8 * t);
    int32_t off = 0;
    for (const auto& r : s.reqs) {
      std::memcpy(h_pairs + 2 * (size_t)off, r->pairs, 8 * (size_t)r->n);
      std::memcpy(h_poses + 14 * (size_t)off, r->poses, 56 * (size_t)r->n);
      off += r->n;
    }
    hipError_t e;
    if ((e = hipMemcpyAsync(s.d_in, s.h_in, kInBytes * t, hipMemcpyHostToDevice, s.stream)) !=
        hipSuccess)
      return e;
    ogjk::SceneArgs a = sa;
    a.pairs = (const int32_t*)s.d_in;
    const ogjk::OutArgs o{(gkSimplex*)s.d_out, (float*)(s.d_out + sizeof(gkSimplex) * t),
                          (float*)(s.d_out + (sizeof(gkSimplex) + 4) * t),
                          s.with_f64 ? s.d_f64 : nullptr};
    if ((e = ogjk::launch_posed(OGJK_MODE_GJK_EPA, a, (const float*)(s.d_in + 8 * t), false, o, total,
                                s.d_ws, s.stream)) != hipSuccess)
      return e;
    if ((e = hipMemcpyAsync(s.h_out, s.d_out, kOutBytes * t, hipMemcpyDeviceToHost, s.stream)) !=
        hipSuccess)
      return e;
    if (s.with_f64 &&
        (e = hipMemcpyAsync(s.h_f64, s.d_f64, 8 * OGJK_F64_RECORD * t, hipMemcpyDeviceToHost,
                            s.stream)) != hipSuccess)
      return e;
    return hipEventRecord(s.done, s.stream);
  }

  void dispatch_loop() {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lock(mu);
    for (;;) {
      Slot* free_slot = nullptr;
      cv_dispatch.wait(lock, [&] {
        if (stop) return true;
        if (pending.empty()) return false;
        for (Slot& s : slots)
          if (!s.busy) {
            free_slot = &s;
            return true;
          }
        return false;
      });
      if (stop) return;
      Slot& s = *free_slot;
      s.reqs.clear();
      s.offs.clear();
      int32_t total = 0;
      while (!pending.empty() && total + pending.front()->n <= capacity) {
        s.offs.push_back(total);
        total += pending.front()->n;
        pending.front()->batch = launches;
        s.reqs.push_back(pending.front());
        pending.pop_front();
      }
      s.busy = true;
      s.launched = false;
      launches++;
      lock.unlock();
      const hipError_t e = launch(s);
      lock.lock();
      s.launched = true;
      if (e != hipSuccess) {
        async_error = e;
        for (auto& r : s.reqs) {
          r->status = OGJK_ERR_HIP;
          r->done = true;
        }
        s.reqs.clear();
        s.busy = false;
        cv_done.notify_all();
        continue;
      }
      cv_complete.notify_all();
    }
  }

  void complete_loop() {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lock(mu);
    for (;;) {
      cv_complete.wait(lock, [&] {
        if (stop) return true;
        for (Slot& s : slots)
          if (s.busy && s.launched && !s.reqs.empty()) return true;
        return false;
      });
      if (stop) return;
      bool progressed = false;
      for (Slot& s : slots) {
        if (!s.busy || !s.launched || s.reqs.empty()) continue;
        lock.unlock();
        const hipError_t q = hipEventQuery(s.done);
        if (q == hipErrorNotReady) {
          lock.lock();
          continue;
        }
        // results of every request of the batch to the callers' buffers
        for (size_t k = 0; k < s.reqs.size(); k++) {
          Request& r = *s.reqs[k];
          const size_t o = (size_t)s.offs[k], n = (size_t)r.n;
          if (q == hipSuccess) {
            const size_t t = (size_t)s.total;
            const gkSimplex* h_simp = (const gkSimplex*)s.h_out;
            const float* h_dist = (const float*)(s.h_out + sizeof(gkSimplex) * t);
            const float* h_nrm = (const float*)(s.h_out + (sizeof(gkSimplex) + 4) * t);
            if (r.simplices) std::memcpy(r.simplices, h_simp + o, sizeof(gkSimplex) * n);
            if (r.distances) std::memcpy(r.distances, h_dist + o, 4 * n);
            if (r.normals) std::memcpy(r.normals, h_nrm + 3 * o, 12 * n);
            if (r.f64) std::memcpy(r.f64, s.h_f64 + OGJK_F64_RECORD * o, 8 * OGJK_F64_RECORD * n);
          }
        }
        const Clock::time_point now = Clock::now();
        lock.lock();
        for (auto& r : s.reqs) {
          r->latency_s = std::chrono::duration<double>(now - r->t_submit).count();
          r->status = q == hipSuccess ? OGJK_OK : OGJK_ERR_HIP;
          r->done = true;
          pairs_done += r->n;
        }
        if (q != hipSuccess) async_error = q;
        s.reqs.clear();
        s.busy = false;
        progressed = true;
      }
      if (progressed) {
        cv_done.notify_all();
        cv_dispatch.notify_all();
      } else {
        lock.unlock();
        std::this_thread::yield();
        lock.lock();
      }
    }
  }
};

extern "C" {

ogjk_status ogjk_server_create(const ogjk_scene* scene, const ogjk_server_config* cfg,
                               ogjk_server* out) {
  if (!out) return ogjk_fail(OGJK_ERR_INVALID, "NULL server handle");
  *out = nullptr;
  if (!scene || !scene->x || !scene->y || !scene->z || !scene->body_off || !scene->body_cnt ||
      !scene->body_r1)
    return ogjk_fail(OGJK_ERR_INVALID, "incomplete scene");
  if (scene->clusters) return ogjk_fail(OGJK_ERR_UNSUPPORTED, "posed bodies are not culled");
  ogjk_server_config c{65536, 4, 0};
  if (cfg) c = *cfg;
  if (c.max_batch_pairs <= 0 || c.slots <= 0 || c.slots > 64)
    return ogjk_fail(OGJK_ERR_INVALID, "max_batch_pairs > 0 and 1 <= slots <= 64 required");
  auto s = std::make_unique<ogjk_server_s>();
  s->capacity = c.max_batch_pairs;
  int nslots = c.slots;
  if (c.memory_budget > 0) {
    const int64_t per = (int64_t)ogjk_server_s::slot_device_bytes(c.max_batch_pairs);
    const int64_t fit = c.memory_budget / per;
    if (fit < 1) return ogjk_fail(OGJK_ERR_NOMEM, "budget below one slot (%lld bytes)",
                                  (long long)per);
    if (fit < nslots) nslots = (int)fit;
  }
  if (hipGetDevice(&s->device) != hipSuccess) return ogjk_fail(OGJK_ERR_HIP, "hipGetDevice");
  s->sa = ogjk::SceneArgs{scene->x, scene->y, scene->z, scene->body_off, scene->body_cnt,
                          scene->body_r1, scene->n_bodies, nullptr};
  s->slots.resize(nslots);
  for (Slot& sl : s->slots) {
    const hipError_t e = s->init_slot(sl);
    if (e != hipSuccess)
      return ogjk_fail(e == hipErrorOutOfMemory ? OGJK_ERR_NOMEM : OGJK_ERR_HIP,
                       hipGetErrorString(e));
  }
  ogjk_server_s* raw = s.get();
  raw->dispatcher = std::thread([raw] { raw->dispatch_loop(); });
  raw->completer = std::thread([raw] { raw->complete_loop(); });
  *out = s.release();
  return OGJK_OK;
}

ogjk_status ogjk_server_submit(ogjk_server s, const int32_t* pairs, const float* poses,
                               int32_t n_pairs, gkSimplex* simplices, float* distances,
                               float* normals, double* f64, int64_t* ticket) {
  if (!s || !ticket) return ogjk_fail(OGJK_ERR_INVALID, "NULL server or ticket");
  if (n_pairs <= 0 || n_pairs > s->capacity)
    return ogjk_fail(OGJK_ERR_INVALID, "request of %d pairs (1..%d per request)", n_pairs,
                     s->capacity);
  if (!pairs || !poses || !distances)
    return ogjk_fail(OGJK_ERR_INVALID, "NULL pairs, poses or distances");
  auto r = std::make_shared<Request>();
  r->pairs = pairs;
  r->poses = poses;
  r->n = n_pairs;
  r->simplices = simplices;
  r->distances = distances;
  r->normals = normals;
  r->f64 = f64;
  r->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lock(s->mu);
    if (s->closing) return ogjk_fail(OGJK_ERR_INVALID, "server is being destroyed");
    r->ticket = s->next_ticket++;
    s->tickets[r->ticket] = r;
    s->pending.push_back(r);
    *ticket = r->ticket;
  }
  s->cv_dispatch.notify_one();
  return OGJK_OK;
}

ogjk_status ogjk_server_wait(ogjk_server s, int64_t ticket, double* latency_s, int64_t* batch) {
  if (!s) return ogjk_fail(OGJK_ERR_INVALID, "NULL server");
  std::shared_ptr<Request> r;
  {
    std::unique_lock<std::mutex> lock(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return ogjk_fail(OGJK_ERR_INVALID, "unknown ticket %lld",
                                                 (long long)ticket);
    r = it->second;
    s->waiters++;
    s->cv_done.wait(lock, [&] { return r->done; });
    s->waiters--;
    s->tickets.erase(ticket);
    // a destroy in progress waits for the last waiter to leave
    if (s->closing && s->waiters == 0) s->cv_done.notify_all();
  }
  if (latency_s) *latency_s = r->latency_s;
  if (batch) *batch = r->batch;
  if (r->cancelled)
    return ogjk_fail(r->status, "request cancelled: the server was destroyed before launching it");
  if (r->status != OGJK_OK) return ogjk_fail(r->status, "request failed on the device");
  return OGJK_OK;
}

ogjk_status ogjk_server_poll(ogjk_server s, int64_t ticket, int* done) {
  if (!s || !done) return ogjk_fail(OGJK_ERR_INVALID, "NULL server or flag");
  std::lock_guard<std::mutex> lock(s->mu);
  auto it = s->tickets.find(ticket);
  if (it == s->tickets.end()) return ogjk_fail(OGJK_ERR_INVALID, "unknown ticket");
  *done = it->second->done ? 1 : 0;
  return OGJK_OK;
}

ogjk_status ogjk_server_info(ogjk_server s, int32_t* slots, int32_t* capacity, int64_t* launches,
                             int64_t* pairs_done) {
  if (!s) return ogjk_fail(OGJK_ERR_INVALID, "NULL server");
  std::lock_guard<std::mutex> lock(s->mu);
  if (slots) *slots = (int32_t)s->slots.size();
  if (capacity) *capacity = s->capacity;
  if (launches) *launches = s->launches;
  if (pairs_done) *pairs_done = s->pairs_done;
  return OGJK_OK;
}

ogjk_status ogjk_server_destroy(ogjk_server s) {
  if (!s) return OGJK_OK;
  s->shutdown();
  const hipError_t e = s->async_error;
  delete s;
  if (e != hipSuccess) return ogjk_fail(OGJK_ERR_HIP, "server: %s", hipGetErrorString(e));
  return OGJK_OK;
}

}  // extern "C"
