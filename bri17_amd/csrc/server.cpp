// Request loop of bri17_server_* (include/bri17_amd.h): continuous batching
// of operator applies and Green solves of mixed grids on one device.
//
// Threads: any number of submitters; one worker per server that owns the
// streams, the batch contexts and the pool of device buffers.  Shared state
// (queue, request records, counters) is guarded by `mtx`; the worker never
// holds it across a HIP call that can block.
//
// Worker loop: (1) retire finished waves (event query; batched waves also
// report their dependency-wait status word), returning their buffers to the
// idle pool; (2) while a stream slot is free, take requests from the head of
// the queue -- one request with fixed-size kernels or a large grid alone, or
// a run of consecutive batchable requests of one dtype; a heavy request
// (large grid) whose turn comes while `max_heavy_waves` heavy waves are in
// flight stays queued and the requests behind it go first -- reserve their HBM
// footprint (idle buffers first, then new allocations, evicting idle buffers
// when the budget requires) and launch them; (3) when nothing is in flight,
// sleep until a submission arrives, else poll.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "batch.hpp"
#include "bri17_amd.h"
#include "capi_internal.hpp"
#include "plan.hpp"

namespace {

using Clock = std::chrono::steady_clock;

constexpr size_t kDefaultSpin = 1 << 24;

// Buffers come in size classes (4 per power of two) so that idle workspaces
// are reused across nearby grid sizes.
size_t size_class(size_t n) {
  if (n <= 4096) return 4096;
  int e = 63 - __builtin_clzll((unsigned long long)n);  // 2^e <= n < 2^(e+1)
  const size_t step = (size_t)1 << (e - 2);
  return (n + step - 1) / step * step;
}

struct Req {
  int64_t id = 0;
  bri17_request r{};
  std::shared_ptr<b17::Plan> plan;
  size_t field_bytes = 0;
  size_t ws_cls = 0, stage_cls = 0;  // stage_cls: class of each staged field (host memory)
  size_t footprint = 0;
  bool solo = false;
  bool heavy = false;
  int state = BRI17_REQ_QUEUED;
  int error = 0;
  Clock::time_point t_submit, t_launch, t_done;
  int64_t wave = -1;
  int wave_size = 0;
  // device buffers held while admitted
  void* ws = nullptr;
  void* d_in = nullptr;
  void* d_out = nullptr;
  void* h_stage = nullptr;  // host-memory requests: pinned staging (in, then out)
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  bri17_batch* batch = nullptr;
  int* h_status = nullptr;  // pinned copy of the wave's dependency-wait status
  bool busy = false;
  bool batched = false;
  bool heavy = false;
  int launch_error = 0;
  std::vector<std::shared_ptr<Req>> reqs;
};

}  // namespace

struct bri17_server {
  int device = 0;
  bri17_server_config cfg{};
  // ---- shared, under mtx
  std::mutex mtx;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Req>> queue;
  std::unordered_map<int64_t, std::shared_ptr<Req>> records;
  std::map<const b17::Plan*, std::shared_ptr<b17::Plan>> plans;
  int64_t next_id = 1, next_wave = 0;
  int64_t n_submitted = 0, n_completed = 0, n_failed = 0, n_rejected = 0, n_cancelled = 0;
  int64_t n_waves = 0, n_batched = 0, n_solo = 0, n_running = 0;
  size_t admitted = 0, cached = 0, peak = 0;
  int64_t peak_waves = 0;
  bool stop = false, trim = false;
  // ---- worker-owned
  std::vector<Slot> slots;
  std::map<size_t, std::vector<void*>> idle;
  std::map<size_t, std::vector<void*>> idle_host;  // pinned staging buffers
  std::thread worker;
};

namespace {

void note_peak(bri17_server* s) { s->peak = std::max(s->peak, s->admitted + s->cached); }

// ---------------------------------------------------------------- pool
// (worker thread only; counters updated under mtx)
void evict_one(bri17_server* s) {
  for (auto it = s->idle.rbegin(); it != s->idle.rend(); ++it) {
    if (it->second.empty()) continue;
    void* p = it->second.back();
    it->second.pop_back();
    (void)hipFree(p);
    std::lock_guard<std::mutex> g(s->mtx);
    s->cached -= it->first;
    return;
  }
}

void evict_all(bri17_server* s) {
  for (auto& kv : s->idle_host) {
    for (void* p : kv.second) (void)hipHostFree(p);
    kv.second.clear();
  }
  for (auto& kv : s->idle) {
    for (void* p : kv.second) (void)hipFree(p);
    std::lock_guard<std::mutex> g(s->mtx);
    s->cached -= kv.first * kv.second.size();
    kv.second.clear();
  }
}

void release(bri17_server* s, void* p, size_t cls) {
  if (!p) return;
  s->idle[cls].push_back(p);
  std::lock_guard<std::mutex> g(s->mtx);
  s->cached += cls;
}

// Reserve the footprint of `q` and bind its device buffers.  false: it does
// not fit now (nothing is held).  A hipMalloc failure with nothing else in
// flight fails the request (q.error set).
bool admit(bri17_server* s, Req& q, bool others_in_flight) {
  struct Need {
    void** slot;
    size_t cls;
  } need[3];
  int nn = 0;
  need[nn++] = {&q.ws, q.ws_cls};
  if (q.r.memory == BRI17_MEM_HOST) {
    need[nn++] = {&q.d_in, q.stage_cls};
    need[nn++] = {&q.d_out, q.stage_cls};
  }
  size_t reused = 0;
  for (int i = 0; i < nn; i++) {
    auto it = s->idle.find(need[i].cls);
    if (it != s->idle.end() && !it->second.empty()) {
      *need[i].slot = it->second.back();
      it->second.pop_back();
      reused += need[i].cls;
    }
  }
  {
    std::lock_guard<std::mutex> g(s->mtx);
    s->cached -= reused;
    s->admitted += q.footprint;
  }
  auto over = [&] {
    std::lock_guard<std::mutex> g(s->mtx);
    return s->admitted + s->cached > s->cfg.hbm_budget_bytes;
  };
  while (over()) {
    bool any = false;
    for (auto& kv : s->idle) any |= !kv.second.empty();
    if (!any) break;
    evict_one(s);
  }
  bool ok = !over();
  for (int i = 0; ok && i < nn; i++) {
    if (*need[i].slot) continue;
    if (hipMalloc(need[i].slot, need[i].cls) != hipSuccess) {
      (void)hipGetLastError();
      evict_all(s);
      if (hipMalloc(need[i].slot, need[i].cls) != hipSuccess) {
        (void)hipGetLastError();
        *need[i].slot = nullptr;
        ok = false;
        if (!others_in_flight) q.error = BRI17_ERR_NOMEM;
      }
    }
  }
  if (!ok) {
    {
      std::lock_guard<std::mutex> g(s->mtx);
      s->admitted -= q.footprint;
    }
    for (int i = 0; i < nn; i++) {
      release(s, *need[i].slot, need[i].cls);
      *need[i].slot = nullptr;
    }
    return false;
  }
  std::lock_guard<std::mutex> g(s->mtx);
  note_peak(s);
  return true;
}

void unadmit(bri17_server* s, Req& q) {
  release(s, q.ws, q.ws_cls);
  release(s, q.d_in, q.stage_cls);
  release(s, q.d_out, q.stage_cls);
  q.ws = q.d_in = q.d_out = nullptr;
  std::lock_guard<std::mutex> g(s->mtx);
  s->admitted -= q.footprint;
}

// Host-memory requests are staged through a pinned buffer: the worker copies
// the input into it, and the copies to and from the device are then
// asynchronous (a copy from or to pageable memory would hold the worker
// until the wave's kernels finish); the result is copied out when the wave
// retires.
void release_host(bri17_server* s, Req& q) {
  if (q.h_stage) s->idle_host[q.stage_cls].push_back(q.h_stage);
  q.h_stage = nullptr;
}

// ---------------------------------------------------------------- waves
int launch_wave(bri17_server* s, Slot& sl) {
  hipStream_t st = sl.stream;
  for (auto& q : sl.reqs) {
    if (q->r.memory != BRI17_MEM_HOST) continue;
    auto& pool = s->idle_host[q->stage_cls];
    if (!pool.empty()) {
      q->h_stage = pool.back();
      pool.pop_back();
    } else if (hipHostMalloc(&q->h_stage, q->stage_cls, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      q->h_stage = nullptr;
      return BRI17_ERR_NOMEM;
    }
    std::memcpy(q->h_stage, q->r.in, q->field_bytes);
    if (hipMemcpyAsync(q->d_in, q->h_stage, q->field_bytes, hipMemcpyHostToDevice, st) != hipSuccess)
      return BRI17_ERR_HIP;
  }
  auto in_of = [](const Req& q) { return q.r.memory == BRI17_MEM_HOST ? q.d_in : q.r.in; };
  auto out_of = [](const Req& q) { return q.r.memory == BRI17_MEM_HOST ? q.d_out : q.r.out; };
  // A wave of one batchable request runs the plan's own three kernels: the
  // same runtime-size arithmetic as the persistent launch (so the same bits),
  // without its table upload and dependency counters.
  if (sl.batched && sl.reqs.size() > 1) {
    std::vector<const b17::Plan*> plans;
    for (auto& q : sl.reqs) plans.push_back(q->plan.get());
    int threads = 0;
    size_t lds = 0;
    b17_batch_shape(plans, threads, lds);
    std::vector<b17::BatchInst> insts;
    insts.reserve(sl.reqs.size());
    for (auto& q : sl.reqs)
      insts.push_back(b17_batch_inst(*q->plan, q->r.mu, q->r.nu, q->r.mode, in_of(*q), out_of(*q),
                                     q->ws, threads, lds));
    *sl.h_status = 0;
    int spin;
    {
      std::lock_guard<std::mutex> g(s->mtx);
      spin = s->cfg.spin_limit;
    }
    bri17_batch_set_spin_limit(sl.batch, spin);
    int rc = b17_batch_launch(sl.batch, sl.reqs[0]->r.dtype, insts, threads, lds, st);
    if (rc) return rc;
    if (hipMemcpyAsync(sl.h_status, b17_batch_ctl(sl.batch) + 1, sizeof(int),
                       hipMemcpyDeviceToHost, st) != hipSuccess)
      return BRI17_ERR_HIP;
  } else {
    const Req& q = *sl.reqs[0];
    b17::Status r = b17::run_pipeline(*q.plan, q.r.mu, q.r.nu, q.r.mode, in_of(q), out_of(q), q.ws, st);
    if (!r.ok()) return r.code;
  }
  for (auto& q : sl.reqs)
    if (q->r.memory == BRI17_MEM_HOST &&
        hipMemcpyAsync(q->h_stage, q->d_out, q->field_bytes, hipMemcpyDeviceToHost, st) != hipSuccess)
      return BRI17_ERR_HIP;
  return hipEventRecord(sl.done, st) == hipSuccess ? BRI17_OK : BRI17_ERR_HIP;
}

void finish_wave(bri17_server* s, Slot& sl, int code) {
  const Clock::time_point t = Clock::now();
  if (!code && sl.batched && sl.reqs.size() > 1 && *sl.h_status)
    code = BRI17_ERR_STATE;  // a dependency wait timed out
  for (auto& q : sl.reqs) {
    if (q->r.memory == BRI17_MEM_HOST && q->h_stage) {
      if (!code) std::memcpy(q->r.out, q->h_stage, q->field_bytes);
      release_host(s, *q);
    }
    unadmit(s, *q);
  }
  {
    std::lock_guard<std::mutex> g(s->mtx);
    for (auto& q : sl.reqs) {
      q->t_done = t;
      q->state = code ? BRI17_REQ_FAILED : BRI17_REQ_DONE;
      q->error = code;
      q->plan.reset();  // the resident copy stays in s->plans
      (code ? s->n_failed : s->n_completed)++;
      s->n_running--;
    }
  }
  sl.reqs.clear();
  sl.busy = false;
  s->cv_done.notify_all();
}

bool retire(bri17_server* s) {
  bool any = false;
  for (auto& sl : s->slots) {
    if (!sl.busy) continue;
    if (sl.launch_error) {
      // work already queued on the stream may still read the buffers
      (void)hipStreamSynchronize(sl.stream);
      (void)hipGetLastError();
      finish_wave(s, sl, sl.launch_error);
      sl.launch_error = 0;
      any = true;
      continue;
    }
    hipError_t e = hipEventQuery(sl.done);
    if (e == hipErrorNotReady) continue;
    finish_wave(s, sl, e == hipSuccess ? BRI17_OK : BRI17_ERR_HIP);
    any = true;
  }
  return any;
}

bool busy_any(const bri17_server* s) {
  for (auto& sl : s->slots)
    if (sl.busy) return true;
  return false;
}

// Take the next wave from the queue into `sl` and launch it: the first
// request that may start now (a heavy one only while fewer than
// max_heavy_waves heavy waves are in flight), with the batchable requests
// right behind it.  false: nothing launched (nothing may start, or the
// request does not fit in the HBM budget now).
bool start_wave(bri17_server* s, Slot& sl) {
  int heavy_in_flight = 0;
  for (auto& o : s->slots) heavy_in_flight += o.busy && o.heavy ? 1 : 0;
  const bool heavy_ok = s->cfg.max_heavy_waves < 0 || heavy_in_flight < s->cfg.max_heavy_waves;
  std::vector<std::shared_ptr<Req>> take;
  size_t at = 0;  // queue position of the first request taken
  {
    std::lock_guard<std::mutex> g(s->mtx);
    while (at < s->queue.size() && s->queue[at]->heavy && !heavy_ok) at++;
    if (at == s->queue.size()) return false;
    const auto first = s->queue[at];
    size_t end = at + 1;
    if (!first->solo) {
      while (end < s->queue.size() && (int)(end - at) < s->cfg.max_batch) {
        const auto& nx = s->queue[end];
        if (nx->solo || nx->r.dtype != first->r.dtype) break;
        end++;
      }
    }
    take.assign(s->queue.begin() + (std::ptrdiff_t)at, s->queue.begin() + (std::ptrdiff_t)end);
    s->queue.erase(s->queue.begin() + (std::ptrdiff_t)at, s->queue.begin() + (std::ptrdiff_t)end);
  }
  // reserve footprints in order; what does not fit goes back where it was
  const bool others = busy_any(s);
  size_t n_ok = 0;
  int fail_code = 0;
  for (; n_ok < take.size(); n_ok++) {
    if (!admit(s, *take[n_ok], others || n_ok > 0)) {
      fail_code = take[n_ok]->error;
      break;
    }
  }
  if (n_ok < take.size()) {
    std::lock_guard<std::mutex> g(s->mtx);
    size_t back_from = n_ok;
    if (fail_code && n_ok == 0) {
      // the request cannot be allocated even with an idle device: fail it
      auto& q = take[0];
      q->state = BRI17_REQ_FAILED;
      q->t_done = Clock::now();
      q->plan.reset();
      s->n_failed++;
      back_from = 1;
      s->cv_done.notify_all();
    }
    // (a cancellation may have shortened the queue since the requests were
    // taken out)
    at = std::min(at, s->queue.size());
    s->queue.insert(s->queue.begin() + (std::ptrdiff_t)at, take.begin() + (std::ptrdiff_t)back_from,
                    take.end());
    take.resize(n_ok);
  }
  if (take.empty()) return fail_code != 0;
  sl.reqs = std::move(take);
  sl.batched = !sl.reqs[0]->solo;
  sl.heavy = sl.reqs[0]->heavy;
  sl.busy = true;
  int64_t in_flight = 0;
  for (auto& o : s->slots) in_flight += o.busy ? 1 : 0;
  {
    std::lock_guard<std::mutex> g(s->mtx);
    s->peak_waves = std::max(s->peak_waves, in_flight);
    const int64_t w = s->next_wave++;
    const Clock::time_point t = Clock::now();
    for (auto& q : sl.reqs) {
      q->state = BRI17_REQ_RUNNING;
      q->t_launch = t;
      q->wave = w;
      q->wave_size = (int)sl.reqs.size();
    }
    s->n_running += (int64_t)sl.reqs.size();
    s->n_waves++;
    (sl.batched ? s->n_batched : s->n_solo)++;
  }
  sl.launch_error = launch_wave(s, sl);
  return true;
}

void worker_main(bri17_server* s) {
  (void)hipSetDevice(s->device);
  for (;;) {
    retire(s);
    bool trim = false;
    {
      std::lock_guard<std::mutex> g(s->mtx);
      trim = s->trim;
      s->trim = false;
    }
    if (trim) evict_all(s);
    for (auto& sl : s->slots) {
      if (sl.busy) continue;
      if (!start_wave(s, sl)) break;
    }
    std::unique_lock<std::mutex> lk(s->mtx);
    if (busy_any(s)) {
      lk.unlock();
      std::this_thread::yield();
      continue;
    }
    if (s->stop && s->queue.empty()) break;
    if (!s->queue.empty()) {
      // the head did not fit with nothing in flight (only possible right
      // after an eviction race); go round again
      continue;
    }
    s->cv_work.wait(lk, [&] { return s->stop || s->trim || !s->queue.empty(); });
  }
  evict_all(s);
}

int validate(const bri17_request* r) {
  if (!r || !r->in || !r->out || r->in == r->out) return BRI17_ERR_INVALID;
  if (r->dim != 2 && r->dim != 3) return BRI17_ERR_INVALID;
  if (r->mode != 0 && r->mode != 1) return BRI17_ERR_INVALID;
  if (r->dtype != BRI17_F32 && r->dtype != BRI17_F64) return BRI17_ERR_INVALID;
  if (r->memory != BRI17_MEM_DEVICE && r->memory != BRI17_MEM_HOST) return BRI17_ERR_INVALID;
  for (int d = 0; d < r->dim; d++)
    if (r->shape[d] < 1 || !(r->L[d] > 0.0)) return BRI17_ERR_INVALID;
  if (!(r->mu > 0.0) || !(r->nu > -1.0 && r->nu < 0.5)) return BRI17_ERR_INVALID;
  return BRI17_OK;
}

void fill_result(const Req& q, bri17_request_result* res) {
  if (!res) return;
  res->state = q.state;
  res->error = q.error;
  const bool launched = q.wave >= 0;
  const bool final = q.state >= BRI17_REQ_DONE;
  res->queued_s = launched ? std::chrono::duration<double>(q.t_launch - q.t_submit).count() : 0.0;
  res->latency_s = final ? std::chrono::duration<double>(q.t_done - q.t_submit).count() : 0.0;
  res->wave = q.wave;
  res->wave_size = q.wave_size;
}

}  // namespace

extern "C" {

int bri17_server_create(const bri17_server_config* cfg, bri17_server** out) {
  if (!out) return b17_err(BRI17_ERR_INVALID, "null argument");
  *out = nullptr;
  bri17_server_config c{};
  if (cfg) c = *cfg;
  if (c.max_batch <= 0) c.max_batch = 64;
  c.max_batch = std::min(c.max_batch, 4096);
  if (c.max_waves <= 0) c.max_waves = 4;
  if (c.max_waves > 16) return b17_err(BRI17_ERR_INVALID, "max_waves is at most 16");
  if (c.solo_nodes == 0) c.solo_nodes = (int64_t)1 << 21;
  if (c.spin_limit <= 0) c.spin_limit = (int)kDefaultSpin;
  if (c.heavy_nodes <= 0) c.heavy_nodes = (int64_t)1 << 22;
  if (c.max_heavy_waves == 0) c.max_heavy_waves = 1;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return b17_err(BRI17_ERR_HIP, "no HIP device");
  if (c.hbm_budget_bytes == 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return b17_err(BRI17_ERR_HIP, "hipMemGetInfo failed");
    c.hbm_budget_bytes = fr / 2;
  }
  auto* s = new (std::nothrow) bri17_server;
  if (!s) return b17_err(BRI17_ERR_NOMEM, "out of host memory");
  s->device = dev;
  s->cfg = c;
  s->slots.resize(c.max_waves);
  int rc = BRI17_OK;
  for (auto& sl : s->slots) {
    if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess ||
        hipHostMalloc((void**)&sl.h_status, sizeof(int), hipHostMallocDefault) != hipSuccess) {
      rc = b17_err(BRI17_ERR_HIP, "stream / event / pinned allocation failed");
      break;
    }
    if ((rc = bri17_batch_create(&sl.batch)) != BRI17_OK) break;
    bri17_batch_set_spin_limit(sl.batch, c.spin_limit);
  }
  if (rc == BRI17_OK) {
    try {
      s->worker = std::thread(worker_main, s);
    } catch (...) {
      rc = b17_err(BRI17_ERR_STATE, "cannot start the worker thread");
    }
  }
  if (rc != BRI17_OK) {
    for (auto& sl : s->slots) {
      if (sl.batch) bri17_batch_destroy(sl.batch);
      if (sl.h_status) (void)hipHostFree(sl.h_status);
      if (sl.done) (void)hipEventDestroy(sl.done);
      if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    delete s;
    return rc;
  }
  *out = s;
  return BRI17_OK;
}

int bri17_server_destroy(bri17_server* s) {
  if (!s) return BRI17_OK;
  {
    std::lock_guard<std::mutex> g(s->mtx);
    // queued requests are cancelled; admitted ones run to completion
    for (auto& q : s->queue) {
      q->state = BRI17_REQ_CANCELLED;
      q->plan.reset();
      s->n_cancelled++;
    }
    s->queue.clear();
    s->stop = true;
  }
  s->cv_work.notify_all();
  s->cv_done.notify_all();
  if (s->worker.joinable()) s->worker.join();
  for (auto& sl : s->slots) {
    bri17_batch_destroy(sl.batch);
    (void)hipHostFree(sl.h_status);
    (void)hipEventDestroy(sl.done);
    (void)hipStreamDestroy(sl.stream);
  }
  delete s;
  return BRI17_OK;
}

int bri17_server_submit(bri17_server* s, const bri17_request* r, int64_t* id) {
  return bri17_server_submit_many(s, r, 1, id);
}

int bri17_server_submit_many(bri17_server* s, const bri17_request* reqs, int n, int64_t* ids) {
  if (!s || !reqs || !ids || n < 1) return b17_err(BRI17_ERR_INVALID, "null argument");
  for (int i = 0; i < n; i++) ids[i] = 0;
  auto reject = [&](int code, const std::string& msg) {
    std::lock_guard<std::mutex> g(s->mtx);
    s->n_rejected += n;
    return b17_err(code, msg.c_str());
  };
  auto where = [&](int i) { return n > 1 ? " (request " + std::to_string(i) + ")" : std::string(); };
  std::vector<std::shared_ptr<Req>> qs((size_t)n);
  int prev = -1;
  (void)hipGetDevice(&prev);
  if (prev != s->device) (void)hipSetDevice(s->device);
  int rc = BRI17_OK;
  std::string msg;
  for (int i = 0; i < n && !rc; i++) {
    const bri17_request* r = reqs + i;
    if (validate(r)) {
      rc = BRI17_ERR_INVALID;
      msg = "invalid request" + where(i);
      break;
    }
    auto q = std::make_shared<Req>();
    q->r = *r;
    for (int d = r->dim; d < 3; d++) {
      q->r.shape[d] = 1;
      q->r.L[d] = 1.0;
    }
    rc = b17_acquire_plan(r->dim, q->r.shape, q->r.L, r->dtype, nullptr, q->plan);
    if (rc) {
      msg = b17::get_error() + where(i);
      break;
    }
    const b17::Plan& pl = *q->plan;
    q->field_bytes = (size_t)pl.dim * (size_t)pl.nnodes * pl.esz;
    q->ws_cls = size_class(pl.ws_bytes);
    q->footprint = q->ws_cls;
    if (r->memory == BRI17_MEM_HOST) {
      q->stage_cls = size_class(q->field_bytes);
      q->footprint += 2 * q->stage_cls;
    } else {
      q->footprint += 2 * q->field_bytes;
    }
    if (q->footprint > s->cfg.hbm_budget_bytes) {
      rc = BRI17_ERR_NOMEM;
      msg = "request footprint exceeds the server's HBM budget" + where(i);
      break;
    }
    q->solo = pl.use_static || (s->cfg.solo_nodes > 0 && pl.nnodes >= s->cfg.solo_nodes);
    q->heavy = pl.nnodes >= s->cfg.heavy_nodes;
    qs[i] = std::move(q);
  }
  if (prev >= 0 && prev != s->device) (void)hipSetDevice(prev);
  if (rc) return reject(rc, msg);
  const Clock::time_point t = Clock::now();
  {
    std::lock_guard<std::mutex> g(s->mtx);
    if (s->stop) return b17_err(BRI17_ERR_STATE, "server is shutting down");
    for (int i = 0; i < n; i++) {
      auto& q = qs[i];
      q->t_submit = t;
      q->id = s->next_id++;
      s->plans.emplace(q->plan.get(), q->plan);
      s->records.emplace(q->id, q);
      s->queue.push_back(q);
      s->n_submitted++;
      ids[i] = q->id;
    }
  }
  s->cv_work.notify_one();
  return BRI17_OK;
}

int bri17_server_poll(bri17_server* s, int64_t id, bri17_request_result* res) {
  if (!s) return b17_err(BRI17_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> g(s->mtx);
  auto it = s->records.find(id);
  if (it == s->records.end()) return b17_err(BRI17_ERR_INVALID, "unknown request id");
  fill_result(*it->second, res);
  if (it->second->state >= BRI17_REQ_DONE) s->records.erase(it);
  return BRI17_OK;
}

int bri17_server_wait(bri17_server* s, int64_t id, double timeout_s, bri17_request_result* res) {
  if (!s) return b17_err(BRI17_ERR_INVALID, "null argument");
  std::unique_lock<std::mutex> lk(s->mtx);
  auto it = s->records.find(id);
  if (it == s->records.end()) return b17_err(BRI17_ERR_INVALID, "unknown request id");
  std::shared_ptr<Req> q = it->second;
  auto final = [&] { return q->state >= BRI17_REQ_DONE; };
  if (timeout_s < 0) {
    s->cv_done.wait(lk, final);
  } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), final)) {
    fill_result(*q, res);
    return b17_err(BRI17_ERR_STATE, "timed out waiting for the request");
  }
  fill_result(*q, res);
  s->records.erase(id);
  return BRI17_OK;
}

int bri17_server_cancel(bri17_server* s, int64_t id) {
  if (!s) return b17_err(BRI17_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> g(s->mtx);
  auto it = std::find_if(s->queue.begin(), s->queue.end(),
                         [&](const std::shared_ptr<Req>& q) { return q->id == id; });
  if (it == s->queue.end()) return b17_err(BRI17_ERR_STATE, "request is not queued");
  (*it)->state = BRI17_REQ_CANCELLED;
  (*it)->t_done = Clock::now();
  (*it)->plan.reset();
  s->queue.erase(it);
  s->n_cancelled++;
  s->cv_done.notify_all();
  return BRI17_OK;
}

int bri17_server_drain(bri17_server* s) {
  if (!s) return b17_err(BRI17_ERR_INVALID, "null argument");
  std::unique_lock<std::mutex> lk(s->mtx);
  s->cv_done.wait(lk, [&] { return s->queue.empty() && s->n_running == 0; });
  return BRI17_OK;
}

int bri17_server_get_stats(bri17_server* s, bri17_server_stats* st) {
  if (!s || !st) return b17_err(BRI17_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> g(s->mtx);
  st->submitted = s->n_submitted;
  st->completed = s->n_completed;
  st->failed = s->n_failed;
  st->rejected = s->n_rejected;
  st->cancelled = s->n_cancelled;
  st->waves = s->n_waves;
  st->batched_waves = s->n_batched;
  st->solo_waves = s->n_solo;
  st->queued = (int64_t)s->queue.size();
  st->running = s->n_running;
  st->hbm_budget = s->cfg.hbm_budget_bytes;
  st->hbm_admitted = s->admitted;
  st->hbm_cached = s->cached;
  st->hbm_peak = s->peak;
  st->plans_resident = (int64_t)s->plans.size();
  st->peak_waves_in_flight = s->peak_waves;
  return BRI17_OK;
}

int bri17_server_set_spin_limit(bri17_server* s, int limit) {
  if (!s || limit < 1) return b17_err(BRI17_ERR_INVALID, "invalid argument");
  std::lock_guard<std::mutex> g(s->mtx);
  s->cfg.spin_limit = limit;
  return BRI17_OK;
}

int bri17_server_trim(bri17_server* s) {
  if (!s) return b17_err(BRI17_ERR_INVALID, "null argument");
  {
    std::lock_guard<std::mutex> g(s->mtx);
    s->trim = true;
  }
  s->cv_work.notify_all();
  return BRI17_OK;
}

}  // extern "C"
