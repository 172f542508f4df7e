// Host side of the persistent batched launch (bri17_batch_* of
// include/bri17_amd.h): builds the instance table and the work queue of a
// batch of operators, uploads them with the launch and tracks completion.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <set>
#include <vector>

#include "batch.hpp"
#include "bri17_amd.h"
#include "capi_internal.hpp"

struct bri17_batch {
  void* d_buf = nullptr;  // device: instances | items | control words
  size_t d_cap = 0;
  void* h_buf = nullptr;  // pinned host staging of the same layout
  size_t h_cap = 0;
  hipEvent_t done = nullptr;  // recorded after the last launch
  bool pending = false;
  int* d_ctl = nullptr;
  int nctl = 0;
  int spin_limit = 1 << 24;
  int last_nitems = 0, last_grid = 0, last_threads = 0, last_dtype = 0;
  int last_nfwd = 0, last_npen = 0, last_n = 0;
  int last_primes = 0;  // prime-radix tier of the launch (fft_prime_tier)
  hipStream_t last_stream = nullptr;
  size_t last_lds = 0, last_off_items = 0;
  int graphs = 0;  // live graphs captured from this context (they replay its tables)
  std::mutex mtx;
};

namespace {

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

int wait_pending(bri17_batch* b) {
  if (b->pending) {
    if (hipEventSynchronize(b->done) != hipSuccess)
      return b17_err(BRI17_ERR_HIP, "hipEventSynchronize(batch) failed");
    b->pending = false;
  }
  return BRI17_OK;
}

}  // namespace

extern "C" {

int bri17_batch_create(bri17_batch** out) {
  if (!out) return b17_err(BRI17_ERR_INVALID, "null argument");
  auto* b = new (std::nothrow) bri17_batch;
  if (!b) return b17_err(BRI17_ERR_NOMEM, "out of host memory");
  if (hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess) {
    delete b;
    return b17_err(BRI17_ERR_HIP, "hipEventCreate failed");
  }
  *out = b;
  return BRI17_OK;
}

int bri17_batch_destroy(bri17_batch* b) {
  if (!b) return BRI17_OK;
  wait_pending(b);
  if (b->d_buf) (void)hipFree(b->d_buf);
  if (b->h_buf) (void)hipHostFree(b->h_buf);
  if (b->done) (void)hipEventDestroy(b->done);
  delete b;
  return BRI17_OK;
}

int bri17_batch_set_spin_limit(bri17_batch* b, int limit) {
  if (!b || limit < 0) return b17_err(BRI17_ERR_INVALID, "invalid argument");
  b->spin_limit = limit;
  return BRI17_OK;
}

int bri17_batch_run(bri17_batch* b, const bri17_batch_item* items, int n, void* stream) {
  if (!b || !items || n < 1) return b17_err(BRI17_ERR_INVALID, "invalid argument");
  if (n > 65535) return b17_err(BRI17_ERR_INVALID, "at most 65535 instances per batch");
  std::lock_guard<std::mutex> g(b->mtx);
  if (b->graphs)
    return b17_err(BRI17_ERR_STATE,
                   "this batch context is captured in a graph, which replays its device tables: "
                   "run other batches on another context, or destroy the graph first");
  const int dtype = items[0].op ? items[0].op->plan->dtype : -1;
  std::set<const bri17_operator*> seen;
  for (int i = 0; i < n; i++) {
    const bri17_batch_item& it = items[i];
    if (!it.op || !it.in || !it.out || it.in == it.out)
      return b17_err(BRI17_ERR_INVALID, "batch item with a null or aliased buffer");
    if (it.mode != 0 && it.mode != 1) return b17_err(BRI17_ERR_INVALID, "mode must be 0 or 1");
    if (it.op->plan->dtype != dtype)
      return b17_err(BRI17_ERR_INVALID, "all operators of a batch must share the dtype");
    if (!seen.insert(it.op).second)
      return b17_err(BRI17_ERR_INVALID,
                     "an operator appears twice in one batch (its workspace would be shared)");
  }
  std::vector<const b17::Plan*> plans((size_t)n);
  for (int i = 0; i < n; i++) plans[i] = items[i].op->plan.get();
  int threads = 0;
  size_t lds = 0;
  b17_batch_shape(plans, threads, lds);
  std::vector<b17::BatchInst> insts((size_t)n);
  for (int i = 0; i < n; i++) {
    bri17_operator* op = items[i].op;
    std::lock_guard<std::mutex> og(op->mtx);
    int rc = b17_ensure_ws(op);
    if (rc) return rc;
    insts[i] = b17_batch_inst(*op->plan, op->mu, op->nu, items[i].mode, items[i].in, items[i].out,
                              op->ws, threads, lds);
  }
  return b17_batch_launch(b, dtype, insts, threads, lds, (hipStream_t)stream);
}

}  // extern "C"

// Workgroups of the batch kernel are sized for items of moderate size, not
// for the largest solo launch: two 512-thread workgroups with 77.5 KiB of
// LDS each per CU (the bodies' 128 VGPRs allow 16 waves per CU), and each
// plan's items re-tiled to fit (b17::batch_geometry).  Measured on the
// 64-instance workload: 1.31 ms per launch, against 1.53 ms with four
// 256-thread workgroups of 39 KiB, 1.71 ms with eight of 128 threads and
// 19 KiB, 1.64 ms with one of 1,024 threads.
void b17_batch_shape(const std::vector<const b17::Plan*>& plans, int& threads, size_t& lds) {
  threads = 512;
  lds = 78 * 1024 - 512;
  for (const b17::Plan* pl : plans) {
    size_t l = 0;
    int t = 0;
    b17::batch_minimum(*pl, l, t);
    threads = std::max(threads, t);
    lds = std::max(lds, l);
  }
  threads = std::min(threads, 1024);
}

b17::BatchInst b17_batch_inst(const b17::Plan& pl, double mu, double nu, int mode, const void* in,
                              void* out, void* ws, int threads, size_t lds) {
  b17::BatchInst I;
  std::memset(&I, 0, sizeof(I));
  I.pc = b17::pencil_for(pl, mu, nu, mode);
  (void)b17::batch_geometry(pl, lds, threads, I.pp, I.na, I.pc);  // fits: lds >= batch_minimum
  I.pc.xcd_remap = 0;
  I.in = in;
  I.out = out;
  I.ws = ws;
  I.nb = I.pc.ntiles;
  return I;
}

int b17_batch_launch(bri17_batch* b, int dtype, const std::vector<b17::BatchInst>& insts,
                     int threads, size_t lds, hipStream_t st) {
  const int n = (int)insts.size();
  // The same instance table as the previous launch on the same stream (same
  // operators, modes, materials and buffers, e.g. one apply per iteration of
  // a solver): its device copy and work queue are still valid, so the launch
  // is enqueued at once -- no rebuild, no upload, and no wait for the
  // previous launch, after which stream order resets the counters.
  if (b->last_nitems && b->d_buf && b->h_buf && n == b->last_n && st == b->last_stream &&
      dtype == b->last_dtype && threads == b->last_threads && lds == b->last_lds &&
      std::memcmp(b->h_buf, insts.data(), sizeof(b17::BatchInst) * n) == 0) {
    char* d = static_cast<char*>(b->d_buf);
    int grid = 0;
    hipError_t e = b17::launch_batch(dtype, b->last_primes,
                                     reinterpret_cast<const b17::BatchInst*>(d),
                                     reinterpret_cast<const int2*>(d + b->last_off_items),
                                     b->last_nitems, b->last_nfwd, b->last_npen, b->d_ctl, b->nctl,
                                     threads, lds, b->spin_limit, st, &grid);
    if (e != hipSuccess) return b17_err(BRI17_ERR_HIP, hipGetErrorString(e));
    if (hipEventRecord(b->done, st) != hipSuccess)
      return b17_err(BRI17_ERR_HIP, "hipEventRecord(batch) failed");
    b->pending = true;
    b->last_grid = grid;
    return BRI17_OK;
  }
  int rc = wait_pending(b);
  if (rc) return rc;
  std::vector<int2> work;
  // queue: every forward plane block, then every pencil tile, then every
  // inverse plane block -- an item depends only on items before it.  Within
  // each kind, instances with the largest items come first, so that the
  // end of each phase is filled with small items.
  auto item_size = [&](int i, int type) -> int64_t {
    const b17::BatchInst& I = insts[i];
    if (type == 1) return (int64_t)I.pc.dim * I.pc.n0 * I.pc.p * I.pc.w;
    const int rows = I.pp.has_col ? I.pp.q * I.pp.g3 : I.pp.q;
    return (int64_t)rows * I.pp.nl;
  };
  std::vector<int> order((size_t)n);
  for (int type = 0; type < 3; type++) {
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return item_size(a, type) > item_size(b, type); });
    for (int i : order) {
      const int cnt = type == 1 ? insts[i].nb : insts[i].na;
      for (int k = 0; k < cnt; k++) work.push_back(make_int2(i | (type << 16), k));
    }
  }
  const int nitems = (int)work.size();
  const int nctl = 4 + 2 * n;
  int n_fwd = 0, n_pen = 0;
  for (int i = 0; i < n; i++) {
    n_fwd += insts[i].na;
    n_pen += insts[i].nb;
  }
  const size_t off_items = align_up(sizeof(b17::BatchInst) * n, 256);
  const size_t off_ctl = align_up(off_items + sizeof(int2) * work.size(), 256);
  const size_t bytes = off_ctl + sizeof(int) * nctl;
  if (bytes > b->d_cap) {
    if (b->d_buf) (void)hipFree(b->d_buf);
    if (b->h_buf) (void)hipHostFree(b->h_buf);
    b->d_buf = b->h_buf = nullptr;
    b->d_cap = b->h_cap = 0;
    if (hipMalloc(&b->d_buf, bytes) != hipSuccess)
      return b17_err(BRI17_ERR_NOMEM, "hipMalloc(batch tables) failed");
    if (hipHostMalloc(&b->h_buf, bytes, hipHostMallocDefault) != hipSuccess)
      return b17_err(BRI17_ERR_NOMEM, "hipHostMalloc(batch staging) failed");
    b->d_cap = b->h_cap = bytes;
  }
  char* h = static_cast<char*>(b->h_buf);
  std::memcpy(h, insts.data(), sizeof(b17::BatchInst) * n);
  std::memcpy(h + off_items, work.data(), sizeof(int2) * work.size());
  if (hipMemcpyAsync(b->d_buf, b->h_buf, off_ctl, hipMemcpyHostToDevice, st) != hipSuccess)
    return b17_err(BRI17_ERR_HIP, "hipMemcpyAsync(batch tables) failed");
  char* d = static_cast<char*>(b->d_buf);
  b->d_ctl = reinterpret_cast<int*>(d + off_ctl);
  b->nctl = nctl;
  int grid = 0;
  int primes = 0;
  for (const b17::BatchInst& I : insts)
    primes = std::max({primes, b17::plane_prime_tier(I.pp), b17::pencil_prime_tier(I.pc)});
  hipError_t e = b17::launch_batch(dtype, primes, reinterpret_cast<const b17::BatchInst*>(d),
                                   reinterpret_cast<const int2*>(d + off_items), nitems, n_fwd,
                                   n_pen, b->d_ctl, nctl, threads, lds, b->spin_limit, st, &grid);
  if (e != hipSuccess) return b17_err(BRI17_ERR_HIP, hipGetErrorString(e));
  if (hipEventRecord(b->done, st) != hipSuccess)
    return b17_err(BRI17_ERR_HIP, "hipEventRecord(batch) failed");
  b->pending = true;
  b->last_nitems = nitems;
  b->last_n = n;
  b->last_stream = st;
  b->last_nfwd = n_fwd;
  b->last_npen = n_pen;
  b->last_grid = grid;
  b->last_threads = threads;
  b->last_lds = lds;
  b->last_dtype = dtype;
  b->last_primes = primes;
  b->last_off_items = off_items;
  return BRI17_OK;
}

int* b17_batch_ctl(bri17_batch* b) { return b ? b->d_ctl : nullptr; }

// A graph replays the context's device tables and every member operator's
// workspace: while it lives, the context runs nothing else (bri17_batch_run
// refuses) and the workspaces are pinned (never reallocated).
int b17_batch_pin(bri17_batch* b, const bri17_batch_item* items, int n) {
  std::lock_guard<std::mutex> g(b->mtx);
  for (int i = 0; i < n; i++) {
    std::lock_guard<std::mutex> og(items[i].op->mtx);
    items[i].op->ws_pinned = true;
  }
  b->graphs++;
  return BRI17_OK;
}

void b17_batch_unpin(bri17_batch* b) {
  std::lock_guard<std::mutex> g(b->mtx);
  if (b->graphs > 0) b->graphs--;
}

// Launch the last batch again from the tables already on the device
// (counter reset + persistent kernel only), e.g. inside a graph capture.
int b17_batch_relaunch(bri17_batch* b, hipStream_t st) {
  if (!b || !b->d_buf || !b->last_nitems) return b17_err(BRI17_ERR_STATE, "no batch to relaunch");
  char* d = static_cast<char*>(b->d_buf);
  int grid = 0;
  hipError_t e = b17::launch_batch(b->last_dtype, b->last_primes,
                                   reinterpret_cast<const b17::BatchInst*>(d),
                                   reinterpret_cast<const int2*>(d + b->last_off_items),
                                   b->last_nitems, b->last_nfwd, b->last_npen, b->d_ctl, b->nctl,
                                   b->last_threads, b->last_lds, b->spin_limit, st, &grid);
  if (e != hipSuccess) return b17_err(BRI17_ERR_HIP, hipGetErrorString(e));
  return BRI17_OK;
}

extern "C" {

int bri17_batch_wait(bri17_batch* b, int* dependency_timeouts) {
  if (!b) return b17_err(BRI17_ERR_INVALID, "null batch");
  std::lock_guard<std::mutex> g(b->mtx);
  int rc = wait_pending(b);
  if (rc) return rc;
  if (dependency_timeouts) {
    *dependency_timeouts = 0;
    if (b->d_ctl && hipMemcpy(dependency_timeouts, b->d_ctl + 1, sizeof(int),
                              hipMemcpyDeviceToHost) != hipSuccess)
      return b17_err(BRI17_ERR_HIP, "hipMemcpy(batch status) failed");
  }
  return BRI17_OK;
}

int bri17_batch_info(const bri17_batch* b, int* work_items, int* workgroups, int* threads,
                     size_t* lds_bytes) {
  if (!b) return b17_err(BRI17_ERR_INVALID, "null batch");
  if (work_items) *work_items = b->last_nitems;
  if (workgroups) *workgroups = b->last_grid;
  if (threads) *threads = b->last_threads;
  if (lds_bytes) *lds_bytes = b->last_lds;
  return BRI17_OK;
}

}  // extern "C"
