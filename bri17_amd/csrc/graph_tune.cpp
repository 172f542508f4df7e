// HIP-graph capture of the operator pipeline and of batched launches, and the
// launch-shape autotuner (bri17_graph_*, bri17_batch_capture, bri17_autotune*
// and bri17_tuning_cache_* of include/bri17_amd.h).  No reference
// counterpart: the reference plans FFTW once per factory with FFTW_ESTIMATE
// (tests/test_bri17.cpp:119-127) and runs eagerly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "bri17_amd.h"
#include "capi_internal.hpp"

struct bri17_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bri17_batch* batch = nullptr;  // batch graphs: the context whose tables they replay
};

namespace {

std::mutex g_cache_mtx;
std::map<std::string, bri17_tuning> g_cache;

std::string cache_key(int dim, const int* shape, int dtype) {
  std::ostringstream o;
  o << dim << ":";
  for (int d = 0; d < dim; d++) o << (d ? "x" : "") << shape[d];
  o << ":" << (dtype == BRI17_F32 ? "f32" : "f64");
  return o.str();
}

// capture `enqueue(stream)` into an instantiated graph
template <typename F>
int capture(F enqueue, bri17_graph** out) {
  hipStream_t st = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
    return b17_err(BRI17_ERR_HIP, "hipStreamCreate failed");
  auto* g = new (std::nothrow) bri17_graph;
  int rc = BRI17_OK;
  if (!g) {
    rc = b17_err(BRI17_ERR_NOMEM, "out of host memory");
  } else if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    rc = b17_err(BRI17_ERR_HIP, "hipStreamBeginCapture failed");
  } else {
    rc = enqueue(st);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(st, &graph);
    if (rc == BRI17_OK && e != hipSuccess) rc = b17_err(BRI17_ERR_HIP, "hipStreamEndCapture failed");
    if (rc == BRI17_OK) {
      g->graph = graph;
      if (hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0) != hipSuccess)
        rc = b17_err(BRI17_ERR_HIP, "hipGraphInstantiate failed");
    } else if (graph) {
      (void)hipGraphDestroy(graph);
    }
  }
  (void)hipStreamDestroy(st);
  if (rc != BRI17_OK) {
    if (g) bri17_graph_destroy(g);
    return rc;
  }
  *out = g;
  return BRI17_OK;
}

}  // namespace

namespace {

// Built-in launch shapes of runtime-size grids, found by the autotuner on an
// MI355X (bri17_amd/tune_defaults.py writes tuned_defaults.inc): for the
// cubic fp32 grids of the serving workload without fixed-size kernels, the
// tuner's narrow pencil tiles (several workgroups per CU) and radix orders
// run 6-36% faster than the planner's defaults.
struct DefaultTuning {
  int dim, shape[3], dtype, plane_split, pencil_width;
  int radices[3][8];
};
const DefaultTuning kDefaults[] = {
#include "tuned_defaults.inc"
    {0, {0, 0, 0}, 0, 0, 0, {{0}, {0}, {0}}}};

bool default_tuning(int dim, const int* shape, int dtype, bri17_tuning* t) {
  for (const DefaultTuning& d : kDefaults) {
    if (d.dim != dim || d.dtype != dtype) continue;
    bool same = true;
    for (int i = 0; i < dim; i++) same &= d.shape[i] == shape[i];
    if (!same) continue;
    std::memset(t, 0, sizeof(*t));
    t->plane_split = d.plane_split;
    t->pencil_width = d.pencil_width;
    t->xcd_remap = -1;
    t->force_generic = 1;
    t->plane_prefetch = -1;
    std::memcpy(t->radices, d.radices, sizeof(t->radices));
    return true;
  }
  return false;
}

}  // namespace

namespace b17 {
// tuning registered for a grid, else the built-in one (used by
// bri17_operator_create)
bool cached_tuning(int dim, const int* shape, int dtype, bri17_tuning* t) {
  {
    std::lock_guard<std::mutex> g(g_cache_mtx);
    auto it = g_cache.find(cache_key(dim, shape, dtype));
    if (it != g_cache.end()) {
      *t = it->second;
      return true;
    }
  }
  return default_tuning(dim, shape, dtype, t);
}
}  // namespace b17

extern "C" {

int bri17_graph_create(bri17_operator* op, int mode, const void* in, void* out,
                       bri17_graph** graph) {
  if (!op || !in || !out || !graph || in == out || mode < 0 || mode > 3)
    return b17_err(BRI17_ERR_INVALID, "invalid argument");
  *graph = nullptr;
  auto call = [&](hipStream_t st) {
    switch (mode) {
      case 0: return bri17_operator_apply(op, in, out, st);
      case 1: return bri17_operator_solve(op, in, out, st);
      case 2: return bri17_operator_strain(op, in, out, st);
      default: return bri17_operator_eigenstress(op, in, out, st);
    }
  };
  // one eager run first: allocates the workspace and sets kernel attributes
  // outside the capture
  int rc = call(nullptr);
  if (rc) return rc;
  if (hipDeviceSynchronize() != hipSuccess) return b17_err(BRI17_ERR_HIP, "device error");
  {
    // the graph holds the workspace address from now on
    std::lock_guard<std::mutex> g(op->mtx);
    op->ws_pinned = true;
  }
  return capture(call, graph);
}

int bri17_batch_capture(bri17_batch* b, const bri17_batch_item* items, int n, bri17_graph** graph) {
  if (!b || !items || !graph) return b17_err(BRI17_ERR_INVALID, "invalid argument");
  *graph = nullptr;
  // eager run: uploads the work queue the graph will replay
  int rc = bri17_batch_run(b, items, n, nullptr);
  if (rc) return rc;
  int to = 0;
  rc = bri17_batch_wait(b, &to);
  if (rc) return rc;
  // the graph holds the context's tables and every member's workspace
  rc = b17_batch_pin(b, items, n);
  if (rc) return rc;
  rc = capture([&](hipStream_t st) { return b17_batch_relaunch(b, st); }, graph);
  if (rc) {
    b17_batch_unpin(b);
    return rc;
  }
  (*graph)->batch = b;
  return BRI17_OK;
}

int bri17_graph_launch(bri17_graph* g, void* stream) {
  if (!g || !g->exec) return b17_err(BRI17_ERR_INVALID, "null graph");
  if (hipGraphLaunch(g->exec, (hipStream_t)stream) != hipSuccess)
    return b17_err(BRI17_ERR_HIP, "hipGraphLaunch failed");
  return BRI17_OK;
}

int bri17_graph_destroy(bri17_graph* g) {
  if (!g) return BRI17_OK;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->batch) b17_batch_unpin(g->batch);
  delete g;
  return BRI17_OK;
}

int bri17_tuning_cache_set(int dim, const int* shape, int dtype, const bri17_tuning* t) {
  if ((dim != 2 && dim != 3) || !shape || !t) return b17_err(BRI17_ERR_INVALID, "invalid argument");
  std::lock_guard<std::mutex> g(g_cache_mtx);
  g_cache[cache_key(dim, shape, dtype)] = *t;
  return BRI17_OK;
}

int bri17_tuning_cache_clear(void) {
  std::lock_guard<std::mutex> g(g_cache_mtx);
  g_cache.clear();
  return BRI17_OK;
}

int bri17_autotune(int dim, const int* shape, const double* L, int dtype, int reps,
                   bri17_tuning* best, double* best_ms, int* tried) {
  if (!shape || !L || !best || reps < 1) return b17_err(BRI17_ERR_INVALID, "invalid argument");
  int64_t nn = 1;
  for (int d = 0; d < dim; d++) nn *= shape[d];
  const size_t esz = dtype == BRI17_F32 ? 4 : 8;
  const size_t bytes = (size_t)dim * nn * esz;
  void *in = nullptr, *out = nullptr;
  if (hipMalloc(&in, bytes) != hipSuccess || hipMalloc(&out, bytes) != hipSuccess) {
    if (in) (void)hipFree(in);
    return b17_err(BRI17_ERR_NOMEM, "hipMalloc(autotune buffers) failed");
  }
  (void)hipMemset(in, 0, bytes);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);

  auto time_one = [&](const bri17_tuning& t, double* ms) -> bool {
    bri17_operator* op = nullptr;
    if (bri17_operator_create_tuned(dim, shape, L, 1.0, 0.3, dtype, &t, &op) != BRI17_OK)
      return false;
    bool ok = true;
    for (int i = 0; i < 2 && ok; i++) ok = bri17_operator_apply(op, in, out, nullptr) == BRI17_OK;
    (void)hipEventRecord(e0, nullptr);
    for (int i = 0; i < reps && ok; i++) ok = bri17_operator_apply(op, in, out, nullptr) == BRI17_OK;
    (void)hipEventRecord(e1, nullptr);
    float t_ms = 0.f;
    if (ok) ok = hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&t_ms, e0, e1) == hipSuccess;
    bri17_operator_destroy(op);
    *ms = t_ms / reps;
    return ok;
  };

  bri17_tuning cur;
  std::memset(&cur, 0, sizeof(cur));
  cur.xcd_remap = -1;
  cur.plane_prefetch = -1;
  double cur_ms = 1e30;
  int count = 0;
  auto consider = [&](const bri17_tuning& t) {
    double ms;
    count++;
    if (time_one(t, &ms) && ms < cur_ms) {
      cur_ms = ms;
      cur = t;
    }
  };
  consider(cur);  // automatic choices (fixed-size kernels where they exist)
  // fixed-size path: persistent plane variants (bits 0, 1); all-components-in-
  // LDS pencil instead of the register-block one (bit 3) and its persistent
  // variant (bit 2)
  for (int pf : {0, 1, 2, 3, 8, 9, 12, 13}) {
    bri17_tuning t = cur;
    t.plane_prefetch = pf;
    consider(t);
  }
  // runtime-size path: split and tile width
  const bri17_tuning base = cur;
  bri17_tuning gbest = base;
  gbest.force_generic = 1;
  double gms = 1e30;
  for (int P : {1, 2, 3, 4, 6, 8, 9, 12, 16}) {
    if (dim == 3 && shape[1] % P) continue;
    if (dim == 2 && P != 1) continue;
    for (int W : {1, 2, 4, 8, 16, 32, 64}) {
      bri17_tuning t = base;
      t.force_generic = 1;
      t.plane_split = dim == 3 ? P : 0;
      t.pencil_width = W;
      double ms;
      count++;
      if (time_one(t, &ms) && ms < gms) {
        gms = ms;
        gbest = t;
      }
    }
  }
  // radix orders of the three roles on the best runtime-size shape
  const int lens[3] = {shape[dim - 1] % 2 ? shape[dim - 1] : shape[dim - 1] / 2,
                       dim == 3 ? shape[1] / std::max(1, gbest.plane_split) : 0, shape[0]};
  for (int role = 0; role < 3; role++) {
    const int n = lens[role];
    if (n < 4) continue;
    std::vector<std::vector<int>> orders;
    for (int r0 : {16, 8, 4, 2}) {
      for (int r3 : {9, 3}) {  // odd part in radix-9 or radix-3 passes
        std::vector<int> o;
        int m = n;
        while (m % r0 == 0 && (int)o.size() < 7) o.push_back(r0), m /= r0;
        for (int r : {8, 4, 2, r3, 3, 5, 7})
          while (m % r == 0 && (int)o.size() < 7) o.push_back(r), m /= r;
        if (m == 1 && !o.empty() &&
            std::find(orders.begin(), orders.end(), o) == orders.end())
          orders.push_back(o);
      }
    }
    const size_t nfwd = orders.size();
    for (size_t i = 0; i < nfwd; i++) {
      std::vector<int> o(orders[i].rbegin(), orders[i].rend());
      if (o != orders[i]) orders.push_back(o);
    }
    for (const auto& o : orders) {
      bri17_tuning t = gbest;
      std::memset(t.radices[role], 0, sizeof(t.radices[role]));
      for (size_t i = 0; i < o.size(); i++) t.radices[role][i] = o[i];
      double ms;
      count++;
      if (time_one(t, &ms) && ms < gms) {
        gms = ms;
        gbest = t;
      }
    }
  }
  if (gms < cur_ms) {
    cur = gbest;
    cur_ms = gms;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(in);
  (void)hipFree(out);
  if (cur_ms >= 1e30) return b17_err(BRI17_ERR_UNSUPPORTED, "no candidate launch shape ran");
  *best = cur;
  if (best_ms) *best_ms = cur_ms;
  if (tried) *tried = count;
  return BRI17_OK;
}

}  // extern "C"
