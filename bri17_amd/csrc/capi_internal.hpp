// Private definitions shared by the C ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <mutex>
#include <vector>

#include "bri17_amd.h"
#include "batch.hpp"
#include "plan.hpp"

struct bri17_operator {
  std::shared_ptr<b17::Plan> plan;
  double mu = 1.0, nu = 0.0;
  void* ws = nullptr;         // half-spectrum workspace (device)
  size_t ws_bytes = 0;
  bool ws_pinned = false;     // captured in a graph: never reallocated
  void* host_in = nullptr;    // staging buffers of the *_host entry points
  void* host_out = nullptr;
  size_t host_bytes = 0;
  std::mutex mtx;             // serialises use of the owned workspace
  std::mutex host_mtx;        // serialises the *_host entry points
  // per-kernel timing (bri17_operator_profile)
  bool profiling = false;
  std::vector<hipEvent_t> events;  // 4 per profiled call
  size_t ev_used = 0;
};

// status helpers: record the message for bri17_last_error and return code
int b17_finish(const b17::Status& s);
int b17_err(int code, const char* msg);
// shared plan of a grid (built on first use, kept while any user holds it);
// `tuning` NULL: the tuning registered for the grid, else the defaults
int b17_acquire_plan(int dim, const int* shape, const double* L, int dtype,
                     const bri17_tuning* tuning, std::shared_ptr<b17::Plan>& out);
// allocate (or grow) the operator's workspace; bytes 0: the plan's (DIM
// components).  Call with op->mtx held.
int b17_ensure_ws(bri17_operator* op, size_t bytes = 0);
// launch the last batch of a context again (tables already on the device)
int b17_batch_relaunch(bri17_batch* b, hipStream_t st);
// workgroup shape of a persistent batched launch over these plans: 256
// threads and 39 KiB of LDS unless an item needs more
void b17_batch_shape(const std::vector<const b17::Plan*>& plans, int& threads, size_t& lds);
// descriptor of one instance of a persistent batched launch whose
// workgroups have `lds` bytes of LDS (from b17_batch_shape)
b17::BatchInst b17_batch_inst(const b17::Plan& pl, double mu, double nu, int mode, const void* in,
                              void* out, void* ws, int threads, size_t lds);
// upload an instance table and launch the persistent kernel on `st` (waits
// for the context's previous launch first); threads / lds are the maxima
// over the instances' plans.  The caller serialises use of `b`.
int b17_batch_launch(bri17_batch* b, int dtype, const std::vector<b17::BatchInst>& insts,
                     int threads, size_t lds, hipStream_t st);
// device address of the control words [head, error, done...] of the last launch
int* b17_batch_ctl(bri17_batch* b);
// a graph of the context's last launch is created / destroyed: pins the
// member operators' workspaces, and refuses bri17_batch_run while any lives
int b17_batch_pin(bri17_batch* b, const bri17_batch_item* items, int n);
void b17_batch_unpin(bri17_batch* b);

namespace b17 {
// tuning registered with bri17_tuning_cache_set for a grid, if any
bool cached_tuning(int dim, const int* shape, int dtype, bri17_tuning* t);
}  // namespace b17
