// Private definitions shared by the C ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <mutex>
#include <vector>

#include "bri17_amd.h"
#include "plan.hpp"

struct bri17_operator {
  std::shared_ptr<b17::Plan> plan;
  double mu = 1.0, nu = 0.0;
  void* ws = nullptr;         // half-spectrum workspace (device)
  void* host_in = nullptr;    // staging buffers of the *_host entry points
  void* host_out = nullptr;
  std::mutex mtx;             // serialises use of the owned workspace
  // per-kernel timing (bri17_operator_profile)
  bool profiling = false;
  std::vector<hipEvent_t> events;  // 4 per profiled call
  size_t ev_used = 0;
};

// status helpers: record the message for bri17_last_error and return code
int b17_finish(const b17::Status& s);
int b17_err(int code, const char* msg);
// allocate the operator's workspace on first use
int b17_ensure_ws(bri17_operator* op);
// launch the last batch of a context again (tables already on the device)
int b17_batch_relaunch(bri17_batch* b, hipStream_t st);

namespace b17 {
// tuning registered with bri17_tuning_cache_set for a grid, if any
bool cached_tuning(int dim, const int* shape, int dtype, bri17_tuning* t);
}  // namespace b17
