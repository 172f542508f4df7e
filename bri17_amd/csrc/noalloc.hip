// The operator call that allocates nothing (bri17_operator_run_noalloc of
// include/bri17_amd.h): the same three-kernel pipeline as
// bri17_operator_apply / _solve / _strain / _eigenstress, with the
// half-spectrum workspace taken from a static device array instead of the
// operator's own (hipMalloc'd) workspace.  For callers that cannot allocate:
// code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc.  No reference counterpart.
//
// Every GPU the library runs this on holds its own copy of `scratch` (64 MiB)
// until the process ends.  In host code `scratch` names a host-side stand-in,
// not device memory: its device address on the current device comes from
// hipGetSymbolAddress, looked up once per device on the first call there and
// kept in `g_scratch`; only that address is ever handed to the kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "bri17_amd.h"
#include "capi_internal.hpp"
#include "plan.hpp"

static __device__ __align__(256) unsigned char scratch[BRI17_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mtx;
void* g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device

// Device address of the scratch array on the current device (looked up on
// the first call on that device).
int scratch_on_current_device(void** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0)
    return b17_err(BRI17_ERR_HIP, "no current HIP device");
  if (dev >= kMaxDevices)
    return b17_err(BRI17_ERR_UNSUPPORTED, "device index beyond the scratch table (64 devices)");
  std::lock_guard<std::mutex> g(g_scratch_mtx);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return b17_err(BRI17_ERR_HIP,
                     (std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)).c_str());
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return BRI17_OK;
}

}  // namespace

extern "C" int bri17_operator_run_noalloc(bri17_operator* op, int mode, const void* in, void* out,
                                          void* stream) {
  if (!op || !in || !out) return b17_err(BRI17_ERR_INVALID, "null argument");
  if (in == out) return b17_err(BRI17_ERR_INVALID, "input and output must not alias");
  if (mode < 0 || mode > 3) return b17_err(BRI17_ERR_INVALID, "mode must be 0, 1, 2 or 3");
  const b17::Plan& pl = *op->plan;
  // the workspace of the mode: the half spectrum of max(in, out) components
  const int comps = std::max(b17::mode_cin(pl.dim, mode), b17::mode_cout(pl.dim, mode));
  const size_t need = pl.ws_bytes / pl.dim * comps;
  if (need > BRI17_NOALLOC_SCRATCH_BYTES)
    return b17_err(BRI17_ERR_NOMEM,
                   ("this grid and mode need a workspace of " + std::to_string(need) +
                    " bytes, more than the static scratch of " +
                    std::to_string((size_t)BRI17_NOALLOC_SCRATCH_BYTES) +
                    " bytes: use bri17_operator_apply / _solve / _strain / _eigenstress")
                       .c_str());
  void* ws = nullptr;
  int rc = scratch_on_current_device(&ws);
  if (rc) return rc;
  // the forward plane kernel writes every workspace entry the pencil and the
  // inverse read, so nothing needs clearing between calls
  std::lock_guard<std::mutex> g(op->mtx);
  return b17_finish(b17::run_pipeline(pl, op->mu, op->nu, mode, in, out, ws, (hipStream_t)stream));
}
