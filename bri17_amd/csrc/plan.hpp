// Host-side planning: radix factorisation, launch geometry, device-resident
// per-axis twiddle and modal-coefficient tables.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "kernels.hpp"

namespace b17 {

// thread-local last-error string of the C ABI
void set_error(const std::string& msg);
const std::string& get_error();

struct Status {
  int code = 0;
  std::string msg;
  bool ok() const { return code == 0; }
};

// Tunables of one plan (chosen by defaults or by the autotuner).
struct Tuning {
  int plane_split = 0;       // middle-axis split P (3D); 0 = automatic
  int pencil_width = 0;      // pencil tile width W; 0 = automatic
  size_t lds_budget = 160 * 1024;
  int xcd_remap = -1;           // -1: off for the fixed-size kernels, on otherwise
  int force_generic = 0;        // 1: never use the fixed-size kernels
  int plane_prefetch = 3;       // fixed-size kernels persistent: bit0 fwd, bit1 inv, bit2 pencil
  std::vector<int> radices[3];  // forced radix orders per FFT length role
};

// Radix order (largest first) for a Stockham transform of length n, or empty
// (any n >= 1: prime factors above 13 take one direct-DFT pass each).
std::vector<int> default_radices(int n);
// want_wave: the stage may run wave-local passes (FftDesc::wave).
bool make_fft_desc(int n, int tws, const std::vector<int>& radices, FftDesc& d,
                   std::string* why, bool want_wave = false);

struct Plan {
  int dim = 0, dtype = 0;
  int shape[3] = {1, 1, 1};
  double L[3] = {1, 1, 1};
  size_t esz = 0;  // sizeof(T)
  int64_t nnodes = 0;
  double cell_volume = 1.0;

  // device tables (owned)
  void* tw[3] = {nullptr, nullptr, nullptr};
  void* coef[3] = {nullptr, nullptr, nullptr};
  void* scoef[3] = {nullptr, nullptr, nullptr};  // strain-displacement factors

  PlaneParams plane{};
  int plane_blocks = 0, plane_threads = 0;
  size_t plane_lds = 0;
  PencilParams pencil{};
  int pencil_threads = 0;
  size_t pencil_lds = 0;
  size_t ws_bytes = 0;
  bool use_static = false;  // fixed-size kernels (kernels_static.hip)
  Tuning tuning;

  ~Plan();
};

Status build_plan(int dim, const int* shape, const double* L, int dtype,
                  const Tuning& tun, std::shared_ptr<Plan>& out);

// Work items of a plan inside the persistent batch kernel: the same
// arithmetic as the plan (radix orders, middle-axis split), with the 2D
// plane row group narrowed and the pencil tile the widest (up to 64
// columns) so that every item fits `lds` bytes, and 3D plane items of split
// 1 grouped by consecutive planes up to what `threads` transform in one
// round.  false: a single row or column does not fit.
bool batch_geometry(const Plan& pl, size_t lds, int threads, PlaneParams& pp, int& plane_blocks,
                    PencilParams& pc);
// Least LDS (bytes) and threads with which every item of the plan can run.
void batch_minimum(const Plan& pl, size_t& lds, int& threads);

// JSON summary of a plan's geometry (for logs and benchmarks).
std::string describe_plan(const Plan& pl);

// Launch the three-kernel pipeline.  mode 0: K u, mode 1: K^-1 f.
// Pencil parameters of a plan for one call: mode (0 apply, 1 solve) and the
// material constants with the mode's scale folded in.
PencilParams pencil_for(const Plan& pl, double mu, double nu, int mode);

// Launch geometry of one mode.  Modes 0 and 1 use the plan's own geometry;
// modes 2 (strain, DIM -> SYM components) and 3 (eigenstress, SYM -> SYM)
// run the runtime-size kernels over their component counts, with the pencil
// tile narrowed to hold max(in, out) components.
struct ModeGeometry {
  PlaneParams fwd, inv;
  int fwd_blocks = 0, inv_blocks = 0;
  int plane_threads = 0;
  size_t plane_lds = 0;
  PencilParams pc{};
  int pencil_threads = 0;
  size_t pencil_lds = 0;
  size_t ws_bytes = 0;
};
Status mode_geometry(const Plan& pl, double mu, double nu, int mode, ModeGeometry& g);

// `ev`, if not null, holds 4 events recorded before the first kernel and after
// each of the three kernels (per-kernel timing).
Status run_pipeline(const Plan& pl, double mu, double nu, int mode, const void* in,
                    void* out, void* ws, hipStream_t st, hipEvent_t* ev = nullptr);

}  // namespace b17
