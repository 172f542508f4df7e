// C ABI of bri17_amd (include/bri17_amd.h): grid helpers, the per-frequency
// Hooke functions on the host, and the device operator.
#include "bri17_amd.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstring>
#include <memory>
#include <mutex>
#include <map>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "capi_internal.hpp"
#include "plan.hpp"

using b17::Plan;
using b17::Status;

int b17_finish(const Status& s) {
  if (!s.ok()) b17::set_error(s.msg);
  return s.code;
}

int b17_err(int code, const char* msg) {
  b17::set_error(msg);
  return code;
}

int b17_acquire_plan(int dim, const int* shape, const double* L, int dtype,
                     const bri17_tuning* tuning, std::shared_ptr<b17::Plan>& out) {
  if (!shape || !L) return b17_err(BRI17_ERR_INVALID, "null argument");
  b17::Tuning tun;
  bri17_tuning cached;
  if (!tuning && (dim == 2 || dim == 3) && b17::cached_tuning(dim, shape, dtype, &cached))
    tuning = &cached;
  if (tuning) {
    tun.plane_split = tuning->plane_split;
    tun.pencil_width = tuning->pencil_width;
    if (tuning->lds_budget_kb > 0) tun.lds_budget = (size_t)tuning->lds_budget_kb * 1024;
    if (tuning->xcd_remap >= 0) tun.xcd_remap = tuning->xcd_remap;
    tun.force_generic = tuning->force_generic;
    if (tuning->plane_prefetch >= 0) tun.plane_prefetch = tuning->plane_prefetch;
    for (int r = 0; r < 3; r++)
      for (int i = 0; i < 8 && tuning->radices[r][i] > 0; i++)
        tun.radices[r].push_back(tuning->radices[r][i]);
  }
  // plans (device twiddle / modal tables + launch geometry) are shared by all
  // live users of the same device, grid, dtype and tuning
  int device = -1;  // without a device build_plan reports the error below
  if (hipGetDevice(&device) != hipSuccess) device = -1;
  std::ostringstream key;
  key << device << "|" << dim << "|" << dtype;
  for (int d = 0; d < dim && d < 3; d++) key << "|" << shape[d] << "|" << std::hexfloat << L[d];
  key << "|" << tun.plane_split << "|" << tun.pencil_width << "|" << tun.lds_budget << "|"
      << tun.xcd_remap << "|" << tun.force_generic << "|" << tun.plane_prefetch;
  for (int r = 0; r < 3; r++) {
    key << "|";
    for (int x : tun.radices[r]) key << x << ",";
  }
  static std::mutex cache_mtx;
  static std::map<std::string, std::weak_ptr<Plan>> cache;
  std::lock_guard<std::mutex> g(cache_mtx);
  auto it = cache.find(key.str());
  if (it != cache.end()) out = it->second.lock();
  if (!out) {
    int rc = b17_finish(b17::build_plan(dim, shape, L, dtype, tun, out));
    if (rc) return rc;
    cache[key.str()] = out;
  }
  return BRI17_OK;
}

namespace {

int finish(const Status& s) { return b17_finish(s); }
int err(int code, const char* msg) { return b17_err(code, msg); }

bool check_grid(int dim, const int* shape) {
  if (dim != 2 && dim != 3) return false;
  if (!shape) return false;
  for (int d = 0; d < dim; d++)
    if (shape[d] < 1) return false;
  return true;
}

constexpr double kPi = 3.141592653589793238462643383279502884;

// bri17.hpp:212-236.  Trigonometry in double (as the reference's unqualified
// cos/sin resolve to the double overloads, bri17.hpp:220-224).
template <typename T>
void modal_B(int D, const int* shape, const T* L, const int* k, std::complex<double>* B) {
  double c[3] = {0, 0, 0}, s[3] = {0, 0, 0}, sum = 0.0;
  for (int i = 0; i < D; i++) {
    double a = kPi * k[i] / shape[i];
    sum += a;
    c[i] = std::cos(a);
    s[i] = std::sin(a) * shape[i] / (double)L[i];
  }
  std::complex<double> pre{-2.0 * std::sin(sum), 2.0 * std::cos(sum)};
  if (D == 2) {
    B[0] = pre * (s[0] * c[1]);
    B[1] = pre * (c[0] * s[1]);
  } else {
    B[0] = pre * (s[0] * c[1] * c[2]);
    B[1] = pre * (c[0] * s[1] * c[2]);
    B[2] = pre * (c[0] * c[1] * s[2]);
  }
}

// bri17.hpp:247-292 with the cancellation-free half-angle phi.
template <typename T>
void modal_K(int D, const int* shape, const T* L, double mu, double nu, const int* k,
             double* K) {
  double phi[3] = {0, 0, 0}, chi[3] = {0, 0, 0}, psi[3] = {0, 0, 0};
  for (int i = 0; i < D; i++) {
    double h = (double)L[i] / shape[i];
    double sh = std::sin(kPi * k[i] / shape[i]);
    double b = 2.0 * kPi * k[i] / shape[i];
    phi[i] = 4.0 * sh * sh / (h * h);
    chi[i] = 1.0 - 2.0 * sh * sh / 3.0;
    psi[i] = std::sin(b) / h;
  }
  double sc = mu / (1.0 - 2.0 * nu);
  if (D == 2) {
    double H0 = phi[0] * chi[1], H1 = chi[0] * phi[1], dg = mu * (H0 + H1);
    K[0] = sc * H0 + dg;
    K[1] = K[2] = sc * psi[0] * psi[1];
    K[3] = sc * H1 + dg;
  } else {
    double H0 = phi[0] * chi[1] * chi[2], H1 = chi[0] * phi[1] * chi[2],
           H2 = chi[0] * chi[1] * phi[2];
    double dg = mu * (H0 + H1 + H2);
    K[0] = sc * H0 + dg;
    K[4] = sc * H1 + dg;
    K[8] = sc * H2 + dg;
    K[1] = K[3] = sc * psi[0] * psi[1] * chi[2];
    K[2] = K[6] = sc * psi[0] * chi[1] * psi[2];
    K[5] = K[7] = sc * chi[0] * psi[1] * psi[2];
  }
}

template <typename T>
int strain_displacement(int dim, const int* shape, const T* L, const int* k, T* B) {
  if (!check_grid(dim, shape) || !L || !k || !B) return err(BRI17_ERR_INVALID, "invalid argument");
  std::complex<double> b[3];
  modal_B(dim, shape, L, k, b);
  for (int i = 0; i < dim; i++) {
    B[2 * i] = (T)b[i].real();
    B[2 * i + 1] = (T)b[i].imag();
  }
  return BRI17_OK;
}

template <typename T>
int stiffness(int dim, const int* shape, const T* L, double mu, double nu, const int* k, T* K) {
  if (!check_grid(dim, shape) || !L || !k || !K) return err(BRI17_ERR_INVALID, "invalid argument");
  double Kr[9];
  modal_K(dim, shape, L, mu, nu, k, Kr);
  for (int i = 0; i < dim * dim; i++) {
    K[2 * i] = (T)Kr[i];
    K[2 * i + 1] = (T)0;
  }
  return BRI17_OK;
}

// bri17.hpp:308-355: rhs = tau_mat conj(B), u = K^-1 rhs (Cholesky, as
// K.llt().solve), eta = sym(B u^T) in Voigt-Mandel order; zero at k = 0.
template <typename T>
int eigenstress(int dim, const int* shape, const T* L, double mu, double nu, const int* k,
                const T* tau, T* eta) {
  if (!check_grid(dim, shape) || !L || !k || !tau || !eta)
    return err(BRI17_ERR_INVALID, "invalid argument");
  using C = std::complex<double>;
  const int D = dim, S = D == 2 ? 3 : 6;
  const double r2 = std::sqrt(2.0);
  bool null_frequency = true;
  for (int d = 0; d < D; d++) null_frequency &= (k[d] == 0);
  if (null_frequency) {
    for (int i = 0; i < 2 * S; i++) eta[i] = (T)0;
    return BRI17_OK;
  }
  C B[3], t[6];
  double K[9];
  modal_B(D, shape, L, k, B);
  modal_K(D, shape, L, mu, nu, k, K);
  for (int i = 0; i < S; i++) t[i] = C((double)tau[2 * i], (double)tau[2 * i + 1]);
  C Tm[3][3];
  if (D == 2) {
    Tm[0][0] = t[0];
    Tm[1][1] = t[1];
    Tm[0][1] = Tm[1][0] = t[2] / r2;
  } else {
    Tm[0][0] = t[0];
    Tm[1][1] = t[1];
    Tm[2][2] = t[2];
    Tm[1][2] = Tm[2][1] = t[3] / r2;
    Tm[0][2] = Tm[2][0] = t[4] / r2;
    Tm[0][1] = Tm[1][0] = t[5] / r2;
  }
  C rhs[3];
  for (int i = 0; i < D; i++) {
    rhs[i] = 0.0;
    for (int j = 0; j < D; j++) rhs[i] += Tm[i][j] * std::conj(B[j]);
  }
  double Lc[3][3] = {{0}};
  for (int j = 0; j < D; j++) {
    double s = K[j * D + j];
    for (int p = 0; p < j; p++) s -= Lc[j][p] * Lc[j][p];
    Lc[j][j] = std::sqrt(s);
    for (int i = j + 1; i < D; i++) {
      double v = K[i * D + j];
      for (int p = 0; p < j; p++) v -= Lc[i][p] * Lc[j][p];
      Lc[i][j] = v / Lc[j][j];
    }
  }
  C y[3], u[3];
  for (int i = 0; i < D; i++) {
    C v = rhs[i];
    for (int p = 0; p < i; p++) v -= Lc[i][p] * y[p];
    y[i] = v / Lc[i][i];
  }
  for (int i = D - 1; i >= 0; i--) {
    C v = y[i];
    for (int p = i + 1; p < D; p++) v -= Lc[p][i] * u[p];
    u[i] = v / Lc[i][i];
  }
  auto E = [&](int a, int b) { return 0.5 * (B[a] * u[b] + u[a] * B[b]); };
  C o[6];
  if (D == 2) {
    o[0] = E(0, 0);
    o[1] = E(1, 1);
    o[2] = r2 * E(0, 1);
  } else {
    o[0] = E(0, 0);
    o[1] = E(1, 1);
    o[2] = E(2, 2);
    o[3] = r2 * E(1, 2);
    o[4] = r2 * E(2, 0);
    o[5] = r2 * E(0, 1);
  }
  for (int i = 0; i < S; i++) {
    eta[2 * i] = (T)o[i].real();
    eta[2 * i + 1] = (T)o[i].imag();
  }
  return BRI17_OK;
}


// workspace bytes of a mode: the half spectrum of max(in, out) components
size_t mode_ws_bytes(const b17::Plan& pl, int mode) {
  const int c = std::max(b17::mode_cin(pl.dim, mode), b17::mode_cout(pl.dim, mode));
  return pl.ws_bytes / pl.dim * c;
}

int run(bri17_operator* op, int mode, const void* in, void* out, void* stream) {
  if (!op || !in || !out) return err(BRI17_ERR_INVALID, "null argument");
  if (in == out) return err(BRI17_ERR_INVALID, "input and output must not alias");
  std::lock_guard<std::mutex> g(op->mtx);
  int rc = b17_ensure_ws(op, mode_ws_bytes(*op->plan, mode));
  if (rc) return rc;
  hipEvent_t* ev = nullptr;
  if (op->profiling) {
    if (op->ev_used + 4 > op->events.size()) {
      for (int i = 0; i < 64; i++) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return err(BRI17_ERR_HIP, "hipEventCreate failed");
        op->events.push_back(e);
      }
    }
    ev = op->events.data() + op->ev_used;
    op->ev_used += 4;
  }
  return finish(b17::run_pipeline(*op->plan, op->mu, op->nu, mode, in, out, op->ws,
                                  (hipStream_t)stream, ev));
}

int run_host(bri17_operator* op, int mode, const void* in, void* out) {
  if (!op || !in || !out) return err(BRI17_ERR_INVALID, "null argument");
  const b17::Plan& pl = *op->plan;
  const size_t field = (size_t)pl.nnodes * pl.esz;
  const size_t in_bytes = b17::mode_cin(pl.dim, mode) * field;
  const size_t out_bytes = b17::mode_cout(pl.dim, mode) * field;
  const size_t need = std::max(in_bytes, out_bytes);
  std::lock_guard<std::mutex> hg(op->host_mtx);
  if (op->host_bytes < need) {
    // synchronous entry points: nothing is in flight on the old buffers
    if (op->host_in) (void)hipFree(op->host_in);
    if (op->host_out) (void)hipFree(op->host_out);
    op->host_in = op->host_out = nullptr;
    op->host_bytes = 0;
    if (hipMalloc(&op->host_in, need) != hipSuccess || hipMalloc(&op->host_out, need) != hipSuccess)
      return err(BRI17_ERR_NOMEM, "hipMalloc(host staging) failed");
    op->host_bytes = need;
  }
  if (hipMemcpy(op->host_in, in, in_bytes, hipMemcpyHostToDevice) != hipSuccess)
    return err(BRI17_ERR_HIP, "hipMemcpy to device failed");
  int rc = run(op, mode, op->host_in, op->host_out, nullptr);
  if (rc) return rc;
  hipError_t e = hipMemcpy(out, op->host_out, out_bytes, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return err(BRI17_ERR_HIP, hipGetErrorString(e));
  return BRI17_OK;
}

}  // namespace

int b17_ensure_ws(bri17_operator* op, size_t bytes) {
  if (bytes == 0) bytes = op->plan->ws_bytes;
  if (op->ws && op->ws_bytes >= bytes) return BRI17_OK;
  if (op->ws && op->ws_pinned)
    return err(BRI17_ERR_STATE,
               "the workspace is captured in a graph and cannot grow: run this mode once "
               "before capturing");
  if (op->ws) {
    // growing (first strain / eigenstress call): earlier launches on any
    // stream may still use the old workspace
    if (hipDeviceSynchronize() != hipSuccess) return err(BRI17_ERR_HIP, "hipDeviceSynchronize failed");
    (void)hipFree(op->ws);
    op->ws = nullptr;
    op->ws_bytes = 0;
  }
  hipError_t e = hipMalloc(&op->ws, bytes);
  if (e != hipSuccess) {
    op->ws = nullptr;
    return err(BRI17_ERR_NOMEM, "hipMalloc(workspace) failed");
  }
  op->ws_bytes = bytes;
  return BRI17_OK;
}

extern "C" {

const char* bri17_last_error(void) { return b17::get_error().c_str(); }
const char* bri17_version(void) { return "0.1.0+gfx950"; }

int bri17_grid_node_at(int dim, const int* shape, const int* idx, int* node) {
  if (!check_grid(dim, shape) || !idx || !node) return err(BRI17_ERR_INVALID, "invalid argument");
  int n = 0;
  for (int d = 0; d < dim; d++) n = n * shape[d] + idx[d];
  *node = n;
  return BRI17_OK;
}

int bri17_grid_cell_nodes(int dim, const int* shape, int cell, int* nodes) {
  if (!check_grid(dim, shape) || !nodes) return err(BRI17_ERR_INVALID, "invalid argument");
  int idx[3];
  int c = cell;
  for (int d = dim - 1; d >= 0; d--) {
    idx[d] = c % shape[d];
    c /= shape[d];
  }
  for (int n = 0; n < (1 << dim); n++) {
    int node = 0;
    for (int d = 0; d < dim; d++) {
      int i = idx[d] + ((n >> (dim - 1 - d)) & 1);
      if (i == shape[d]) i = 0;
      node = node * shape[d] + i;
    }
    nodes[n] = node;
  }
  return BRI17_OK;
}

int bri17_modal_strain_displacement_f64(int dim, const int* shape, const double* L,
                                        const int* k, double* B) {
  return strain_displacement<double>(dim, shape, L, k, B);
}
int bri17_modal_strain_displacement_f32(int dim, const int* shape, const float* L,
                                        const int* k, float* B) {
  return strain_displacement<float>(dim, shape, L, k, B);
}
int bri17_modal_stiffness_f64(int dim, const int* shape, const double* L, double mu, double nu,
                              const int* k, double* K) {
  return stiffness<double>(dim, shape, L, mu, nu, k, K);
}
int bri17_modal_stiffness_f32(int dim, const int* shape, const float* L, float mu, float nu,
                              const int* k, float* K) {
  return stiffness<float>(dim, shape, L, mu, nu, k, K);
}
int bri17_modal_eigenstress_to_opposite_strain_f64(int dim, const int* shape, const double* L,
                                                   double mu, double nu, const int* k,
                                                   const double* tau, double* eta) {
  return eigenstress<double>(dim, shape, L, mu, nu, k, tau, eta);
}
int bri17_modal_eigenstress_to_opposite_strain_f32(int dim, const int* shape, const float* L,
                                                   float mu, float nu, const int* k,
                                                   const float* tau, float* eta) {
  return eigenstress<float>(dim, shape, L, mu, nu, k, tau, eta);
}

int bri17_operator_create(int dim, const int* shape, const double* L, double mu, double nu,
                          int dtype, bri17_operator** out) {
  return bri17_operator_create_tuned(dim, shape, L, mu, nu, dtype, nullptr, out);
}

int bri17_operator_create_tuned(int dim, const int* shape, const double* L, double mu,
                                double nu, int dtype, const bri17_tuning* tuning,
                                bri17_operator** out) {
  if (!out || !shape || !L) return err(BRI17_ERR_INVALID, "null argument");
  *out = nullptr;
  std::shared_ptr<Plan> pl;
  int rc = b17_acquire_plan(dim, shape, L, dtype, tuning, pl);
  if (rc) return rc;
  auto* op = new (std::nothrow) bri17_operator;
  if (!op) return err(BRI17_ERR_NOMEM, "out of host memory");
  op->plan = pl;
  op->mu = mu;
  op->nu = nu;
  *out = op;
  return BRI17_OK;
}

int bri17_operator_describe(const bri17_operator* op, char* buf, size_t len) {
  if (!op || !buf || len == 0) return err(BRI17_ERR_INVALID, "null argument");
  std::string s = b17::describe_plan(*op->plan);
  size_t n = std::min(len - 1, s.size());
  std::memcpy(buf, s.data(), n);
  buf[n] = '\0';
  return BRI17_OK;
}

int bri17_operator_destroy(bri17_operator* op) {
  if (!op) return BRI17_OK;
  for (hipEvent_t e : op->events) (void)hipEventDestroy(e);
  if (op->ws) (void)hipFree(op->ws);
  if (op->host_in) (void)hipFree(op->host_in);
  if (op->host_out) (void)hipFree(op->host_out);
  delete op;
  return BRI17_OK;
}

int bri17_operator_set_material(bri17_operator* op, double mu, double nu) {
  if (!op) return err(BRI17_ERR_INVALID, "null operator");
  std::lock_guard<std::mutex> g(op->mtx);
  op->mu = mu;
  op->nu = nu;
  return BRI17_OK;
}

int bri17_operator_info(const bri17_operator* op, int64_t* num_dofs, size_t* workspace_bytes) {
  if (!op) return err(BRI17_ERR_INVALID, "null operator");
  if (num_dofs) *num_dofs = (int64_t)op->plan->dim * op->plan->nnodes;
  if (workspace_bytes) *workspace_bytes = op->plan->ws_bytes;
  return BRI17_OK;
}

int bri17_operator_profile(bri17_operator* op, int enable) {
  if (!op) return err(BRI17_ERR_INVALID, "null operator");
  std::lock_guard<std::mutex> g(op->mtx);
  op->profiling = enable != 0;
  op->ev_used = 0;
  return BRI17_OK;
}

int bri17_operator_profile_read(bri17_operator* op, int* calls, double* ms) {
  if (!op || !calls || !ms) return err(BRI17_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> g(op->mtx);
  ms[0] = ms[1] = ms[2] = 0.0;
  *calls = (int)(op->ev_used / 4);
  for (size_t c = 0; c < op->ev_used; c += 4) {
    if (hipEventSynchronize(op->events[c + 3]) != hipSuccess)
      return err(BRI17_ERR_HIP, "hipEventSynchronize failed");
    for (int i = 0; i < 3; i++) {
      float t = 0.f;
      if (hipEventElapsedTime(&t, op->events[c + i], op->events[c + i + 1]) != hipSuccess)
        return err(BRI17_ERR_HIP, "hipEventElapsedTime failed");
      ms[i] += t;
    }
  }
  op->ev_used = 0;
  return BRI17_OK;
}

int bri17_operator_apply(bri17_operator* op, const void* u, void* Ku, void* stream) {
  return run(op, 0, u, Ku, stream);
}
int bri17_operator_solve(bri17_operator* op, const void* f, void* u, void* stream) {
  return run(op, 1, f, u, stream);
}
int bri17_operator_strain(bri17_operator* op, const void* u, void* eps, void* stream) {
  return run(op, 2, u, eps, stream);
}
int bri17_operator_eigenstress(bri17_operator* op, const void* tau, void* eta, void* stream) {
  return run(op, 3, tau, eta, stream);
}
int bri17_operator_strain_host(bri17_operator* op, const void* u, void* eps) {
  return run_host(op, 2, u, eps);
}
int bri17_operator_eigenstress_host(bri17_operator* op, const void* tau, void* eta) {
  return run_host(op, 3, tau, eta);
}
int bri17_operator_apply_host(bri17_operator* op, const void* u, void* Ku) {
  return run_host(op, 0, u, Ku);
}
int bri17_operator_solve_host(bri17_operator* op, const void* f, void* u) {
  return run_host(op, 1, f, u);
}

}  // extern "C"
