// Host-side checks of the C ABI run under AddressSanitizer and
// UndefinedBehaviorSanitizer (host code only; the device code of the library
// is built as usual).  Exercises, on the CPU:
//   - the grid helpers (bri17.hpp:78-158) and the per-frequency Hooke
//     functions (bri17.hpp:212-236, 247-292, 308-355) over many k, both
//     dtypes and dimensions, against closed-form invariants (K^(0) = 0,
//     K^ symmetric and K^(-k) = K^(k), B^ of modulus 2 |sin a / h| products);
//   - argument validation (null pointers, bad dimensions, aliasing), which
//     must return an error code and set bri17_last_error;
//   - the planning and error paths that run before any device work: the
//     tuning cache, operator / batch / server creation.  Without a GPU these
//     fail with BRI17_ERR_HIP (or succeed on a GPU host); either way no
//     memory error may occur.
// With the argument "device" it also runs operators, a batch and the
// request server (several submitting threads) on the GPU.
// Exit status 0 on success.
#include <atomic>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include <hip/hip_runtime.h>

#include "bri17_amd.h"

static int g_fail = 0;
#define CHECK(cond)                                                          \
  do {                                                                       \
    if (!(cond)) {                                                           \
      std::printf("FAIL %s:%d: %s (%s)\n", __FILE__, __LINE__, #cond,        \
                  bri17_last_error());                                       \
      g_fail++;                                                              \
    }                                                                        \
  } while (0)

template <typename T>
static void per_frequency(int dim) {
  const int shape[3] = {6, 9, 10};
  const T L[3] = {T(1.1), T(0.7), T(1.9)};
  const T mu = T(1.3), nu = T(0.27);
  const int S = dim == 2 ? 3 : 6;
  std::mt19937 rng(7);
  for (int it = 0; it < 400; it++) {
    int k[3], mk[3];
    for (int d = 0; d < 3; d++) {
      k[d] = (int)(rng() % (3 * shape[d])) - shape[d];
      mk[d] = -k[d];
    }
    if (it == 0) k[0] = k[1] = k[2] = mk[0] = mk[1] = mk[2] = 0;
    T K[18], Km[18], B[6], tau[12], eta[12];
    int rc;
    if constexpr (sizeof(T) == 8) {
      rc = bri17_modal_stiffness_f64(dim, shape, L, mu, nu, k, K);
      rc |= bri17_modal_stiffness_f64(dim, shape, L, mu, nu, mk, Km);
      rc |= bri17_modal_strain_displacement_f64(dim, shape, L, k, B);
    } else {
      rc = bri17_modal_stiffness_f32(dim, shape, L, mu, nu, k, K);
      rc |= bri17_modal_stiffness_f32(dim, shape, L, mu, nu, mk, Km);
      rc |= bri17_modal_strain_displacement_f32(dim, shape, L, k, B);
    }
    CHECK(rc == BRI17_OK);
    const T tol = sizeof(T) == 8 ? T(1e-12) : T(1e-5);
    T kmax = 0;
    for (int i = 0; i < 2 * dim * dim; i++) kmax = std::max(kmax, std::fabs(K[i]));
    for (int i = 0; i < dim; i++)
      for (int j = 0; j < dim; j++) {
        CHECK(K[2 * (i * dim + j) + 1] == T(0));                          // real block
        CHECK(std::fabs(K[2 * (i * dim + j)] - K[2 * (j * dim + i)]) <= tol * kmax);  // symmetric
        CHECK(std::fabs(K[2 * (i * dim + j)] - Km[2 * (i * dim + j)]) <= tol * kmax);  // even in k
      }
    if (it == 0) CHECK(kmax == T(0));  // K^(0) = 0
    for (int i = 0; i < S; i++) {
      tau[2 * i] = T((int)(rng() % 2001) - 1000) / T(1000);
      tau[2 * i + 1] = T((int)(rng() % 2001) - 1000) / T(1000);
    }
    if constexpr (sizeof(T) == 8)
      rc = bri17_modal_eigenstress_to_opposite_strain_f64(dim, shape, L, mu, nu, k, tau, eta);
    else
      rc = bri17_modal_eigenstress_to_opposite_strain_f32(dim, shape, L, mu, nu, k, tau, eta);
    CHECK(rc == BRI17_OK);
    for (int i = 0; i < 2 * S; i++) CHECK(std::isfinite((double)eta[i]));
    if (it == 0)
      for (int i = 0; i < 2 * S; i++) CHECK(eta[i] == T(0));  // null frequency (bri17.hpp:322-339)
  }
}

static void grid_helpers() {
  const int shape[3] = {3, 4, 5};
  int nodes[8], node = -1;
  for (int dim = 2; dim <= 3; dim++) {
    int n = 1;
    for (int d = 0; d < dim; d++) n *= shape[d];
    for (int cell = 0; cell < n; cell++) {
      CHECK(bri17_grid_cell_nodes(dim, shape, cell, nodes) == BRI17_OK);
      for (int c = 0; c < (1 << dim); c++) CHECK(nodes[c] >= 0 && nodes[c] < n);
      CHECK(nodes[0] == cell);
    }
    const int idx[3] = {2, 3, 4};
    CHECK(bri17_grid_node_at(dim, shape, idx, &node) == BRI17_OK);
    CHECK(node == (dim == 2 ? 2 * 4 + 3 : (2 * 4 + 3) * 5 + 4));
  }
}

static void validation() {
  const int shape[3] = {4, 4, 4}, bad[3] = {4, 0, 4}, k[3] = {1, 2, 3};
  const double L[3] = {1, 1, 1};
  double K[18], B[6], tau[12] = {0}, eta[12];
  int nodes[8];
  CHECK(bri17_grid_cell_nodes(4, shape, 0, nodes) == BRI17_ERR_INVALID);
  CHECK(std::strlen(bri17_last_error()) > 0);
  CHECK(bri17_grid_cell_nodes(3, bad, 0, nodes) == BRI17_ERR_INVALID);
  CHECK(bri17_grid_cell_nodes(3, shape, 0, nullptr) == BRI17_ERR_INVALID);
  CHECK(bri17_grid_node_at(3, nullptr, k, nodes) == BRI17_ERR_INVALID);
  CHECK(bri17_modal_stiffness_f64(1, shape, L, 1, 0.3, k, K) == BRI17_ERR_INVALID);
  CHECK(bri17_modal_stiffness_f64(3, shape, nullptr, 1, 0.3, k, K) == BRI17_ERR_INVALID);
  CHECK(bri17_modal_stiffness_f64(3, shape, L, 1, 0.3, nullptr, K) == BRI17_ERR_INVALID);
  CHECK(bri17_modal_strain_displacement_f64(3, shape, L, k, nullptr) == BRI17_ERR_INVALID);
  CHECK(bri17_modal_strain_displacement_f64(3, bad, L, k, B) == BRI17_ERR_INVALID);
  CHECK(bri17_modal_eigenstress_to_opposite_strain_f64(3, shape, L, 1, 0.3, k, nullptr, eta) ==
        BRI17_ERR_INVALID);
  CHECK(bri17_modal_eigenstress_to_opposite_strain_f64(2, shape, L, 1, 0.3, k, tau, nullptr) ==
        BRI17_ERR_INVALID);
  // operator creation: bad grids are refused before any device call
  bri17_operator* op = nullptr;
  const int prime[3] = {4, 4, 65537};  // rows beyond a workgroup's LDS and threads
  CHECK(bri17_operator_create(4, shape, L, 1, 0.3, BRI17_F32, &op) == BRI17_ERR_INVALID);
  CHECK(bri17_operator_create(3, prime, L, 1, 0.3, BRI17_F32, &op) != BRI17_OK);
  CHECK(bri17_operator_create(3, shape, L, 1, 0.3, 7, &op) != BRI17_OK);
  CHECK(bri17_operator_create(3, shape, L, 1, 0.3, BRI17_F32, nullptr) == BRI17_ERR_INVALID);
  CHECK(op == nullptr);
  CHECK(bri17_operator_apply(nullptr, K, B, nullptr) == BRI17_ERR_INVALID);
  CHECK(bri17_operator_run_noalloc(nullptr, 0, K, B, nullptr) == BRI17_ERR_INVALID);
  CHECK(bri17_operator_destroy(nullptr) == BRI17_OK);
  // tuning cache
  bri17_tuning t;
  std::memset(&t, 0, sizeof(t));
  t.xcd_remap = -1;
  t.plane_prefetch = -1;
  t.radices[0][0] = 4;
  CHECK(bri17_tuning_cache_set(3, shape, BRI17_F32, &t) == BRI17_OK);
  CHECK(bri17_tuning_cache_set(5, shape, BRI17_F32, &t) == BRI17_ERR_INVALID);
  CHECK(bri17_tuning_cache_clear() == BRI17_OK);
  // batches and the server
  bri17_batch* b = nullptr;
  CHECK(bri17_batch_create(nullptr) == BRI17_ERR_INVALID);
  if (bri17_batch_create(&b) == BRI17_OK) {
    CHECK(bri17_batch_run(b, nullptr, 1, nullptr) == BRI17_ERR_INVALID);
    bri17_batch_item it{nullptr, 0, K, B};
    CHECK(bri17_batch_run(b, &it, 1, nullptr) == BRI17_ERR_INVALID);
    CHECK(bri17_batch_set_spin_limit(b, -1) == BRI17_ERR_INVALID);
    CHECK(bri17_batch_destroy(b) == BRI17_OK);
  }
  bri17_server* srv = nullptr;
  CHECK(bri17_server_create(nullptr, nullptr) == BRI17_ERR_INVALID);
  bri17_server_config cfg;
  std::memset(&cfg, 0, sizeof(cfg));
  cfg.hbm_budget_bytes = 1 << 20;
  if (bri17_server_create(&cfg, &srv) == BRI17_OK) {
    bri17_request r;
    std::memset(&r, 0, sizeof(r));
    r.dim = 3;
    r.shape[0] = r.shape[1] = r.shape[2] = 64;
    r.L[0] = r.L[1] = r.L[2] = 1.0;
    r.mu = 1.0;
    r.nu = 0.3;
    r.in = K;
    r.out = B;
    int64_t id = -1;
    CHECK(bri17_server_submit(srv, &r, &id) == BRI17_ERR_NOMEM);  // footprint above 1 MiB
    r.nu = 0.5;
    CHECK(bri17_server_submit(srv, &r, &id) == BRI17_ERR_INVALID);
    CHECK(bri17_server_destroy(srv) == BRI17_OK);
  }
}

// GPU section: operators of both dtypes, a batch, and the server fed by
// several threads.  Results are compared with each other (same kernels).
static void device_work() {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    std::printf("no device\n");
    g_fail++;
    return;
  }
  const int shapes[3][3] = {{32, 48, 40}, {64, 64, 64}, {36, 27, 30}};
  const double L[3] = {1, 1.2, 0.8};
  std::vector<bri17_operator*> ops;
  std::vector<void*> bufs;
  const size_t maxn = 3 * 64 * 64 * 64;
  std::vector<float> h(maxn), a(maxn), c(maxn);
  for (size_t i = 0; i < maxn; i++) h[i] = (float)std::sin(0.37 * (double)i);
  for (int g = 0; g < 3; g++) {
    bri17_operator* op = nullptr;
    CHECK(bri17_operator_create(3, shapes[g], L, 1.1, 0.3, BRI17_F32, &op) == BRI17_OK);
    if (!op) return;
    ops.push_back(op);
    const size_t n = (size_t)3 * shapes[g][0] * shapes[g][1] * shapes[g][2];
    void *din = nullptr, *dout = nullptr;
    CHECK(hipMalloc(&din, n * 4) == hipSuccess && hipMalloc(&dout, n * 4) == hipSuccess);
    CHECK(hipMemcpy(din, h.data(), n * 4, hipMemcpyHostToDevice) == hipSuccess);
    bufs.push_back(din);
    bufs.push_back(dout);
    CHECK(bri17_operator_apply_host(op, h.data(), a.data()) == BRI17_OK);
    CHECK(bri17_operator_apply(op, din, dout, nullptr) == BRI17_OK);
    CHECK(hipMemcpy(c.data(), dout, n * 4, hipMemcpyDeviceToHost) == hipSuccess);
    CHECK(std::memcmp(a.data(), c.data(), n * 4) == 0);
    CHECK(bri17_operator_run_noalloc(op, 1, din, dout, nullptr) == BRI17_OK);
    CHECK(bri17_operator_solve_host(op, h.data(), a.data()) == BRI17_OK);
    CHECK(hipMemcpy(c.data(), dout, n * 4, hipMemcpyDeviceToHost) == hipSuccess);
    CHECK(std::memcmp(a.data(), c.data(), n * 4) == 0);
    char desc[4096];
    CHECK(bri17_operator_describe(op, desc, sizeof(desc)) == BRI17_OK);
    CHECK(bri17_operator_describe(op, desc, 8) == BRI17_OK && std::strlen(desc) == 7);
  }
  bri17_batch* b = nullptr;
  CHECK(bri17_batch_create(&b) == BRI17_OK);
  bri17_batch_item items[3];
  for (int g = 0; g < 3; g++) items[g] = {ops[g], g & 1, bufs[2 * g], bufs[2 * g + 1]};
  CHECK(bri17_batch_run(b, items, 3, nullptr) == BRI17_OK);
  int to = 1;
  CHECK(bri17_batch_wait(b, &to) == BRI17_OK && to == 0);
  bri17_graph* graph = nullptr;
  CHECK(bri17_batch_capture(b, items, 3, &graph) == BRI17_OK);
  CHECK(bri17_batch_run(b, items, 3, nullptr) == BRI17_ERR_STATE);
  CHECK(bri17_graph_launch(graph, nullptr) == BRI17_OK);
  CHECK(hipDeviceSynchronize() == hipSuccess);
  CHECK(bri17_graph_destroy(graph) == BRI17_OK);
  CHECK(bri17_batch_destroy(b) == BRI17_OK);
  // server: four submitting threads, host-memory requests
  bri17_server* srv = nullptr;
  bri17_server_config cfg;
  std::memset(&cfg, 0, sizeof(cfg));
  cfg.max_batch = 8;
  cfg.max_waves = 3;
  CHECK(bri17_server_create(&cfg, &srv) == BRI17_OK);
  std::atomic<int> bad{0};
  std::vector<std::thread> th;
  for (int t = 0; t < 4; t++)
    th.emplace_back([&, t] {
      const int g = t % 3;
      const size_t n = (size_t)3 * shapes[g][0] * shapes[g][1] * shapes[g][2];
      std::vector<float> out(n), ref(n);
      if (bri17_operator_apply_host(ops[g], h.data(), ref.data()) != BRI17_OK) bad++;
      for (int rep = 0; rep < 6; rep++) {
        bri17_request r;
        std::memset(&r, 0, sizeof(r));
        r.dim = 3;
        for (int d = 0; d < 3; d++) {
          r.shape[d] = shapes[g][d];
          r.L[d] = L[d];
        }
        r.mu = 1.1;
        r.nu = 0.3;
        r.mode = 0;
        r.dtype = BRI17_F32;
        r.memory = BRI17_MEM_HOST;
        r.in = h.data();
        r.out = out.data();
        int64_t id = -1;
        bri17_request_result res;
        if (bri17_server_submit(srv, &r, &id) != BRI17_OK ||
            bri17_server_wait(srv, id, 60.0, &res) != BRI17_OK || res.state != BRI17_REQ_DONE ||
            std::memcmp(out.data(), ref.data(), n * 4) != 0)
          bad++;
      }
    });
  for (auto& x : th) x.join();
  CHECK(bad.load() == 0);
  bri17_server_stats st;
  CHECK(bri17_server_get_stats(srv, &st) == BRI17_OK && st.completed == 24);
  CHECK(bri17_server_destroy(srv) == BRI17_OK);
  for (void* p : bufs) (void)hipFree(p);
  for (bri17_operator* op : ops) CHECK(bri17_operator_destroy(op) == BRI17_OK);
}

int main(int argc, char** argv) {
  grid_helpers();
  per_frequency<double>(2);
  per_frequency<double>(3);
  per_frequency<float>(2);
  per_frequency<float>(3);
  validation();
  const bool device = argc > 1 && std::strcmp(argv[1], "device") == 0;
  if (device) device_work();
  std::printf("%s (%d failures)\n", g_fail ? "FAILED" : "OK", g_fail);
  std::fflush(stdout);
  // After device work, leave without running the HIP runtime's teardown:
  // at process exit it returns device-side allocations through ASan's
  // device-allocator hooks after the runtime is marked unloaded, which ASan
  // reports as an internal CHECK failure (a runtime / sanitizer interaction
  // at exit, after every check above has run).
  if (device) std::_Exit(g_fail ? 1 : 0);
  return g_fail ? 1 : 0;
}
