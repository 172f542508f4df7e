/* The CPU oracle (oracle/bri17_oracle.c, test infrastructure) run under
 * AddressSanitizer and UndefinedBehaviorSanitizer on small 2D and 3D grids,
 * odd and even sides, one and several OpenMP threads.  Besides the absence of
 * memory errors it checks invariants the oracle's own tests also pin:
 * K const = 0, <v, K u> = <K v, u>, K G f = f - mean f, the element-by-element
 * product equal to the spectral apply, FFT forward then backward = |N| x,
 * and the eigenstress operator's zero mean.  Exit status 0 on success. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bri17_oracle.h"

static int g_fail = 0;
#define CHECK(cond, what, val)                                               \
  do {                                                                       \
    if (!(cond)) {                                                           \
      printf("FAIL %s: %s (%g)\n", what, #cond, (double)(val));              \
      g_fail++;                                                              \
    }                                                                        \
  } while (0)

static double dot(const double* a, const double* b, int64_t n) {
  double s = 0;
  for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

static double nrm(const double* a, int64_t n) { return sqrt(dot(a, a, n)); }

static void run_grid(int dim, const int* shape, const double* L, int nthreads) {
  ora_grid g;
  memset(&g, 0, sizeof(g));
  g.dim = dim;
  for (int d = 0; d < dim; d++) {
    g.shape[d] = shape[d];
    g.L[d] = L[d];
  }
  const int64_t N = ora_grid_size(&g);
  const int64_t n = dim * N;
  const int S = dim == 2 ? 3 : 6;
  const double mu = 1.3, nu = 0.27;
  double* u = malloc(sizeof(double) * n);
  double* v = malloc(sizeof(double) * n);
  double* Ku = malloc(sizeof(double) * n);
  double* Kv = malloc(sizeof(double) * n);
  double* w = malloc(sizeof(double) * n);
  double* c = malloc(sizeof(double) * 2 * n);
  double* tau = malloc(sizeof(double) * S * N);
  double* eta = malloc(sizeof(double) * S * N);
  double* eps = malloc(sizeof(double) * S * N);
  ora_fill_uniform(11, n, -1.0, 1.0, u);
  ora_fill_normal(12, n, v);
  ora_fill_uniform(13, S * N, -1.0, 1.0, tau);

  /* K const = 0 */
  for (int64_t i = 0; i < n; i++) w[i] = 1.0 + (double)(i / N);
  double im = 0;
  ora_apply(&g, mu, nu, w, Ku, &im, nthreads);
  CHECK(nrm(Ku, n) <= 1e-11 * nrm(w, n), "K const", nrm(Ku, n));
  /* symmetry */
  ora_apply(&g, mu, nu, u, Ku, &im, nthreads);
  ora_apply(&g, mu, nu, v, Kv, NULL, nthreads);
  const double a = dot(v, Ku, n), b = dot(Kv, u, n);
  CHECK(fabs(a - b) <= 1e-11 * (fabs(a) + fabs(b) + 1e-300), "symmetry", a - b);
  /* element-by-element product */
  double h[3], Ke[24 * 24];
  for (int d = 0; d < dim; d++) h[d] = L[d] / shape[d];
  ora_element_stiffness(dim, h, mu, nu, Ke);
  ora_ebe_apply(&g, Ke, u, w, nthreads);
  for (int64_t i = 0; i < n; i++) w[i] -= Ku[i];
  CHECK(nrm(w, n) <= 1e-12 * nrm(Ku, n), "EBE", nrm(w, n));
  /* K G f = f - mean f */
  ora_solve(&g, mu, nu, v, w, nthreads);
  ora_apply(&g, mu, nu, w, Kv, NULL, nthreads);
  for (int comp = 0; comp < dim; comp++) {
    double m = 0;
    for (int64_t i = 0; i < N; i++) m += v[comp * N + i];
    m /= (double)N;
    for (int64_t i = 0; i < N; i++) Kv[comp * N + i] -= v[comp * N + i] - m;
  }
  CHECK(nrm(Kv, n) <= 1e-11 * nrm(v, n), "K G f", nrm(Kv, n));
  /* FFT round trip */
  for (int64_t i = 0; i < n; i++) {
    c[2 * i] = u[i];
    c[2 * i + 1] = v[i];
  }
  ora_fft(dim, shape, c, -1, nthreads);
  ora_fft(dim, shape, c, +1, nthreads);
  double e = 0;
  for (int64_t i = 0; i < N; i++)
    e = fmax(e, fabs(c[2 * i] / (double)N - u[i]) + fabs(c[2 * i + 1] / (double)N - v[i]));
  CHECK(e <= 1e-12, "FFT round trip", e);
  /* strain and eigenstress run (values pinned by the Python-side tests) */
  ora_strain(&g, u, eps, nthreads);
  ora_eigenstress_strain(&g, mu, nu, tau, eta, nthreads);
  for (int comp = 0; comp < S; comp++) {
    double m = 0, s = 0;
    for (int64_t i = 0; i < N; i++) {
      m += eta[comp * N + i];
      s += fabs(eta[comp * N + i]);
    }
    CHECK(fabs(m) <= 1e-10 * (s + 1e-300), "eigenstress mean", m);
  }
  double Be[6 * 24];
  ora_element_strain_displacement(dim, h, Be);
  ora_ebe_strain(&g, Be, u, tau, nthreads);
  for (int64_t i = 0; i < S * N; i++) tau[i] -= eps[i];
  CHECK(nrm(tau, S * N) <= 1e-12 * nrm(eps, S * N), "EBE strain", nrm(tau, S * N));
  ora_ebe_strain_transpose(&g, Be, eps, w, nthreads);
  int64_t nodes[8];
  ora_cell_nodes(&g, N - 1, nodes);
  for (int i = 0; i < (1 << dim); i++) CHECK(nodes[i] >= 0 && nodes[i] < N, "cell nodes", nodes[i]);
  free(u);
  free(v);
  free(Ku);
  free(Kv);
  free(w);
  free(c);
  free(tau);
  free(eta);
  free(eps);
}

int main(void) {
  const int s2[2] = {6, 5}, s3[3] = {3, 4, 5}, t3[3] = {8, 6, 7};
  const double L2[2] = {1.0, 1.4}, L3[3] = {3.3, 4.8, 6.5}, M3[3] = {1.0, 0.8, 1.2};
  for (int nt = 1; nt <= 3; nt += 2) {
    run_grid(2, s2, L2, nt);
    run_grid(3, s3, L3, nt);
    run_grid(3, t3, M3, nt);
  }
  printf("%s (%d failures)\n", g_fail ? "FAILED" : "OK", g_fail);
  return g_fail ? 1 : 0;
}
