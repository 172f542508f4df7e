/*
 * bri17 CPU oracle -- TEST INFRASTRUCTURE ONLY (see bri17_oracle.h).
 *
 * fp64 restatement of sbrisard/bri17 v0.1, written from the formulas:
 *   include/bri17/bri17.hpp      CartesianGrid, Hooke per-k functions
 *   tests/test_bri17.cpp         apply / strain harnesses, assembly
 *   reference/Q4_Q8_element.mac  element matrices
 *   sphinx/theory.rst            DFT conventions, k = 0 convention
 * The FFT is a recursive mixed-radix decimation-in-time Cooley-Tukey with a
 * generic prime butterfly -- deliberately a different algorithm from the
 * Stockham kernels of the product, so the two cannot share a mistake.
 */
#include "bri17_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
  double re, im;
} cplx;

static inline cplx cmul(cplx a, cplx b) {
  cplx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
  return r;
}
static inline cplx cadd(cplx a, cplx b) {
  cplx r = {a.re + b.re, a.im + b.im};
  return r;
}
static inline cplx cconj(cplx a) {
  cplx r = {a.re, -a.im};
  return r;
}
static inline cplx cscale(cplx a, double s) {
  cplx r = {a.re * s, a.im * s};
  return r;
}

static const double PI = 3.14159265358979323846264338327950288;

int ora_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

int64_t ora_grid_size(const ora_grid* g) {
  int64_t n = 1;
  for (int d = 0; d < g->dim; d++) n *= g->shape[d];
  return n;
}

/* bri17.hpp:127-158.  Corner n has per-axis offset bit (n >> (DIM-1-d)) & 1,
 * i.e. the last axis is the fastest bit: 2D (i,j),(i,j+1),(i+1,j),(i+1,j+1). */
void ora_cell_nodes(const ora_grid* g, int64_t cell, int64_t* nodes) {
  int D = g->dim;
  int idx[3];
  int64_t c = cell;
  for (int d = D - 1; d >= 0; d--) {
    idx[d] = (int)(c % g->shape[d]);
    c /= g->shape[d];
  }
  for (int n = 0; n < (1 << D); n++) {
    int64_t node = 0;
    for (int d = 0; d < D; d++) {
      int bit = (n >> (D - 1 - d)) & 1;
      int i = idx[d] + bit;
      if (i == g->shape[d]) i = 0;
      node = node * g->shape[d] + i;
    }
    nodes[n] = node;
  }
}

/* bri17.hpp:247-292, literally: phi = 2(1-cos b)/h^2, chi = (2+cos b)/3,
 * psi = sin b / h, scaling = mu/(1-2nu). */
static void modal_stiffness_real(const ora_grid* g, double mu, double nu,
                                 const int* k, double* K) {
  int D = g->dim;
  double phi[3] = {0}, chi[3] = {0}, psi[3] = {0};
  for (int i = 0; i < D; i++) {
    double h = g->L[i] / g->shape[i];
    double beta = 2.0 * PI * k[i] / g->shape[i];
    phi[i] = 2.0 * (1.0 - cos(beta)) / h / h;
    chi[i] = (2.0 + cos(beta)) / 3.0;
    psi[i] = sin(beta) / h;
  }
  double scaling = mu / (1.0 - 2.0 * nu);
  if (D == 2) {
    double H00 = phi[0] * chi[1];
    double H11 = chi[0] * phi[1];
    double Kd = mu * (H00 + H11);
    K[0] = scaling * H00 + Kd;
    K[1] = scaling * psi[0] * psi[1];
    K[2] = K[1];
    K[3] = scaling * H11 + Kd;
  } else {
    double H00 = phi[0] * chi[1] * chi[2];
    double H11 = chi[0] * phi[1] * chi[2];
    double H22 = chi[0] * chi[1] * phi[2];
    double Kd = mu * (H00 + H11 + H22);
    K[0] = scaling * H00 + Kd;
    K[1] = scaling * psi[0] * psi[1] * chi[2];
    K[2] = scaling * psi[0] * chi[1] * psi[2];
    K[3] = K[1];
    K[4] = scaling * H11 + Kd;
    K[5] = scaling * chi[0] * psi[1] * psi[2];
    K[6] = K[2];
    K[7] = K[5];
    K[8] = scaling * H22 + Kd;
  }
}

void ora_modal_stiffness(const ora_grid* g, double mu, double nu, const int* k,
                         double* K) {
  double Kr[9];
  int D = g->dim;
  modal_stiffness_real(g, mu, nu, k, Kr);
  for (int i = 0; i < D * D; i++) {
    K[2 * i] = Kr[i];
    K[2 * i + 1] = 0.0;
  }
}

/* bri17.hpp:212-236: alpha = pi k/N, c = cos alpha, s = sin alpha N/L,
 * prefactor = (-2 sin sum alpha, 2 cos sum alpha). */
static void modal_B(const ora_grid* g, const int* k, cplx* B) {
  int D = g->dim;
  double c[3] = {0}, s[3] = {0}, sum_alpha = 0.0;
  for (int i = 0; i < D; i++) {
    double alpha = PI * k[i] / g->shape[i];
    sum_alpha += alpha;
    c[i] = cos(alpha);
    s[i] = sin(alpha) * g->shape[i] / g->L[i];
  }
  cplx pre = {-2.0 * sin(sum_alpha), 2.0 * cos(sum_alpha)};
  if (D == 2) {
    B[0] = cscale(pre, s[0] * c[1]);
    B[1] = cscale(pre, c[0] * s[1]);
  } else {
    B[0] = cscale(pre, s[0] * c[1] * c[2]);
    B[1] = cscale(pre, c[0] * s[1] * c[2]);
    B[2] = cscale(pre, c[0] * c[1] * s[2]);
  }
}

void ora_modal_strain_displacement(const ora_grid* g, const int* k, double* B) {
  cplx b[3];
  modal_B(g, k, b);
  for (int i = 0; i < g->dim; i++) {
    B[2 * i] = b[i].re;
    B[2 * i + 1] = b[i].im;
  }
}

/* Cholesky K = L L^T of a real SPD DIM x DIM matrix and solve L L^T x = b for
 * complex b (the K.llt().solve of bri17.hpp:341; K^ is real). */
static void llt_solve(int D, const double* K, const cplx* b, cplx* x) {
  double Lm[3][3] = {{0}};
  for (int j = 0; j < D; j++) {
    double s = K[j * D + j];
    for (int p = 0; p < j; p++) s -= Lm[j][p] * Lm[j][p];
    Lm[j][j] = sqrt(s);
    for (int i = j + 1; i < D; i++) {
      double t = K[i * D + j];
      for (int p = 0; p < j; p++) t -= Lm[i][p] * Lm[j][p];
      Lm[i][j] = t / Lm[j][j];
    }
  }
  cplx y[3];
  for (int i = 0; i < D; i++) {
    cplx t = b[i];
    for (int p = 0; p < i; p++) t = cadd(t, cscale(y[p], -Lm[i][p]));
    y[i] = cscale(t, 1.0 / Lm[i][i]);
  }
  for (int i = D - 1; i >= 0; i--) {
    cplx t = y[i];
    for (int p = i + 1; p < D; p++) t = cadd(t, cscale(x[p], -Lm[p][i]));
    x[i] = cscale(t, 1.0 / Lm[i][i]);
  }
}

/* bri17.hpp:308-355. */
void ora_modal_eigenstress_to_opposite_strain(const ora_grid* g, double mu,
                                              double nu, const int* k,
                                              const double* tau, double* eta) {
  int D = g->dim;
  int sym = D == 2 ? 3 : 6;
  const double r2 = sqrt(2.0);
  int null_frequency = 1;
  for (int d = 0; d < D; d++) null_frequency &= (k[d] == 0);
  if (null_frequency) {
    for (int i = 0; i < 2 * sym; i++) eta[i] = 0.0;
    return;
  }
  cplx B[3], t[6];
  double K[9];
  modal_B(g, k, B);
  modal_stiffness_real(g, mu, nu, k, K);
  for (int i = 0; i < sym; i++) {
    t[i].re = tau[2 * i];
    t[i].im = tau[2 * i + 1];
  }
  cplx T[3][3];
  if (D == 2) {
    T[0][0] = t[0];
    T[1][1] = t[1];
    T[0][1] = T[1][0] = cscale(t[2], 1.0 / r2);
  } else {
    T[0][0] = t[0];
    T[1][1] = t[1];
    T[2][2] = t[2];
    T[1][2] = T[2][1] = cscale(t[3], 1.0 / r2);
    T[0][2] = T[2][0] = cscale(t[4], 1.0 / r2);
    T[0][1] = T[1][0] = cscale(t[5], 1.0 / r2);
  }
  cplx rhs[3], u[3];
  for (int i = 0; i < D; i++) {
    cplx s = {0.0, 0.0};
    for (int j = 0; j < D; j++) s = cadd(s, cmul(T[i][j], cconj(B[j])));
    rhs[i] = s;
  }
  llt_solve(D, K, rhs, u);
  cplx E[3][3];
  for (int i = 0; i < D; i++)
    for (int j = 0; j < D; j++)
      E[i][j] = cscale(cadd(cmul(B[i], u[j]), cmul(u[i], B[j])), 0.5);
  cplx out[6];
  if (D == 2) {
    out[0] = E[0][0];
    out[1] = E[1][1];
    out[2] = cscale(E[0][1], r2);
  } else {
    out[0] = E[0][0];
    out[1] = E[1][1];
    out[2] = E[2][2];
    out[3] = cscale(E[1][2], r2);
    out[4] = cscale(E[2][0], r2);
    out[5] = cscale(E[0][1], r2);
  }
  for (int i = 0; i < sym; i++) {
    eta[2 * i] = out[i].re;
    eta[2 * i + 1] = out[i].im;
  }
}

/* ------------------------------------------------------------------------ */
/* FFT: recursive decimation in time, generic prime radix.                   */

typedef struct {
  int n;
  int nfac;
  int fac[64];
  cplx* w; /* w[m] = exp(sign 2 pi i m / n) */
} fft_plan1;

static void plan1_init(fft_plan1* p, int n, int sign) {
  p->n = n;
  p->nfac = 0;
  int m = n;
  /* radix 4 first (fewer combine passes), then primes ascending */
  while (m % 4 == 0) {
    p->fac[p->nfac++] = 4;
    m /= 4;
  }
  for (int q = 2; m > 1; q++) {
    while (m % q == 0) {
      p->fac[p->nfac++] = q;
      m /= q;
    }
  }
  p->w = (cplx*)malloc(sizeof(cplx) * (size_t)n);
  /* twiddles evaluated in long double and rounded once, so that e.g.
   * cos(2 pi/3) is exactly -0.5 and conjugate symmetry is preserved */
  const long double PIL = 3.141592653589793238462643383279502884L;
  for (int i = 0; i < n; i++) {
    long double a = 2.0L * PIL * (long double)i / (long double)n;
    p->w[i].re = (double)cosl(a);
    p->w[i].im = sign * (double)sinl(a);
  }
}

static void plan1_free(fft_plan1* p) { free(p->w); }

/* out[k] = sum_j in[j*is] w_n^{jk}, n = current length, wstride = N/n */
static void fft_rec(const fft_plan1* P, int n, const cplx* in, int64_t is,
                    cplx* out, int wstride, const int* fac) {
  if (n == 1) {
    out[0] = in[0];
    return;
  }
  int p = fac[0];
  int m = n / p;
  for (int r = 0; r < p; r++)
    fft_rec(P, m, in + r * is, is * p, out + (int64_t)r * m, wstride * p,
            fac + 1);
  cplx tloc[16];
  cplx* t = p <= 16 ? tloc : (cplx*)malloc(sizeof(cplx) * (size_t)p);
  const int root = wstride * m; /* w_p = w_N^{N/p} */
  for (int k = 0; k < m; k++) {
    for (int r = 0; r < p; r++) {
      cplx y = out[(int64_t)r * m + k];
      t[r] = (r == 0 || k == 0) ? y : cmul(y, P->w[(int64_t)r * k * wstride]);
    }
    for (int q = 0; q < p; q++) {
      cplx acc = t[0];
      for (int r = 1; r < p; r++)
        acc = cadd(acc, cmul(t[r], P->w[(int64_t)((r * q) % p) * root]));
      out[(int64_t)q * m + k] = acc;
    }
  }
  if (t != tloc) free(t);
}

void ora_fft(int dim, const int* shape, double* data, int sign, int nthreads) {
  cplx* x = (cplx*)data;
  int64_t total = 1;
  for (int d = 0; d < dim; d++) total *= shape[d];
  for (int d = 0; d < dim; d++) {
    int n = shape[d];
    if (n == 1) continue;
    int64_t inner = 1;
    for (int e = d + 1; e < dim; e++) inner *= shape[e];
    int64_t nlines = total / n;
    fft_plan1 P;
    plan1_init(&P, n, sign);
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
    {
      cplx* a = (cplx*)malloc(sizeof(cplx) * (size_t)n);
      cplx* b = (cplx*)malloc(sizeof(cplx) * (size_t)n);
#pragma omp for schedule(static)
      for (int64_t l = 0; l < nlines; l++) {
        int64_t outer = l / inner, in_i = l % inner;
        cplx* base = x + outer * n * inner + in_i;
        for (int j = 0; j < n; j++) a[j] = base[(int64_t)j * inner];
        fft_rec(&P, n, a, 1, b, 1, P.fac);
        for (int j = 0; j < n; j++) base[(int64_t)j * inner] = b[j];
      }
      free(a);
      free(b);
    }
    plan1_free(&P);
  }
}

static void unravel(const ora_grid* g, int64_t i, int* k) {
  for (int d = g->dim - 1; d >= 0; d--) {
    k[d] = (int)(i % g->shape[d]);
    i /= g->shape[d];
  }
}

static double cell_volume(const ora_grid* g) {
  double v = 1.0;
  for (int d = 0; d < g->dim; d++) v *= g->L[d] / g->shape[d];
  return v;
}

/* tests/test_bri17.cpp:56-107 */
void ora_apply(const ora_grid* g, double mu, double nu, const double* u,
               double* Ku, double* imag_max, int nthreads) {
  int D = g->dim;
  int64_t N = ora_grid_size(g);
  cplx* uh = (cplx*)malloc(sizeof(cplx) * (size_t)(D * N));
  cplx* Kh = (cplx*)malloc(sizeof(cplx) * (size_t)(D * N));
  int nt = nthreads > 0 ? nthreads : 1;
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int64_t i = 0; i < D * N; i++) {
    uh[i].re = u[i];
    uh[i].im = 0.0;
  }
  for (int c = 0; c < D; c++) ora_fft(D, g->shape, (double*)(uh + c * N), -1, nt);
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int64_t i = 0; i < N; i++) {
    int k[3];
    double K[9];
    unravel(g, i, k);
    modal_stiffness_real(g, mu, nu, k, K);
    for (int r = 0; r < D; r++) {
      cplx s = {0.0, 0.0};
      for (int c = 0; c < D; c++) s = cadd(s, cscale(uh[c * N + i], K[r * D + c]));
      Kh[r * N + i] = s;
    }
  }
  for (int c = 0; c < D; c++) ora_fft(D, g->shape, (double*)(Kh + c * N), +1, nt);
  double corr = cell_volume(g) / (double)N;
  double im = 0.0;
  for (int64_t i = 0; i < D * N; i++) {
    Ku[i] = Kh[i].re * corr;
    double a = fabs(Kh[i].im * corr);
    if (a > im) im = a;
  }
  if (imag_max) *imag_max = im;
  free(uh);
  free(Kh);
}

void ora_solve(const ora_grid* g, double mu, double nu, const double* f,
               double* u, int nthreads) {
  int D = g->dim;
  int64_t N = ora_grid_size(g);
  int nt = nthreads > 0 ? nthreads : 1;
  cplx* fh = (cplx*)malloc(sizeof(cplx) * (size_t)(D * N));
  for (int64_t i = 0; i < D * N; i++) {
    fh[i].re = f[i];
    fh[i].im = 0.0;
  }
  for (int c = 0; c < D; c++) ora_fft(D, g->shape, (double*)(fh + c * N), -1, nt);
  double hv = cell_volume(g);
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int64_t i = 0; i < N; i++) {
    if (i == 0) {
      for (int c = 0; c < D; c++) fh[c * N].re = fh[c * N].im = 0.0;
      continue;
    }
    int k[3];
    double K[9];
    unravel(g, i, k);
    modal_stiffness_real(g, mu, nu, k, K);
    cplx b[3], x[3];
    for (int c = 0; c < D; c++) b[c] = fh[c * N + i];
    llt_solve(D, K, b, x);
    for (int c = 0; c < D; c++) fh[c * N + i] = cscale(x[c], 1.0 / hv);
  }
  for (int c = 0; c < D; c++) ora_fft(D, g->shape, (double*)(fh + c * N), +1, nt);
  for (int64_t i = 0; i < D * N; i++) u[i] = fh[i].re / (double)N;
  free(fh);
}

/* tests/test_bri17.cpp:194-245 */
void ora_strain(const ora_grid* g, const double* u, double* eps, int nthreads) {
  int D = g->dim;
  int S = D == 2 ? 3 : 6;
  int64_t N = ora_grid_size(g);
  int nt = nthreads > 0 ? nthreads : 1;
  const double r2 = sqrt(2.0);
  cplx* uh = (cplx*)malloc(sizeof(cplx) * (size_t)(D * N));
  cplx* eh = (cplx*)malloc(sizeof(cplx) * (size_t)(S * N));
  for (int64_t i = 0; i < D * N; i++) {
    uh[i].re = u[i];
    uh[i].im = 0.0;
  }
  for (int c = 0; c < D; c++) ora_fft(D, g->shape, (double*)(uh + c * N), -1, nt);
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int64_t i = 0; i < N; i++) {
    int k[3];
    cplx B[3], U[3];
    unravel(g, i, k);
    modal_B(g, k, B);
    for (int c = 0; c < D; c++) U[c] = uh[c * N + i];
#define EPS(a, b) cscale(cadd(cmul(B[a], U[b]), cmul(U[a], B[b])), 0.5)
    if (D == 2) {
      eh[i] = EPS(0, 0);
      eh[N + i] = EPS(1, 1);
      eh[2 * N + i] = cscale(EPS(0, 1), r2);
    } else {
      eh[i] = EPS(0, 0);
      eh[N + i] = EPS(1, 1);
      eh[2 * N + i] = EPS(2, 2);
      eh[3 * N + i] = cscale(EPS(1, 2), r2);
      eh[4 * N + i] = cscale(EPS(2, 0), r2);
      eh[5 * N + i] = cscale(EPS(0, 1), r2);
    }
#undef EPS
  }
  for (int c = 0; c < S; c++) ora_fft(D, g->shape, (double*)(eh + c * N), +1, nt);
  for (int64_t i = 0; i < S * N; i++) eps[i] = eh[i].re / (double)N;
  free(uh);
  free(eh);
}

void ora_eigenstress_strain(const ora_grid* g, double mu, double nu,
                            const double* tau, double* eta, int nthreads) {
  int D = g->dim;
  int S = D == 2 ? 3 : 6;
  int64_t N = ora_grid_size(g);
  int nt = nthreads > 0 ? nthreads : 1;
  cplx* th = (cplx*)malloc(sizeof(cplx) * (size_t)(S * N));
  for (int64_t i = 0; i < S * N; i++) {
    th[i].re = tau[i];
    th[i].im = 0.0;
  }
  for (int c = 0; c < S; c++) ora_fft(D, g->shape, (double*)(th + c * N), -1, nt);
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int64_t i = 0; i < N; i++) {
    int k[3];
    double t[12], e[12];
    unravel(g, i, k);
    for (int c = 0; c < S; c++) {
      t[2 * c] = th[c * N + i].re;
      t[2 * c + 1] = th[c * N + i].im;
    }
    ora_modal_eigenstress_to_opposite_strain(g, mu, nu, k, t, e);
    for (int c = 0; c < S; c++) {
      th[c * N + i].re = e[2 * c];
      th[c * N + i].im = e[2 * c + 1];
    }
  }
  for (int c = 0; c < S; c++) ora_fft(D, g->shape, (double*)(th + c * N), +1, nt);
  for (int64_t i = 0; i < S * N; i++) eta[i] = th[i].re / (double)N;
  free(th);
}

/* ------------------------------------------------------------------------ */
/* Element matrices (reference/Q4_Q8_element.mac).                           */

/* Node n of the element: per-axis offset bit (n >> (D-1-d)) & 1 (mac:44-51,
 * the list product makes the last coordinate the fastest). */
static double shape_grad(int D, const double* h, int n, int d, const double* xi) {
  double v = 1.0;
  for (int e = 0; e < D; e++) {
    int bit = (n >> (D - 1 - e)) & 1;
    if (e == d)
      v *= (bit ? 1.0 : -1.0) / h[e];
    else
      v *= bit ? xi[e] : 1.0 - xi[e];
  }
  return v;
}

/* U = 1/2 int [lambda tr(eps)^2 + 2 mu eps:eps]  (mac:78-91) ->
 * Ke[(c,n),(c',n')] = int lambda dN_n/dx_c dN_n'/dx_c'
 *                   + mu delta_cc' grad N_n . grad N_n'
 *                   + mu dN_n/dx_c' dN_n'/dx_c,
 * with lambda = 2 mu nu / (1 - 2 nu) (mac:144). */
void ora_element_stiffness(int D, const double* h, double mu, double nu,
                           double* Ke) {
  int nn = 1 << D, nd = D * nn;
  double lambda = 2.0 * mu * nu / (1.0 - 2.0 * nu);
  const double g0 = 0.5 - 0.5 / sqrt(3.0), g1 = 0.5 + 0.5 / sqrt(3.0);
  for (int i = 0; i < nd * nd; i++) Ke[i] = 0.0;
  int npts = 1 << D;
  double vol = 1.0;
  for (int d = 0; d < D; d++) vol *= h[d];
  for (int q = 0; q < npts; q++) {
    double xi[3];
    for (int d = 0; d < D; d++) xi[d] = ((q >> d) & 1) ? g1 : g0;
    double w = vol / npts;
    double G[8][3];
    for (int n = 0; n < nn; n++)
      for (int d = 0; d < D; d++) G[n][d] = shape_grad(D, h, n, d, xi);
    for (int c = 0; c < D; c++)
      for (int n = 0; n < nn; n++)
        for (int c2 = 0; c2 < D; c2++)
          for (int n2 = 0; n2 < nn; n2++) {
            double v = lambda * G[n][c] * G[n2][c2] + mu * G[n][c2] * G[n2][c];
            if (c == c2) {
              double dot = 0.0;
              for (int d = 0; d < D; d++) dot += G[n][d] * G[n2][d];
              v += mu * dot;
            }
            Ke[(c * nn + n) * nd + c2 * nn + n2] += w * v;
          }
  }
}

/* Cell-averaged strain, Voigt-Mandel rows (mac:159-174). */
void ora_element_strain_displacement(int D, const double* h, double* Be) {
  int nn = 1 << D, nd = D * nn;
  int S = D == 2 ? 3 : 6;
  static const int first3[6] = {0, 1, 2, 1, 2, 0}, second3[6] = {0, 1, 2, 2, 0, 1};
  static const int first2[3] = {0, 1, 0}, second2[3] = {0, 1, 1};
  const int* fi = D == 2 ? first2 : first3;
  const int* se = D == 2 ? second2 : second3;
  for (int r = 0; r < S; r++) {
    int i = fi[r], j = se[r];
    double w = (i == j) ? 1.0 : sqrt(2.0);
    for (int c = 0; c < D; c++)
      for (int n = 0; n < nn; n++) {
        /* average of dN_n/dx_d over the cell: (+-1/h_d) * (1/2)^(D-1) */
        double v = 0.0;
        if (c == i) {
          int bit = (n >> (D - 1 - j)) & 1;
          v += 0.5 * (bit ? 1.0 : -1.0) / h[j];
        }
        if (c == j) {
          int bit = (n >> (D - 1 - i)) & 1;
          v += 0.5 * (bit ? 1.0 : -1.0) / h[i];
        }
        v *= ldexp(1.0, -(D - 1));
        Be[r * nd + c * nn + n] = w * v;
      }
  }
}

/* node index -> cell that has this node as its corner nl */
static int64_t corner_cell(const ora_grid* g, const int* idx, int nl) {
  int D = g->dim;
  int64_t cell = 0;
  for (int d = 0; d < D; d++) {
    int bit = (nl >> (D - 1 - d)) & 1;
    int i = idx[d] - bit;
    if (i < 0) i += g->shape[d];
    cell = cell * g->shape[d] + i;
  }
  return cell;
}

void ora_ebe_apply(const ora_grid* g, const double* Ke, const double* u,
                   double* Ku, int nthreads) {
  int D = g->dim, nn = 1 << D, nd = D * nn;
  int64_t N = ora_grid_size(g);
  int nt = nthreads > 0 ? nthreads : 1;
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int64_t node = 0; node < N; node++) {
    int idx[3];
    unravel(g, node, idx);
    double acc[3] = {0.0, 0.0, 0.0};
    for (int nl = 0; nl < nn; nl++) {
      int64_t cell = corner_cell(g, idx, nl);
      int64_t cn[8];
      ora_cell_nodes(g, cell, cn);
      for (int c = 0; c < D; c++) {
        const double* row = Ke + (size_t)(c * nn + nl) * nd;
        double s = 0.0;
        for (int c2 = 0; c2 < D; c2++)
          for (int n2 = 0; n2 < nn; n2++) s += row[c2 * nn + n2] * u[c2 * N + cn[n2]];
        acc[c] += s;
      }
    }
    for (int c = 0; c < D; c++) Ku[c * N + node] = acc[c];
  }
}

void ora_ebe_strain(const ora_grid* g, const double* Be, const double* u,
                    double* eps, int nthreads) {
  int D = g->dim, nn = 1 << D, nd = D * nn, S = D == 2 ? 3 : 6;
  int64_t N = ora_grid_size(g);
  int nt = nthreads > 0 ? nthreads : 1;
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int64_t cell = 0; cell < N; cell++) {
    int64_t cn[8];
    ora_cell_nodes(g, cell, cn);
    for (int r = 0; r < S; r++) {
      double s = 0.0;
      for (int c = 0; c < D; c++)
        for (int n = 0; n < nn; n++) s += Be[r * nd + c * nn + n] * u[c * N + cn[n]];
      eps[r * N + cell] = s;
    }
  }
}

void ora_ebe_strain_transpose(const ora_grid* g, const double* Be,
                              const double* tau, double* f, int nthreads) {
  int D = g->dim, nn = 1 << D, nd = D * nn, S = D == 2 ? 3 : 6;
  int64_t N = ora_grid_size(g);
  int nt = nthreads > 0 ? nthreads : 1;
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int64_t node = 0; node < N; node++) {
    int idx[3];
    unravel(g, node, idx);
    double acc[3] = {0.0, 0.0, 0.0};
    for (int nl = 0; nl < nn; nl++) {
      int64_t cell = corner_cell(g, idx, nl);
      for (int c = 0; c < D; c++)
        for (int r = 0; r < S; r++) acc[c] += Be[r * nd + c * nn + nl] * tau[r * N + cell];
    }
    for (int c = 0; c < D; c++) f[c * N + node] = acc[c];
  }
}

/* ------------------------------------------------------------------------ */
/* SplitMix64 counter generator.                                             */

static inline uint64_t splitmix(uint64_t seed, uint64_t i) {
  uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static inline double u53(uint64_t z) { return (double)(z >> 11) * (1.0 / 9007199254740992.0); }

void ora_fill_uniform(uint64_t seed, int64_t n, double lo, double hi, double* out) {
  for (int64_t i = 0; i < n; i++) out[i] = lo + (hi - lo) * u53(splitmix(seed, (uint64_t)i));
}

void ora_fill_normal(uint64_t seed, int64_t n, double* out) {
  for (int64_t i = 0; i < n; i++) {
    double a = u53(splitmix(seed, 2 * (uint64_t)i));
    double b = u53(splitmix(seed, 2 * (uint64_t)i + 1));
    out[i] = sqrt(-2.0 * log(1.0 - a)) * cos(2.0 * PI * b);
  }
}
