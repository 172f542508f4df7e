/*
 * bri17 CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, fp64 restatement of the reference algorithms of sbrisard/bri17
 * (v0.1).  Only the test suite, the smoke check and the CPU-baseline leg of
 * the benchmark load this library, and only as the checker; the product
 * (libbri17_amd.so) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.  Parity status:
 *   ora_element_stiffness / ora_element_strain_displacement:
 *       pinned by the Maxima Ke / Be fixtures (tests/test_bri17.cpp:344-356,
 *       372-531, 547-552, 570-598) at rtol 1e-15 / atol 1e-14.
 *   ora_modal_stiffness + ora_fft + ora_apply:
 *       pinned by the reference's own dense-operator test (operator probed with
 *       unit vectors vs. assembled Ke, tests/test_bri17.cpp:130-150,153-174,
 *       360,535), by a brute-force DFT and numpy.fft, and by invariants.
 *   ora_modal_strain_displacement + ora_strain:
 *       pinned by the reference's strain-displacement test (:194-245,295-316,
 *       558,604) with the Be fixtures.
 *   ora_solve: pinned by K.G.f = f - mean(f) with K from the pinned apply and
 *       by the EBE product.
 *   ora_modal_eigenstress_to_opposite_strain / ora_eigenstress_strain:
 *       pinned by an independent real-space route (EBE Be^T tau load, pinned
 *       Green solve, EBE strain) -- no reference fixture exists for it.
 */
#ifndef BRI17_ORACLE_H
#define BRI17_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Grid: shape[d] cells along axis d, side L[d]; nodes row-major, last axis
 * fastest (include/bri17/bri17.hpp:34-93). */
typedef struct {
  int dim;
  int shape[3];
  double L[3];
} ora_grid;

int64_t ora_grid_size(const ora_grid* g);

/* bri17.hpp:127-158 -- periodic cell corners in binary order. */
void ora_cell_nodes(const ora_grid* g, int64_t cell, int64_t* nodes);

/* bri17.hpp:247-292 -- K is DIM*DIM interleaved complex (re, im). */
void ora_modal_stiffness(const ora_grid* g, double mu, double nu, const int* k,
                         double* K);

/* bri17.hpp:212-236 -- B is DIM interleaved complex. */
void ora_modal_strain_displacement(const ora_grid* g, const int* k, double* B);

/* bri17.hpp:308-355 -- tau, eta: SYM interleaved complex (Voigt-Mandel). */
void ora_modal_eigenstress_to_opposite_strain(const ora_grid* g, double mu,
                                              double nu, const int* k,
                                              const double* tau, double* eta);

/* Unnormalised multi-dimensional complex DFT in place (FFTW sign convention:
 * sign=-1 forward, +1 backward), as used at tests/test_bri17.cpp:119-127.
 * data: interleaved complex, C order. */
void ora_fft(int dim, const int* shape, double* data, int sign, int nthreads);

/* tests/test_bri17.cpp:56-107: Ku = (|h|/|N|) IDFT(K^ . DFT(u)), full c2c
 * spectrum, per-k modal stiffness recomputed.  u, Ku: real, (DIM, N...) C
 * order.  If imag_max != NULL it receives max |imag| before the real part is
 * taken (the caller-side check of :140). */
void ora_apply(const ora_grid* g, double mu, double nu, const double* u,
               double* Ku, double* imag_max, int nthreads);

/* Green solve K u = f with the reference conventions: u^_k = K^_k^{-1} f^_k /
 * |h| for k != 0, u^_0 = 0 (sphinx/theory.rst:208-212), u = IDFT(u^)/|N|.
 * The per-k solve is a Cholesky (LLT) factorisation, as at bri17.hpp:341. */
void ora_solve(const ora_grid* g, double mu, double nu, const double* f,
               double* u, int nthreads);

/* tests/test_bri17.cpp:194-245: cell-averaged strains, Voigt-Mandel rows
 * [xx,yy,(zz),sqrt2 yz, sqrt2 zx, sqrt2 xy] (2D: [xx,yy,sqrt2 xy]).
 * eps: (SYM, N...) real. */
void ora_strain(const ora_grid* g, const double* u, double* eps, int nthreads);

/* Field-level eigenstress operator: tau (SYM, N...) real cellwise constant ->
 * eta = -eps (SYM, N...) real, by DFT, bri17.hpp:308-355 per k, IDFT/|N|
 * (the per-k loop of python/demo.py:33-40, completed by the inverse DFT). */
void ora_eigenstress_strain(const ora_grid* g, double mu, double nu,
                            const double* tau, double* eta, int nthreads);

/* Q4/H8 element matrices (reference/Q4_Q8_element.mac:44-91,159-174) by
 * 2-point Gauss quadrature, exact for (tri)bilinear shape functions.
 * Ke: (DIM*2^DIM)^2 row-major, DOFs component-major.  Be: SYM x DIM*2^DIM. */
void ora_element_stiffness(int dim, const double* h, double mu, double nu,
                           double* Ke);
void ora_element_strain_displacement(int dim, const double* h, double* Be);

/* Matrix-free element-by-element products (the assembly of
 * tests/test_bri17.cpp:153-174 and :295-316 applied to a vector).
 * Gather formulation, race free. */
void ora_ebe_apply(const ora_grid* g, const double* Ke, const double* u,
                   double* Ku, int nthreads);
void ora_ebe_strain(const ora_grid* g, const double* Be, const double* u,
                    double* eps, int nthreads);
/* Transpose of ora_ebe_strain: f = B^T tau (nodal force of a cellwise
 * eigenstress, sphinx/theory.rst:183-205 in real space, up to sign). */
void ora_ebe_strain_transpose(const ora_grid* g, const double* Be,
                              const double* tau, double* f, int nthreads);

/* Counter-based generator shared by tests and benchmark inputs:
 * SplitMix64(seed, index) -> uniform in [lo, hi) / standard normal. */
void ora_fill_uniform(uint64_t seed, int64_t n, double lo, double hi,
                      double* out);
void ora_fill_normal(uint64_t seed, int64_t n, double* out);

int ora_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
