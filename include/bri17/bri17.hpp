/**
 * bri17 on MI355X: drop-in C++ header for sbrisard/bri17 v0.1.
 *
 * Same namespace, class templates and member signatures as the reference
 * header (include/bri17/bri17.hpp:34-362): CartesianGrid<T, DIM> and
 * Hooke<T, DIM>.  The per-frequency methods are served by libbri17_amd.so
 * through its C ABI (include/bri17_amd.h) instead of inline code, and no
 * Eigen is needed.  Added: StiffnessMatrix<T, DIM>, the operator K u and the
 * Green solve K u = f on device pointers (the reference only has the test
 * harness StiffnessMatrixFactory::compute_Ku, tests/test_bri17.cpp:56-107).
 *
 * Differences from the reference, all deliberate:
 *  - get_node_at is constrained with `requires` instead of static_assert, so
 *    the class templates can be explicitly instantiated;
 *  - failures of the library (only possible in StiffnessMatrix) throw
 *    std::runtime_error; the reference reports no errors.
 *
 * Link with -lbri17_amd.
 */
#pragma once

#include <array>
#include <complex>
#include <concepts>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <iostream>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <string>
#include <typeinfo>

#include "bri17_amd.h"

namespace bri17 {

namespace detail {
inline void check(int rc) {
  if (rc != BRI17_OK) throw std::runtime_error(std::string("bri17_amd: ") + bri17_last_error());
}
}  // namespace detail

/** A rectangular grid with fixed spacing in each direction
 * (include/bri17/bri17.hpp:34-159). */
template <typename T, int DIM>
  requires(std::floating_point<T> && ((DIM == 2) || (DIM == 3)))
class CartesianGrid {
 public:
  /** Number of nodes per cell: `2 ** DIM`. */
  static constexpr int num_nodes_per_cell = 1 << DIM;
  /** Number of cells in each direction. */
  std::array<int, DIM> const shape;
  /** Size of the grid in each direction. */
  std::array<T, DIM> const L;
  /** Total number of cells. */
  int const size;

  CartesianGrid(std::array<int, DIM> shape, std::array<T, DIM> L)
      : shape{shape},
        L{L},
        size{std::reduce(shape.cbegin(), shape.cend(), int{1}, std::multiplies())} {}

  std::string repr() const {
    std::ostringstream stream;
    stream << "CartesianGrid<" << typeid(T).name() << "," << DIM << ">{shape={";
    for (auto n : shape) stream << n << ",";
    stream << "},L={";
    for (auto x : L) stream << x << ",";
    stream << "}}";
    return stream.str();
  }

  /** Index of the node at [i, j] (row-major). */
  int get_node_at(int i, int j) const
    requires(DIM == 2)
  {
    return i * shape[1] + j;
  }

  /** Index of the node at [i, j, k] (row-major). */
  int get_node_at(int i, int j, int k) const
    requires(DIM == 3)
  {
    return (i * shape[1] + j) * shape[2] + k;
  }

  /** Periodic corners of a cell, in the reference's binary order. */
  std::array<int, num_nodes_per_cell> get_cell_nodes(int cell) const {
    std::array<int, num_nodes_per_cell> nodes;
    detail::check(bri17_grid_cell_nodes(DIM, shape.data(), cell, nodes.data()));
    return nodes;
  }
};

template <typename T, int DIM>
std::ostream &operator<<(std::ostream &os, const CartesianGrid<T, DIM> &grid) {
  return os << grid.repr();
}

/** Per-frequency operators of [Bri17] (include/bri17/bri17.hpp:176-356). */
template <typename T, int DIM>
  requires(std::floating_point<T> && ((DIM == 2) || (DIM == 3)))
class Hooke {
 public:
  T const mu;
  T const nu;
  CartesianGrid<T, DIM> const grid;

  Hooke(T mu, T nu, CartesianGrid<T, DIM> &grid) : mu{mu}, nu{nu}, grid{grid} {};

  /** Same text as the reference (which leaves the brace open and appends a
   * newline, bri17.hpp:197-202). */
  std::string repr() const {
    std::ostringstream stream;
    stream << "Hooke<" << typeid(T).name() << "," << DIM << ">{mu=" << mu << ",nu=" << nu
           << ",grid=" << grid << std::endl;
    return stream.str();
  }

  /** B^(k), DIM entries (bri17.hpp:212-236). */
  void modal_strain_displacement(int const *k, std::complex<T> *B) const {
    if constexpr (std::is_same_v<T, float>) {
      detail::check(bri17_modal_strain_displacement_f32(DIM, grid.shape.data(), grid.L.data(), k,
                                                        reinterpret_cast<float *>(B)));
    } else {
      std::array<double, DIM> L;
      for (int d = 0; d < DIM; d++) L[d] = (double)grid.L[d];
      std::complex<double> b[DIM];
      detail::check(bri17_modal_strain_displacement_f64(DIM, grid.shape.data(), L.data(), k,
                                                        reinterpret_cast<double *>(b)));
      for (int d = 0; d < DIM; d++) B[d] = std::complex<T>(b[d]);
    }
  }

  /** K^(k), DIM x DIM entries, row-major (bri17.hpp:247-292). */
  void modal_stiffness(int const *k, std::complex<T> *K) const {
    if constexpr (std::is_same_v<T, float>) {
      detail::check(bri17_modal_stiffness_f32(DIM, grid.shape.data(), grid.L.data(), mu, nu, k,
                                              reinterpret_cast<float *>(K)));
    } else {
      std::array<double, DIM> L;
      for (int d = 0; d < DIM; d++) L[d] = (double)grid.L[d];
      std::complex<double> kk[DIM * DIM];
      detail::check(bri17_modal_stiffness_f64(DIM, grid.shape.data(), L.data(), (double)mu,
                                              (double)nu, k, reinterpret_cast<double *>(kk)));
      for (int d = 0; d < DIM * DIM; d++) K[d] = std::complex<T>(kk[d]);
    }
  }

  /** Opposite of the strain induced by the eigenstress tau^(k), Voigt-Mandel
   * order (bri17.hpp:308-355). */
  void modal_eigenstress_to_opposite_strain(int const *k, std::complex<T> const *tau,
                                            std::complex<T> *eta) const {
    constexpr int sym = DIM == 2 ? 3 : 6;
    if constexpr (std::is_same_v<T, float>) {
      detail::check(bri17_modal_eigenstress_to_opposite_strain_f32(
          DIM, grid.shape.data(), grid.L.data(), mu, nu, k,
          reinterpret_cast<float const *>(tau), reinterpret_cast<float *>(eta)));
    } else {
      std::array<double, DIM> L;
      for (int d = 0; d < DIM; d++) L[d] = (double)grid.L[d];
      std::complex<double> t[sym], e[sym];
      for (int i = 0; i < sym; i++) t[i] = std::complex<double>(tau[i]);
      detail::check(bri17_modal_eigenstress_to_opposite_strain_f64(
          DIM, grid.shape.data(), L.data(), (double)mu, (double)nu, k,
          reinterpret_cast<double const *>(t), reinterpret_cast<double *>(e)));
      for (int i = 0; i < sym; i++) eta[i] = std::complex<T>(e[i]);
    }
  }
};

template <typename T, int DIM>
std::ostream &operator<<(std::ostream &os, const Hooke<T, DIM> &hooke) {
  return os << hooke.repr();
}

/**
 * The matrix-free periodic stiffness operator of a Hooke material on its
 * grid, on the GPU.  Fields are DIM*|N| values of T in C order
 * (DIM, N0, ..., N_{DIM-1}), the layout of the reference harness
 * (tests/test_bri17.cpp:66-67).
 */
template <typename T, int DIM>
  requires((std::is_same_v<T, float> || std::is_same_v<T, double>) && ((DIM == 2) || (DIM == 3)))
class StiffnessMatrix {
 public:
  explicit StiffnessMatrix(Hooke<T, DIM> const &hooke) {
    std::array<double, DIM> L;
    for (int d = 0; d < DIM; d++) L[d] = (double)hooke.grid.L[d];
    detail::check(bri17_operator_create(DIM, hooke.grid.shape.data(), L.data(), (double)hooke.mu,
                                        (double)hooke.nu,
                                        std::is_same_v<T, float> ? BRI17_F32 : BRI17_F64, &op_));
  }
  ~StiffnessMatrix() { bri17_operator_destroy(op_); }
  StiffnessMatrix(StiffnessMatrix const &) = delete;
  StiffnessMatrix &operator=(StiffnessMatrix const &) = delete;

  /** Number of degrees of freedom, DIM * |N|. */
  std::int64_t num_dofs() const {
    std::int64_t n = 0;
    detail::check(bri17_operator_info(op_, &n, nullptr));
    return n;
  }

  /** Ku = K u on device pointers, asynchronous on `stream` (a hipStream_t). */
  void apply(T const *u, T *Ku, void *stream = nullptr) const {
    detail::check(bri17_operator_apply(op_, u, Ku, stream));
  }

  /** u with K u = f - mean(f), mean(u) = 0, on device pointers. */
  void solve(T const *f, T *u, void *stream = nullptr) const {
    detail::check(bri17_operator_solve(op_, f, u, stream));
  }

  /** Strain eps = B u of a nodal displacement field (tests/test_bri17.cpp:194-238):
   *  u (DIM, N...) -> eps (SYM, N...) cell field, Voigt-Mandel order. */
  void strain(T const *u, T *eps, void *stream = nullptr) const {
    detail::check(bri17_operator_strain(op_, u, eps, stream));
  }

  /** Opposite of the strain induced by an eigenstress cell field:
   *  Hooke::modal_eigenstress_to_opposite_strain at every frequency. */
  void eigenstress_to_opposite_strain(T const *tau, T *eta, void *stream = nullptr) const {
    detail::check(bri17_operator_eigenstress(op_, tau, eta, stream));
  }

  /** Host-pointer variants (synchronous). */
  void apply_host(T const *u, T *Ku) const { detail::check(bri17_operator_apply_host(op_, u, Ku)); }
  void solve_host(T const *f, T *u) const { detail::check(bri17_operator_solve_host(op_, f, u)); }
  void strain_host(T const *u, T *eps) const {
    detail::check(bri17_operator_strain_host(op_, u, eps));
  }
  void eigenstress_to_opposite_strain_host(T const *tau, T *eta) const {
    detail::check(bri17_operator_eigenstress_host(op_, tau, eta));
  }

  bri17_operator *handle() const { return op_; }

 private:
  bri17_operator *op_ = nullptr;
};

}  // namespace bri17
