// The reference's "Global assembly tests" (tests/test_bri17.cpp:318-605),
// rewritten against the drop-in header include/bri17/bri17.hpp.
//
// Usage:
//   test_dropin host <golden_dir>    grid and per-frequency checks (no GPU);
//                                    prints per-k values for the Python side
//   test_dropin device <golden_dir>  2D / 3D stiffness and strain sections:
//                                    the operators probed column by column on
//                                    the GPU against the assembled Maxima Ke
//                                    and Be.
// Exit status 0 on success.
#include <cmath>
#include <complex>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "bri17/bri17.hpp"

static std::vector<double> read_values(const std::string &path) {
  std::ifstream f(path);
  std::vector<double> v;
  double x;
  while (f >> x) v.push_back(x);
  return v;
}

// tests/test_bri17.cpp:9-28 (rows x cols, row-major)
static int assert_equal(const std::vector<double> &expected, const std::vector<double> &actual,
                        int rows, int cols, double rtol, double atol, const char *what) {
  int bad = 0;
  for (int i = 0; i < rows * cols; i++) {
    double tol = rtol * std::fabs(expected[i]) + atol;
    if (!(std::fabs(actual[i] - expected[i]) <= tol)) {
      if (bad < 5)
        std::printf("%s [%d, %d]: expected = %.17g, actual = %.17g\n", what, i / cols, i % cols,
                    expected[i], actual[i]);
      bad++;
    }
  }
  std::printf("%s: %d of %d entries out of tolerance\n", what, bad, rows * cols);
  return bad;
}

// tests/test_bri17.cpp:153-174
template <int DIM>
static std::vector<double> assemble(const bri17::CartesianGrid<double, DIM> &grid,
                                    const std::vector<double> &Ke) {
  const int nn = grid.num_nodes_per_cell, ne = nn * DIM, n = DIM * grid.size;
  std::vector<double> K((size_t)n * n, 0.0);
  for (int cell = 0; cell < grid.size; cell++) {
    auto nodes = grid.get_cell_nodes(cell);
    for (int ie = 0; ie < ne; ie++) {
      int i = nodes[ie % nn] + grid.size * (ie / nn);
      for (int je = 0; je < ne; je++) {
        int j = nodes[je % nn] + grid.size * (je / nn);
        K[(size_t)i * n + j] += Ke[(size_t)ie * ne + je];
      }
    }
  }
  return K;
}

// StiffnessMatrixFactory::run (tests/test_bri17.cpp:130-150) on the device
template <int DIM>
static std::vector<double> probe(const bri17::Hooke<double, DIM> &hooke) {
  bri17::StiffnessMatrix<double, DIM> K{hooke};
  const int n = (int)K.num_dofs();
  std::vector<double> u(n, 0.0), Ku(n), M((size_t)n * n);
  for (int j = 0; j < n; j++) {
    u[j] = 1.0;
    K.apply_host(u.data(), Ku.data());
    for (int i = 0; i < n; i++) M[(size_t)i * n + j] = Ku[i];
    u[j] = 0.0;
  }
  return M;
}

// tests/test_bri17.cpp:295-316
template <int DIM>
static std::vector<double> assemble_strain(const bri17::CartesianGrid<double, DIM> &grid,
                                           const std::vector<double> &Be) {
  const int nn = grid.num_nodes_per_cell, ne = nn * DIM, S = DIM * (DIM + 1) / 2;
  const int rows = S * grid.size, cols = DIM * grid.size;
  std::vector<double> B((size_t)rows * cols, 0.0);
  for (int cell = 0; cell < grid.size; cell++) {
    auto nodes = grid.get_cell_nodes(cell);
    for (int il = 0; il < S; il++)
      for (int je = 0; je < ne; je++) {
        int j = nodes[je % nn] + grid.size * (je / nn);
        B[(size_t)(il * grid.size + cell) * cols + j] += Be[(size_t)il * ne + je];
      }
  }
  return B;
}

// StrainDisplacementMatrixFactory::run (tests/test_bri17.cpp:263-285) on the device
template <int DIM>
static std::vector<double> probe_strain(const bri17::Hooke<double, DIM> &hooke) {
  bri17::StiffnessMatrix<double, DIM> K{hooke};
  const int n = (int)K.num_dofs(), rows = DIM * (DIM + 1) / 2 * hooke.grid.size;
  std::vector<double> u(n, 0.0), eps(rows), M((size_t)rows * n);
  for (int j = 0; j < n; j++) {
    u[j] = 1.0;
    K.strain_host(u.data(), eps.data());
    for (int i = 0; i < rows; i++) M[(size_t)i * n + j] = eps[i];
    u[j] = 0.0;
  }
  return M;
}

int main(int argc, char **argv) {
  if (argc < 3) {
    std::printf("usage: %s host|device <golden_dir>\n", argv[0]);
    return 2;
  }
  const std::string mode = argv[1], dir = argv[2];
  // tests/test_bri17.cpp:319-333
  const double mu = 5.6, nu = 0.3;
  bri17::CartesianGrid<double, 2> grid2{{3, 4}, {3 * 1.1, 4 * 1.2}};
  bri17::CartesianGrid<double, 3> grid3{{3, 4, 5}, {3 * 1.1, 4 * 1.2, 5 * 1.3}};
  bri17::Hooke<double, 2> hooke2{mu, nu, grid2};
  bri17::Hooke<double, 3> hooke3{mu, nu, grid3};

  if (mode == "host") {
    std::printf("repr %s\n", grid3.repr().c_str());
    std::printf("node %d %d\n", grid2.get_node_at(2, 3), grid3.get_node_at(2, 3, 4));
    for (int cell : {0, 17, 59}) {
      auto nodes = grid3.get_cell_nodes(cell);
      std::printf("cell %d", cell);
      for (int x : nodes) std::printf(" %d", x);
      std::printf("\n");
    }
    for (int k0 = 0; k0 < 3; k0++)
      for (int k2 = 0; k2 < 5; k2 += 2) {
        int k[3] = {k0, 1, k2};
        std::complex<double> K[9], B[3];
        hooke3.modal_stiffness(k, K);
        hooke3.modal_strain_displacement(k, B);
        std::printf("K %d %d %d", k[0], k[1], k[2]);
        for (auto z : K) std::printf(" %.17g %.17g", z.real(), z.imag());
        std::printf("\nB %d %d %d", k[0], k[1], k[2]);
        for (auto z : B) std::printf(" %.17g %.17g", z.real(), z.imag());
        std::printf("\n");
      }
    bri17::CartesianGrid<float, 3> gf{{3, 4, 5}, {3.3f, 4.8f, 6.5f}};
    bri17::Hooke<float, 3> hf{5.6f, 0.3f, gf};
    int k[3] = {1, 2, 3};
    std::complex<float> Kf[9];
    hf.modal_stiffness(k, Kf);
    std::printf("Kf %.9g %.9g\n", Kf[0].real(), Kf[1].real());
    return 0;
  }

  int bad = 0;
  {
    auto K_exp = assemble<2>(grid2, read_values(dir + "/Ke2.txt"));
    auto K_act = probe<2>(hooke2);
    bad += assert_equal(K_exp, K_act, 2 * grid2.size, 2 * grid2.size, 1e-15, 1e-14,
                        "2D stiffness matrix");
  }
  {
    auto K_exp = assemble<3>(grid3, read_values(dir + "/Ke3.txt"));
    auto K_act = probe<3>(hooke3);
    bad += assert_equal(K_exp, K_act, 3 * grid3.size, 3 * grid3.size, 1e-15, 1e-14,
                        "3D stiffness matrix");
  }
  // tests/test_bri17.cpp:538-605 (rtol 1e-15, atol 1e-14, :558,604)
  {
    auto B_exp = assemble_strain<2>(grid2, read_values(dir + "/Be2.txt"));
    auto B_act = probe_strain<2>(hooke2);
    bad += assert_equal(B_exp, B_act, 3 * grid2.size, 2 * grid2.size, 1e-15, 1e-14,
                        "2D strain-displacement matrix");
  }
  {
    auto B_exp = assemble_strain<3>(grid3, read_values(dir + "/Be3.txt"));
    auto B_act = probe_strain<3>(hooke3);
    bad += assert_equal(B_exp, B_act, 6 * grid3.size, 3 * grid3.size, 1e-15, 1e-14,
                        "3D strain-displacement matrix");
  }
  return bad ? 1 : 0;
}
