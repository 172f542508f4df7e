// Kernel parameter blocks and launchers of the periodic stiffness operator.
//
// Pipeline for a field u of shape (DIM, N0, ..., NL) in C order (the layout of
// tests/test_bri17.cpp:66-67,81-83):
//
//   plane_fwd  : R2C along the last axis (+ in 3D the first Cooley-Tukey stage
//                of the middle axis), writes the half-spectrum workspace
//   pencil     : remaining middle-axis stage, FFT along axis 0, the modal
//                DIM x DIM block (K^ multiply or K^-1 solve, bri17.hpp:247-292),
//                inverse FFT along axis 0, inverse middle-axis stage
//   plane_inv  : inverse middle-axis stage + C2R along the last axis
//
// Workspace layout (complex T): ws[c][i0][p][ka][k] with middle index
// i1 = p*Q + ka (3D) and k in [0, rs), rs = NL/2+1; 2D: ws[c][i0][k].
// The fixed-size kernels (kernels_static.hip) store k < NL/2 in rows of
// NL/2 entries and the Nyquist entries k = NL/2 in a region of their own
// after them (same total size).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bri17_types.hpp"

namespace b17 {

// Per-axis modal coefficient table entry (bri17.hpp:255-264), precomputed in
// fp64 on the host with the half-angle form phi = 4 sin^2(pi k/N)/h^2.
template <typename T>
struct ModalCoef {
  T phi, chi, psi, pad;
};

// Per-axis strain-displacement factors (bri17.hpp:212-236): with
// a = pi k / N and h = L / N, c = cos a, sn = sin a, sh = sin a / h.  The
// vector B^(k) = 2i exp(i sum a) (sh_0 c_1 c_2, c_0 sh_1 c_2, c_0 c_1 sh_2)
// (2D: (sh_0 c_1, c_0 sh_1)) is built from DIM entries, exp(i sum a) as the
// product of the (c + i sn).
template <typename T>
struct StrainCoef {
  T c, sn, sh, pad;
};

// Pencil modes: 0 K u, 1 Green solve (DIM -> DIM components); 2 strain of a
// displacement field (DIM -> SYM Voigt-Mandel components); 3 opposite strain
// induced by an eigenstress field (SYM -> SYM).
template <int DIM, int MODE>
struct ModeComps {
  static constexpr int sym = DIM == 2 ? 3 : 6;
  static constexpr int cin = MODE == 3 ? sym : DIM;
  static constexpr int cout = MODE >= 2 ? sym : DIM;
  static constexpr int cmax = cin > cout ? cin : cout;
};
__host__ __device__ constexpr int mode_cin(int dim, int mode) {
  return mode == 3 ? (dim == 2 ? 3 : 6) : dim;
}
__host__ __device__ constexpr int mode_cout(int dim, int mode) {
  return mode >= 2 ? (dim == 2 ? 3 : 6) : dim;
}

struct PlaneParams {
  int dim;
  int nl;        // row length (last axis)
  int nlh;       // nl/2 + 1 stored spectrum entries per row
  int even;      // nl even: packed half-length complex FFT
  int m;         // row FFT length: nl/2 (even) or nl (odd)
  int rsl;       // LDS row stride (complex elements)
  int q;         // rows per workgroup
  int p;         // middle-axis split factor (1 in 2D)
  int n1;        // 3D: middle axis length; 2D: rows per workgroup
  int has_col;   // 3D: do the Q-point middle-axis stage
  int rs;        // workspace row stride (complex elements)
  int nrows;     // 2D: total number of rows (components * N0); 3D unused
  int g3;        // 3D, split 1: consecutive planes per work item (1 outside
                 // the batch kernel)
  int nplanes;   // 3D with g3 > 1: planes of the field (components * N0)
  FastDiv mdiv;  // division by m
  FastDiv hdiv;  // division by m/2+1 (pairs per row)
  FastDiv nlhdiv;
  FftDesc row;   // length m
  FftDesc col;   // length q (3D)
  const void* tw_row;  // exp(-2 pi i j / nl), j < nl
  const void* tw_col;  // exp(-2 pi i j / n1), j < n1 (3D)
};

struct PencilParams {
  int dim;
  int n0;
  int p;        // middle-axis split factor
  int q;        // N1 / p (3D); 1 (2D)
  int rs;       // workspace row stride
  int nlh;      // valid entries per row
  int f;        // flat extent of one p-block: q * rs
  int w;        // tile width (contiguous complex elements)
  int ntiles;
  int nyq_tiles;    // fixed-size layout: the last nyq_tiles tiles walk the
                    // Nyquist region (0: Nyquist entries in-row)
  int64_t nyq_off;  // element offset of the Nyquist region
  int mode;     // 0: K u, 1: K^-1 f, 2: strain B u, 3: eigenstress -> -strain
  int xcd_remap;
  double mu, s;   // material constants with the apply / solve scale folded in
                  // (modes 1 and 3: also divided by the block normalisation)
  double kinv;    // modes 1 and 3: factor of the normalised block's inverse
  double scale;   // |h|/|N| (apply) or 1/(|h||N|) (solve), for reference
  double bscale;  // modes 2 and 3: factor of the strain output (1/|N|, 2/|N|)
  FastDiv wdiv, pwdiv, rsdiv;
  FftDesc ax0;
  FftDesc pfft;         // P-point transform of a split p without a fused radix
                        // kernel (twiddle stride q into tw1)
  const void* tw0;      // exp(-2 pi i j / n0)
  const void* tw1;      // exp(-2 pi i j / n1) (3D)
  const void* coef[3];  // ModalCoef<T>[N_d] per axis
  const void* scoef[3]; // StrainCoef<T>[N_d] per axis
};

// Splits whose P-point DFT is one fused radix butterfly per column.
__host__ __device__ constexpr bool fused_split(int p) {
  return p == 1 || p == 2 || p == 3 || p == 4 || p == 5 || p == 7 || p == 8 || p == 16;
}

// Prime-radix tier of a transform: 0 none, 1 a radix-11 or radix-13 pass, 2
// a direct-DFT pass of a prime above 16.  Each tier runs in its own
// instantiation of the runtime-size kernels (template parameter PR of their
// bodies), so that those passes do not raise the register budget of the
// others.
__host__ __device__ inline int fft_prime_tier(const FftDesc& d) {
  int t = 0;
  for (int i = 0; i < d.npass; i++) {
    const int r = d.pass[i].r;
    if (r > 16) t = 2;
    else if ((r == 11 || r == 13) && t < 1) t = 1;
  }
  return t;
}
__host__ __device__ inline int plane_prime_tier(const PlaneParams& pp) {
  const int a = fft_prime_tier(pp.row), b = pp.has_col ? fft_prime_tier(pp.col) : 0;
  return a > b ? a : b;
}
__host__ __device__ inline int pencil_prime_tier(const PencilParams& pc) {
  const int a = fft_prime_tier(pc.ax0);
  const int b = pc.p > 1 && !fused_split(pc.p) ? fft_prime_tier(pc.pfft) : 0;
  return a > b ? a : b;
}

// Entries (complex T) of the twiddle tables a work item copies into LDS after
// its data: plane items the row table (nl) and in 3D the middle-axis table
// (n1); pencil tiles the axis-0 table (n0) and, for a split P without a fused
// butterfly, the middle-axis table (P q = n1).
__host__ __device__ inline int plane_table_entries(const PlaneParams& pp) {
  return pp.nl + (pp.has_col ? pp.n1 : 0);
}
__host__ __device__ inline int pencil_table_entries(const PencilParams& pc) {
  return pc.n0 + ((pc.p > 1 && !fused_split(pc.p)) ? pc.p * pc.q : 0);
}

template <typename T>
hipError_t launch_plane_fwd(const T* u, void* ws, const PlaneParams& pp,
                            int nblocks, int nthreads, size_t lds,
                            hipStream_t st);
template <typename T>
hipError_t launch_plane_inv(const void* ws, T* out, const PlaneParams& pp,
                            int nblocks, int nthreads, size_t lds,
                            hipStream_t st);
template <typename T>
hipError_t launch_pencil(void* ws, const PencilParams& pp, int nthreads,
                         size_t lds, hipStream_t st);

}  // namespace b17
