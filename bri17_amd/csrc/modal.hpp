// The modal DIM x DIM block of the periodic stiffness operator at one
// frequency (include/bri17/bri17.hpp:255-288), shared by the generic and the
// fixed-size pencil kernels.
//
// The scalar factor of the apply (|h|/|N|) or of the Green solve
// (1/(|h||N|)) is folded into the material constants by the host:
//   apply: mu' = mu c, s' = s c          ->  y = K'(mu', s') x = c K x
//   solve: mu' = mu / c, s' = s / c      ->  y = K'^-1 x = c K^-1 x
// with s = mu / (1 - 2 nu) = lambda + mu.
#pragma once

#include "fft_core.hpp"
#include "kernels.hpp"

namespace b17 {

// y = K x (MODE 0) or y = K^-1 x (MODE 1, y = 0 when `null`) for the real
// symmetric 3x3 block K = [[K00,K01,K02],[K01,K11,K12],[K02,K12,K22]].
template <typename T, int MODE>
__device__ __forceinline__ void sym3_apply(T K00, T K11, T K22, T K01, T K02, T K12, bool null,
                                           cpx<T>* x) {
  if constexpr (MODE == 0) {
    const cpx<T> y0 = x[0] * K00 + x[1] * K01 + x[2] * K02;
    const cpx<T> y1 = x[0] * K01 + x[1] * K11 + x[2] * K12;
    const cpx<T> y2 = x[0] * K02 + x[1] * K12 + x[2] * K22;
    x[0] = y0;
    x[1] = y1;
    x[2] = y2;
  } else {
    if (null) {
      x[0] = x[1] = x[2] = cpx<T>{T(0), T(0)};
      return;
    }
    // adjugate of the SPD block (closed-form inverse)
    const T A00 = K11 * K22 - K12 * K12;
    const T A01 = K02 * K12 - K01 * K22;
    const T A02 = K01 * K12 - K02 * K11;
    const T A11 = K00 * K22 - K02 * K02;
    const T A12 = K01 * K02 - K00 * K12;
    const T A22 = K00 * K11 - K01 * K01;
    const T r = T(1) / (K00 * A00 + K01 * A01 + K02 * A02);
    const cpx<T> y0 = (x[0] * A00 + x[1] * A01 + x[2] * A02) * r;
    const cpx<T> y1 = (x[0] * A01 + x[1] * A11 + x[2] * A12) * r;
    const cpx<T> y2 = (x[0] * A02 + x[1] * A12 + x[2] * A22) * r;
    x[0] = y0;
    x[1] = y1;
    x[2] = y2;
  }
}

template <typename T, int MODE>
__device__ __forceinline__ void sym2_apply(T K00, T K11, T K01, bool null, cpx<T>* x) {
  if constexpr (MODE == 0) {
    const cpx<T> y0 = x[0] * K00 + x[1] * K01;
    const cpx<T> y1 = x[0] * K01 + x[1] * K11;
    x[0] = y0;
    x[1] = y1;
  } else {
    if (null) {
      x[0] = x[1] = cpx<T>{T(0), T(0)};
      return;
    }
    const T r = T(1) / (K00 * K11 - K01 * K01);
    const cpx<T> y0 = (x[0] * K11 - x[1] * K01) * r;
    const cpx<T> y1 = (x[1] * K00 - x[0] * K01) * r;
    x[0] = y0;
    x[1] = y1;
  }
}

// Block at k = (k0, k1, kl) from the per-axis tables phi, chi, psi:
//   H_ii = phi_i prod_{m != i} chi_m,   K_ii = s H_ii + mu sum_m H_mm,
//   K_ij = s psi_i psi_j prod_{m not in {i,j}} chi_m   (2D: no chi product).
template <typename T, int DIM, int MODE, typename CP>
__device__ __forceinline__ void modal_block(const CP* coef, int k0, int k1,
                                            int kl, T mu, T s, cpx<T>* x) {
  if constexpr (DIM == 2) {
    const ModalCoef<T> a = coef[0][k0], b = coef[1][kl];
    const T H0 = a.phi * b.chi, H1 = a.chi * b.phi;
    const T dg = mu * (H0 + H1);
    sym2_apply<T, MODE>(s * H0 + dg, s * H1 + dg, s * a.psi * b.psi, k0 == 0 && kl == 0, x);
  } else {
    const ModalCoef<T> a = coef[0][k0], b = coef[1][k1], c = coef[2][kl];
    const T H0 = a.phi * b.chi * c.chi;
    const T H1 = a.chi * b.phi * c.chi;
    const T H2 = a.chi * b.chi * c.phi;
    const T dg = mu * (H0 + H1 + H2);
    sym3_apply<T, MODE>(s * H0 + dg, s * H1 + dg, s * H2 + dg, s * a.psi * b.psi * c.chi,
                        s * a.psi * b.chi * c.psi, s * a.chi * b.psi * c.psi,
                        k0 == 0 && k1 == 0 && kl == 0, x);
  }
}

}  // namespace b17
