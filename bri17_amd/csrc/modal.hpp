// The modal DIM x DIM block of the periodic stiffness operator at one
// frequency, shared by the generic and the fixed-size pencil kernels.
#pragma once

#include "fft_core.hpp"
#include "kernels.hpp"

namespace b17 {

// The modal DIM x DIM block at frequency k (bri17.hpp:255-288), applied to x
// (MODE 0: y = scale K^ x) or solved against x (MODE 1: y = scale K^-1 x,
// y = 0 at k = 0).  K^ is real symmetric.
template <typename T, int DIM, int MODE>
__device__ __forceinline__ void modal_block(const ModalCoef<T>* const* coef, int k0,
                                            int k1, int kl, T mu, T s, T scale,
                                            cpx<T>* x) {
  if constexpr (DIM == 2) {
    const ModalCoef<T> a = coef[0][k0], b = coef[1][kl];
    const T H0 = a.phi * b.chi, H1 = a.chi * b.phi;
    const T dg = mu * (H0 + H1);
    const T K00 = s * H0 + dg, K11 = s * H1 + dg, K01 = s * a.psi * b.psi;
    if constexpr (MODE == 0) {
      const cpx<T> y0 = (x[0] * K00 + x[1] * K01) * scale;
      const cpx<T> y1 = (x[0] * K01 + x[1] * K11) * scale;
      x[0] = y0;
      x[1] = y1;
    } else {
      if (k0 == 0 && kl == 0) {
        x[0] = {T(0), T(0)};
        x[1] = {T(0), T(0)};
        return;
      }
      const T r = scale / (K00 * K11 - K01 * K01);
      const cpx<T> y0 = (x[0] * K11 - x[1] * K01) * r;
      const cpx<T> y1 = (x[1] * K00 - x[0] * K01) * r;
      x[0] = y0;
      x[1] = y1;
    }
  } else {
    const ModalCoef<T> a = coef[0][k0], b = coef[1][k1], c = coef[2][kl];
    const T H0 = a.phi * b.chi * c.chi;
    const T H1 = a.chi * b.phi * c.chi;
    const T H2 = a.chi * b.chi * c.phi;
    const T dg = mu * (H0 + H1 + H2);
    const T K00 = s * H0 + dg, K11 = s * H1 + dg, K22 = s * H2 + dg;
    const T K01 = s * a.psi * b.psi * c.chi;
    const T K02 = s * a.psi * b.chi * c.psi;
    const T K12 = s * a.chi * b.psi * c.psi;
    if constexpr (MODE == 0) {
      const cpx<T> y0 = (x[0] * K00 + x[1] * K01 + x[2] * K02) * scale;
      const cpx<T> y1 = (x[0] * K01 + x[1] * K11 + x[2] * K12) * scale;
      const cpx<T> y2 = (x[0] * K02 + x[1] * K12 + x[2] * K22) * scale;
      x[0] = y0;
      x[1] = y1;
      x[2] = y2;
    } else {
      if (k0 == 0 && k1 == 0 && kl == 0) {
        x[0] = x[1] = x[2] = {T(0), T(0)};
        return;
      }
      // adjugate of the symmetric SPD block
      const T A00 = K11 * K22 - K12 * K12;
      const T A01 = K02 * K12 - K01 * K22;
      const T A02 = K01 * K12 - K02 * K11;
      const T A11 = K00 * K22 - K02 * K02;
      const T A12 = K01 * K02 - K00 * K12;
      const T A22 = K00 * K11 - K01 * K01;
      const T r = scale / (K00 * A00 + K01 * A01 + K02 * A02);
      const cpx<T> y0 = (x[0] * A00 + x[1] * A01 + x[2] * A02) * r;
      const cpx<T> y1 = (x[0] * A01 + x[1] * A11 + x[2] * A12) * r;
      const cpx<T> y2 = (x[0] * A02 + x[1] * A12 + x[2] * A22) * r;
      x[0] = y0;
      x[1] = y1;
      x[2] = y2;
    }
  }
}

}  // namespace b17
