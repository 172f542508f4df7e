// The modal DIM x DIM block of the periodic stiffness operator at one
// frequency (include/bri17/bri17.hpp:255-288), shared by the generic and the
// fixed-size pencil kernels.
//
// The scalar factor of the apply (|h|/|N|) or of the Green solve
// (1/(|h||N|)) is folded into the material constants by the host:
//   apply: mu' = mu c, s' = s c          ->  y = K'(mu', s') x = c K x
//   solve: mu' = mu / (c k), s' = s / (c k), kinv = 1 / k
//                                        ->  y = kinv K'^-1 x = c K^-1 x
// with s = mu / (1 - 2 nu) = lambda + mu.  The normalisation of the
// inverted modes, k = (mu / c + s / c) 4 / h_min^2, brings the block's
// entries to O(1) whatever the units of mu and L: its determinant (cubic in
// the entries) would otherwise leave the fp32 range for mu above about 1e7 on
// a 256^3 unit cube.  The eigenstress mode (3) is normalised the same way.
#pragma once

#include "fft_core.hpp"
#include "kernels.hpp"

namespace b17 {

// 1/d for the block inverse.  fp32: the hardware reciprocal (v_rcp_f32,
// 1 ulp) refined by one Newton-Raphson step, e = 1 - d r, r' = r + r e (both
// fused), which leaves an error within about 1 ulp without the
// scale/fixup sequence of an IEEE division.  d is the determinant of an SPD
// block away from k = 0 (normal range; k = 0 is handled before).  fp64: the
// division.
template <typename T>
__device__ __forceinline__ T block_recip(T d) {
  if constexpr (sizeof(T) == 4) {
    const float r = __builtin_amdgcn_rcpf(d);
    return __builtin_fmaf(r, __builtin_fmaf(-d, r, 1.0f), r);
  } else {
    return T(1) / d;
  }
}

// y = K x (MODE 0) or y = kinv K^-1 x (MODE 1, y = 0 when `null`) for the
// real symmetric 3x3 block K = [[K00,K01,K02],[K01,K11,K12],[K02,K12,K22]].
template <typename T, int MODE>
__device__ __forceinline__ void sym3_apply(T K00, T K11, T K22, T K01, T K02, T K12, bool null,
                                           cpx<T>* x, T kinv = T(1)) {
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
    const T r = block_recip(K00 * A00 + K01 * A01 + K02 * A02) * kinv;
    const cpx<T> y0 = (x[0] * A00 + x[1] * A01 + x[2] * A02) * r;
    const cpx<T> y1 = (x[0] * A01 + x[1] * A11 + x[2] * A12) * r;
    const cpx<T> y2 = (x[0] * A02 + x[1] * A12 + x[2] * A22) * r;
    x[0] = y0;
    x[1] = y1;
    x[2] = y2;
  }
}

template <typename T, int MODE>
__device__ __forceinline__ void sym2_apply(T K00, T K11, T K01, bool null, cpx<T>* x,
                                           T kinv = T(1)) {
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
    const T r = block_recip(K00 * K11 - K01 * K01) * kinv;
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
                                            int kl, T mu, T s, cpx<T>* x, T kinv = T(1)) {
  if constexpr (DIM == 2) {
    const ModalCoef<T> a = coef[0][k0], b = coef[1][kl];
    const T H0 = a.phi * b.chi, H1 = a.chi * b.phi;
    const T dg = mu * (H0 + H1);
    sym2_apply<T, MODE>(s * H0 + dg, s * H1 + dg, s * a.psi * b.psi, k0 == 0 && kl == 0, x, kinv);
  } else {
    const ModalCoef<T> a = coef[0][k0], b = coef[1][k1], c = coef[2][kl];
    const T H0 = a.phi * b.chi * c.chi;
    const T H1 = a.chi * b.phi * c.chi;
    const T H2 = a.chi * b.chi * c.phi;
    const T dg = mu * (H0 + H1 + H2);
    sym3_apply<T, MODE>(s * H0 + dg, s * H1 + dg, s * H2 + dg, s * a.psi * b.psi * c.chi,
                        s * a.psi * b.chi * c.psi, s * a.chi * b.psi * c.psi,
                        k0 == 0 && k1 == 0 && kl == 0, x, kinv);
  }
}

// exp(i sum a) and the real factors beta of B^ = 2i exp(i sum a) beta at k.
template <typename T, int DIM, typename SP>
__device__ __forceinline__ cpx<T> strain_factors(const SP* sc, int k0, int k1, int kl, T* beta) {
  if constexpr (DIM == 2) {
    const StrainCoef<T> a = sc[0][k0], b = sc[1][kl];
    beta[0] = a.sh * b.c;
    beta[1] = a.c * b.sh;
    return cmul(cpx<T>{a.c, a.sn}, cpx<T>{b.c, b.sn});
  } else {
    const StrainCoef<T> a = sc[0][k0], b = sc[1][k1], c = sc[2][kl];
    beta[0] = a.sh * b.c * c.c;
    beta[1] = a.c * b.sh * c.c;
    beta[2] = a.c * b.c * c.sh;
    return cmul(cmul(cpx<T>{a.c, a.sn}, cpx<T>{b.c, b.sn}), cpx<T>{c.c, c.sn});
  }
}

// Voigt-Mandel order of the symmetric tensor components
// ([xx, yy, sqrt2 xy] / [xx, yy, zz, sqrt2 yz, sqrt2 zx, sqrt2 xy],
// tests/test_bri17.cpp:209-228, bri17.hpp:324-353)
template <int DIM>
struct Voigt {
  static constexpr int n = DIM == 2 ? 3 : 6;
  static constexpr int a[6] = {0, 1, DIM == 2 ? 0 : 2, 1, 2, 0};
  static constexpr int b[6] = {0, 1, DIM == 2 ? 1 : 2, 2, 0, 1};
  __host__ __device__ static constexpr bool shear(int i) { return DIM == 2 ? i == 2 : i >= 3; }
};

// Mode 2: eps^ = sym(B^ u^) (tests/test_bri17.cpp:196-228) times bscale, in
// Voigt-Mandel order.  With B^ = 2i E beta: eps_ab = i E (beta_a u_b + beta_b u_a).
template <typename T, int DIM, typename SP>
__device__ __forceinline__ void strain_block(const SP* sc, int k0, int k1, int kl, T bscale,
                                             cpx<T>* x) {
  constexpr T r2 = T(1.4142135623730950488016887242096981L);
  T beta[3];
  const cpx<T> E = strain_factors<T, DIM>(sc, k0, k1, kl, beta);
  const cpx<T> Z = {-E.y * bscale, E.x * bscale};  // i E bscale
  cpx<T> u[DIM];
#pragma unroll
  for (int c = 0; c < DIM; c++) u[c] = x[c];
#pragma unroll
  for (int i = 0; i < Voigt<DIM>::n; i++) {
    const int a = Voigt<DIM>::a[i], b = Voigt<DIM>::b[i];
    const cpx<T> w = u[b] * beta[a] + u[a] * beta[b];
    x[i] = cmul(Z, Voigt<DIM>::shear(i) ? w * r2 : w);
  }
}

// Mode 3: bri17.hpp:308-355 at k, times bscale.  rhs = tau_mat conj(B^),
// u = K^-1 rhs, eta = sym(B^ u); with B^ = 2i E beta and |E| = 1:
// eta_ab = 2 (beta_a v_b + beta_b v_a), v = K^-1 (tau_mat beta).  Zero at k = 0.
// The host passes the normalised block (mu / k, s / k) and kinv = 1 / k.
template <typename T, int DIM, typename CP, typename SP>
__device__ __forceinline__ void eigenstress_block(const CP* coef, const SP* sc, int k0, int k1,
                                                  int kl, T mu, T s, T bscale, T kinv, cpx<T>* x) {
  constexpr T r2 = T(1.4142135623730950488016887242096981L);
  constexpr T ir2 = T(0.70710678118654752440084436210484903L);
  constexpr int S = Voigt<DIM>::n;
  const bool null = DIM == 2 ? (k0 == 0 && kl == 0) : (k0 == 0 && k1 == 0 && kl == 0);
  T beta[3];
  (void)strain_factors<T, DIM>(sc, k0, k1, kl, beta);
  // v = tau_mat beta
  cpx<T> t[3][3];
#pragma unroll
  for (int i = 0; i < S; i++) {
    const int a = Voigt<DIM>::a[i], b = Voigt<DIM>::b[i];
    const cpx<T> v = Voigt<DIM>::shear(i) ? x[i] * ir2 : x[i];
    t[a][b] = v;
    t[b][a] = v;
  }
  cpx<T> v[DIM];
#pragma unroll
  for (int a = 0; a < DIM; a++) {
    cpx<T> acc = t[a][0] * beta[0];
#pragma unroll
    for (int b = 1; b < DIM; b++) acc = acc + t[a][b] * beta[b];
    v[a] = acc;
  }
  modal_block<T, DIM, 1>(coef, k0, k1, kl, mu, s, v, kinv);
  if (null) {
#pragma unroll
    for (int i = 0; i < S; i++) x[i] = cpx<T>{T(0), T(0)};
    return;
  }
  const T f = T(2) * bscale;
#pragma unroll
  for (int i = 0; i < S; i++) {
    const int a = Voigt<DIM>::a[i], b = Voigt<DIM>::b[i];
    const cpx<T> w = v[b] * beta[a] + v[a] * beta[b];
    x[i] = Voigt<DIM>::shear(i) ? w * (f * r2) : w * f;
  }
}

}  // namespace b17
