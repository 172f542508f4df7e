// HIP kernels of the matrix-free periodic stiffness operator (gfx950).
//
// Reference behaviour replaced: the FFT -> per-k K^(k).u^(k) -> inverse FFT
// -> |h|/|N| scaling of StiffnessMatrixFactory::compute_Ku
// (tests/test_bri17.cpp:56-107) with the modal stiffness of
// Hooke::modal_stiffness (include/bri17/bri17.hpp:247-292), and the Green
// solve with the k = 0 convention of sphinx/theory.rst:208-212.
//
// The transforms exploit that u is real and K^(-k) = K^(k): only the half
// spectrum along the last axis is formed (R2C / C2R), which is exact.
#include "fft_core.hpp"
#include "kernels.hpp"
#include "kernels_generic.hpp"

namespace b17 {


// Plain launches, one workgroup per plane block / pencil tile.  PR: the
// prime-radix tier compiled in (fft_prime_tier), chosen from the plan.
template <typename T, int PR>
__global__ void __launch_bounds__(1024)
    k_plane_fwd(const T* __restrict__ u, cpx<T>* __restrict__ ws, const PlaneParams pp) {
  extern __shared__ __align__(16) unsigned char smem[];
  plane_fwd_body<T, PR>((gptr<const T>)u, (gptr<cpx<T>>)ws, pp, blockIdx.x,
                        reinterpret_cast<cpx<T>*>(smem));
}

template <typename T, int PR>
__global__ void __launch_bounds__(1024)
    k_plane_inv(const cpx<T>* __restrict__ ws, T* __restrict__ out, const PlaneParams pp) {
  extern __shared__ __align__(16) unsigned char smem[];
  plane_inv_body<T, PR>((gptr<const cpx<T>>)ws, (gptr<T>)out, pp, blockIdx.x,
                        reinterpret_cast<cpx<T>*>(smem));
}

template <typename T, int DIM, int MODE, int PR>
__global__ void __launch_bounds__(1024) k_pencil(cpx<T>* __restrict__ ws, const PencilParams pp) {
  extern __shared__ __align__(16) unsigned char smem[];
  int t = blockIdx.x;
  if (pp.xcd_remap) {
    // consecutive tiles share L2 lines: keep them on one XCD (blocks are
    // dealt round-robin over the 8 XCDs)
    const int n = gridDim.x, per = n / 8, rem = n % 8;
    const int x = t % 8, idx = t / 8;
    t = x * per + (x < rem ? x : rem) + idx;
  }
  pencil_body<T, DIM, MODE, PR>((gptr<cpx<T>>)ws, pp, t, reinterpret_cast<cpx<T>*>(smem));
}

// ------------------------------------------------------------------ launchers

template <typename K>
static hipError_t set_lds(K kernel, size_t lds) {
  if (lds > 65536)
    return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  return hipSuccess;
}

template <typename T>
hipError_t launch_plane_fwd(const T* u, void* ws, const PlaneParams& pp, int nblocks,
                            int nthreads, size_t lds, hipStream_t st) {
  const int pr = plane_prime_tier(pp);
  auto k = pr == 2 ? k_plane_fwd<T, 2> : pr == 1 ? k_plane_fwd<T, 1> : k_plane_fwd<T, 0>;
  hipError_t e = set_lds(k, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3(nblocks), dim3(nthreads), lds, st, u,
                     reinterpret_cast<cpx<T>*>(ws), pp);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_plane_inv(const void* ws, T* out, const PlaneParams& pp, int nblocks,
                            int nthreads, size_t lds, hipStream_t st) {
  const int pr = plane_prime_tier(pp);
  auto k = pr == 2 ? k_plane_inv<T, 2> : pr == 1 ? k_plane_inv<T, 1> : k_plane_inv<T, 0>;
  hipError_t e = set_lds(k, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3(nblocks), dim3(nthreads), lds, st,
                     reinterpret_cast<const cpx<T>*>(ws), out, pp);
  return hipGetLastError();
}

template <typename T, int DIM, int MODE>
static hipError_t launch_pencil_t(void* ws, const PencilParams& pp, int nthreads, size_t lds,
                                  hipStream_t st) {
  const int pr = pencil_prime_tier(pp);
  auto k = pr == 2   ? k_pencil<T, DIM, MODE, 2>
           : pr == 1 ? k_pencil<T, DIM, MODE, 1>
                     : k_pencil<T, DIM, MODE, 0>;
  hipError_t e = set_lds(k, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3(pp.ntiles), dim3(nthreads), lds, st,
                     reinterpret_cast<cpx<T>*>(ws), pp);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_pencil(void* ws, const PencilParams& pp, int nthreads, size_t lds,
                         hipStream_t st) {
  switch (pp.dim * 4 + pp.mode) {
    case 8: return launch_pencil_t<T, 2, 0>(ws, pp, nthreads, lds, st);
    case 9: return launch_pencil_t<T, 2, 1>(ws, pp, nthreads, lds, st);
    case 10: return launch_pencil_t<T, 2, 2>(ws, pp, nthreads, lds, st);
    case 11: return launch_pencil_t<T, 2, 3>(ws, pp, nthreads, lds, st);
    case 12: return launch_pencil_t<T, 3, 0>(ws, pp, nthreads, lds, st);
    case 13: return launch_pencil_t<T, 3, 1>(ws, pp, nthreads, lds, st);
    case 14: return launch_pencil_t<T, 3, 2>(ws, pp, nthreads, lds, st);
    case 15: return launch_pencil_t<T, 3, 3>(ws, pp, nthreads, lds, st);
    default: return hipErrorInvalidValue;
  }
}

template hipError_t launch_plane_fwd<float>(const float*, void*, const PlaneParams&, int, int,
                                            size_t, hipStream_t);
template hipError_t launch_plane_fwd<double>(const double*, void*, const PlaneParams&, int, int,
                                             size_t, hipStream_t);
template hipError_t launch_plane_inv<float>(const void*, float*, const PlaneParams&, int, int,
                                            size_t, hipStream_t);
template hipError_t launch_plane_inv<double>(const void*, double*, const PlaneParams&, int, int,
                                             size_t, hipStream_t);
template hipError_t launch_pencil<float>(void*, const PencilParams&, int, size_t, hipStream_t);
template hipError_t launch_pencil<double>(void*, const PencilParams&, int, size_t, hipStream_t);

}  // namespace b17
