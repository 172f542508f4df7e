// Persistent batched launch over many operators (see batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace b17 {

// Device-side descriptor of one batch instance.
struct BatchInst {
  PlaneParams pp;
  PencilParams pc;  // mu and s already folded with the mode's scale
  const void* in;
  void* out;
  void* ws;
  int na;  // plane blocks (forward and inverse each)
  int nb;  // pencil tiles
};

// Work item encoding: x = inst | type << 16 (0 plane fwd, 1 pencil,
// 2 plane inv), y = block / tile index within the instance.  n_fwd, n_pen:
// forward items and pencil tiles of all instances.
// ctl: [head, error, forward items done, pencil tiles done, done(2 per
// instance)], nctl = 4 + 2 * ninst.
// primes: the highest prime-radix tier of the instances (fft_prime_tier; the
// kernel instantiation that holds those passes is launched).
hipError_t launch_batch(int dtype, int primes, const BatchInst* insts, const int2* items,
                        int nitems, int n_fwd, int n_pen, int* ctl, int nctl, int threads,
                        size_t lds, int spin_limit, hipStream_t st, int* grid_out);

}  // namespace b17
