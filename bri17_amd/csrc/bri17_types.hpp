// Plain-data descriptors shared between host planning code and kernels.
#pragma once

#include <stdint.h>

namespace b17 {

// n / d == (umulhi(n, m) + n) >> s for 0 <= n < 2^31 (round-up multiplier).
struct FastDiv {
  uint32_t d, m, s;
};

inline FastDiv make_fastdiv(uint32_t d) {
  FastDiv f;
  f.d = d;
  uint32_t s = 0;
  while ((1ull << s) < d) s++;
  f.s = s;
  f.m = (uint32_t)(((1ull << 32) * ((1ull << s) - d)) / d + 1);
  return f;
}

constexpr int kMaxPass = 12;

// One Stockham pass: radix r, span ns (product of earlier radices),
// nr = n / r butterflies per sequence, twm = n / (ns r).
struct PassDesc {
  int r, ns, nr, twm;
  FastDiv nsdiv;
};

// A length-n transform done by tpt threads per sequence.  Twiddles are read
// as tw[m * tws] from a table of exp(-2 pi i m / (n tws)).  wave: tpt is a
// power of two up to 64 and a sequence's threads are lanes of one wave, which
// orders the passes' LDS exchanges without workgroup barriers.
struct FftDesc {
  int n, npass, tpt, tws, wave;
  PassDesc pass[kMaxPass];
};

}  // namespace b17
