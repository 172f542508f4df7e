// Device-side FFT engine for gfx950: Stockham autosort passes on sequences
// staged in LDS, radix 2/3/4/5/7/8/9/11/13/16 butterflies held in VGPRs.
//
// Replaces the FFTW c2c transforms of the reference apply harness
// (tests/test_bri17.cpp:57,95, plans :119-127).  Conventions follow FFTW /
// sphinx/theory.rst:60-73: forward = exp(-2 pi i k n / N), unnormalised;
// inverse = exp(+2 pi i k n / N), unnormalised.
//
// A workgroup transforms many sequences at once.  Each sequence of length N
// is handled by `tpt` threads; for a pass of radix R a thread owns
// ceil((N/R)/tpt) butterflies.  A pass reads every butterfly's R inputs into
// registers, applies twiddles and the radix-R DFT, crosses a barrier and
// writes the R outputs in Stockham order (in place), so the result lands in
// natural order with no bit reversal.  Twiddles come from a table of
// exp(-2 pi i m / Nt) evaluated in fp64 on the host and rounded to T.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bri17_types.hpp"

namespace b17 {

// Pointer into global memory.  Pointers that a kernel reads from memory (an
// instance table, the tables of a parameter block) are otherwise generic and
// their loads and stores compile to flat_* instructions; global_* ones do
// not wait on LDS traffic.
// (The host pass of hipcc parses device code too and only sees plain pointers.)
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
#else
template <typename T>
using gptr = T*;
#endif

template <typename T>
struct cpx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) {
  return {a.x + b.x, a.y + b.y};
}
template <typename T>
__device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) {
  return {a.x - b.x, a.y - b.y};
}
template <typename T>
__device__ __forceinline__ cpx<T> operator*(cpx<T> a, T s) {
  return {a.x * s, a.y * s};
}
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename T>
__device__ __forceinline__ cpx<T> cmulc(cpx<T> a, cpx<T> b) {
  return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename T>
__device__ __forceinline__ cpx<T> conj(cpx<T> a) {
  return {a.x, -a.y};
}
// multiply by -i (forward) or +i (inverse)
template <bool INV, typename T>
__device__ __forceinline__ cpx<T> rot(cpx<T> a) {
  return INV ? cpx<T>{-a.y, a.x} : cpx<T>{a.y, -a.x};
}
// multiply by the constant exp(-+ i theta) given (c, s) = (cos, sin) theta
template <bool INV, typename T>
__device__ __forceinline__ cpx<T> crot(cpx<T> a, T c, T s) {
  // forward: (c - i s) a ; inverse: (c + i s) a
  return INV ? cpx<T>{c * a.x - s * a.y, c * a.y + s * a.x}
             : cpx<T>{c * a.x + s * a.y, c * a.y - s * a.x};
}

__device__ __forceinline__ uint32_t fdiv(uint32_t n, FastDiv d) {
  return (__umulhi(n, d.m) + n) >> d.s;
}

// Orders one wave's LDS accesses around it for the compiler (a wavefront-
// scope fence and wave barrier; the hardware executes a wave's LDS
// instructions in order).  An exchange whose writers and readers all lie in
// one wave needs no workgroup barrier.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Barrier between the read and write halves of a Stockham pass: a wave
// barrier when every thread of each sequence lies in one wave (`wave`,
// workgroup-uniform), else a workgroup barrier.
__device__ __forceinline__ void pass_sync(bool wave) {
  if (wave)
    wave_lds_sync();
  else
    __syncthreads();
}

// Threads of a stage of equal-length sequences of descriptor D, D.tpt per
// sequence.  D.wave (a power-of-two tpt up to 64): a sequence's threads are
// lanes of one wave, with the sequence index fastest across the lanes of a
// wave, and the stage's passes synchronise per wave.  Otherwise, if
// `seq_fastest`, the sequence index is fastest across all threads (LDS
// accesses of neighbouring threads hit neighbouring sequences, for
// sequences strided in LDS), else a sequence's threads are consecutive.
// S: sequences per round of the workgroup.
struct SeqThreads {
  int sl, lane, S;
  bool wave;
};
__device__ __forceinline__ SeqThreads seq_threads(int tid, int nthr, const FftDesc& D,
                                                  bool seq_fastest) {
  SeqThreads m;
  const int tpt = D.tpt;
  m.wave = D.wave != 0;
  if (m.wave) {
    const int spw = 64 / tpt;  // sequences per wave
    const int li = tid & 63;
    m.sl = (tid >> 6) * spw + li % spw;
    m.lane = li / spw;
    m.S = (nthr >> 6) * spw;
  } else if (seq_fastest) {
    m.S = nthr / tpt;
    m.sl = tid % m.S;
    m.lane = tid / m.S;
  } else {
    m.sl = tid / tpt;
    m.lane = tid - m.sl * tpt;
    m.S = nthr / tpt;
  }
  return m;
}

// ---------------------------------------------------------------- butterflies

template <bool INV, typename T>
__device__ __forceinline__ void bfly2(cpx<T>& a0, cpx<T>& a1) {
  cpx<T> t = a0 - a1;
  a0 = a0 + a1;
  a1 = t;
}

template <bool INV, typename T>
__device__ __forceinline__ void bfly4(cpx<T>& a0, cpx<T>& a1, cpx<T>& a2,
                                      cpx<T>& a3) {
  cpx<T> t0 = a0 + a2, t1 = a0 - a2, t2 = a1 + a3, t3 = rot<INV>(a1 - a3);
  a0 = t0 + t2;
  a2 = t0 - t2;
  a1 = t1 + t3;
  a3 = t1 - t3;
}

template <bool INV, typename T>
__device__ __forceinline__ void bfly3(cpx<T>* v) {
  const T s = T(0.86602540378443864676372317075294);
  cpx<T> t = v[1] + v[2];
  cpx<T> m = v[0] - t * T(0.5);
  cpx<T> d = rot<INV>(v[1] - v[2]) * s;
  v[0] = v[0] + t;
  v[1] = m + d;
  v[2] = m - d;
}

template <bool INV, typename T>
__device__ __forceinline__ void bfly5(cpx<T>* v) {
  const T c1 = T(0.30901699437494742410229341718282);
  const T c2 = T(-0.80901699437494742410229341718282);
  const T s1 = T(0.95105651629515357211643933337938);
  const T s2 = T(0.58778525229247312916870595463907);
  cpx<T> t1 = v[1] + v[4], t2 = v[2] + v[3], t3 = v[1] - v[4], t4 = v[2] - v[3];
  cpx<T> b1 = v[0] + t1 * c1 + t2 * c2;
  cpx<T> b2 = v[0] + t1 * c2 + t2 * c1;
  cpx<T> e1 = rot<INV>(t3 * s1 + t4 * s2);
  cpx<T> e2 = rot<INV>(t3 * s2 - t4 * s1);
  v[0] = v[0] + t1 + t2;
  v[1] = b1 + e1;
  v[4] = b1 - e1;
  v[2] = b2 + e2;
  v[3] = b2 - e2;
}

template <bool INV, typename T>
__device__ __forceinline__ void bfly7(cpx<T>* v) {
  const T c1 = T(0.62348980185873353052500488400424);
  const T c2 = T(-0.22252093395631440428890256449679);
  const T c3 = T(-0.90096886790241912623610231950745);
  const T s1 = T(0.78183148246802980870844452667406);
  const T s2 = T(0.97492791218182360701813168299393);
  const T s3 = T(0.43388373911755812047576833284836);
  cpx<T> p1 = v[1] + v[6], p2 = v[2] + v[5], p3 = v[3] + v[4];
  cpx<T> m1 = v[1] - v[6], m2 = v[2] - v[5], m3 = v[3] - v[4];
  cpx<T> b1 = v[0] + p1 * c1 + p2 * c2 + p3 * c3;
  cpx<T> b2 = v[0] + p1 * c2 + p2 * c3 + p3 * c1;
  cpx<T> b3 = v[0] + p1 * c3 + p2 * c1 + p3 * c2;
  cpx<T> e1 = rot<INV>(m1 * s1 + m2 * s2 + m3 * s3);
  cpx<T> e2 = rot<INV>(m1 * s2 - m2 * s3 - m3 * s1);
  cpx<T> e3 = rot<INV>(m1 * s3 - m2 * s1 + m3 * s2);
  v[0] = v[0] + p1 + p2 + p3;
  v[1] = b1 + e1;
  v[6] = b1 - e1;
  v[2] = b2 + e2;
  v[5] = b2 - e2;
  v[3] = b3 + e3;
  v[4] = b3 - e3;
}

// radix 8 = 2 x 4: DFT4 on even / odd samples, twiddle w8^k, DFT2.
template <bool INV, typename T>
__device__ __forceinline__ void bfly8(cpx<T>* v) {
  const T c = T(0.70710678118654752440084436210485);
  bfly4<INV>(v[0], v[2], v[4], v[6]);
  bfly4<INV>(v[1], v[3], v[5], v[7]);
  cpx<T> o1 = crot<INV>(v[3], c, c);
  cpx<T> o2 = rot<INV>(v[5]);
  cpx<T> o3 = crot<INV>(v[7], -c, c);
  cpx<T> y[8];
  y[0] = v[0] + v[1];
  y[4] = v[0] - v[1];
  y[1] = v[2] + o1;
  y[5] = v[2] - o1;
  y[2] = v[4] + o2;
  y[6] = v[4] - o2;
  y[3] = v[6] + o3;
  y[7] = v[6] - o3;
#pragma unroll
  for (int i = 0; i < 8; i++) v[i] = y[i];
}

// radix 9 = 3 x 3: DFT3 over q of x[3q+p], twiddle w9^(p ka), DFT3 over p.
template <bool INV, typename T>
__device__ __forceinline__ void bfly9(cpx<T>* v) {
  const T c1 = T(0.76604444311897803520239265055541667);   // cos(2 pi/9)
  const T s1 = T(0.64278760968653932632264340990726344);   // sin(2 pi/9)
  const T c2 = T(0.17364817766693034885171662676931480);   // cos(4 pi/9)
  const T s2 = T(0.98480775301220805936674302458952301);   // sin(4 pi/9)
  const T c4 = T(-0.93969262078590838405410927732473146);  // cos(8 pi/9)
  const T s4 = T(0.34202014332566873304409961468225958);   // sin(8 pi/9)
  cpx<T> y[3][3];  // y[p][ka]
#pragma unroll
  for (int p = 0; p < 3; p++) {
    y[p][0] = v[p];
    y[p][1] = v[p + 3];
    y[p][2] = v[p + 6];
    bfly3<INV>(y[p]);
  }
  y[1][1] = crot<INV>(y[1][1], c1, s1);  // e = 1
  y[1][2] = crot<INV>(y[1][2], c2, s2);  // e = 2
  y[2][1] = crot<INV>(y[2][1], c2, s2);  // e = 2
  y[2][2] = crot<INV>(y[2][2], c4, s4);  // e = 4
#pragma unroll
  for (int ka = 0; ka < 3; ka++) {
    cpx<T> z[3] = {y[0][ka], y[1][ka], y[2][ka]};
    bfly3<INV>(z);
    v[ka] = z[0];
    v[ka + 3] = z[1];
    v[ka + 6] = z[2];
  }
}

// radix 16 = 4 x 4: DFT4 over q of x[4q+p], twiddle w16^(p ka), DFT4 over p.
template <bool INV, typename T>
__device__ __forceinline__ void bfly16(cpx<T>* v) {
  const T c1 = T(0.92387953251128675612818318939679);  // cos(pi/8)
  const T s1 = T(0.38268343236508977172845998403040);  // sin(pi/8)
  const T c2 = T(0.70710678118654752440084436210485);  // cos(pi/4)
#pragma unroll
  for (int p = 0; p < 4; p++) bfly4<INV>(v[p], v[p + 4], v[p + 8], v[p + 12]);
  // Y[p][ka] sits at v[p + 4 ka]; twiddle by w16^(p ka)
  v[5] = crot<INV>(v[5], c1, s1);     // p1 ka1: e=1
  v[9] = crot<INV>(v[9], c2, c2);     // p1 ka2: e=2
  v[13] = crot<INV>(v[13], s1, c1);   // p1 ka3: e=3
  v[6] = crot<INV>(v[6], c2, c2);     // p2 ka1: e=2
  v[10] = rot<INV>(v[10]);            // p2 ka2: e=4
  v[14] = crot<INV>(v[14], -c2, c2);  // p2 ka3: e=6
  v[7] = crot<INV>(v[7], s1, c1);     // p3 ka1: e=3
  v[11] = crot<INV>(v[11], -c2, c2);  // p3 ka2: e=6
  v[15] = crot<INV>(v[15], -c1, -s1); // p3 ka3: e=9
  cpx<T> y[16];
#pragma unroll
  for (int ka = 0; ka < 4; ka++) {
    cpx<T> a0 = v[4 * ka], a1 = v[4 * ka + 1], a2 = v[4 * ka + 2], a3 = v[4 * ka + 3];
    bfly4<INV>(a0, a1, a2, a3);
    y[ka] = a0;
    y[ka + 4] = a1;
    y[ka + 8] = a2;
    y[ka + 12] = a3;
  }
#pragma unroll
  for (int i = 0; i < 16; i++) v[i] = y[i];
}

// Odd prime radix R = 11, 13: the inputs pair as (k, R - k), k = 1..H,
// H = (R - 1)/2, with t_k = v_k + v_{R-k} and d_k = v_k - v_{R-k}.  Output m
// and R - m share A_m = v_0 + sum_k t_k cos(2 pi k m / R) and
// B_m = sum_k d_k sin(2 pi k m / R): X_m = A_m -+ i B_m, X_{R-m} = A_m +- i B_m
// (forward / inverse).  The index k m is reduced mod R at compile time.
template <int R>
struct PrimeRoots;
template <>
struct PrimeRoots<11> {  // cos, sin (2 pi j / 11), j = 1..5
  static constexpr double c[6] = {1.0,
                                  0.841253532831181168861811648919367718,
                                  0.415415013001886425529274149229623204,
                                  -0.142314838273285140443792668616369669,
                                  -0.654860733945285064056925072466293553,
                                  -0.959492973614497389890368057066327699};
  static constexpr double s[6] = {0.0,
                                  0.540640817455597582107635954318691695,
                                  0.909631995354518371411715383079028460,
                                  0.989821441880932732376092037776718787,
                                  0.755749574354258283774035843972344420,
                                  0.281732556841429697711417915346616899};
};
template <>
struct PrimeRoots<13> {  // cos, sin (2 pi j / 13), j = 1..6
  static constexpr double c[7] = {1.0,
                                  0.885456025653209895900375522015098879,
                                  0.568064746731155802511807559127516625,
                                  0.120536680255323053349067687452543582,
                                  -0.354604887042535625969637892600018474,
                                  -0.748510748171101098634630599701351384,
                                  -0.970941817426052027156982276293789227};
  static constexpr double s[7] = {0.0,
                                  0.464723172043768545656015335133104778,
                                  0.822983865893656394579617423439381991,
                                  0.992708874098053992800751649492520179,
                                  0.935016242685414823439784599837830729,
                                  0.663122658240795202376785492666766280,
                                  0.239315664287557767148753726260211895};
};

template <int R, bool INV, typename T>
__device__ __forceinline__ void bfly_prime(cpx<T>* v) {
  constexpr int H = (R - 1) / 2;
  cpx<T> t[H], d[H];
#pragma unroll
  for (int k = 1; k <= H; k++) {
    t[k - 1] = v[k] + v[R - k];
    d[k - 1] = v[k] - v[R - k];
  }
  cpx<T> y[R];
  y[0] = v[0];
#pragma unroll
  for (int k = 0; k < H; k++) y[0] = y[0] + t[k];
#pragma unroll
  for (int m = 1; m <= H; m++) {
    cpx<T> a = v[0], b = {T(0), T(0)};
#pragma unroll
    for (int k = 1; k <= H; k++) {
      const int j = (k * m) % R;
      const T c = T(PrimeRoots<R>::c[j <= H ? j : R - j]);
      const T sn = j <= H ? T(PrimeRoots<R>::s[j]) : T(-PrimeRoots<R>::s[R - j]);
      a = a + t[k - 1] * c;
      b = b + d[k - 1] * sn;
    }
    const cpx<T> e = rot<INV>(b);
    y[m] = a + e;
    y[R - m] = a - e;
  }
#pragma unroll
  for (int i = 0; i < R; i++) v[i] = y[i];
}

template <int R, bool INV, typename T>
__device__ __forceinline__ void bfly(cpx<T>* v) {
  if constexpr (R == 2)
    bfly2<INV>(v[0], v[1]);
  else if constexpr (R == 3)
    bfly3<INV>(v);
  else if constexpr (R == 4)
    bfly4<INV>(v[0], v[1], v[2], v[3]);
  else if constexpr (R == 5)
    bfly5<INV>(v);
  else if constexpr (R == 7)
    bfly7<INV>(v);
  else if constexpr (R == 8)
    bfly8<INV>(v);
  else if constexpr (R == 9)
    bfly9<INV>(v);
  else if constexpr (R == 16)
    bfly16<INV>(v);
  else if constexpr (R == 11 || R == 13)
    bfly_prime<R, INV>(v);
}

// ------------------------------------------------------------- LDS sequences

// Where sequence s of a workgroup tile lives in LDS: element n of sequence s
// is at base(s) + n * es (units of complex elements).
struct SeqLayout {
  int es;        // element stride
  int w;         // sequences per "row group"
  int gstride;   // LDS distance between row groups
  FastDiv wdiv;  // division by w
  __device__ __forceinline__ int base(int s) const {
    uint32_t g = fdiv((uint32_t)s, wdiv);
    return (int)g * gstride + (s - (int)g * w);
  }
};

// One Stockham pass of radix R over the sequences this thread works on.
// `lane` in [0, tpt) indexes the thread within its sequence.
template <typename T, int R, bool INV, int MAXB, typename TWP>
__device__ __forceinline__ void fft_pass(cpx<T>* buf, TWP tw,
                                         const PassDesc& P, int N, int tws,
                                         int base, int es, int lane, int tpt,
                                         bool active, bool wave) {
  const int NR = P.nr;
  const int Ns = P.ns;
  const int twm = P.twm * tws;
  cpx<T> v[MAXB][R];
  int ob[MAXB];
  if (active) {
#pragma unroll
    for (int i = 0; i < MAXB; i++) {
      const int j = lane + i * tpt;
      if (j < NR) {
        const cpx<T>* src = buf + base + j * es;
#pragma unroll
        for (int k = 0; k < R; k++) v[i][k] = src[k * NR * es];
        const int q = (int)fdiv((uint32_t)j, P.nsdiv);
        const int jm = j - q * Ns;
        if (Ns > 1) {
#pragma unroll
          for (int k = 1; k < R; k++) {
            cpx<T> w = tw[jm * k * twm];
            v[i][k] = INV ? cmulc(v[i][k], w) : cmul(v[i][k], w);
          }
        }
        bfly<R, INV>(v[i]);
        ob[i] = base + (q * Ns * R + jm) * es;
      }
    }
  }
  pass_sync(wave);
  if (active) {
#pragma unroll
    for (int i = 0; i < MAXB; i++) {
      const int j = lane + i * tpt;
      if (j < NR) {
#pragma unroll
        for (int k = 0; k < R; k++) buf[ob[i] + k * Ns * es] = v[i][k];
      }
    }
  }
  pass_sync(wave);
}

// Outputs a thread holds in a direct-DFT pass (the planner keeps n / tpt at
// or below it for transforms with such a pass).
constexpr int kMaxDft = 8;

// Pass of a prime radix R above 16 (17, 19, ..., any prime): a direct DFT
// per butterfly.  Output q of butterfly j (item e = q NR + j of the pass) is
// sum_r w_R^(q r) t_r, t_r the twiddled input r; a thread forms its items
// e = lane + i tpt one after the other (at most MAXE = kMaxDft, the
// planner's bound for such passes), keeps them until the pass's
// barrier and writes them in Stockham order.  The roots come from the
// transform's table: w_R^s is entry s (N / R) tws.  O(R) operations per
// entry: slower than the butterflies, for sides that have such a factor.
template <typename T, bool INV, int MAXE, typename TWP>
__device__ __forceinline__ void dft_pass(cpx<T>* buf, TWP tw, const PassDesc& P, int N, int tws,
                                         int base, int es, int lane, int tpt, bool active,
                                         bool wave) {
  const int R = P.r, NR = P.nr, Ns = P.ns;
  const int twm = P.twm * tws;     // inter-pass twiddle step
  const int rstep = (N / R) * tws;  // table entry of w_R
  cpx<T> o[MAXE];
  int op[MAXE];
  if (active) {
#pragma unroll
    for (int i = 0; i < MAXE; i++) {
      const int e = lane + i * tpt;
      if (e < N) {
        const int q = e / NR, j = e - q * NR;
        const int qb = (int)fdiv((uint32_t)j, P.nsdiv), jm = j - qb * Ns;
        const cpx<T>* src = buf + base + j * es;
        cpx<T> acc = src[0];
        int s = 0;           // q r mod R
        int t = 0;           // jm r twm
#pragma unroll 1
        for (int r = 1; r < R; r++) {
          s += q;
          if (s >= R) s -= R;
          t += jm * twm;
          cpx<T> x = src[r * NR * es];
          if (Ns > 1) x = INV ? cmulc(x, tw[t]) : cmul(x, tw[t]);
          const cpx<T> w = tw[s * rstep];
          acc = acc + (INV ? cmulc(x, w) : cmul(x, w));
        }
        o[i] = acc;
        op[i] = base + (qb * Ns * R + jm + q * Ns) * es;
      }
    }
  }
  pass_sync(wave);
  if (active) {
#pragma unroll
    for (int i = 0; i < MAXE; i++) {
      const int e = lane + i * tpt;
      if (e < N) buf[op[i]] = o[i];
    }
  }
  pass_sync(wave);
}

// Full FFT of length D.n on the LDS sequence at `base` (stride es).  Every
// thread of the workgroup (`wave`: of the wave, whose threads then hold
// whole sequences; see seq_threads) must call this (barriers inside);
// `active` marks threads that own a sequence.
// `tw` is a pointer to const cpx<T> in any address space.  PR: the passes of
// prime radices compiled in: 0 none, 1 the radix-11 and radix-13
// butterflies, 2 also the direct DFTs of primes above 16 (they hold more
// values at once; kernels without them keep the smaller register budget).
template <typename T, bool INV, int MAXE, int PR = 0, typename TWP>
__device__ __forceinline__ void lds_fft(cpx<T>* buf, TWP tw,
                                        const FftDesc& D, int base, int es,
                                        int lane, bool active, bool wave = false) {
  for (int p = 0; p < D.npass; p++) {
    const PassDesc& P = D.pass[p];
    switch (P.r) {
      case 2:
        fft_pass<T, 2, INV, MAXE / 2>(buf, tw, P, D.n, D.tws, base, es, lane, D.tpt, active, wave);
        break;
      case 3:
        fft_pass<T, 3, INV, MAXE / 3>(buf, tw, P, D.n, D.tws, base, es, lane, D.tpt, active, wave);
        break;
      case 4:
        fft_pass<T, 4, INV, MAXE / 4>(buf, tw, P, D.n, D.tws, base, es, lane, D.tpt, active, wave);
        break;
      case 5:
        fft_pass<T, 5, INV, MAXE / 5>(buf, tw, P, D.n, D.tws, base, es, lane, D.tpt, active, wave);
        break;
      case 7:
        fft_pass<T, 7, INV, MAXE / 7>(buf, tw, P, D.n, D.tws, base, es, lane, D.tpt, active, wave);
        break;
      case 8:
        fft_pass<T, 8, INV, MAXE / 8>(buf, tw, P, D.n, D.tws, base, es, lane, D.tpt, active, wave);
        break;
      case 9:
        fft_pass<T, 9, INV, MAXE / 9>(buf, tw, P, D.n, D.tws, base, es, lane, D.tpt, active, wave);
        break;
      case 16:
        fft_pass<T, 16, INV, MAXE / 16>(buf, tw, P, D.n, D.tws, base, es, lane, D.tpt, active, wave);
        break;
      case 11:
        if constexpr (PR >= 1)
          fft_pass<T, 11, INV, MAXE / 11>(buf, tw, P, D.n, D.tws, base, es, lane, D.tpt, active, wave);
        break;
      case 13:
        if constexpr (PR >= 1)
          fft_pass<T, 13, INV, MAXE / 13>(buf, tw, P, D.n, D.tws, base, es, lane, D.tpt, active, wave);
        break;
      default:
        if constexpr (PR >= 2)
          dft_pass<T, INV, kMaxDft>(buf, tw, P, D.n, D.tws, base, es, lane, D.tpt, active, wave);
        break;
    }
  }
}

}  // namespace b17
