// Compile-time-size fast path of the periodic stiffness operator (3D grids
// whose sides factor as two radix passes, e.g. 64, 128, 256).
//
// Same pipeline and workspace layout as kernels.hip (plane R2C + first
// middle-axis stage -> pencil: last middle-axis stage, axis-0 FFT, modal
// block, inverse -> plane inverse + C2R), but every transform is a two-pass
// Stockham FFT held in registers: the first pass reads global memory straight
// into VGPRs, one LDS exchange feeds the second pass, and the row -> column
// transposition is the only other trip through LDS.  With all sizes known at
// compile time the index arithmetic folds to shifts and constants.
//
// The plane kernels are persistent (one workgroup per CU slot looping over
// planes): the global loads of the next plane are issued into the registers
// freed by the first FFT pass, so they are in flight while the rest of the
// current plane is computed.
#include "fft_core.hpp"
#include "kernels.hpp"
#include "kernels_static.hpp"
#include "modal.hpp"

namespace b17 {

// one pad element every 32 complex entries: breaks the power-of-two strides
// of the exchanges across LDS banks
__device__ __forceinline__ int pd(int i) { return i + (i >> 5); }
__host__ __device__ constexpr int pd_size(int n) { return n + (n >> 5) + 1; }

template <typename T>
__device__ __forceinline__ T fmat(T a, T b, T c) {
  if constexpr (sizeof(T) == 4) return __builtin_fmaf(a, b, c);
  else return __builtin_fma(a, b, c);
}

// The two half-length real-transform splits with w = w_{2M}^k:
//   R2C: X[k] = (A + conj B)/2 - i w (A - conj B)/2,  A = Z[k], B = Z[M-k]
//   C2R: Z[k] = (A + conj B) + i conj(w) (A - conj B), A = X[k], B = X[M-k]
// regrouped per input, four fused operations per output part, with the
// per-entry constants p = 1 + w.y, m = 1 - w.y, h = w.x (halved for the R2C,
// which the forward plane kernel writes out per component):
//   R2C: X.x = A.x p + B.x m + (A.y + B.y) h,  X.y = A.y p - B.y m - (A.x - B.x) h
//   C2R: Z.x = A.x p + B.x m - (A.y + B.y) h,  Z.y = A.y p - B.y m + (A.x - B.x) h
template <typename T>
__device__ __forceinline__ cpx<T> s_pre_c2r_pmh(cpx<T> A, cpx<T> B, T p, T m, T h) {
  return {fmat(A.x, p, fmat(B.x, m, fmat(-A.y, h, -(B.y * h)))),
          fmat(A.y, p, fmat(-B.y, m, fmat(A.x, h, -(B.x * h))))};
}

// wave_lds_sync (fft_core.hpp) orders the plane kernels' row exchanges (the
// M/R1 threads of a row are consecutive lanes of one wave) and column
// exchanges (see col_cs) without a workgroup barrier.

// Column-stage thread mapping of the plane kernels: the TPC threads of a
// column are lanes of one wave (64 / TPC consecutive columns per wave, column
// fastest), so the column exchange is ordered per wave.  A global access of
// a wave then covers TPC rows of 64 / TPC columns (64-byte segments for fp32
// at TPC = 8), whose other halves the neighbouring wave of the same CU
// accesses at the same time.
template <int TPC>
__device__ __forceinline__ int col_cs(int tl) {
  constexpr int CPW = 64 / TPC;
  return (tl / 64) * CPW + (tl % CPW);
}
template <int TPC>
__device__ __forceinline__ int col_l(int tl) {
  constexpr int CPW = 64 / TPC;
  return (tl % 64) / CPW;
}

// A copy of v the optimiser cannot see through: indices derived from it are
// recomputed inside a persistent loop instead of being hoisted into hundreds
// of live registers.
__device__ __forceinline__ int opaque(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

// One complex entry stored through a buffer resource.  A lane whose byte
// offset lies beyond the resource stores nothing, so a per-lane condition
// needs no branch around the store: with every global access of a persistent
// loop outside divergent branches, the compiler's count of memory operations
// in flight at the loop head stays exact, and the wait for the prefetched
// loads does not also wait for the previous work item's stores.
constexpr uint32_t kBufferNone = 0x80000000u;  // beyond every resource used here
template <typename T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(cpx<T>* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(base, 0, (int)bytes, 0x00020000);
}
template <typename T>
__device__ __forceinline__ void buffer_store(cpx<T> v, __amdgpu_buffer_rsrc_t r, uint32_t off) {
  if constexpr (sizeof(T) == 4) {
    typedef unsigned int u2 __attribute__((ext_vector_type(2)));
    const u2 d = {__float_as_uint(v.x), __float_as_uint(v.y)};
    __builtin_amdgcn_raw_buffer_store_b64(d, r, (int)off, 0, 0);
  } else {
    typedef unsigned int u4 __attribute__((ext_vector_type(4)));
    const unsigned long long a = __double_as_longlong(v.x), b = __double_as_longlong(v.y);
    const u4 d = {(unsigned)a, (unsigned)(a >> 32), (unsigned)b, (unsigned)(b >> 32)};
    __builtin_amdgcn_raw_buffer_store_b128(d, r, (int)off, 0, 0);
  }
}

// s_waitcnt immediate that waits until at most n vector memory operations of
// the wave are in flight (gfx9 encoding: vmcnt in bits 3:0 and 15:14; expcnt
// and lgkmcnt left at their no-wait maxima).  The compiler's wait insertion
// takes such an explicit wait into account.
__host__ __device__ constexpr int vmcnt_wait(int n) {
  return ((n & 15) | ((n >> 4) << 14)) | (7 << 4) | (15 << 8);
}

// One complex entry loaded through a buffer resource (0 for a lane whose
// offset lies beyond the resource: no memory access, no branch).
template <typename T>
__device__ __forceinline__ cpx<T> buffer_load(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  if constexpr (sizeof(T) == 4) {
    typedef unsigned int u2 __attribute__((ext_vector_type(2)));
    const u2 d = __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0);
    return {__uint_as_float(d.x), __uint_as_float(d.y)};
  } else {
    typedef unsigned int u4 __attribute__((ext_vector_type(4)));
    const u4 d = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
    return {__longlong_as_double((long long)(((unsigned long long)d.y << 32) | d.x)),
            __longlong_as_double((long long)(((unsigned long long)d.w << 32) | d.z))};
  }
}

template <typename T>
__device__ __forceinline__ void copy_table(cpx<T>* dst, const cpx<T>* src, int n, int tid,
                                           int nthr) {
  for (int i = tid; i < n; i += nthr) dst[i] = src[i];
}

// ------------------------------------------------------------------ plane fwd
//
// Work item (c, i0, p): rows i1 = q P + p, q < Q, of M complex (= 2M reals).
// Row FFT M = R1 x R2 by M/R1 threads per row; column stage Q = C1 x C2 by
// Q/C1 threads per column on M column sequences (columns 0 and M, both real
// after the row transform, are packed into one complex column).
//
// Workspace: the spectrum rows (c, i0, p, ka) hold k < M, M entries each, so
// every row starts on an M-entry boundary; the Nyquist entries k = M lie in a
// region of their own, nyq[(c, i0, p, ka)], after the rows.

template <typename T, int M, int R1, int Q, int P>
__device__ __forceinline__ void plane_load_rows(cpx<T>* v, const T* __restrict__ u, int item,
                                                int tid) {
  constexpr int TPR = M / R1;
  const int plane = item / P, p = item - (item / P) * P;
  const int q = tid / TPR, j = tid - (tid / TPR) * TPR;
  const cpx<T>* src =
      reinterpret_cast<const cpx<T>*>(u + ((int64_t)plane * (P * Q) + q * P + p) * (2 * M));
  // The field is read once: non-temporal loads leave the last-level cache
  // (256 MB) to the half-spectrum workspace, which the pencil and inverse
  // kernels then read from there.
#pragma unroll
  for (int k = 0; k < R1; k++) {
    const T* sp = reinterpret_cast<const T*>(src + j + k * TPR);
    v[k] = cpx<T>{__builtin_nontemporal_load(sp), __builtin_nontemporal_load(sp + 1)};
  }
}

// LDS index of (row, col) in the plane kernels' rows of RS entries, padded by
// one entry every 16 (a padded row of 128 entries is 136 = 8 mod 16 entries
// long).  With 8-byte entries a quarter wave's 16 lanes are conflict-free
// when their indices differ mod 16: the row passes (lanes j of rows q, q+1
// at columns 16 j + k, or j + 8 b + 16 k) and the column passes (16
// consecutive columns from a multiple of 16) all are.  Callers split the
// index into a lane part and a compile-time part so that the latter becomes
// an instruction offset: row and col terms are independent, and col >> 4
// splits when the constant part of col is a multiple of 16 or cannot carry
// into bit 4.
template <int RS>
__device__ __forceinline__ constexpr int prow(int row) {
  static_assert(RS % 16 == 0, "padded rows of 16-entry multiples");
  return row * (RS + RS / 16);
}
__device__ __forceinline__ constexpr int pcol(int col) { return col + (col >> 4); }
// entries of Q padded rows of RS entries (one spare entry past the end)
__host__ __device__ constexpr int prow_size(int q, int rs) { return q * (rs + rs / 16) + 1; }

template <typename T, int M, int R1, int R2, int Q, int C1, int C2, int P, bool PF, bool PACK>
__global__ void __launch_bounds__(Q * M / R1)
    k_plane_fwd_s(const T* __restrict__ u, cpx<T>* __restrict__ ws, cpx<T>* __restrict__ nyq,
                  const cpx<T>* __restrict__ g_twr, const cpx<T>* __restrict__ g_twc, int nitems) {
  constexpr int NT = Q * M / R1;
  constexpr int TPR = M / R1;
  constexpr int TPC = Q / C1;
  constexpr int N1 = P * Q;
  constexpr int RS = M;  // row stride (before padding)
  static_assert(R1 * R2 == M && C1 * C2 == Q && R1 == C1, "plane geometry");
  static_assert(R1 % R2 == 0 && C1 % C2 == 0, "second pass must divide the first");
  static_assert(R1 == 16 && M % 16 == 0, "LDS offsets split into lane + constant parts");
  static_assert(64 % TPR == 0, "the threads of a row lie in one wave");
  static_assert(64 % TPC == 0 && M % (64 / TPC) == 0 && NT % 64 == 0,
                "the threads of a column lie in one wave");
  static_assert((M & (M - 1)) == 0 && (Q & (Q - 1)) == 0, "power-of-two plane sides");
  extern __shared__ __align__(16) unsigned char smem[];
  cpx<T>* buf = reinterpret_cast<cpx<T>*>(smem);
  cpx<T>* twr = buf + prow_size(Q, RS);  // w_{2M}^j, j < 2M
  cpx<T>* twc = twr + 2 * M;             // w_{N1}^j, j < N1
  cpx<T>* scr = twc + N1;                // packed-column scratch, Q entries
  cpx<T>* scn = scr + Q;                 // Nyquist entries of the item, Q entries
  cpx<T>* tw2 = scn + Q;                 // row pass 2 twiddles w_M^(jj k) at [k R1 + jj]
  const int tid = threadIdx.x;
  const __amdgpu_buffer_rsrc_t rn = buffer_rsrc(nyq, (uint32_t)(nitems * Q * sizeof(cpx<T>)));

  copy_table(twr, g_twr, 2 * M, tid, NT);
  copy_table(twc, g_twc, N1, tid, NT);
  for (int i = tid; i < R2 * R1; i += NT) tw2[i] = g_twr[2 * (i % R1) * (i / R1)];
  __syncthreads();

  cpx<T> v[R1];
  int item = blockIdx.x;
  if (PF && item < nitems) {
    plane_load_rows<T, M, R1, Q, P>(v, u, item, tid);
    __builtin_amdgcn_s_waitcnt(vmcnt_wait(0));
  }
  auto process = [&](const int item) {
    const int tl = opaque(tid);
    if (!PF) plane_load_rows<T, M, R1, Q, P>(v, u, item, tl);
    const int q = tl / TPR, j = tl - (tl / TPR) * TPR;
    const int cs = col_cs<TPC>(tl), l = col_l<TPC>(tl);
    const int plane = item / P;
    const int p = item - plane * P;
    // ---- rows: pass 1 on the loaded registers; (q, j R1 + k), k < R1
    {
      bfly<R1, false>(v);
      cpx<T>* bp = buf + prow<RS>(q) + pcol(j * R1);
#pragma unroll
      for (int k = 0; k < R1; k++) bp[k] = v[k];
    }
    // prefetch the next work item into the now free registers
    if (PF && item + (int)gridDim.x < nitems)
      plane_load_rows<T, M, R1, Q, P>(v, u, item + gridDim.x, tl);
    // the row's TPR threads are lanes of one wave: no workgroup barrier
    wave_lds_sync();
    // pass 2: R1/R2 butterflies per thread, natural-order output;
    // (q, jj + k R1) with jj = j + b TPR < R1
    cpx<T> w[R1 / R2][R2];
    cpx<T>* rp = buf + prow<RS>(q) + j;
#pragma unroll
    for (int b = 0; b < R1 / R2; b++) {
      const int jj = j + b * TPR;
#pragma unroll
      for (int k = 0; k < R2; k++) {
        w[b][k] = rp[b * TPR + pcol(k * R1)];
        if (k) w[b][k] = cmul(w[b][k], tw2[k * R1 + jj]);
      }
      bfly<R2, false>(w[b]);
    }
    // in place: a thread writes back exactly the entries it read (no barrier)
#pragma unroll
    for (int b = 0; b < R1 / R2; b++)
#pragma unroll
      for (int k = 0; k < R2; k++) rp[b * TPR + pcol(k * R1)] = w[b][k];
    __syncthreads();

    // ---- columns: pass 1 reads Z of two columns and forms the spectrum;
    // rows qq = l + kk TPC of columns cs and M - cs.  Lane cs = 0 reads column
    // 0 twice: the same formula gives X[q][0] = Re Z + Im Z (real), and the
    // lane packs X0 + i XM with X[q][M] = Re Z - Im Z (no divergent path).
    cpx<T> x[C1];
    {
      const cpx<T>* cp = buf + prow<RS>(l) + pcol(cs);
      const cpx<T>* cm = buf + prow<RS>(l) + pcol((M - cs) & (M - 1));
      const cpx<T> wc = twr[cs];
      const T wp = T(0.5) * (T(1) + wc.y), wm = T(0.5) * (T(1) - wc.y), wh = T(0.5) * wc.x;
      cpx<T> A[C1], B[C1];
#pragma unroll
      for (int kk = 0; kk < C1; kk++) {
        A[kk] = cp[prow<RS>(kk * TPC)];
        B[kk] = cm[prow<RS>(kk * TPC)];
      }
#pragma unroll
      for (int kk = 0; kk < C1; kk++) {
        const cpx<T> a = A[kk], b = B[kk];
        // written per component: returned as a pair, the same operations
        // were scheduled into 128 VGPRs instead of 98
        x[kk].x = fmat(a.x, wp, fmat(b.x, wm, fmat(a.y, wh, b.y * wh)));
        x[kk].y = fmat(a.y, wp, fmat(-b.y, wm, fmat(-a.x, wh, b.x * wh)));
        x[kk].y = cs == 0 ? a.x - a.y : x[kk].y;
      }
    }
    bfly<C1, false>(x);
    __syncthreads();
    // column exchange layout: element m of sub-sequence l of column cs at row
    // l C1 + (m ^ (l & 1)) (the swizzle keeps both the writes and the reads
    // of a quarter wave on distinct banks)
    {
      const int lb = l & 1;
      cpx<T>* ce = buf + prow<M>(l * C1 + lb) + pcol(cs);
      cpx<T>* co = buf + prow<M>(l * C1 - lb) + pcol(cs);
#pragma unroll
      for (int kk = 0; kk < C1; kk++) ((kk & 1) ? co : ce)[prow<M>(kk)] = x[kk];
    }
    // the column's TPC threads are lanes of one wave
    wave_lds_sync();
    cpx<T> y[C1 / C2][C2];
    {
      const cpx<T>* ce = buf + prow<M>(l) + pcol(cs);
      const cpx<T>* co = buf + prow<M>(l ^ 1) + pcol(cs);
#pragma unroll
      for (int b = 0; b < C1 / C2; b++) {
        const int jj = l + b * TPC;
#pragma unroll
        for (int k = 0; k < C2; k++) {
          y[b][k] = ((k & 1) ? co : ce)[prow<M>(b * TPC + k * C1)];
          if (k) y[b][k] = cmul(y[b][k], twc[P * jj * k]);
        }
        bfly<C2, false>(y[b]);
      }
    }
    // packed column: Y0 = (Z[ka] + conj Z[-ka])/2, YM = -i (Z[ka] - conj Z[-ka])/2
    // (with PACK the packed column itself is stored, in the k = 0 slot)
    if (!PACK && cs == 0) {
#pragma unroll
      for (int b = 0; b < C1 / C2; b++)
#pragma unroll
        for (int k = 0; k < C2; k++) scr[l + b * TPC + k * C1] = y[b][k];
    }
    if (!PACK) __syncthreads();
    // lanes cs = 0 unpack: Y0 replaces the packed value, YM (twiddled) is
    // staged in LDS for the Nyquist region
    if (!PACK && cs == 0) {
#pragma unroll
      for (int b = 0; b < C1 / C2; b++)
#pragma unroll
        for (int k = 0; k < C2; k++) {
          const int ka = l + b * TPC + k * C1;
          const cpx<T> val = y[b][k];
          const cpx<T> zn = scr[(Q - ka) % Q];
          y[b][k] = {T(0.5) * (val.x + zn.x), T(0.5) * (val.y - zn.y)};
          cpx<T> ym = {T(0.5) * (val.y + zn.y), T(0.5) * (zn.x - val.x)};
          if (P > 1 && p) ym = cmul(ym, twc[p * ka]);
          scn[ka] = ym;
        }
    }
    if (P > 1 && p) {
#pragma unroll
      for (int b = 0; b < C1 / C2; b++)
#pragma unroll
        for (int k = 0; k < C2; k++) y[b][k] = cmul(y[b][k], twc[p * (l + b * TPC + k * C1)]);
    }
    cpx<T>* dst = ws + ((int64_t)plane * N1 + p * Q) * M + cs;
#pragma unroll
    for (int b = 0; b < C1 / C2; b++)
#pragma unroll
      for (int k = 0; k < C2; k++) dst[(l + b * TPC + k * C1) * M] = y[b][k];
    __syncthreads();  // LDS is rewritten by the next work item
    if constexpr (!PACK) {
      // the item's Nyquist entries: thread tid < Q stores entry tid (scn is
      // rewritten only after several barriers of the next item)
      buffer_store(scn[tid & (Q - 1)], rn,
                   tid < Q ? (uint32_t)(((plane * N1 + p * Q) + tid) * sizeof(cpx<T>)) : kBufferNone);
    }
    // The next item's loads were issued before this item's stores (C1, and
    // one Nyquist store): waiting for them here leaves the stores in flight
    // across the loop head.
    static_assert(C1 + 1 < 64, "vmcnt range");
    if (PF) __builtin_amdgcn_s_waitcnt(vmcnt_wait(PACK ? C1 : C1 + 1));
  };
  for (; item < nitems; item += gridDim.x) process(item);
}

// ------------------------------------------------------------------ plane inv

// Loads of work item (c, i0, p): column entries (ka = l + kk TPC, cs) into x,
// and Nyquist entry ka = tid into xn for threads tid < Q (no branch).
template <typename T, int M, int C1, int Q, int P, bool PACK>
__device__ __forceinline__ void plane_load_cols(cpx<T>* x, cpx<T>& xn, const cpx<T>* __restrict__ ws,
                                                __amdgpu_buffer_rsrc_t rn, int item, int tid) {
  constexpr int TPC = Q / C1;
  const int plane = item / P, p = item - (item / P) * P;
  const int cs = col_cs<TPC>(tid), l = col_l<TPC>(tid);
  const cpx<T>* src = ws + ((int64_t)plane * (P * Q) + p * Q) * M;
#pragma unroll
  for (int kk = 0; kk < C1; kk++) x[kk] = src[(l + kk * TPC) * M + cs];
  if constexpr (!PACK)
    xn = buffer_load<T>(rn, tid < Q ? (uint32_t)((plane * (P * Q) + p * Q + tid) * sizeof(cpx<T>))
                                    : kBufferNone);
}

template <typename T, int M, int R1, int R2, int Q, int C1, int C2, int P, bool PF, bool PACK>
__global__ void __launch_bounds__(Q * M / R1)
    k_plane_inv_s(const cpx<T>* __restrict__ ws, const cpx<T>* __restrict__ nyq,
                  T* __restrict__ out, const cpx<T>* __restrict__ g_twr,
                  const cpx<T>* __restrict__ g_twc, int nitems) {
  constexpr int NT = Q * M / R1;
  constexpr int TPR = M / R1;
  constexpr int TPC = Q / C1;
  constexpr int N1 = P * Q;
  static_assert(R1 * R2 == M && C1 * C2 == Q && R1 == C1, "plane geometry");
  static_assert(R1 % R2 == 0 && C1 % C2 == 0, "second pass must divide the first");
  static_assert(R1 == 16 && 16 % TPR == 0 && M % 16 == 0,
                "LDS offsets split into lane + constant parts");
  static_assert(64 % TPR == 0, "the threads of a row lie in one wave");
  static_assert(64 % TPC == 0 && M % (64 / TPC) == 0 && NT % 64 == 0,
                "the threads of a column lie in one wave");
  extern __shared__ __align__(16) unsigned char smem[];
  cpx<T>* buf = reinterpret_cast<cpx<T>*>(smem);  // rows of M entries (padded)
  cpx<T>* nyl = buf + prow_size(Q, M);             // X[q][M] of every row
  cpx<T>* tpm = nyl + Q;                           // (1 + w.y, 1 - w.y), (w.x, 0) of w_{2M}^mm
  cpx<T>* twc = tpm + 2 * M;
  cpx<T>* nin = twc + N1;                          // the item's Nyquist entries
  cpx<T>* tw2 = nin + Q;                           // row pass 2 twiddles, [k R1 + jj]
  const int tid = threadIdx.x;
  const __amdgpu_buffer_rsrc_t rn =
      buffer_rsrc(const_cast<cpx<T>*>(nyq), (uint32_t)(nitems * Q * sizeof(cpx<T>)));

  for (int i = tid; i < M; i += NT) {
    const cpx<T> w = g_twr[i];
    tpm[2 * i] = {T(1) + w.y, T(1) - w.y};
    tpm[2 * i + 1] = {w.x, T(0)};
  }
  copy_table(twc, g_twc, N1, tid, NT);
  for (int i = tid; i < R2 * R1; i += NT) tw2[i] = g_twr[2 * (i % R1) * (i / R1)];

  cpx<T> x[C1], xn;
  int item = blockIdx.x;
  if (PF && item < nitems) {
    plane_load_cols<T, M, C1, Q, P, PACK>(x, xn, ws, rn, item, tid);
    __builtin_amdgcn_s_waitcnt(vmcnt_wait(0));
    if (!PACK && tid < Q) nin[tid] = xn;
  }
  __syncthreads();
  for (; item < nitems; item += gridDim.x) {
    const int tl = opaque(tid);
    if (!PF) {
      plane_load_cols<T, M, C1, Q, P, PACK>(x, xn, ws, rn, item, tl);
      if (!PACK) {
        if (tl < Q) nin[tl] = xn;
        __syncthreads();
      }
    }
    const int cs = col_cs<TPC>(tl), l = col_l<TPC>(tl);
    const int q = tl / TPR, j = tl - (tl / TPR) * TPR;
    const int plane = item / P;
    const int p = item - plane * P;
    // ---- columns: inverse Q-point stage
#pragma unroll
    for (int kk = 0; kk < C1; kk++) {
      cpx<T> val = x[kk];
      if constexpr (!PACK) {
        const cpx<T> xm = nin[l + kk * TPC];  // same address across the wave
        val = cs == 0 ? cpx<T>{val.x - xm.y, val.y + xm.x} : val;  // X0 + i XM (both Hermitian)
      }
      if (P > 1 && p) val = cmulc(val, twc[p * (l + kk * TPC)]);
      x[kk] = val;
    }
    bfly<C1, true>(x);
    {
      // column exchange: element m of sub-sequence l of column cs at row
      // l C1 + (m ^ (l & 1))
      const int lb = l & 1;
      cpx<T>* ce = buf + prow<M>(l * C1 + lb) + pcol(cs);
      cpx<T>* co = buf + prow<M>(l * C1 - lb) + pcol(cs);
#pragma unroll
      for (int kk = 0; kk < C1; kk++) ((kk & 1) ? co : ce)[prow<M>(kk)] = x[kk];
    }
    // prefetch the next work item
    if (PF && item + (int)gridDim.x < nitems)
      plane_load_cols<T, M, C1, Q, P, PACK>(x, xn, ws, rn, item + gridDim.x, tl);
    // the column's TPC threads are lanes of one wave
    wave_lds_sync();
    cpx<T> y[C1 / C2][C2];
    cpx<T>* cp = buf + prow<M>(l) + pcol(cs);
    {
      const cpx<T>* co = buf + prow<M>(l ^ 1) + pcol(cs);
#pragma unroll
      for (int b = 0; b < C1 / C2; b++) {
        const int jj = l + b * TPC;
#pragma unroll
        for (int k = 0; k < C2; k++) {
          y[b][k] = ((k & 1) ? co : (const cpx<T>*)cp)[prow<M>(b * TPC + k * C1)];
          if (k) y[b][k] = cmulc(y[b][k], twc[P * jj * k]);
        }
        bfly<C2, true>(y[b]);
      }
    }
    // the write-back below goes to natural rows, which the wave's other
    // threads read above (swizzled)
    wave_lds_sync();
    // rows q = jj + C1 k, natural order, written back into the column (its
    // entries are read and written by the lanes of one wave); unpack columns
    // 0 and M
#pragma unroll
    for (int b = 0; b < C1 / C2; b++)
#pragma unroll
      for (int k = 0; k < C2; k++) {
        const cpx<T> val = y[b][k];
        cp[prow<M>(b * TPC + k * C1)] = cs == 0 ? cpx<T>{val.x, T(0)} : val;
      }
    if (cs == 0) {
#pragma unroll
      for (int b = 0; b < C1 / C2; b++)
#pragma unroll
        for (int k = 0; k < C2; k++) nyl[l + b * TPC + k * C1] = {y[b][k].y, T(0)};
    }
    __syncthreads();

    // ---- rows: C2R split, then the inverse M-point transform
    cpx<T> v[R1];
    {
      const cpx<T>* rp = buf + prow<M>(q);
      // entry mm = j + k TPR pairs with M - mm; mm = 0 (lane j = 0, k = 0)
      // pairs with X[q][M], and the same formula with w = 1 gives
      // Z[0] = (X0 + XM) + i (X0 - XM) for X0, XM real
#pragma unroll
      for (int k = 0; k < R1; k++) {
        const int mm = j + k * TPR;
        const cpx<T> A = rp[j + pcol(k * TPR)];
        const cpx<T>* bp = rp + pcol(M - mm);
        if (k == 0) bp = j == 0 ? nyl + q : bp;
        const cpx<T> pm = tpm[2 * mm], h = tpm[2 * mm + 1];
        v[k] = s_pre_c2r_pmh(A, *bp, pm.x, pm.y, h.x);
      }
    }
    bfly<R1, true>(v);
    // row q's entries are read and rewritten by the TPR lanes of one wave
    wave_lds_sync();
    {
      cpx<T>* bp = buf + prow<M>(q) + pcol(j * R1);
#pragma unroll
      for (int k = 0; k < R1; k++) bp[k] = v[k];
    }
    wave_lds_sync();
    cpx<T> w[R1 / R2][R2];
    const cpx<T>* rp = buf + prow<M>(q) + j;
#pragma unroll
    for (int b = 0; b < R1 / R2; b++) {
      const int jj = j + b * TPR;
#pragma unroll
      for (int k = 0; k < R2; k++) {
        w[b][k] = rp[b * TPR + pcol(k * R1)];
        if (k) w[b][k] = cmulc(w[b][k], tw2[k * R1 + jj]);
      }
      bfly<R2, true>(w[b]);
    }
    cpx<T>* dst = reinterpret_cast<cpx<T>*>(out + ((int64_t)plane * N1 + q * P + p) * (2 * M));
#pragma unroll
    for (int b = 0; b < R1 / R2; b++)
#pragma unroll
      for (int k = 0; k < R2; k++) dst[j + b * TPR + k * R1] = w[b][k];
    if (PF) {
      // The next item's loads were issued before this item's R1 stores:
      // waiting for them leaves the stores in flight across the loop head.
      static_assert(R1 < 64, "vmcnt range");
      __builtin_amdgcn_s_waitcnt(vmcnt_wait(R1));
      if (!PACK && tl < Q) nin[tl] = xn;
    }
    __syncthreads();  // LDS is rewritten by the next work item
  }
}

// ---------------------------------------------------------------------- pencil
//
// Tile (ka, f0..f0+W): all components c, all i0, all split indices p.  The
// thread (c, w, l) owns i0 = l + k TPA (k < A0) of sequence (c, w) for every
// p: it performs the split DFT over p and the first axis-0 pass in registers.

template <typename T, int N0, int A0, int P, int W>
__device__ __forceinline__ void pencil_load(cpx<T> (&v)[P][A0], const cpx<T>* __restrict__ ws,
                                            int F, int c, int l, int fw) {
  const cpx<T>* src = ws + (int64_t)c * N0 * P * F + fw;
  constexpr int TPA = N0 / A0;
#pragma unroll
  for (int k = 0; k < A0; k++)
#pragma unroll
    for (int pp_ = 0; pp_ < P; pp_++) v[pp_][k] = src[((int64_t)(l + k * TPA) * P + pp_) * F];
}

// Pencil tiles of the fixed-size layout: tiles [0, ntiles - nyq_tiles) walk
// the spectrum rows (F = Q M columns per row group (c, i0, p)), the last
// nyq_tiles tiles the Nyquist region (F = Q columns, one per ka).
struct STile {
  int64_t off;  // element offset of the region
  int F;        // columns per row group
  int f0;       // first column of the tile
  bool nyq;
};
__device__ __forceinline__ STile s_tile(const PencilParams& pp, int t, int W) {
  const int ta = pp.ntiles - pp.nyq_tiles;
  if (t < ta) return {0, pp.f, t * W, false};
  return {pp.nyq_off, pp.q, (t - ta) * W, true};
}
// (ka, k) of column f of a tile: k = f mod M in the rows, k = M (= pp.rs) in
// the Nyquist region
__device__ __forceinline__ void s_freq(const PencilParams& pp, bool nyq, int f, int& ka, int& kl) {
  if (nyq) {
    ka = f;
    kl = pp.rs;
  } else {
    ka = (int)fdiv((uint32_t)f, pp.rsdiv);
    kl = f - ka * pp.rs;
  }
}

// Persistent variant (PF): one workgroup per CU slot walks its tiles, and the
// loads of its next tile are issued into the registers freed by the first
// axis-0 pass, so they are in flight while the rest of the tile is computed.
// Tiles stay XCD-local: XCD x works through a contiguous range of tiles.
template <typename T, int N0, int A0, int B0, int P, int W, int MODE, bool PF>
__global__ void __launch_bounds__(3 * W * (N0 / A0))
    k_pencil_s(cpx<T>* __restrict__ ws, const PencilParams pp) {
  constexpr int TPA = N0 / A0;
  constexpr int NT = 3 * W * TPA;
  constexpr int ROWS = 3 * P * N0;  // LDS rows of W entries: ((c P + kb) N0 + n0)
  static_assert(A0 * B0 == N0 && A0 % B0 == 0, "pencil geometry");
  static_assert(NT % (P * W) == 0, "block mapping");
  extern __shared__ __align__(16) unsigned char smem[];
  cpx<T>* buf = reinterpret_cast<cpx<T>*>(smem);
  cpx<T>* tw = buf + pd_size(ROWS * W);
  const ModalCoef<T>* coef0 = reinterpret_cast<const ModalCoef<T>*>(pp.coef[0]);
  const ModalCoef<T>* coef1 = reinterpret_cast<const ModalCoef<T>*>(pp.coef[1]);
  const ModalCoef<T>* coef2 = reinterpret_cast<const ModalCoef<T>*>(pp.coef[2]);
  const int tid = threadIdx.x;
  const int T_ = pp.ntiles;

  // tiles of this workgroup: first + j * stride < cnt, offset by base
  int base = 0, cnt = T_, first = blockIdx.x, stride = gridDim.x;
  if (pp.xcd_remap) {
    // consecutive tiles share L2 lines: keep them on one XCD (workgroups are
    // dealt round-robin over the 8 XCDs)
    const int xx = blockIdx.x % 8, per = T_ / 8, rem = T_ % 8;
    base = xx * per + (xx < rem ? xx : rem);
    cnt = per + (xx < rem ? 1 : 0);
    first = blockIdx.x / 8;
    stride = PF ? (int)gridDim.x / 8 : cnt;  // one tile per workgroup without PF
    if (PF && gridDim.x % 8) {
      base = 0, cnt = T_, first = blockIdx.x, stride = gridDim.x;
    }
  }
  if (first >= cnt) return;

  copy_table(tw, reinterpret_cast<const cpx<T>*>(pp.tw0), N0, tid, NT);

  // tile -> region and columns; fw: this lane's column
  auto column = [&](int tile, STile& st, int& wv, int& fw) {
    st = s_tile(pp, tile, W);
    wv = st.F - st.f0 < W ? st.F - st.f0 : W;
    const int w = tid % W;
    // out-of-range columns of the last tile read a valid address and are never stored
    fw = w < wv ? st.f0 + w : st.F - 1;
  };

  cpx<T> v[P][A0];
  {
    STile st;
    int wv, fw;
    column(base + first, st, wv, fw);
    pencil_load<T, N0, A0, P, W>(v, ws + st.off, st.F, tid / (W * TPA), (tid / W) % TPA, fw);
  }
  for (int j = first; j < cnt; j += stride) {
    const int tl = opaque(tid);
    const int t = base + j;
    const int w = tl % W;
    const int l = (tl / W) % TPA;
    const int c = tl / (W * TPA);
    STile st;
    int wv, fw;
    column(t, st, wv, fw);
    const int F = st.F, f0 = st.f0;
    const bool wok = w < wv;
    if (PF && j > first) __syncthreads();  // the previous tile's last LDS reads are done

    // split DFT over p + pass 1
    if constexpr (P > 1) {
#pragma unroll
      for (int k = 0; k < A0; k++) {
        cpx<T> col[P];
#pragma unroll
        for (int pp_ = 0; pp_ < P; pp_++) col[pp_] = v[pp_][k];
        bfly<P, false>(col);
#pragma unroll
        for (int pp_ = 0; pp_ < P; pp_++) v[pp_][k] = col[pp_];
      }
    }
#pragma unroll
    for (int kb = 0; kb < P; kb++) {
      bfly<A0, false>(v[kb]);
#pragma unroll
      for (int k = 0; k < A0; k++) buf[pd(((c * P + kb) * N0 + l * A0 + k) * W + w)] = v[kb][k];
    }
    // prefetch the next tile into the now free registers
    if (PF && j + stride < cnt) {
      STile nst;
      int nwv, nfw;
      column(t + stride, nst, nwv, nfw);
      pencil_load<T, N0, A0, P, W>(v, ws + nst.off, nst.F, c, l, nfw);
    }
    __syncthreads();
    // pass 2 -> natural order k0 = jj + A0 k
    {
      cpx<T> y[P][A0 / B0][B0];
#pragma unroll
      for (int kb = 0; kb < P; kb++)
#pragma unroll
        for (int b = 0; b < A0 / B0; b++) {
          const int jj = l + b * TPA;
#pragma unroll
          for (int k = 0; k < B0; k++) {
            y[kb][b][k] = buf[pd(((c * P + kb) * N0 + jj + k * A0) * W + w)];
            if (k) y[kb][b][k] = cmul(y[kb][b][k], tw[jj * k]);
          }
          bfly<B0, false>(y[kb][b]);
        }
      __syncthreads();
#pragma unroll
      for (int kb = 0; kb < P; kb++)
#pragma unroll
        for (int b = 0; b < A0 / B0; b++)
#pragma unroll
          for (int k = 0; k < B0; k++)
            buf[pd(((c * P + kb) * N0 + l + b * TPA + k * A0) * W + w)] = y[kb][b][k];
    }
    __syncthreads();

    // modal block: a thread keeps (kb, w) fixed and walks k0, so the factors
    // of axes 1 and 2 are formed once
    {
      constexpr int KT = NT / (P * W);  // threads per (kb, w)
      const int bw = tl % W;
      const int bk = (tl / W) % KT;
      const int bkb = tl / (W * KT);
      int ka, kl;
      s_freq(pp, st.nyq, f0 + bw, ka, kl);
      if (bw < wv) {
        const int k1 = ka + pp.q * bkb;
        const T mu = T(pp.mu), s = T(pp.s), kinv = T(pp.kinv);
        const ModalCoef<T> b = coef1[k1], cc = coef2[kl];
        const T chichi = b.chi * cc.chi, phichi = b.phi * cc.chi, chiphi = b.chi * cc.phi;
        const T s_psichi = s * b.psi * cc.chi, s_chipsi = s * b.chi * cc.psi;
        const T s_psipsi = s * b.psi * cc.psi;
        const bool kz = k1 == 0 && kl == 0;
        for (int k0 = bk; k0 < N0; k0 += KT) {
          const ModalCoef<T> a = coef0[k0];
          const T H0 = a.phi * chichi, H1 = a.chi * phichi, H2 = a.chi * chiphi;
          const T dg = mu * (H0 + H1 + H2);
          cpx<T> xx[3];
#pragma unroll
          for (int q3 = 0; q3 < 3; q3++) xx[q3] = buf[pd(((q3 * P + bkb) * N0 + k0) * W + bw)];
          sym3_apply<T, MODE>(s * H0 + dg, s * H1 + dg, s * H2 + dg, a.psi * s_psichi,
                              a.psi * s_chipsi, a.chi * s_psipsi, kz && k0 == 0, xx, kinv);
#pragma unroll
          for (int q3 = 0; q3 < 3; q3++) buf[pd(((q3 * P + bkb) * N0 + k0) * W + bw)] = xx[q3];
        }
      }
    }
    __syncthreads();

    // inverse pass 1 (inputs k0 = l + k TPA)
    {
      cpx<T> z[P][A0];
#pragma unroll
      for (int kb = 0; kb < P; kb++) {
#pragma unroll
        for (int k = 0; k < A0; k++) z[kb][k] = buf[pd(((c * P + kb) * N0 + l + k * TPA) * W + w)];
        bfly<A0, true>(z[kb]);
      }
      __syncthreads();
#pragma unroll
      for (int kb = 0; kb < P; kb++)
#pragma unroll
        for (int k = 0; k < A0; k++) buf[pd(((c * P + kb) * N0 + l * A0 + k) * W + w)] = z[kb][k];
    }
    __syncthreads();
    // inverse pass 2 -> i0 = jj + A0 k, then the inverse split DFT and store
    {
      cpx<T> y[P][A0 / B0][B0];
#pragma unroll
      for (int kb = 0; kb < P; kb++)
#pragma unroll
        for (int b = 0; b < A0 / B0; b++) {
          const int jj = l + b * TPA;
#pragma unroll
          for (int k = 0; k < B0; k++) {
            y[kb][b][k] = buf[pd(((c * P + kb) * N0 + jj + k * A0) * W + w)];
            if (k) y[kb][b][k] = cmulc(y[kb][b][k], tw[jj * k]);
          }
          bfly<B0, true>(y[kb][b]);
        }
      if constexpr (P > 1) {
#pragma unroll
        for (int b = 0; b < A0 / B0; b++)
#pragma unroll
          for (int k = 0; k < B0; k++) {
            cpx<T> col[P];
#pragma unroll
            for (int kb = 0; kb < P; kb++) col[kb] = y[kb][b][k];
            bfly<P, true>(col);
#pragma unroll
            for (int kb = 0; kb < P; kb++) y[kb][b][k] = col[kb];
          }
      }
      if (wok) {
        cpx<T>* dst = ws + st.off + (int64_t)c * N0 * P * F + f0 + w;
#pragma unroll
        for (int b = 0; b < A0 / B0; b++)
#pragma unroll
          for (int k = 0; k < B0; k++) {
            const int i0 = l + b * TPA + k * A0;
#pragma unroll
            for (int pp_ = 0; pp_ < P; pp_++) dst[((int64_t)i0 * P + pp_) * F] = y[pp_][b][k];
          }
      }
    }
  }
}

// ---------------------------------------------------------- pencil, N0 = R^2
//
// For N0 = R x R the axis-0 transform is one radix-R pass in registers, an
// LDS transposition, and a second radix-R pass, and the output of the second
// pass holds k0 = l + R j: exactly the input set of the inverse's first pass.
// The components are therefore transformed one after the other through
// one-component LDS buffers (P N0 W entries; two of them alternate), the
// modal block is applied in registers (a thread holds all DIM components of
// its frequencies), and each component is transformed back and stored.  The
// split index p is a lane bit (partner lane = lane ^ W): the P = 2 split DFT
// is a lane exchange.

template <typename T, int R, int W, bool INV>
__device__ __forceinline__ void pencil_r_fft(cpx<T> (&v)[R], cpx<T>* buf, const cpx<T>* tw,
                                             int rowbase, int l, int w) {
  // registers (index k of n = l + R k) -> radix R -> LDS [p][m][l] -> twiddle
  // w_N0^(i l') -> radix R -> registers (index j of n' = l' + R j)
  bfly<R, INV>(v);
#pragma unroll
  for (int m = 0; m < R; m++) buf[pd((rowbase + m) * R * W + l * W + w)] = v[m];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < R; i++) {
    cpx<T> x = buf[pd((rowbase + l) * R * W + i * W + w)];
    if (i) x = INV ? cmulc(x, tw[i * l]) : cmul(x, tw[i * l]);
    v[i] = x;
  }
  bfly<R, INV>(v);
}

// Two-point DFT over the split index held by lanes w and w + 16 of each
// 32-lane group (P = 2, W = 16).  Forward: lane p holds x_p on entry and
// y_p on exit; inverse: unnormalised, the same network.
template <typename T, int P, int W, bool INV>
__device__ __forceinline__ cpx<T> split_dft(cpx<T> x, int p) {
  static_assert(P == 2 && W == 16 && sizeof(T) == 4, "lane exchange of 16-lane rows, fp32");
  (void)p;
  // v_permlane16_swap(a, b) gives the even row (a_even, a_odd) and the odd
  // row (b_even, b_odd): swapping (x.x, x.y) hands the even row both real
  // parts and the odd row both imaginary parts, each row forms their sum and
  // difference, and swapping (sum, difference) gives the even row x_0 + x_1
  // and the odd row x_0 - x_1.  The two-point DFT is its own inverse up to a
  // factor 2, so both directions are the same exchange.
  const auto r1 = __builtin_amdgcn_permlane16_swap(__float_as_uint(x.x), __float_as_uint(x.y),
                                                   false, false);
  const T a = __uint_as_float(r1[0]), b = __uint_as_float(r1[1]);
  const auto r2 = __builtin_amdgcn_permlane16_swap(__float_as_uint(a + b), __float_as_uint(a - b),
                                                   false, false);
  return {__uint_as_float(r2[0]), __uint_as_float(r2[1])};
}

// Packed DC / Nyquist columns.  Along the last axis the entries k = 0 and
// k = M of a row are real after the row transform; the plane kernels keep
// them packed as one complex column Z = X0 + i XM through the middle-axis
// transform, in the k = 0 slot of each row, so the workspace holds exactly
// Q M columns per row group and the pencil kernel runs Q M / W tiles.  Over
// the full middle axis Y0(k1) = (Z(k1) + conj Z(-k1)) / 2 and
// YM(k1) = -i (Z(k1) - conj Z(-k1)) / 2 are the k = 0 and k = M columns; with
// the split P = 2 the partner of k1 = ka + Q kb is -k1 = (Q - ka) + Q (1 - kb).
// The first segment of each row is therefore walked in paired tiles: lanes
// w < W/2 hold row a, lanes w >= W/2 row Q - a (row Q/2 for a = 0), and the
// k = 0 lanes exchange Z with their partner lane after the split DFT (lane A
// keeps Y0(k1), lane B YM(k1) of the same k1) and exchange back before the
// inverse split DFT (Z(k1) = Y0 + i YM, Z(-k1) = conj Y0 + i conj YM).  For
// k1 = 0 and k1 = Q (row 0, lane w = 0) Y0 and YM are real sequences along
// axis 0 packed in one lane: after the axis-0 transform their spectra are
// separated with the partner frequency -k0 through LDS and receive the
// blocks of k2 = 0 and k2 = M.
template <int PM, int W>
struct PackedTile {
  int row, k;          // the lane's column: row ka, entry k
  int k1, kl;          // frequencies of the lane's modal block
  int xmask;           // partner lane = lane ^ xmask (0: none)
  bool role_a;         // keeps Y0 (else YM)
  bool real_pair;      // k1 in {0, Q}, k2 in {0, M} packed in one lane
};
template <int PM, int W>
__device__ __forceinline__ PackedTile<PM, W> packed_tile(int t, int w, int kb, int q) {
  constexpr int SEG = PM / W;  // tiles per row
  PackedTile<PM, W> r;
  const int row = t / SEG, sg = t - (t / SEG) * SEG;
  r.xmask = 0;
  r.role_a = true;
  r.real_pair = false;
  if (sg) {
    r.row = row;
    r.k = sg * W + w;
  } else {
    const int a = row >> 1, h = row & 1;
    const bool hi = w >= W / 2;
    r.row = hi ? (a ? q - a : q / 2) : a;
    r.k = h * (W / 2) + (w & (W / 2 - 1));
  }
  r.k1 = r.row + q * kb;
  r.kl = r.k;
  if (!sg && !(row & 1) && r.k == 0) {
    const int a = row >> 1;
    if (a) {
      r.xmask = W + W / 2;  // (l, kb, 0) <-> (l, 1 - kb, W/2)
      r.role_a = w == 0;
      if (!r.role_a) {
        r.k1 = a + q * (1 - kb);
        r.kl = PM;
      }
    } else if (w) {
      r.xmask = W;  // row Q/2: (l, kb) <-> (l, 1 - kb)
      r.role_a = kb == 0;
      r.k1 = q / 2;
      r.kl = kb ? PM : 0;
    } else {
      r.real_pair = true;
    }
  }
  return r;
}

template <typename T>
__device__ __forceinline__ cpx<T> shfl_c(cpx<T> x, int src) {
  return cpx<T>{__shfl(x.x, src), __shfl(x.y, src)};
}

template <typename T, int R, int P, int W, int MODE, int PM>
__global__ void __launch_bounds__(W * R * P)
    k_pencil_r(cpx<T>* __restrict__ ws, const PencilParams pp) {
  constexpr int N0 = R * R;
  static_assert(P == 2 && W == 16 && PM % W == 0, "packed columns with the split P = 2");
  extern __shared__ __align__(16) unsigned char smem[];
  cpx<T>* buf = reinterpret_cast<cpx<T>*>(smem);
  // Two buffers alternate over the six transforms of a tile (3 components
  // forward, 3 inverse): a buffer is rewritten two transforms after it was
  // read, and the barrier of the transform between orders the two, so each
  // transform needs one barrier.
  constexpr int BS = pd_size(P * N0 * W);
  cpx<T>* tw = buf + 2 * BS;
  const ModalCoef<T>* coef0 = reinterpret_cast<const ModalCoef<T>*>(pp.coef[0]);
  const ModalCoef<T>* coef1 = reinterpret_cast<const ModalCoef<T>*>(pp.coef[1]);
  const ModalCoef<T>* coef2 = reinterpret_cast<const ModalCoef<T>*>(pp.coef[2]);
  const int tid = threadIdx.x;
  int t = blockIdx.x;
  if (pp.xcd_remap) {
    // consecutive tiles share L2 lines: keep them on one XCD
    const int n = gridDim.x, per = n / 8, rem = n % 8;
    const int xx = t % 8, idx = t / 8;
    t = xx * per + (xx < rem ? xx : rem) + idx;
  }
  {
    // The paired tiles of even rows (tile 2 a SEG, a < Q/2) carry the lane
    // exchanges and take longer: they are dispatched first, so that the
    // workgroup slots they hold take fewer of the other tiles.
    constexpr int G = 2 * (PM / W);  // one exchange tile in every G tiles
    const int nx = (int)gridDim.x / G;
    t = t < nx ? t * G : ((t - nx) / (G - 1)) * G + (t - nx) % (G - 1) + 1;
  }
  const int F = pp.f;  // Q M columns per row group (c, i0, p)
  const int q = pp.q;
  const int w = tid % W;
  const int p = (tid / W) % P;
  const int l = tid / (W * P);
  const int kb = p;  // split frequency this lane holds after the split DFT
  const PackedTile<PM, W> pt = packed_tile<PM, W>(t, w, kb, q);
  const int lane = tid & 63;
  const int src = lane ^ pt.xmask;
  const bool paired = t % (PM / W) == 0 && ((t / (PM / W)) & 1) == 0;  // tiles with k = 0 lanes
  copy_table(tw, reinterpret_cast<const cpx<T>*>(pp.tw0), N0, tid, W * R * P);
  // axis-0 block coefficients in LDS (read once per frequency, 16 per
  // thread: LDS latency instead of a chain of L2 loads)
  ModalCoef<T>* cf0 = reinterpret_cast<ModalCoef<T>*>(tw + N0);
  for (int i = tid; i < N0; i += W * R * P) cf0[i] = coef0[i];
  // (c, i0 = l + R k, p) lies at (l P + p) F + f + (c N0 + R k) P F.  The
  // lane's pointer is re-derived from an opaque copy of its column per
  // component, so the compiler does not precompute (and keep) every address.
  const int fl = pt.row * PM + pt.k;
  auto lane_ptr = [&](int c) {
    const int tl = opaque(tid);
    const int ll = tl / (W * P), lp = (tl / W) % P;
    return ws + (int64_t)(ll * P + lp) * F + opaque(fl) + (int64_t)c * N0 * P * F;
  };

  cpx<T> X[3][R];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const cpx<T>* srcp = lane_ptr(c);
#pragma unroll
    for (int k = 0; k < R; k++) X[c][k] = srcp[(int64_t)(R * k) * P * F];
#pragma unroll
    for (int k = 0; k < R; k++) X[c][k] = split_dft<T, P, W, false>(X[c][k], p);
    if (paired) {
      // unpack: lane A keeps Y0 = (Z_A + conj Z_B) / 2, lane B keeps
      // YM = -i (Z_A - conj Z_B) / 2 (Z_B = Z(-k1)); exchanges in groups of
      // four, then selects (no divergent branch per value)
#pragma unroll
      for (int g = 0; g < R; g += 4) {
        cpx<T> o[4];
#pragma unroll
        for (int k = 0; k < 4; k++) o[k] = shfl_c(X[c][g + k], src);
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const cpx<T> x = X[c][g + k];
          const cpx<T> ya = {T(0.5) * (x.x + o[k].x), T(0.5) * (x.y - o[k].y)};
          const cpx<T> yb = {T(0.5) * (o[k].y + x.y), T(0.5) * (x.x - o[k].x)};
          const cpx<T> y = pt.role_a ? ya : yb;
          X[c][g + k] = pt.xmask ? y : x;
        }
      }
    }
    pencil_r_fft<T, R, W, false>(X[c], buf + (c & 1) * BS, tw, kb * R, l, w);  // (its barrier also covers tw)
  }

  const T mu = T(pp.mu), s = T(pp.s), kinv = T(pp.kinv);
  cpx<T>* E = buf + BS;  // free until the first inverse transform
  if (t == 0) {
    // row 0, k = 0: the spectra of the packed real pair, before the block
    if (pt.real_pair) {
#pragma unroll
      for (int c = 0; c < 3; c++)
#pragma unroll
        for (int j = 0; j < R; j++) E[(c * 2 + kb) * N0 + l + R * j] = X[c][j];
    }
    __syncthreads();
  }

  // modal block at k0 = l + R j of the lane's (k1, k2)
  {
    const ModalCoef<T> cc = coef2[pt.kl];
    const ModalCoef<T> b = coef1[pt.k1];
    const T chichi = b.chi * cc.chi, phichi = b.phi * cc.chi, chiphi = b.chi * cc.phi;
    const T s_psichi = s * b.psi * cc.chi, s_chipsi = s * b.chi * cc.psi;
    const T s_psipsi = s * b.psi * cc.psi;
    const bool kz = pt.k1 == 0 && pt.kl == 0;
#pragma unroll
    for (int j = 0; j < R; j++) {
      const int k0 = l + R * j;
      const ModalCoef<T> a = cf0[k0];
      const T H0 = a.phi * chichi, H1 = a.chi * phichi, H2 = a.chi * chiphi;
      const T dg = mu * (H0 + H1 + H2);
      cpx<T> xx[3] = {X[0][j], X[1][j], X[2][j]};
      sym3_apply<T, MODE>(s * H0 + dg, s * H1 + dg, s * H2 + dg, a.psi * s_psichi,
                          a.psi * s_chipsi, a.chi * s_psipsi, kz && k0 == 0, xx, kinv);
      X[0][j] = xx[0];
      X[1][j] = xx[1];
      X[2][j] = xx[2];
    }
  }

  if (t == 0) {
    if (pt.real_pair) {
      // F = Y0^ + i YM^ with Y0^, YM^ Hermitian in k0: Y0^ = (F + conj F(-k0))/2,
      // YM^ = -i (F - conj F(-k0))/2; out = B(k2 = 0) Y0^ + i B(k2 = M) YM^
      auto block = [&](int k0, int kl, bool zero, cpx<T>* xx) {
        const ModalCoef<T> a = coef0[k0], b = coef1[q * kb], cc = coef2[kl];
        const T H0 = a.phi * (b.chi * cc.chi), H1 = a.chi * (b.phi * cc.chi);
        const T H2 = a.chi * (b.chi * cc.phi);
        const T dg = mu * (H0 + H1 + H2);
        sym3_apply<T, MODE>(s * H0 + dg, s * H1 + dg, s * H2 + dg, a.psi * (s * b.psi * cc.chi),
                            a.psi * (s * b.chi * cc.psi), a.chi * (s * b.psi * cc.psi), zero, xx,
                            kinv);
      };
#pragma unroll
      for (int j = 0; j < R; j++) {
        const int k0 = l + R * j;
        cpx<T> y0[3], ym[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
          const cpx<T> f = E[(c * 2 + kb) * N0 + k0];
          const cpx<T> g = E[(c * 2 + kb) * N0 + ((N0 - k0) & (N0 - 1))];
          y0[c] = {T(0.5) * (f.x + g.x), T(0.5) * (f.y - g.y)};
          ym[c] = {T(0.5) * (f.y + g.y), T(0.5) * (g.x - f.x)};
        }
        block(k0, 0, kb == 0 && k0 == 0, y0);
        block(k0, PM, false, ym);
#pragma unroll
        for (int c = 0; c < 3; c++) X[c][j] = {y0[c].x - ym[c].y, y0[c].y + ym[c].x};
      }
    }
    __syncthreads();  // E lies in the buffer of the first inverse transform
  }

  // inverse: per component, the two passes, the repack, the inverse split
  // DFT, the store
#pragma unroll
  for (int c = 0; c < 3; c++) {
    pencil_r_fft<T, R, W, true>(X[c], buf + ((c + 1) & 1) * BS, tw, kb * R, l, w);
    if (paired) {
      // lane A: Z = Y0 + i YM; lane B: Z(-k1) = conj Y0 + i conj YM
#pragma unroll
      for (int g = 0; g < R; g += 4) {
        cpx<T> o[4];
#pragma unroll
        for (int k = 0; k < 4; k++) o[k] = shfl_c(X[c][g + k], src);
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const cpx<T> x = X[c][g + k];
          const cpx<T> za = {x.x - o[k].y, x.y + o[k].x}, zb = {o[k].x + x.y, x.x - o[k].y};
          const cpx<T> z = pt.role_a ? za : zb;
          X[c][g + k] = pt.xmask ? z : x;
        }
      }
    }
#pragma unroll
    for (int k = 0; k < R; k++) X[c][k] = split_dft<T, P, W, true>(X[c][k], p);
    cpx<T>* dst = lane_ptr(c);
#pragma unroll
    for (int k = 0; k < R; k++) dst[(int64_t)(R * k) * P * F] = X[c][k];
  }
}

// ------------------------------------------------------------------ dispatch

template <typename K>
static hipError_t set_lds_s(K kernel, size_t lds) {
  if (lds > 65536)
    return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  return hipSuccess;
}

template <typename T, int M, int R1, int R2, int Q, int C1, int C2, int P>
struct PlaneS {
  static size_t lds_fwd() {
    return (prow_size(Q, M) + 2 * M + P * Q + 2 * Q + R1 * R2) * sizeof(cpx<T>);
  }
  static size_t lds_inv() {
    return (prow_size(Q, M) + Q + 2 * M + P * Q + Q + R1 * R2) * sizeof(cpx<T>);
  }
  static constexpr int threads() { return Q * M / R1; }
};

template <typename T, int N0, int A0, int B0, int P, int W>
struct PencilS {
  static size_t lds() { return (pd_size(3 * P * N0 * W) + N0) * sizeof(cpx<T>); }
  static constexpr int threads() { return 3 * W * (N0 / A0); }
};

// Tiles of width W over the fixed-size layout (rows, then Nyquist region).
static void set_tiles(PencilParams& pc, int W) {
  pc.nyq_tiles = (pc.q + W - 1) / W;
  pc.ntiles = (pc.f + W - 1) / W + pc.nyq_tiles;
}

// The register-block pencil (and with it the packed DC / Nyquist columns)
// serves fp32 grids with N0 = R^2, split P = 2 and 16-column tiles.
template <typename T, int N0, int A0, int B0, int P, int M>
constexpr bool packed_case() {
  return A0 == B0 && A0 * B0 == N0 && P == 2 && sizeof(T) == 4 && M % 16 == 0;
}

// Pencil launch of a fixed-size case: the register-block pencil over packed
// columns where it applies unless bit 3 of `pf` asks for the
// all-components-in-LDS pencil; bit 2: persistent variant of the latter.
template <typename T, int N0, int A0, int B0, int P, int W, int M>
static hipError_t launch_pencil_static(cpx<T>* ws, PencilParams pc, int mode, int pf,
                                       hipStream_t st) {
  if constexpr (packed_case<T, N0, A0, B0, P, M>()) {
    if (!(pf & 8)) {
      constexpr int WR = 16;  // 128-byte rows
      auto kp = mode == 0 ? k_pencil_r<T, A0, P, WR, 0, M> : k_pencil_r<T, A0, P, WR, 1, M>;
      const size_t lds = (2 * pd_size(P * N0 * WR) + N0) * sizeof(cpx<T>) + N0 * sizeof(ModalCoef<T>);
      hipError_t e = set_lds_s(kp, lds);
      if (e != hipSuccess) return e;
      pc.nyq_tiles = 0;
      pc.ntiles = pc.f / WR;
      hipLaunchKernelGGL(kp, dim3(pc.ntiles), dim3(WR * A0 * P), lds, st, ws, pc);
      return hipGetLastError();
    }
  }
  set_tiles(pc, W);
  using KS = PencilS<T, N0, A0, B0, P, W>;
  auto kp = mode == 0 ? ((pf & 4) ? k_pencil_s<T, N0, A0, B0, P, W, 0, true>
                                  : k_pencil_s<T, N0, A0, B0, P, W, 0, false>)
                      : ((pf & 4) ? k_pencil_s<T, N0, A0, B0, P, W, 1, true>
                                  : k_pencil_s<T, N0, A0, B0, P, W, 1, false>);
  hipError_t e = set_lds_s(kp, KS::lds());
  if (e != hipSuccess) return e;
  const int gp = (pf & 4) ? resident_blocks(kp, KS::threads(), KS::lds(), pc.ntiles) : pc.ntiles;
  hipLaunchKernelGGL(kp, dim3(gp), dim3(KS::threads()), KS::lds(), st, ws, pc);
  return hipGetLastError();
}

// Supported fixed-size geometries:
//   dtype, N0, N1, N2,
//   plane: M = N2/2 = R1 R2, Q = N1/P = C1 C2, split P,
//   pencil: N0 = A0 B0, tile width W.
#define B17_STATIC_CASES(X)                                                           \
  X(float, 256, 256, 256, /*plane*/ 128, 16, 8, 128, 16, 8, 2, /*pencil*/ 16, 16, 8)  \
  X(float, 128, 128, 128, /*plane*/ 64, 16, 4, 128, 16, 8, 1, /*pencil*/ 16, 8, 16)   \
  X(float, 64, 64, 64, /*plane*/ 32, 16, 2, 64, 16, 4, 1, /*pencil*/ 16, 4, 32)       \
  X(double, 128, 128, 128, /*plane*/ 64, 16, 4, 128, 16, 8, 1, /*pencil*/ 16, 8, 8)   \
  X(double, 64, 64, 64, /*plane*/ 32, 16, 2, 64, 16, 4, 1, /*pencil*/ 16, 4, 16)

// persistent plane grid: workgroups resident at once on the device
template <typename K>
static int resident_blocks(K kernel, int threads, size_t lds, int nitems) {
  static int ncu = 0;
  if (!ncu) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      ncu = 256;
  }
  int per = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, reinterpret_cast<const void*>(kernel),
                                                   threads, lds) != hipSuccess || per < 1)
    per = 1;
  const int n = ncu * per;
  return n < nitems ? n : nitems;
}

bool static_supported(int dtype, int n0, int n1, int n2, StaticGeom* g) {
#define B17_MATCH(T, A, B, C, M, R1, R2, Q, C1, C2, P, A0, B0, W)             \
  if (dtype == (sizeof(T) == 4 ? 0 : 1) && n0 == A && n1 == B && n2 == C) {  \
    if (g) {                                                                 \
      g->p = P;                                                              \
      g->w = W;                                                              \
      g->plane_threads = PlaneS<T, M, R1, R2, Q, C1, C2, P>::threads();      \
      g->plane_lds_fwd = PlaneS<T, M, R1, R2, Q, C1, C2, P>::lds_fwd();      \
      g->plane_lds_inv = PlaneS<T, M, R1, R2, Q, C1, C2, P>::lds_inv();      \
      g->pencil_threads = PencilS<T, A, A0, B0, P, W>::threads();            \
      g->pencil_lds = PencilS<T, A, A0, B0, P, W>::lds();                    \
    }                                                                        \
    return true;                                                             \
  }
  B17_STATIC_CASES(B17_MATCH)
#undef B17_MATCH
  return false;
}

std::string static_describe(int dtype, int n0, int n1, int n2, int pf) {
#define B17_DESC(T, A, B, C, M, R1, R2, Q, C1, C2, P, A0, B0, W)                                  \
  if (dtype == (sizeof(T) == 4 ? 0 : 1) && n0 == A && n1 == B && n2 == C) {                       \
    const bool reg = packed_case<T, A, A0, B0, P, M>() && !(pf & 8);                              \
    const int wr = reg ? 16 : W;                                                                  \
    return std::string("{\"plane_split\":") + std::to_string(P) +                                 \
           ",\"plane_threads\":" + std::to_string(Q * M / R1) +                                  \
           ",\"plane_forward_persistent\":" + ((pf & 1) ? "true" : "false") +                   \
           ",\"plane_inverse_persistent\":" + ((pf & 2) ? "true" : "false") +                   \
           ",\"pencil\":\"" + (reg ? "register-block" : "lds") + "\",\"dc_nyquist\":\"" +      \
           (reg ? "packed" : "separate") + "\",\"pencil_width\":" +                              \
           std::to_string(wr) + ",\"pencil_threads\":" +                                        \
           std::to_string(reg ? wr * A0 * P : 3 * W * (A / A0)) + "}";                            \
  }
  B17_STATIC_CASES(B17_DESC)
#undef B17_DESC
  return std::string();
}

hipError_t launch_static(int dtype, int n0, int n1, int n2, int mode, const void* in, void* out,
                         void* ws, const PencilParams& pc_in, const void* twr, const void* twc,
                         int pf, hipStream_t st, hipEvent_t* ev) {
  hipError_t e = hipErrorInvalidValue;
  auto mark = [&](int i) {
    if (ev && e == hipSuccess) e = hipEventRecord(ev[i], st);
  };
#define B17_RUN(T, A, B, C, M, R1, R2, Q, C1, C2, P, A0, B0, W)                                   \
  if (dtype == (sizeof(T) == 4 ? 0 : 1) && n0 == A && n1 == B && n2 == C) {                       \
    using PS = PlaneS<T, M, R1, R2, Q, C1, C2, P>;                                                \
    /* packed DC / Nyquist columns where the register-block pencil runs */                        \
    constexpr bool PK = packed_case<T, A, A0, B0, P, M>();                                        \
    const bool pack = PK && !(pf & 8);                                                            \
    auto kf = pack ? ((pf & 1) ? k_plane_fwd_s<T, M, R1, R2, Q, C1, C2, P, true, PK>              \
                               : k_plane_fwd_s<T, M, R1, R2, Q, C1, C2, P, false, PK>)            \
                   : ((pf & 1) ? k_plane_fwd_s<T, M, R1, R2, Q, C1, C2, P, true, false>           \
                               : k_plane_fwd_s<T, M, R1, R2, Q, C1, C2, P, false, false>);        \
    auto ki = pack ? ((pf & 2) ? k_plane_inv_s<T, M, R1, R2, Q, C1, C2, P, true, PK>              \
                               : k_plane_inv_s<T, M, R1, R2, Q, C1, C2, P, false, PK>)            \
                   : ((pf & 2) ? k_plane_inv_s<T, M, R1, R2, Q, C1, C2, P, true, false>           \
                               : k_plane_inv_s<T, M, R1, R2, Q, C1, C2, P, false, false>);        \
    e = set_lds_s(kf, PS::lds_fwd());                                                             \
    if (e == hipSuccess) e = set_lds_s(ki, PS::lds_inv());                                        \
    const int nitems = 3 * A * P;                                                                 \
    /* rows of M entries, then (unless packed) the Nyquist region */                              \
    PencilParams pc = pc_in;                                                                      \
    pc.rs = M;                                                                                    \
    pc.rsdiv = make_fastdiv(M);                                                                   \
    pc.f = Q * M;                                                                                 \
    pc.nyq_off = (int64_t)3 * A * B * M;                                                          \
    cpx<T>* nyq = (cpx<T>*)ws + pc.nyq_off;                                                       \
    const int gf = (pf & 1) ? resident_blocks(kf, PS::threads(), PS::lds_fwd(), nitems) : nitems; \
    const int gi = (pf & 2) ? resident_blocks(ki, PS::threads(), PS::lds_inv(), nitems) : nitems; \
    mark(0);                                                                                      \
    if (e == hipSuccess) {                                                                        \
      hipLaunchKernelGGL(kf, dim3(gf), dim3(PS::threads()), PS::lds_fwd(), st, (const T*)in,      \
                         (cpx<T>*)ws, nyq, (const cpx<T>*)twr, (const cpx<T>*)twc, nitems);       \
      e = hipGetLastError();                                                                      \
    }                                                                                             \
    mark(1);                                                                                      \
    if (e == hipSuccess) {                                                                        \
      e = launch_pencil_static<T, A, A0, B0, P, W, M>((cpx<T>*)ws, pc, mode, pf, st);            \
    }                                                                                             \
    mark(2);                                                                                      \
    if (e == hipSuccess) {                                                                        \
      hipLaunchKernelGGL(ki, dim3(gi), dim3(PS::threads()), PS::lds_inv(), st,                    \
                         (const cpx<T>*)ws, (const cpx<T>*)nyq, (T*)out, (const cpx<T>*)twr,      \
                         (const cpx<T>*)twc, nitems);                                             \
      e = hipGetLastError();                                                                      \
    }                                                                                             \
    mark(3);                                                                                      \
    return e;                                                                                     \
  }
  B17_STATIC_CASES(B17_RUN)
#undef B17_RUN
  return e;
}

}  // namespace b17
