// Device bodies of the runtime-size kernels: one workgroup processes one
// plane block (R2C / C2R along the last axis + the middle-axis stage) or one
// pencil tile (axis-0 FFT, modal block, inverse).  Used by the one-launch-per-
// stage kernels of kernels.hip and by the persistent batch kernel of
// batch.hip.  Every loop is written for any workgroup size nthr >= the
// threads-per-sequence of the transforms involved.
#pragma once

#include "fft_core.hpp"
#include "kernels.hpp"
#include "modal.hpp"

namespace b17 {

constexpr int kMaxE = 16;  // max complex elements per thread per FFT pass

// global loads a thread keeps in flight when staging (fewer for fp64: VGPRs)
template <typename T>
constexpr int kLoads = sizeof(T) == 4 ? 16 : 8;

// Copy `tot` elements with up to LB loads in flight per thread: all loads of
// a batch are issued before any of their values is consumed, so a staging
// loop is bound by bandwidth instead of one memory latency per element.
template <int LB, typename V, typename Get, typename Put>
__device__ __forceinline__ void staged_copy(int tot, int tid, int nthr, Get get, Put put) {
  for (int e0 = tid; e0 < tot; e0 += LB * nthr) {
    V v[LB];
#pragma unroll
    for (int i = 0; i < LB; i++) {
      const int e = e0 + i * nthr;
      if (e < tot) v[i] = get(e);
    }
#pragma unroll
    for (int i = 0; i < LB; i++) {
      const int e = e0 + i * nthr;
      if (e < tot) put(e, v[i]);
    }
  }
}

// The field is read once (forward) and written once (inverse): non-temporal
// accesses leave the last-level cache to the half-spectrum workspace, which
// the next sweep reads from there.
template <typename T>
__device__ __forceinline__ cpx<T> nt_load_c(gptr<const cpx<T>> p) {
  const gptr<const T> q = (gptr<const T>)p;
  return cpx<T>{__builtin_nontemporal_load(q), __builtin_nontemporal_load(q + 1)};
}
template <typename T>
__device__ __forceinline__ void nt_store_c(gptr<cpx<T>> p, cpx<T> v) {
  const gptr<T> q = (gptr<T>)p;
  __builtin_nontemporal_store(v.x, q);
  __builtin_nontemporal_store(v.y, q + 1);
}

template <typename T>
__device__ __forceinline__ cpx<T> post_r2c(cpx<T> A, cpx<T> B, cpx<T> w) {
  // X[k] = (A + conj B)/2 - i w (A - conj B)/2, A = Z[k], B = Z[m-k]
  cpx<T> e = {T(0.5) * (A.x + B.x), T(0.5) * (A.y - B.y)};
  cpx<T> d = {T(0.5) * (A.x - B.x), T(0.5) * (A.y + B.y)};
  cpx<T> o = {d.y, -d.x};  // -i d
  return e + cmul(w, o);
}

template <typename T>
__device__ __forceinline__ cpx<T> pre_c2r(cpx<T> A, cpx<T> B, cpx<T> w) {
  // Z[k] = (A + conj B) + i conj(w) (A - conj B), A = X[k], B = X[m-k]
  cpx<T> e = {A.x + B.x, A.y - B.y};
  cpx<T> d = {A.x - B.x, A.y + B.y};
  cpx<T> o = cmulc(d, w);
  return {e.x - o.y, e.y + o.x};
}

// Twiddle tables of a work item, copied from global memory into LDS after
// the item's data: the row table exp(-2 pi i j / nl), j < nl, then (3D) the
// middle-axis table exp(-2 pi i j / n1), j < n1 (plane items; the host sizes
// the LDS with plane_table_entries), or the axis-0 table exp(-2 pi i j / n0)
// and, for splits run as a Stockham transform, the middle-axis table (pencil
// tiles; pencil_table_entries).  Every FFT pass reads its twiddles from
// there: an LDS read instead of a dependent global load whose latency each
// pass would otherwise wait out.
template <typename T>
__device__ __forceinline__ cpx<T>* plane_tables(cpx<T>* lds, const PlaneParams& pp) {
  return lds + pp.g3 * pp.q * pp.rsl;
}

template <typename T>
__device__ __forceinline__ void copy_tables(const PlaneParams& pp, cpx<T>* twr, cpx<T>* twc,
                                            int tid, int nthr) {
  const gptr<const cpx<T>> gr = (gptr<const cpx<T>>)pp.tw_row;
  for (int i = tid; i < pp.nl; i += nthr) twr[i] = gr[i];
  if (pp.has_col) {
    const gptr<const cpx<T>> gc = (gptr<const cpx<T>>)pp.tw_col;
    for (int i = tid; i < pp.n1; i += nthr) twc[i] = gc[i];
  }
}

// ------------------------------------------------------------------ plane fwd

template <typename T, int PR = 0>
__device__ __forceinline__ void plane_fwd_body(gptr<const T> __restrict__ u,
                                               gptr<cpx<T>> __restrict__ ws, const PlaneParams& pp,
                                               int blk, cpx<T>* lds) {
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int grp = blk / pp.p;
  const int p = blk - grp * pp.p;
  // 3D items of the batch kernel may hold g3 consecutive planes (split 1:
  // their rows are consecutive in memory and in LDS); elsewhere g3 = 1
  const int plane = grp * pp.g3;
  int nrows = pp.q;
  if (!pp.has_col) {
    const int rem = pp.nrows - plane * pp.n1;
    nrows = rem < pp.q ? rem : pp.q;
  } else if (pp.g3 > 1) {
    const int np = pp.nplanes - plane;
    nrows = pp.q * (np < pp.g3 ? np : pp.g3);
  }
  const int64_t row0 = (int64_t)plane * pp.n1 + p;
  const int pst = pp.p;
  // twiddle tables in LDS after the rows (visible after the first barrier)
  cpx<T>* const twr = plane_tables<T>(lds, pp);
  cpx<T>* const twc = twr + pp.nl;
  copy_tables<T>(pp, twr, twc, tid, nthr);

  // 1. rows -> LDS.  Even lengths: a row is read as m complex pairs.  Odd
  //    lengths: rows 2j and 2j+1 are read as the real and imaginary parts of
  //    one complex row in slot 2j (one transform for two rows).
  {
    const int m = pp.m, rsl = pp.rsl;
    if (pp.even) {
      staged_copy<kLoads<T>, cpx<T>>(nrows * m, tid, nthr, [&](int e) {
        const int qq = (int)fdiv((uint32_t)e, pp.mdiv);
        return nt_load_c<T>((gptr<const cpx<T>>)(u + (row0 + (int64_t)qq * pst) * pp.nl) + (e - qq * m));
      }, [&](int e, cpx<T> v) {
        const int qq = (int)fdiv((uint32_t)e, pp.mdiv);
        lds[qq * rsl + (e - qq * m)] = v;
      });
    } else {
      staged_copy<kLoads<T>, cpx<T>>(((nrows + 1) / 2) * m, tid, nthr, [&](int e) {
        const int j = (int)fdiv((uint32_t)e, pp.mdiv);
        const int n = e - j * m;
        const gptr<const T> ra = u + (row0 + (int64_t)(2 * j) * pst) * pp.nl;
        return cpx<T>{__builtin_nontemporal_load(ra + n),
                      2 * j + 1 < nrows ? __builtin_nontemporal_load(ra + (int64_t)pst * pp.nl + n) : T(0)};
      }, [&](int e, cpx<T> v) {
        const int j = (int)fdiv((uint32_t)e, pp.mdiv);
        lds[2 * j * rsl + (e - j * m)] = v;
      });
    }
  }
  __syncthreads();

  // 2. row FFTs (a row's threads within one wave where possible)
  {
    const int nseq = pp.even ? nrows : (nrows + 1) / 2;
    const int sst = pp.even ? pp.rsl : 2 * pp.rsl;
    const SeqThreads m = seq_threads(tid, nthr, pp.row, false);
    for (int r0 = 0; r0 < nseq; r0 += m.S) {
      const int s = r0 + m.sl;
      const bool act = m.sl < m.S && s < nseq;
      lds_fft<T, false, kMaxE, PR>(lds, twr, pp.row, act ? s * sst : 0, 1, m.lane, act, m.wave);
    }
    if (m.wave) __syncthreads();
  }

  // 3. split the packed transform into the nl/2+1 spectrum entries
  if (pp.even) {
    const int m = pp.m, h = m / 2 + 1, tot = nrows * h;
    for (int e = tid; e < tot; e += nthr) {
      const int qq = (int)fdiv((uint32_t)e, pp.hdiv);
      const int k = e - qq * h;
      cpx<T>* row = lds + qq * pp.rsl;
      if (k == 0) {
        const cpx<T> z = row[0];
        row[0] = {z.x + z.y, T(0)};
        row[m] = {z.x - z.y, T(0)};
      } else {
        const int k2 = m - k;
        const cpx<T> A = row[k], B = row[k2];
        row[k] = post_r2c(A, B, twr[k]);
        if (k2 != k) row[k2] = post_r2c(B, A, twr[k2]);
      }
    }
    __syncthreads();
  } else {
    // odd length: Z = X + i Y of rows 2j, 2j+1 -> X[k] = (Z[k] + conj Z[N-k]) / 2
    // into slot 2j, Y[k] = (Z[k] - conj Z[N-k]) / 2i into slot 2j+1, k <= N/2
    // (position k < nlh is read and written by its own thread only; N-k >= nlh)
    const int N = pp.m, h = pp.nlh, tot = ((nrows + 1) / 2) * h;
    for (int e = tid; e < tot; e += nthr) {
      const int j = (int)fdiv((uint32_t)e, pp.nlhdiv);
      const int k = e - j * h;
      cpx<T>* z = lds + 2 * j * pp.rsl;
      const cpx<T> A = z[k], B = k ? z[N - k] : A;
      z[k] = cpx<T>{T(0.5) * (A.x + B.x), T(0.5) * (A.y - B.y)};
      if (2 * j + 1 < nrows) z[pp.rsl + k] = cpx<T>{T(0.5) * (A.y + B.y), T(0.5) * (B.x - A.x)};
    }
    __syncthreads();
  }

  // 4. 3D: Q-point stage of the middle axis on every column, then the split
  //    twiddle w_N1^(p ka) is applied on the way out
  if (pp.has_col) {
    const SeqThreads m = seq_threads(tid, nthr, pp.col, true);
    const int ncol = (pp.g3 > 1 ? nrows / pp.q : 1) * pp.nlh;  // columns of every plane
    for (int c0 = 0; c0 < ncol; c0 += m.S) {
      const int s = c0 + m.sl;
      const bool act = m.sl < m.S && s < ncol;
      const int g = act ? (int)fdiv((uint32_t)s, pp.nlhdiv) : 0;
      const int base = act ? g * pp.q * pp.rsl + (s - g * pp.nlh) : 0;
      lds_fft<T, false, kMaxE, PR>(lds, twc, pp.col, base, pp.rsl, m.lane, act, m.wave);
    }
    if (m.wave) __syncthreads();
  }

  // 5. store
  {
    const int tot = nrows * pp.nlh, nlh = pp.nlh;
    gptr<cpx<T>> dst = ws + ((int64_t)plane * pp.n1 + p * pp.q) * pp.rs;
    staged_copy<kLoads<T>, cpx<T>>(tot, tid, nthr, [&](int e) {
      const int ka = (int)fdiv((uint32_t)e, pp.nlhdiv);
      cpx<T> v = lds[ka * pp.rsl + (e - ka * nlh)];
      if (p > 0) v = cmul(v, twc[p * ka]);
      return v;
    }, [&](int e, cpx<T> v) {
      const int ka = (int)fdiv((uint32_t)e, pp.nlhdiv);
      dst[(int64_t)ka * pp.rs + (e - ka * nlh)] = v;
    });
  }
}

// ------------------------------------------------------------------ plane inv

template <typename T, int PR = 0>
__device__ __forceinline__ void plane_inv_body(gptr<const cpx<T>> __restrict__ ws,
                                               gptr<T> __restrict__ out, const PlaneParams& pp,
                                               int blk, cpx<T>* lds) {
  const gptr<const cpx<T>> twc_g = (gptr<const cpx<T>>)pp.tw_col;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int grp = blk / pp.p;
  const int p = blk - grp * pp.p;
  // 3D items of the batch kernel may hold g3 consecutive planes (split 1:
  // their rows are consecutive in memory and in LDS); elsewhere g3 = 1
  const int plane = grp * pp.g3;
  int nrows = pp.q;
  if (!pp.has_col) {
    const int rem = pp.nrows - plane * pp.n1;
    nrows = rem < pp.q ? rem : pp.q;
  } else if (pp.g3 > 1) {
    const int np = pp.nplanes - plane;
    nrows = pp.q * (np < pp.g3 ? np : pp.g3);
  }
  const int64_t row0 = (int64_t)plane * pp.n1 + p;
  const int pst = pp.p;
  // twiddle tables in LDS after the rows (visible after the first barrier)
  cpx<T>* const twr = plane_tables<T>(lds, pp);
  cpx<T>* const twc = twr + pp.nl;
  copy_tables<T>(pp, twr, twc, tid, nthr);

  // 1. load (undo the split twiddle, read from global memory: the LDS copy
  //    of the table is not visible before the barrier)
  {
    const int tot = nrows * pp.nlh, nlh = pp.nlh;
    gptr<const cpx<T>> src = ws + ((int64_t)plane * pp.n1 + p * pp.q) * pp.rs;
    staged_copy<kLoads<T>, cpx<T>>(tot, tid, nthr, [&](int e) {
      const int ka = (int)fdiv((uint32_t)e, pp.nlhdiv);
      return src[(int64_t)ka * pp.rs + (e - ka * nlh)];
    }, [&](int e, cpx<T> v) {
      const int ka = (int)fdiv((uint32_t)e, pp.nlhdiv);
      if (p > 0) v = cmulc(v, twc_g[p * ka]);
      lds[ka * pp.rsl + (e - ka * nlh)] = v;
    });
  }
  __syncthreads();

  // 2. 3D: inverse Q-point stage of the middle axis
  if (pp.has_col) {
    const SeqThreads m = seq_threads(tid, nthr, pp.col, true);
    const int ncol = (pp.g3 > 1 ? nrows / pp.q : 1) * pp.nlh;  // columns of every plane
    for (int c0 = 0; c0 < ncol; c0 += m.S) {
      const int s = c0 + m.sl;
      const bool act = m.sl < m.S && s < ncol;
      const int g = act ? (int)fdiv((uint32_t)s, pp.nlhdiv) : 0;
      const int base = act ? g * pp.q * pp.rsl + (s - g * pp.nlh) : 0;
      lds_fft<T, true, kMaxE, PR>(lds, twc, pp.col, base, pp.rsl, m.lane, act, m.wave);
    }
    if (m.wave) __syncthreads();
  }

  // 3. rebuild the transform the row FFT needs
  if (pp.even) {
    const int m = pp.m, h = m / 2 + 1, tot = nrows * h;
    for (int e = tid; e < tot; e += nthr) {
      const int qq = (int)fdiv((uint32_t)e, pp.hdiv);
      const int k = e - qq * h;
      cpx<T>* row = lds + qq * pp.rsl;
      if (k == 0) {
        const cpx<T> A = row[0], B = row[m];
        const cpx<T> ee = {A.x + B.x, A.y - B.y};
        const cpx<T> d = {A.x - B.x, A.y + B.y};
        row[0] = {ee.x - d.y, ee.y + d.x};
      } else {
        const int k2 = m - k;
        const cpx<T> A = row[k], B = row[k2];
        row[k] = pre_c2r(A, B, twr[k]);
        if (k2 != k) row[k2] = pre_c2r(B, A, twr[k2]);
      }
    }
  } else {
    // odd length: the half spectra X, Y of rows 2j, 2j+1 become the full
    // spectrum of x + i y in slot 2j: Z[k] = X[k] + i Y[k] and
    // Z[N-k] = conj X[k] + i conj Y[k] (k >= 1); the k = 0 entries are real
    const int N = pp.nl, h = pp.nlh, tot = ((nrows + 1) / 2) * h;
    for (int e = tid; e < tot; e += nthr) {
      const int j = (int)fdiv((uint32_t)e, pp.nlhdiv);
      const int k = e - j * h;
      cpx<T>* z = lds + 2 * j * pp.rsl;
      const cpx<T> X = z[k];
      const cpx<T> Y = 2 * j + 1 < nrows ? z[pp.rsl + k] : cpx<T>{T(0), T(0)};
      if (k) {
        z[k] = cpx<T>{X.x - Y.y, X.y + Y.x};
        z[N - k] = cpx<T>{X.x + Y.y, Y.x - X.y};
      } else {
        z[0] = cpx<T>{X.x, Y.x};
      }
    }
  }
  __syncthreads();

  // 4. inverse row FFTs (odd lengths: one per pair of rows)
  {
    const int nseq = pp.even ? nrows : (nrows + 1) / 2;
    const int sst = pp.even ? pp.rsl : 2 * pp.rsl;
    const SeqThreads m = seq_threads(tid, nthr, pp.row, false);
    for (int r0 = 0; r0 < nseq; r0 += m.S) {
      const int s = r0 + m.sl;
      const bool act = m.sl < m.S && s < nseq;
      lds_fft<T, true, kMaxE, PR>(lds, twr, pp.row, act ? s * sst : 0, 1, m.lane, act, m.wave);
    }
    if (m.wave) __syncthreads();
  }

  // 5. store real rows (odd lengths: rows 2j and 2j+1 from the real and
  //    imaginary parts of slot 2j)
  {
    const int m = pp.m, rsl = pp.rsl;
    if (pp.even) {
      staged_copy<kLoads<T>, cpx<T>>(nrows * m, tid, nthr, [&](int e) {
        const int qq = (int)fdiv((uint32_t)e, pp.mdiv);
        return lds[qq * rsl + (e - qq * m)];
      }, [&](int e, cpx<T> v) {
        const int qq = (int)fdiv((uint32_t)e, pp.mdiv);
        nt_store_c<T>((gptr<cpx<T>>)(out + (row0 + (int64_t)qq * pst) * pp.nl) + (e - qq * m), v);
      });
    } else {
      staged_copy<kLoads<T>, cpx<T>>(((nrows + 1) / 2) * m, tid, nthr, [&](int e) {
        const int j = (int)fdiv((uint32_t)e, pp.mdiv);
        return lds[2 * j * rsl + (e - j * m)];
      }, [&](int e, cpx<T> v) {
        const int j = (int)fdiv((uint32_t)e, pp.mdiv);
        const int n = e - j * m;
        const gptr<T> ra = out + (row0 + (int64_t)(2 * j) * pst) * pp.nl;
        __builtin_nontemporal_store(v.x, ra + n);
        if (2 * j + 1 < nrows) __builtin_nontemporal_store(v.y, ra + (int64_t)pst * pp.nl + n);
      });
    }
  }
}

// ---------------------------------------------------------------------- pencil

template <typename T, int R, bool INV>
__device__ __forceinline__ void pdft_r(cpx<T>* lds, int W, FastDiv wdiv, int items,
                                       int tid, int nthr) {
  for (int e = tid; e < items; e += nthr) {
    const int outer = (int)fdiv((uint32_t)e, wdiv);
    const int w = e - outer * W;
    cpx<T>* b = lds + outer * R * W + w;
    cpx<T> v[R];
#pragma unroll
    for (int k = 0; k < R; k++) v[k] = b[k * W];
    bfly<R, INV>(v);
#pragma unroll
    for (int k = 0; k < R; k++) b[k * W] = v[k];
  }
}

// P-point DFT over the split index p for every (c, i0, w) of the tile (the
// caller synchronises before the results are read).
template <typename T, bool INV>
__device__ __forceinline__ void pdft(cpx<T>* lds, int P, int W, FastDiv wdiv,
                                     int nouter, int tid, int nthr) {
  const int items = nouter * W;
  switch (P) {
    case 2: pdft_r<T, 2, INV>(lds, W, wdiv, items, tid, nthr); break;
    case 3: pdft_r<T, 3, INV>(lds, W, wdiv, items, tid, nthr); break;
    case 4: pdft_r<T, 4, INV>(lds, W, wdiv, items, tid, nthr); break;
    case 5: pdft_r<T, 5, INV>(lds, W, wdiv, items, tid, nthr); break;
    case 7: pdft_r<T, 7, INV>(lds, W, wdiv, items, tid, nthr); break;
    case 8: pdft_r<T, 8, INV>(lds, W, wdiv, items, tid, nthr); break;
    case 16: pdft_r<T, 16, INV>(lds, W, wdiv, items, tid, nthr); break;
    default: break;
  }
}

// P-point DFT for splits without a fused butterfly (P = 6, 9, 12, 27, ...):
// a Stockham transform per (c, i0, w) column of the tile, P elements at
// stride W, twiddles exp(-2 pi i j / P) = tw1[j q].
template <typename T, bool INV, int PR>
__device__ __forceinline__ void pdft_stockham(cpx<T>* lds, const cpx<T>* tw1, const PencilParams& pp,
                                              int nouter, int tid, int nthr) {
  const int W = pp.w, P = pp.p;
  const SeqThreads m = seq_threads(tid, nthr, pp.pfft, true);
  const int nseq = nouter * W;
  for (int s0 = 0; s0 < nseq; s0 += m.S) {
    const int s = s0 + m.sl;
    const bool act = m.sl < m.S && s < nseq;
    const int outer = act ? (int)fdiv((uint32_t)s, pp.wdiv) : 0;
    const int base = act ? outer * P * W + (s - outer * W) : 0;
    lds_fft<T, INV, kMaxE, PR>(lds, tw1, pp.pfft, base, W, m.lane, act, m.wave);
  }
}

template <typename T, int DIM, int MODE, int PR = 0>
__device__ __forceinline__ void pencil_body(gptr<cpx<T>> __restrict__ ws, const PencilParams& pp, int t,
                                            cpx<T>* lds) {
  constexpr int CIN = ModeComps<DIM, MODE>::cin, COUT = ModeComps<DIM, MODE>::cout;
  constexpr int CMAX = ModeComps<DIM, MODE>::cmax;
  const gptr<const ModalCoef<T>> coef[3] = {(gptr<const ModalCoef<T>>)pp.coef[0],
                                            (gptr<const ModalCoef<T>>)pp.coef[1],
                                            (gptr<const ModalCoef<T>>)pp.coef[2]};
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int W = pp.w, P = pp.p, N0 = pp.n0, PW = P * W;
  const int f0 = t * W;
  const int wv = pp.f - f0 < W ? pp.f - f0 : W;
  const int cst = N0 * PW;  // LDS distance between components
  // twiddle tables in LDS after the tile (visible after the first barrier)
  cpx<T>* const tw0 = lds + CMAX * cst;
  cpx<T>* const tw1 = tw0 + N0;
  {
    const gptr<const cpx<T>> g0 = (gptr<const cpx<T>>)pp.tw0;
    for (int i = tid; i < N0; i += nthr) tw0[i] = g0[i];
    if (P > 1 && !fused_split(P)) {
      const gptr<const cpx<T>> g1 = (gptr<const cpx<T>>)pp.tw1;
      for (int i = tid; i < P * pp.q; i += nthr) tw1[i] = g1[i];
    }
  }

  // load the tile: element (c, i0, p, w) at ((c N0 + i0) P + p) W + w
  staged_copy<kLoads<T>, cpx<T>>(CIN * N0 * PW, tid, nthr, [&](int e) {
    const int rest = (int)fdiv((uint32_t)e, pp.wdiv);
    const int w = e - rest * W;
    cpx<T> v = {T(0), T(0)};
    if (w < wv) v = ws[(int64_t)rest * pp.f + f0 + w];
    return v;
  }, [&](int e, cpx<T> v) { lds[e] = v; });
  __syncthreads();
  if (P > 1) {
    if (fused_split(P))
      pdft<T, false>(lds, P, W, pp.wdiv, CIN * N0, tid, nthr);
    else
      pdft_stockham<T, false, PR>(lds, tw1, pp, CIN * N0, tid, nthr);
    __syncthreads();  // the axis-0 passes read columns other threads wrote
  }

  const SeqThreads m = seq_threads(tid, nthr, pp.ax0, true);
  for (int s0 = 0; s0 < CIN * PW; s0 += m.S) {
    const int s = s0 + m.sl;
    const bool act = m.sl < m.S && s < CIN * PW;
    const int g = act ? (int)fdiv((uint32_t)s, pp.pwdiv) : 0;
    const int base = act ? g * cst + (s - g * PW) : 0;
    lds_fft<T, false, kMaxE, PR>(lds, tw0, pp.ax0, base, PW, m.lane, act, m.wave);
  }
  if (m.wave) __syncthreads();

  {
    const T mu = T(pp.mu), sc = T(pp.s), kinv = T(pp.kinv);
    const int items = N0 * PW;
    for (int e = tid; e < items; e += nthr) {
      const int k0 = (int)fdiv((uint32_t)e, pp.pwdiv);
      const int r = e - k0 * PW;
      const int kb = (int)fdiv((uint32_t)r, pp.wdiv);
      const int w = r - kb * W;
      if (w >= wv) continue;
      const int f = f0 + w;
      const int ka = (int)fdiv((uint32_t)f, pp.rsdiv);
      const int kl = f - ka * pp.rs;
      if (kl >= pp.nlh) continue;
      const int k1 = ka + pp.q * kb;
      cpx<T> x[ModeComps<DIM, MODE>::cmax];
#pragma unroll
      for (int c = 0; c < CIN; c++) x[c] = lds[c * cst + e];
      if constexpr (MODE <= 1) {
        modal_block<T, DIM, MODE>(coef, k0, k1, kl, mu, sc, x, kinv);
      } else {
        const gptr<const StrainCoef<T>> scf[3] = {(gptr<const StrainCoef<T>>)pp.scoef[0],
                                                  (gptr<const StrainCoef<T>>)pp.scoef[1],
                                                  (gptr<const StrainCoef<T>>)pp.scoef[2]};
        if constexpr (MODE == 2)
          strain_block<T, DIM>(scf, k0, k1, kl, T(pp.bscale), x);
        else
          eigenstress_block<T, DIM>(coef, scf, k0, k1, kl, mu, sc, T(pp.bscale), kinv, x);
      }
#pragma unroll
      for (int c = 0; c < COUT; c++) lds[c * cst + e] = x[c];
    }
  }
  __syncthreads();

  for (int s0 = 0; s0 < COUT * PW; s0 += m.S) {
    const int s = s0 + m.sl;
    const bool act = m.sl < m.S && s < COUT * PW;
    const int g = act ? (int)fdiv((uint32_t)s, pp.pwdiv) : 0;
    const int base = act ? g * cst + (s - g * PW) : 0;
    lds_fft<T, true, kMaxE, PR>(lds, tw0, pp.ax0, base, PW, m.lane, act, m.wave);
  }
  if (m.wave) __syncthreads();
  if (P > 1) {
    if (fused_split(P))
      pdft<T, true>(lds, P, W, pp.wdiv, COUT * N0, tid, nthr);
    else
      pdft_stockham<T, true, PR>(lds, tw1, pp, COUT * N0, tid, nthr);
    __syncthreads();  // the store reads entries other threads wrote
  }

  staged_copy<kLoads<T>, cpx<T>>(COUT * N0 * PW, tid, nthr, [&](int e) { return lds[e]; },
                                 [&](int e, cpx<T> v) {
    const int rest = (int)fdiv((uint32_t)e, pp.wdiv);
    const int w = e - rest * W;
    if (w < wv) ws[(int64_t)rest * pp.f + f0 + w] = v;
  });
}

}  // namespace b17
