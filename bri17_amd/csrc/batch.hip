// One persistent launch for a batch of operators of different grids,
// dimensions and materials (no reference counterpart: the reference applies
// one operator per call on the CPU, tests/test_bri17.cpp:56-107).
//
// Every instance contributes three kinds of work items -- plane blocks
// (R2C + middle-axis stage), pencil tiles (axis-0 FFT + modal block +
// inverse) and inverse plane blocks -- to one queue, ordered so that an item
// only ever depends on items that precede it.  Workgroups take items from an
// atomic head counter; a pencil item waits until all plane blocks of its
// instance are done, an inverse plane item until every pencil tile of the
// launch is.  (Waiting per instance let inverse items start a little
// earlier, but each of those waits ended in an acquire that invalidates an
// XCD's L2 while the last pencils still run: 5% slower on a 64-instance
// launch.)
//
// Publication.  Making a workgroup's workspace writes visible to the other
// XCDs takes a device-scope release, which writes back this XCD's whole L2;
// one per item kept the L2s flushing all the time (forward plane items
// spent half their time in their loads, stores and the release).  So a
// workgroup publishes its finished items in groups: it records them in a
// short list and releases once for the whole list -- when the list is full,
// before any dependency wait, and at the end.  Besides the per-instance
// counters, two global counters count the finished forward items and
// pencil tiles of all instances; once a workgroup has seen one of them
// complete and acquired after that, its later items of the next kind skip
// the wait and the acquire (which invalidates this XCD's L2).  Inverse
// items wait on the global pencil counter only, so a workgroup acquires for
// them once.
//
// Progress.  A workgroup never waits while holding unpublished items, and
// forward items never wait, so the lowest unpublished item is always
// published eventually: the queue drains.  Counters are polled with relaxed
// loads (they read past this XCD's L2 without invalidating it) and followed
// by one acquire.  The wait is bounded: past the spin limit the kernel
// records an error and moves on.
#include "batch.hpp"
#include "kernels_generic.hpp"

namespace b17 {

// finished items a workgroup publishes with one release
constexpr int kPend = 8;

template <typename T, int PR>
__global__ void __launch_bounds__(1024)
    k_batch(const BatchInst* __restrict__ insts, const int2* __restrict__ items, int nitems,
            int n_fwd, int n_pen, int* __restrict__ ctl, int spin_limit) {
  extern __shared__ __align__(16) unsigned char smem[];
  cpx<T>* lds = reinterpret_cast<cpx<T>*>(smem);
  __shared__ int s_item[2];
  __shared__ int s_pend[kPend];  // counters of the unpublished items (thread 0's list)
  const int tid = threadIdx.x;
  int* head = ctl;
  int* err = ctl + 1;
  int* all = ctl + 2;   // finished forward items, finished pencil tiles (every instance)
  int* done = ctl + 4;  // per instance: finished forward items, finished pencil tiles
  int npend = 0;        // workgroup-uniform
  bool seen_fwd = false, seen_pen = false;  // thread 0: acquired after `all` completed
  // every wave waits for its own stores, the barrier makes that true of the
  // whole workgroup, then thread 0 releases once and bumps the counters
  auto publish = [&]() {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (tid == 0) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      int nf = 0, np = 0;
      for (int i = 0; i < npend; i++) {
        const int c = s_pend[i];
        __hip_atomic_fetch_add(done + c, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (c & 1)
          np++;
        else
          nf++;
      }
      if (nf) __hip_atomic_fetch_add(all, nf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (np) __hip_atomic_fetch_add(all + 1, np, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    npend = 0;
  };
  if (tid == 0) s_item[0] = atomicAdd(head, 1);
  __syncthreads();
  // the item index is workgroup-uniform: keep it (and everything addressed
  // through it) in scalar registers
  int cur = __builtin_amdgcn_readfirstlane(s_item[0]);
  int slot = 0;
  for (;;) {
    if (cur >= nitems) break;
    // claim the next item while this one runs (items are claimed in queue
    // order)
    if (tid == 0) s_item[slot ^ 1] = atomicAdd(head, 1);
    const int2 w = items[cur];
    const int inst = w.x & 0xffff;
    const int type = w.x >> 16;
    const int idx = w.y;
    const BatchInst* I = insts + inst;
    if (type > 0) {
      if (npend) publish();
      if (tid == 0) {
        bool& seen = type == 1 ? seen_fwd : seen_pen;
        if (!seen) {
          const int total = type == 1 ? n_fwd : n_pen;
          if (__hip_atomic_load(all + (type - 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >=
              total) {
            seen = true;
          } else {
            // pencils: this instance's forward items; inverse items: every
            // pencil tile of the launch
            const int need = type == 1 ? I->na : total;
            int* cnt = type == 1 ? done + 2 * inst : all + 1;
            int n = 0;
            while (__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
              if (++n > spin_limit) {
                __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
              }
              __builtin_amdgcn_s_sleep(4);
            }
            if (type == 2) seen = true;
          }
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
      }
      __syncthreads();
    }
    if (type == 0) {
      plane_fwd_body<T, PR>((gptr<const T>)I->in, (gptr<cpx<T>>)I->ws, I->pp, idx, lds);
    } else if (type == 2) {
      plane_inv_body<T, PR>((gptr<const cpx<T>>)I->ws, (gptr<T>)I->out, I->pp, idx, lds);
    } else {
      gptr<cpx<T>> ws = (gptr<cpx<T>>)I->ws;
      switch (I->pc.dim * 2 + I->pc.mode) {
        case 4: pencil_body<T, 2, 0, PR>(ws, I->pc, idx, lds); break;
        case 5: pencil_body<T, 2, 1, PR>(ws, I->pc, idx, lds); break;
        case 6: pencil_body<T, 3, 0, PR>(ws, I->pc, idx, lds); break;
        default: pencil_body<T, 3, 1, PR>(ws, I->pc, idx, lds); break;
      }
    }
    __syncthreads();  // the LDS is the next item's
    if (type < 2) {
      if (tid == 0) s_pend[npend] = 2 * inst + type;
      npend++;
      if (npend == kPend) publish();
    }
    slot ^= 1;
    cur = __builtin_amdgcn_readfirstlane(s_item[slot]);
  }
  if (npend) publish();
}

__global__ void k_batch_reset(int* __restrict__ ctl, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ctl[i] = 0;
}

hipError_t launch_batch(int dtype, int primes, const BatchInst* insts, const int2* items,
                        int nitems, int n_fwd, int n_pen, int* ctl, int nctl, int threads,
                        size_t lds, int spin_limit, hipStream_t st, int* grid_out) {
  hipLaunchKernelGGL(k_batch_reset, dim3((nctl + 255) / 256), dim3(256), 0, st, ctl, nctl);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  auto kf = primes == 2 ? k_batch<float, 2> : primes == 1 ? k_batch<float, 1> : k_batch<float, 0>;
  auto kd = primes == 2 ? k_batch<double, 2> : primes == 1 ? k_batch<double, 1> : k_batch<double, 0>;
  const void* fn = dtype == 0 ? reinterpret_cast<const void*>(kf) : reinterpret_cast<const void*>(kd);
  if (lds > 65536) {
    e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  int dev = 0, ncu = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    ncu = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, threads, lds) != hipSuccess || per < 1)
    per = 1;
  int grid = ncu * per;
  if (grid > nitems) grid = nitems;
  if (grid < 1) grid = 1;
  if (grid_out) *grid_out = grid;
  if (dtype == 0)
    hipLaunchKernelGGL(kf, dim3(grid), dim3(threads), lds, st, insts, items, nitems, n_fwd, n_pen,
                       ctl, spin_limit);
  else
    hipLaunchKernelGGL(kd, dim3(grid), dim3(threads), lds, st, insts, items, nitems, n_fwd, n_pen,
                       ctl, spin_limit);
  return hipGetLastError();
}

}  // namespace b17
