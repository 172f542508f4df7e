// One persistent launch for a batch of operators of different grids,
// dimensions and materials (no reference counterpart: the reference applies
// one operator per call on the CPU, tests/test_bri17.cpp:56-107).
//
// Every instance contributes three kinds of work items -- plane blocks
// (R2C + middle-axis stage), pencil tiles (axis-0 FFT + modal block +
// inverse) and inverse plane blocks -- to one queue, ordered so that an item
// only ever depends on items that precede it.  Workgroups take items from an
// atomic head counter; a pencil item waits until all plane blocks of its
// instance are done, an inverse plane item until all its pencil tiles are.
// A waiting workgroup only waits on items already taken by running
// workgroups, so the queue always drains.  Completion counters are published
// with one device-scope release per item and polled with relaxed loads
// followed by one acquire, which makes the producers' workspace writes
// visible across XCDs without flushing caches on every poll.  The wait is
// bounded: past the spin limit the kernel records an error and moves on.
#include "batch.hpp"
#include "kernels_generic.hpp"

namespace b17 {

template <typename T>
__global__ void __launch_bounds__(1024)
    k_batch(const BatchInst* __restrict__ insts, const int2* __restrict__ items, int nitems,
            int* __restrict__ ctl, int spin_limit) {
  extern __shared__ __align__(16) unsigned char smem[];
  cpx<T>* lds = reinterpret_cast<cpx<T>*>(smem);
  __shared__ int s_item[2];
  const int tid = threadIdx.x;
  int* head = ctl;
  int* err = ctl + 1;
  int* done = ctl + 2;
  if (tid == 0) s_item[0] = atomicAdd(head, 1);
  __syncthreads();
  // the item index is workgroup-uniform: keep it (and everything addressed
  // through it) in scalar registers
  int cur = __builtin_amdgcn_readfirstlane(s_item[0]);
  int slot = 0;
  for (;;) {
    if (cur >= nitems) break;
    // claim the next item while this one runs (items are claimed in queue
    // order, so the lowest unfinished item is always some workgroup's
    // current one and its dependencies are finished)
    if (tid == 0) s_item[slot ^ 1] = atomicAdd(head, 1);
    const int2 w = items[cur];
    const int inst = w.x & 0xffff;
    const int type = w.x >> 16;
    const int idx = w.y;
    const BatchInst* I = insts + inst;
    if (type > 0) {
      if (tid == 0) {
        const int need = type == 1 ? I->na : I->nb;
        int* cnt = done + 2 * inst + (type - 1);
        int n = 0;
        // poll with relaxed device-scope loads (they read past this XCD's
        // L2 without invalidating it), then acquire once: the waves of a
        // workgroup share this CU's caches
        while (__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
          if (++n > spin_limit) {
            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
          __builtin_amdgcn_s_sleep(4);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      }
      __syncthreads();
    }
    if (type == 0) {
      plane_fwd_body<T>((gptr<const T>)I->in, (gptr<cpx<T>>)I->ws, I->pp, idx, lds);
    } else if (type == 2) {
      plane_inv_body<T>((gptr<const cpx<T>>)I->ws, (gptr<T>)I->out, I->pp, idx, lds);
    } else {
      gptr<cpx<T>> ws = (gptr<cpx<T>>)I->ws;
      switch (I->pc.dim * 2 + I->pc.mode) {
        case 4: pencil_body<T, 2, 0>(ws, I->pc, idx, lds); break;
        case 5: pencil_body<T, 2, 1>(ws, I->pc, idx, lds); break;
        case 6: pencil_body<T, 3, 0>(ws, I->pc, idx, lds); break;
        default: pencil_body<T, 3, 1>(ws, I->pc, idx, lds); break;
      }
    }
    // every wave waits until its own stores are complete (in its XCD's L2),
    // then one device-scope release (L2 write-back) publishes the workgroup's
    // results to the other XCDs
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (tid == 0 && type < 2)
      __hip_atomic_fetch_add(done + 2 * inst + type, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    slot ^= 1;
    cur = __builtin_amdgcn_readfirstlane(s_item[slot]);
  }
}

__global__ void k_batch_reset(int* __restrict__ ctl, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ctl[i] = 0;
}

hipError_t launch_batch(int dtype, const BatchInst* insts, const int2* items, int nitems, int* ctl,
                        int nctl, int threads, size_t lds, int spin_limit, hipStream_t st,
                        int* grid_out) {
  hipLaunchKernelGGL(k_batch_reset, dim3((nctl + 255) / 256), dim3(256), 0, st, ctl, nctl);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const void* fn = dtype == 0 ? reinterpret_cast<const void*>(k_batch<float>)
                              : reinterpret_cast<const void*>(k_batch<double>);
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
    hipLaunchKernelGGL(k_batch<float>, dim3(grid), dim3(threads), lds, st, insts, items, nitems,
                       ctl, spin_limit);
  else
    hipLaunchKernelGGL(k_batch<double>, dim3(grid), dim3(threads), lds, st, insts, items, nitems,
                       ctl, spin_limit);
  return hipGetLastError();
}

}  // namespace b17
