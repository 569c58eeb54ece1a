// cg_batch_dev.hpp -- device side shared by the batch kernels: work-item
// selection for the persistent and the grid launch modes.
#pragma once

#include <hip/hip_runtime.h>

#include "cg_batch.hpp"
#include "cg_tiles.hpp"

namespace cg {

// Work items are split into 8 contiguous, cost-balanced segments, one per
// XCD (workgroups are dispatched round-robin over the XCDs, so workgroup b
// runs on XCD b % 8): neighbouring tiles -- which share stencil halos --
// are then processed at the same time on the same L2.  A workgroup takes
// items from its own segment's counter and, once that is exhausted, from
// the other segments', so the load still balances.  Returns -1 when every
// segment is exhausted; every workgroup gets there, so the grid drains.
__device__ __forceinline__ int dequeue(int *ctr8, const int *seg, int *s_item) {
  if (threadIdx.x == 0) {
    int it = -1;
    const int home = blockIdx.x & 7;
    for (int k = 0; k < 8 && it < 0; ++k) {
      const int sg = (home + k) & 7;
      const int n = seg[sg + 1] - seg[sg];
      if (n <= 0 || __hip_atomic_load(&ctr8[sg], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= n)
        continue;
      const int i = atomicAdd(&ctr8[sg], 1);
      if (i < n)
        it = seg[sg] + i;
    }
    *s_item = it;
  }
  __syncthreads();
  // wave-uniform: the item's descriptor is then read with scalar loads and
  // held in SGPRs, as a kernel argument would be
  const int it = __builtin_amdgcn_readfirstlane(*s_item);
  __syncthreads();
  return it;
}

// Static mode: a resident grid (a multiple of 8 workgroups) in which the
// workgroups of XCD x take the items of segment x in a fixed round-robin:
// round k of workgroup b is slot b/8 + k * (grid/8).  Neighbouring tiles
// then start their z-march together in every round, as the tiles of one
// launch per transfer do, and share their halo rows in L2 (a workgroup
// taking whichever item is next when it finishes starts neighbours at
// staggered times, and their halo rows leave L2 in between).  Returns -1
// after the workgroup's last round.
__device__ __forceinline__ int static_item(const int *__restrict__ seg, int k) {
  const int x = blockIdx.x & 7;
  const int slot = (int)(blockIdx.x >> 3) + k * (int)(gridDim.x >> 3);
  const int i = seg[x] + slot;
  return i < seg[x + 1] ? i : -1;
}

// Grid mode: one workgroup per item, no loop and no counters.  Workgroup b
// runs on XCD b % 8 and takes slot b / 8 of that XCD's segment; segments
// shorter than the longest leave their last workgroups without an item.
__device__ __forceinline__ int grid_item(const int *__restrict__ seg) {
  const int x = blockIdx.x & 7, slot = blockIdx.x >> 3;
  const int i = seg[x] + slot;
  return i < seg[x + 1] ? i : -1;
}

} // namespace cg
