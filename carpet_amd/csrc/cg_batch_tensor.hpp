// cg_batch_tensor.hpp -- batched tensor prolongation kernels (grid and
// persistent modes), one launch template per line kind and tile shape;
// instantiated in cg_batch_{vlag,clag,eno,veno}_{w,n,s,t}.hip so they compile in
// parallel.
#pragma once

#include "cg_batch_dev.hpp"

namespace cg {

// the small and slim tile shapes (2, 3) are instantiated for FP64 storage
// only
template <int KIND, int SH> constexpr bool has_mixed() { return !fp64_only(KIND) && SH < 2; }

// 512-thread workgroups of orders up to 5 are held to 128 VGPRs: two
// workgroups (4 waves per SIMD) per CU instead of one
template <class L, int NTL, int CX, int CY, int IPT> constexpr int persist_waves() {
  return (2 * CX * CY / IPT >= 512 && L::R <= 6) ? 4 : 1;
}

// every tile shape loads each plane's window one plane ahead of its use:
// two planes ahead was no faster on the narrow, small and slim tiles (the
// 2000-region step equal, the 2000-box hierarchy step 0.6% slower) and
// slower on the wide ones (the 512^3 patch 17.0 ms against 15.8 ms)
constexpr int BATCH_PF = 1;

// NTL: levels of the passes; NSRC: source levels (NSRC > NTL: combined at
// the window store)
template <class L, int NTL, int CX, int CY, int IPT, int XQ, int PR, int NSRC>
__global__ __launch_bounds__(2 * CX * CY / IPT)
__attribute__((amdgpu_waves_per_eu(persist_waves<L, NTL, CX, CY, IPT>()))) void k_persist_tensor(
    const TensorArgs *__restrict__ descs, const Item *__restrict__ items,
    const int *__restrict__ seg, int *counter, int stat) {
  __shared__ __align__(16) double smem[tile_lds<L, NTL, CX, CY, XQ>()];
  __shared__ int s_item;
  for (int k = 0;; ++k) {
    // stat: fixed rounds over the XCD's segment, else the work counters
    const int it = stat ? static_item(seg, k) : dequeue(counter, seg, &s_item);
    if (it < 0)
      return;
    const Item w = items[it];
    const TensorArgs a = descs[w.desc];
    tensor_tile<L, NTL, CX, CY, IPT, PR, BATCH_PF, XQ, pipe_fits<L, NTL, CX, CY, XQ>(), NSRC>(a, (int)w.tile,
                                                                                          smem);
  }
}

template <class L, int NTL, int CX, int CY, int IPT, int XQ, int PR, int NSRC>
__global__ __launch_bounds__(2 * CX * CY / IPT) void k_grid_tensor(
    const TensorArgs *__restrict__ descs, const Item *__restrict__ items,
    const int *__restrict__ seg) {
  __shared__ __align__(16) double smem[tile_lds<L, NTL, CX, CY, XQ>()];
  const int it = grid_item(seg);
  if (it < 0)
    return;
  const Item w = items[it];
  tensor_tile<L, NTL, CX, CY, IPT, PR, BATCH_PF, XQ, pipe_fits<L, NTL, CX, CY, XQ>(), NSRC>(descs[w.desc],
                                                                                            (int)w.tile, smem);
}

template <class L, int NSRC, int PR, int SH>
int batch_tensor_go(const BatchLaunch &l, const TensorArgs *td, const Item *items,
                    const int *seg, int *ctr, hipStream_t st) {
  constexpr int NTL = pass_levels(NSRC, PR);
  using T = TileShape<NTL, SH, PR>;
  const dim3 block(tile_threads<NTL, SH, PR>());
  if (l.wpc == -1)
    hipLaunchKernelGGL((k_grid_tensor<L, NTL, T::CX, T::CY, T::IPT, T::XQ, PR, NSRC>), dim3(l.grid), block,
                       0, st, td, items, seg);
  else
    hipLaunchKernelGGL((k_persist_tensor<L, NTL, T::CX, T::CY, T::IPT, T::XQ, PR, NSRC>), dim3(l.grid),
                       block, 0, st, td, items, seg, ctr, l.wpc == -2 ? 1 : 0);
  return 0;
}

template <int KIND, int SH>
int batch_tensor_kind(const BatchLaunch &l, const TensorArgs *td, const Item *items,
                      const int *seg, int *ctr, hipStream_t st) {
  return with_order<KIND>(l.order, [&](auto tag) -> int {
    using L = decltype(tag);
    auto by_ntl = [&](auto prc) -> int {
      constexpr int PR = decltype(prc)::value;
      switch (l.ntl) {
      case 1: return batch_tensor_go<L, 1, PR, SH>(l, td, items, seg, ctr, st);
      case 2: return batch_tensor_go<L, 2, PR, SH>(l, td, items, seg, ctr, st);
      case 3: return batch_tensor_go<L, 3, PR, SH>(l, td, items, seg, ctr, st);
      }
      return -1;
    };
    if (l.prec == 0)
      return by_ntl(std::integral_constant<int, 0>{});
    if constexpr (has_mixed<KIND, SH>()) {
      if (l.prec == 1)
        return by_ntl(std::integral_constant<int, 1>{});
      if (l.prec == 2)
        return by_ntl(std::integral_constant<int, 2>{});
    }
    return -1;
  });
}

// persistent kernel of a launch class (for the resident-capacity query)
template <int KIND, int SH> const void *batch_tensor_fn_kind(const BatchLaunch &l, int *threads) {
  const void *fn = nullptr;
  with_order<KIND>(l.order, [&](auto tag) -> int {
    using L = decltype(tag);
    auto pick = [&](auto ntlc, auto prc) {
      constexpr int NSRC = decltype(ntlc)::value;
      constexpr int PR = decltype(prc)::value;
      constexpr int NTL = pass_levels(NSRC, PR);
      using T = TileShape<NTL, SH, PR>;
      *threads = tile_threads<NTL, SH, PR>();
      fn = reinterpret_cast<const void *>(&k_persist_tensor<L, NTL, T::CX, T::CY, T::IPT, T::XQ, PR, NSRC>);
    };
    auto by_prec = [&](auto ntlc) {
      if (l.prec == 0)
        pick(ntlc, std::integral_constant<int, 0>{});
      if constexpr (has_mixed<KIND, SH>()) {
        if (l.prec == 1)
          pick(ntlc, std::integral_constant<int, 1>{});
        if (l.prec == 2)
          pick(ntlc, std::integral_constant<int, 2>{});
      }
    };
    switch (l.ntl) {
    case 1: by_prec(std::integral_constant<int, 1>{}); break;
    case 2: by_prec(std::integral_constant<int, 2>{}); break;
    case 3: by_prec(std::integral_constant<int, 3>{}); break;
    }
    return 0;
  });
  return fn;
}

} // namespace cg
