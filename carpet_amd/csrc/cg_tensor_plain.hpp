// cg_tensor_plain.hpp -- per-call tensor prolongation launches, one
// template per line kind; instantiated in cg_tensor_{vlag,clag,eno}.hip so
// the three kinds compile in parallel.
#pragma once

#include <hip/hip_runtime.h>

#include "cg_tiles.hpp"

namespace cg {

// NTL: levels of the passes; NSRC: source levels (NSRC > NTL: combined at
// the window store)
// 512-thread workgroups of orders up to 5 held to 128 VGPRs, two per CU
// (the O=5, 3-level tile otherwise takes 130 and runs one per CU)
template <class L, int NTL, int CX, int CY, int IPT, int XQ, int PR, int NSRC>
__global__ __launch_bounds__(2 * CX * CY / IPT) __attribute__((amdgpu_waves_per_eu(
    (2 * CX * CY / IPT >= 512 && L::R <= 6) ? 4 : 1))) void k_tensor(TensorArgs a) {
  __shared__ __align__(16) double smem[tile_lds<L, NTL, CX, CY, XQ>()];
  tensor_tile<L, NTL, CX, CY, IPT, PR, 1, XQ, pipe_fits<L, NTL, CX, CY, XQ>(), NSRC>(
      a, xcd_remap(blockIdx.x, gridDim.x), smem);
}

template <class L, int NSRC, int PR>
int launch_tensor_t(const TensorArgs &a0, const TileCfg &cfg0, hipStream_t s) {
  constexpr int NTL = pass_levels(NSRC, PR);
  using T = TileShape<NTL, 0, PR>;
  TensorArgs a = a0;
  TileCfg cfg = cfg0;
  cfg.cx = T::CX;
  cfg.cy = T::CY;
  const long long nblk = tensor_tiles(a, cfg);
  if (nblk <= 0 || nblk > 0x7fffffffLL)
    return -1;
  hipLaunchKernelGGL((k_tensor<L, NTL, T::CX, T::CY, T::IPT, T::XQ, PR, NSRC>), dim3((unsigned)nblk),
                     dim3(tile_threads<NTL, 0, PR>()), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// prec: precision mode (Prec in cg_tiles.hpp); the FP32-storage modes
// (1, 2) exist for the Lagrange kinds only
template <int KIND>
int launch_tensor_kind(int order, int ntl, int prec, const TensorArgs &a, const TileCfg &cfg,
                       hipStream_t s) {
  return with_order<KIND>(order, [&](auto tag) -> int {
    using L = decltype(tag);
    auto by_ntl = [&](auto prc) -> int {
      constexpr int PR = decltype(prc)::value;
      switch (ntl) {
      case 1: return launch_tensor_t<L, 1, PR>(a, cfg, s);
      case 2: return launch_tensor_t<L, 2, PR>(a, cfg, s);
      case 3: return launch_tensor_t<L, 3, PR>(a, cfg, s);
      }
      return -1;
    };
    if (prec == 0)
      return by_ntl(std::integral_constant<int, 0>{});
    if constexpr (!fp64_only(KIND)) {
      if (prec == 1)
        return by_ntl(std::integral_constant<int, 1>{});
      if (prec == 2)
        return by_ntl(std::integral_constant<int, 2>{});
    }
    return -1;
  });
}

} // namespace cg
