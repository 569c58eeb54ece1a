// cg_batch_kernels.hip -- persistent kernels of the batched level step.
//
// The transfers of one phase that share a kernel become one launch of a
// persistent grid (occupancy x CUs workgroups).  Each workgroup takes work
// items -- (descriptor, tile) pairs, sorted by decreasing cost -- from a
// device counter until the list is exhausted, so thousands of uneven
// regions and variables run as one launch without per-region launch
// overhead and the biggest tiles start first.  The loop ends for every
// workgroup once the counter passes the item count, so the grid always
// drains.  Phases are separate launches in stream order (a kernel boundary
// makes one phase's results visible to the next across all XCD L2s), and
// the counters are cleared by a small kernel at the start of the step, so
// the whole step can be captured as one hipGraph.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "cg_batch.hpp"
#include "cg_tiles.hpp"

namespace cg {

// pointwise work item: PT consecutive destination points (x fastest)
constexpr int PT = 4096;

struct Item {
  unsigned desc;
  unsigned tile;
};

__device__ __forceinline__ int dequeue(int *counter, int *s_item) {
  if (threadIdx.x == 0)
    *s_item = atomicAdd(counter, 1);
  __syncthreads();
  const int it = *s_item;
  __syncthreads();
  return it;
}

template <class L, int NTL, int CX, int CY, int IPT>
__global__ __launch_bounds__(2 * CX * CY / IPT) void k_persist_tensor(
    const TensorArgs *__restrict__ descs, const Item *__restrict__ items,
    int nitems, int *counter) {
  __shared__ double smem[TensorLds<L, NTL, CX, CY>::TOTAL];
  __shared__ int s_item;
  for (;;) {
    const int it = dequeue(counter, &s_item);
    if (it >= nitems)
      return;
    const Item w = items[it];
    const TensorArgs a = descs[w.desc];
    tensor_tile<L, NTL, CX, CY, IPT>(a, (int)w.tile, smem);
  }
}

template <int OP>
__global__ __launch_bounds__(256) void k_persist_point(
    const PointArgs *__restrict__ descs, const Item *__restrict__ items,
    int nitems, int *counter) {
  __shared__ int s_item;
  for (;;) {
    const int it = dequeue(counter, &s_item);
    if (it >= nitems)
      return;
    const Item w = items[it];
    const PointArgs a = descs[w.desc];
    const unsigned nx = (unsigned)a.n[0];
    const unsigned nxy = nx * (unsigned)a.n[1];
    const unsigned tot = nxy * (unsigned)a.n[2];
    const unsigned e0 = w.tile * (unsigned)PT;
#pragma unroll 4
    for (int m = 0; m < PT / 256; ++m) {
      const unsigned e = e0 + m * 256u + threadIdx.x;
      if (e >= tot)
        break;
      const unsigned k = e / nxy;
      const unsigned r = e - k * nxy;
      const unsigned j = r / nx;
      const unsigned i = r - j * nx;
      point_elem<OP>(a, (int)i, (int)j, (int)k);
    }
  }
}

__global__ void k_zero_counters(int *c, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    c[i] = 0;
}

// ---------------------------------------------------------------------------
namespace {

struct Group {
  int phase, is_tensor, kind, order, ntl, op, prec;
  std::vector<int> members; // transfer indices
  bool same(const Group &g) const {
    return phase == g.phase && is_tensor == g.is_tensor && kind == g.kind &&
           order == g.order && ntl == g.ntl && op == g.op && prec == g.prec;
  }
};

template <class L, int NTL> const void *persist_tensor_fn() {
  using T = TileShape<NTL>;
  return reinterpret_cast<const void *>(&k_persist_tensor<L, NTL, T::CX, T::CY, T::IPT>);
}

const void *persist_point_fn(int op) {
  switch (op) {
  case PO_COPY: return reinterpret_cast<const void *>(&k_persist_point<PO_COPY>);
  case PO_RESTRICT_VC: return reinterpret_cast<const void *>(&k_persist_point<PO_RESTRICT_VC>);
  case PO_RESTRICT_CC: return reinterpret_cast<const void *>(&k_persist_point<PO_RESTRICT_CC>);
  case PO_INTERP: return reinterpret_cast<const void *>(&k_persist_point<PO_INTERP>);
  case PO_TVD: return reinterpret_cast<const void *>(&k_persist_point<PO_TVD>);
  }
  return nullptr;
}

const void *launch_fn(const BatchLaunch &l, int *threads) {
  if (!l.is_tensor) {
    *threads = 256;
    return persist_point_fn(l.op);
  }
  const void *fn = nullptr;
  with_line(l.kind, l.order, [&](auto tag) -> int {
    using L = decltype(tag);
    switch (l.ntl) {
    case 1: fn = persist_tensor_fn<L, 1>(); *threads = 2 * TileShape<1>::CX * TileShape<1>::CY / TileShape<1>::IPT; break;
    case 2: fn = persist_tensor_fn<L, 2>(); *threads = 2 * TileShape<2>::CX * TileShape<2>::CY / TileShape<2>::IPT; break;
    case 3: fn = persist_tensor_fn<L, 3>(); *threads = 2 * TileShape<3>::CX * TileShape<3>::CY / TileShape<3>::IPT; break;
    }
    return 0;
  });
  return fn;
}

} // namespace

void batch_device_free(BatchDevice &dev) {
  if (dev.blob)
    (void)hipFree(dev.blob);
  if (dev.counters)
    (void)hipFree(dev.counters);
  dev.blob = nullptr;
  dev.counters = nullptr;
  dev.launches.clear();
  dev.blob_bytes = 0;
  dev.ncounters = 0;
}

int batch_device_build(BatchDevice &dev, const std::vector<Transfer> &tr,
                       const std::vector<int> &phase, const std::vector<int> &prec,
                       int tile_cfg) {
  batch_device_free(dev);
  // group transfers by (phase, kernel)
  std::vector<Group> groups;
  for (size_t i = 0; i < tr.size(); ++i) {
    Group g{};
    g.phase = phase[i];
    g.prec = prec[i];
    if (tr[i].is_prolong && tr[i].pr.kind != PK_TVD) {
      g.is_tensor = 1;
      g.kind = tr[i].pr.kind;
      g.order = tr[i].pr.order;
      g.ntl = tr[i].pr.ntl;
      g.op = -1;
    } else {
      g.is_tensor = 0;
      g.kind = g.order = -1;
      g.ntl = 1;
      g.op = tr[i].is_prolong ? PO_TVD : tr[i].pt.op;
    }
    auto it = std::find_if(groups.begin(), groups.end(), [&](const Group &x) { return x.same(g); });
    if (it == groups.end()) {
      groups.push_back(g);
      it = groups.end() - 1;
    }
    it->members.push_back((int)i);
  }
  std::stable_sort(groups.begin(), groups.end(),
                   [](const Group &a, const Group &b) { return a.phase < b.phase; });

  int dev_id = 0, ncu = 0;
  hipError_t he = hipGetDevice(&dev_id);
  if (he != hipSuccess)
    return hip_error(he, "hipGetDevice");
  he = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev_id);
  if (he != hipSuccess)
    return hip_error(he, "hipDeviceGetAttribute");

  // host image of the blob: per launch, descriptors then items (8-byte aligned)
  std::vector<char> host;
  auto append = [&](const void *p, size_t n) {
    size_t off = (host.size() + 15) / 16 * 16;
    host.resize(off + n);
    if (n)
      std::memcpy(host.data() + off, p, n);
    return off;
  };
  for (const Group &g : groups) {
    if (g.prec != CG_PREC_F64)
      return set_error(CG_ERR_OPERATOR, "mixed precision is not available in the batch yet");
    BatchLaunch l{};
    l.phase = g.phase;
    l.is_tensor = g.is_tensor;
    l.kind = g.kind;
    l.order = g.order;
    l.ntl = g.ntl;
    l.op = g.op;
    l.prec = g.prec;
    std::vector<Item> items;
    std::vector<double> cost;
    if (g.is_tensor) {
      std::vector<TensorArgs> descs;
      for (int m : g.members) {
        TensorArgs a = tr[m].pr.t;
        TileCfg cfg;
        cfg.cx = g.ntl == 1 ? TileShape<1>::CX : TileShape<3>::CX;
        cfg.cy = g.ntl == 1 ? TileShape<1>::CY : TileShape<3>::CY;
        cfg.zchunk = tile_cfg < 0 ? 16 : tile_cfg;
        const long long nt = tensor_tiles(a, cfg);
        const unsigned d = (unsigned)descs.size();
        descs.push_back(a);
        for (long long t = 0; t < nt; ++t) {
          items.push_back(Item{d, (unsigned)t});
          // cost ~ fine points of the tile
          const int tz = (int)(t / ((long long)a.ntiles[0] * a.ntiles[1]));
          const int nz = std::min(a.zchunk, a.nq[2] - tz * a.zchunk);
          cost.push_back((double)nz * std::min(cfg.cx, a.nq[0]) * std::min(cfg.cy, a.nq[1]));
        }
      }
      l.desc_off = append(descs.data(), descs.size() * sizeof(TensorArgs));
    } else {
      std::vector<PointArgs> descs;
      for (int m : g.members) {
        const PointArgs &a = tr[m].is_prolong ? tr[m].pr.p : tr[m].pt.a;
        const long long tot = (long long)a.n[0] * a.n[1] * a.n[2];
        if (tot >= (1LL << 32))
          return set_error(CG_ERR_ARG, "region too large for the batch");
        const unsigned d = (unsigned)descs.size();
        descs.push_back(a);
        const long long nt = (tot + PT - 1) / PT;
        for (long long t = 0; t < nt; ++t) {
          items.push_back(Item{d, (unsigned)t});
          cost.push_back((double)std::min<long long>(PT, tot - t * PT));
        }
      }
      l.desc_off = append(descs.data(), descs.size() * sizeof(PointArgs));
    }
    // largest tiles first
    std::vector<int> ord(items.size());
    for (size_t i = 0; i < ord.size(); ++i)
      ord[i] = (int)i;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    std::vector<Item> sorted(items.size());
    for (size_t i = 0; i < ord.size(); ++i)
      sorted[i] = items[ord[i]];
    l.item_off = append(sorted.data(), sorted.size() * sizeof(Item));
    l.nitems = (int)sorted.size();
    l.counter = (int)dev.launches.size();
    int threads = 0;
    const void *fn = launch_fn(l, &threads);
    if (!fn)
      return set_error(CG_ERR_ORDER, "no persistent kernel for this transfer class");
    int per_cu = 0;
    he = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, 0);
    if (he != hipSuccess)
      return hip_error(he, "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    l.grid = std::max(1, std::min(l.nitems, std::max(1, per_cu) * ncu));
    dev.launches.push_back(l);
  }
  dev.ncounters = (int)dev.launches.size();
  if (host.empty())
    return CG_OK;
  he = hipMalloc(&dev.blob, host.size());
  if (he != hipSuccess)
    return set_error(CG_ERR_NOMEM, "batch descriptors: %s", hipGetErrorString(he));
  dev.blob_bytes = host.size();
  he = hipMemcpy(dev.blob, host.data(), host.size(), hipMemcpyHostToDevice);
  if (he != hipSuccess)
    return hip_error(he, "batch descriptor upload");
  he = hipMalloc(&dev.counters, sizeof(int) * std::max(1, dev.ncounters));
  if (he != hipSuccess)
    return set_error(CG_ERR_NOMEM, "batch counters: %s", hipGetErrorString(he));
  return CG_OK;
}

int batch_device_run(BatchDevice &dev, hipStream_t st) {
  if (dev.launches.empty())
    return CG_OK;
  hipLaunchKernelGGL(k_zero_counters, dim3((dev.ncounters + 255) / 256), dim3(256), 0, st,
                     dev.counters, dev.ncounters);
  for (const BatchLaunch &l : dev.launches) {
    const void *descs = dev.blob + l.desc_off;
    const Item *items = reinterpret_cast<const Item *>(dev.blob + l.item_off);
    int *ctr = dev.counters + l.counter;
    int rc = 0;
    if (l.is_tensor) {
      rc = with_line(l.kind, l.order, [&](auto tag) -> int {
        using L = decltype(tag);
        auto go = [&](auto ntlc) {
          constexpr int NTL = decltype(ntlc)::value;
          using T = TileShape<NTL>;
          hipLaunchKernelGGL((k_persist_tensor<L, NTL, T::CX, T::CY, T::IPT>), dim3(l.grid),
                             dim3(2 * T::CX * T::CY / T::IPT), 0, st,
                             static_cast<const TensorArgs *>(descs), items, l.nitems, ctr);
          return 0;
        };
        switch (l.ntl) {
        case 1: return go(std::integral_constant<int, 1>{});
        case 2: return go(std::integral_constant<int, 2>{});
        case 3: return go(std::integral_constant<int, 3>{});
        }
        return -1;
      });
    } else {
      const PointArgs *pd = static_cast<const PointArgs *>(descs);
      switch (l.op) {
      case PO_COPY: hipLaunchKernelGGL(k_persist_point<PO_COPY>, dim3(l.grid), dim3(256), 0, st, pd, items, l.nitems, ctr); break;
      case PO_RESTRICT_VC: hipLaunchKernelGGL(k_persist_point<PO_RESTRICT_VC>, dim3(l.grid), dim3(256), 0, st, pd, items, l.nitems, ctr); break;
      case PO_RESTRICT_CC: hipLaunchKernelGGL(k_persist_point<PO_RESTRICT_CC>, dim3(l.grid), dim3(256), 0, st, pd, items, l.nitems, ctr); break;
      case PO_INTERP: hipLaunchKernelGGL(k_persist_point<PO_INTERP>, dim3(l.grid), dim3(256), 0, st, pd, items, l.nitems, ctr); break;
      case PO_TVD: hipLaunchKernelGGL(k_persist_point<PO_TVD>, dim3(l.grid), dim3(256), 0, st, pd, items, l.nitems, ctr); break;
      default: rc = -1;
      }
    }
    if (rc)
      return set_error(CG_ERR_ORDER, "no persistent kernel for this transfer class");
    hipError_t he = hipGetLastError();
    if (he != hipSuccess)
      return hip_error(he, "persistent launch");
  }
  return CG_OK;
}

} // namespace cg
