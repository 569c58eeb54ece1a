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

// default work items: pointwise = PT_DEFAULT consecutive destination points
// (x fastest); tensor = one x/y tile over ZCHUNK_DEFAULT coarse z bases
constexpr int PT_DEFAULT = 4096;
constexpr int ZCHUNK_DEFAULT = 16;

struct Item {
  unsigned desc;
  unsigned tile;
};

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
  const int it = *s_item;
  __syncthreads();
  return it;
}

template <class L, int NTL, int CX, int CY, int IPT, typename TS>
__global__ __launch_bounds__(2 * CX * CY / IPT) void k_persist_tensor(
    const TensorArgs *__restrict__ descs, const Item *__restrict__ items,
    const int *__restrict__ seg, int *counter) {
  __shared__ double smem[TensorLds<L, NTL, CX, CY>::TOTAL];
  __shared__ int s_item;
  for (;;) {
    const int it = dequeue(counter, seg, &s_item);
    if (it < 0)
      return;
    const Item w = items[it];
    const TensorArgs a = descs[w.desc];
    tensor_tile<L, NTL, CX, CY, IPT, TS>(a, (int)w.tile, smem);
  }
}

template <int OP>
__global__ __launch_bounds__(256) void k_persist_point(
    const PointArgs *__restrict__ descs, const Item *__restrict__ items,
    const int *__restrict__ seg, int *counter, int pt) {
  __shared__ int s_item;
  for (;;) {
    const int it = dequeue(counter, seg, &s_item);
    if (it < 0)
      return;
    const Item w = items[it];
    const PointArgs a = descs[w.desc];
    const unsigned nx = (unsigned)a.n[0];
    const unsigned nxy = nx * (unsigned)a.n[1];
    const unsigned tot = nxy * (unsigned)a.n[2];
    const unsigned e0 = w.tile * (unsigned)pt;
#pragma unroll 4
    for (int m = 0; m < pt / 256; ++m) {
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

template <class L, int NTL> const void *persist_tensor_fn(int prec) {
  using T = TileShape<NTL>;
  if (prec)
    return reinterpret_cast<const void *>(&k_persist_tensor<L, NTL, T::CX, T::CY, T::IPT, float>);
  return reinterpret_cast<const void *>(&k_persist_tensor<L, NTL, T::CX, T::CY, T::IPT, double>);
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
    case 1: fn = persist_tensor_fn<L, 1>(l.prec); *threads = 2 * TileShape<1>::CX * TileShape<1>::CY / TileShape<1>::IPT; break;
    case 2: fn = persist_tensor_fn<L, 2>(l.prec); *threads = 2 * TileShape<2>::CX * TileShape<2>::CY / TileShape<2>::IPT; break;
    case 3: fn = persist_tensor_fn<L, 3>(l.prec); *threads = 2 * TileShape<3>::CX * TileShape<3>::CY / TileShape<3>::IPT; break;
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

static int ilog2(long long v) {
  int r = 0;
  while (v > 1) {
    v >>= 1;
    ++r;
  }
  return r;
}

int batch_device_build(BatchDevice &dev, const std::vector<Transfer> &tr,
                       const std::vector<int> &phase, const std::vector<int> &prec,
                       int tile_cfg, const TuneMap *tuned) {
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
    if (g.prec != CG_PREC_F64 && (!g.is_tensor || g.kind == LK_ENO))
      return set_error(CG_ERR_OPERATOR, "mixed precision applies to Lagrange prolongation only");
    BatchLaunch l{};
    l.phase = g.phase;
    l.is_tensor = g.is_tensor;
    l.kind = g.kind;
    l.order = g.order;
    l.ntl = g.ntl;
    l.op = g.op;
    l.prec = g.prec;
    // key of the tuner cache: kernel class + box-size histogram bucket
    // (log2 of the median destination points per transfer) + count bucket
    {
      std::vector<long long> pts;
      for (int m : g.members)
        pts.push_back(tr[m].dst_points());
      std::nth_element(pts.begin(), pts.begin() + pts.size() / 2, pts.end());
      char key[128];
      snprintf(key, sizeof key, "%s k%d o%d t%d p%d op%d h%d n%d", g.is_tensor ? "tensor" : "point",
               g.kind, g.order, g.ntl, g.prec, g.op, ilog2(pts[pts.size() / 2]),
               ilog2((long long)g.members.size()));
      l.key = key;
    }
    l.zchunk = tile_cfg < 0 ? ZCHUNK_DEFAULT : tile_cfg;
    l.pt = PT_DEFAULT;
    if (tuned) {
      auto it = tuned->find(l.key);
      if (it != tuned->end()) {
        l.zchunk = it->second.first;
        l.pt = it->second.second;
      }
    }
    // persistent grid: occupancy x CUs, a multiple of the 8 XCDs
    int threads = 0;
    const void *fn = launch_fn(l, &threads);
    if (!fn)
      return set_error(CG_ERR_ORDER, "no persistent kernel for this transfer class");
    int per_cu = 0;
    he = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, 0);
    if (he != hipSuccess)
      return hip_error(he, "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    const int full_grid = std::max(8, (std::max(1, per_cu) * ncu) / 8 * 8);
    // transfers largest first; a transfer's items in spatial order (x tile
    // fastest, then y, then z chunk), so consecutive items are neighbours
    std::vector<int> members = g.members;
    std::stable_sort(members.begin(), members.end(), [&](int x, int y) {
      return tr[x].dst_points() > tr[y].dst_points();
    });
    std::vector<Item> items;
    std::vector<double> cost;
    if (g.is_tensor) {
      TileCfg cfg;
      cfg.cx = g.ntl == 1 ? TileShape<1>::CX : TileShape<3>::CX;
      cfg.cy = g.ntl == 1 ? TileShape<1>::CY : TileShape<3>::CY;
      // z chunk: whole columns (no halo planes) when that already gives the
      // grid a few items per workgroup, else the longest chunk that does
      if (tile_cfg < 0 && !(tuned && tuned->count(l.key))) {
        l.zchunk = 0;
        for (int zc : {0, 64, 32, 16, 8}) {
          l.zchunk = zc;
          long long tiles = 0;
          for (int m : members) {
            TensorArgs a = tr[m].pr.t;
            cfg.zchunk = zc;
            tiles += tensor_tiles(a, cfg);
          }
          if (tiles >= 4LL * full_grid)
            break;
        }
      }
      std::vector<TensorArgs> descs;
      for (int m : members) {
        TensorArgs a = tr[m].pr.t;
        cfg.zchunk = l.zchunk;
        const long long nt = tensor_tiles(a, cfg);
        const unsigned d = (unsigned)descs.size();
        descs.push_back(a);
        for (long long t = 0; t < nt; ++t) {
          items.push_back(Item{d, (unsigned)t});
          // cost ~ coarse planes x bases of the tile
          const int tz = (int)(t / ((long long)a.ntiles[0] * a.ntiles[1]));
          const int nz = std::min(a.zchunk, a.nq[2] - tz * a.zchunk) + 4;
          cost.push_back((double)nz * std::min(cfg.cx, a.nq[0]) * std::min(cfg.cy, a.nq[1]));
        }
      }
      l.desc_off = append(descs.data(), descs.size() * sizeof(TensorArgs));
    } else {
      std::vector<PointArgs> descs;
      for (int m : members) {
        const PointArgs &a = tr[m].is_prolong ? tr[m].pr.p : tr[m].pt.a;
        const long long tot = (long long)a.n[0] * a.n[1] * a.n[2];
        if (tot >= (1LL << 32))
          return set_error(CG_ERR_ARG, "region too large for the batch");
        const unsigned d = (unsigned)descs.size();
        descs.push_back(a);
        const long long nt = (tot + l.pt - 1) / l.pt;
        for (long long t = 0; t < nt; ++t) {
          items.push_back(Item{d, (unsigned)t});
          cost.push_back((double)std::min<long long>(l.pt, tot - t * l.pt));
        }
      }
      l.desc_off = append(descs.data(), descs.size() * sizeof(PointArgs));
    }
    l.item_off = append(items.data(), items.size() * sizeof(Item));
    l.nitems = (int)items.size();
    // 8 contiguous segments of equal cost, one per XCD
    {
      double total = 0;
      for (double c : cost)
        total += c;
      int segb[9];
      segb[0] = 0;
      double acc = 0;
      int k = 1;
      for (int i = 0; i < l.nitems && k < 8; ++i) {
        acc += cost[i];
        while (k < 8 && acc >= total * k / 8.0)
          segb[k++] = i + 1;
      }
      while (k < 9)
        segb[k++] = l.nitems;
      segb[8] = l.nitems;
      l.seg_off = append(segb, sizeof segb);
    }
    l.counter = (int)dev.launches.size();
    l.grid = std::min(full_grid, std::max(8, (l.nitems + 7) / 8 * 8));
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
  he = hipMalloc(&dev.counters, sizeof(int) * 8 * std::max(1, dev.ncounters));
  if (he != hipSuccess)
    return set_error(CG_ERR_NOMEM, "batch counters: %s", hipGetErrorString(he));
  return CG_OK;
}

static int run_launch(BatchDevice &dev, const BatchLaunch &l, hipStream_t st);

int batch_device_run(BatchDevice &dev, hipStream_t st) {
  if (dev.launches.empty())
    return CG_OK;
  hipLaunchKernelGGL(k_zero_counters, dim3((8 * dev.ncounters + 255) / 256), dim3(256), 0, st,
                     dev.counters, 8 * dev.ncounters);
  for (const BatchLaunch &l : dev.launches) {
    int e = run_launch(dev, l, st);
    if (e)
      return e;
  }
  return CG_OK;
}

int batch_device_run_one(BatchDevice &dev, int idx, hipStream_t st) {
  if (idx < 0 || idx >= (int)dev.launches.size())
    return set_error(CG_ERR_ARG, "no launch %d", idx);
  hipLaunchKernelGGL(k_zero_counters, dim3((8 * dev.ncounters + 255) / 256), dim3(256), 0, st,
                     dev.counters, 8 * dev.ncounters);
  return run_launch(dev, dev.launches[idx], st);
}

static int run_launch(BatchDevice &dev, const BatchLaunch &l, hipStream_t st) {
  {
    const void *descs = dev.blob + l.desc_off;
    const Item *items = reinterpret_cast<const Item *>(dev.blob + l.item_off);
    const int *seg = reinterpret_cast<const int *>(dev.blob + l.seg_off);
    int *ctr = dev.counters + 8 * l.counter;
    int rc = 0;
    if (l.is_tensor) {
      rc = with_line(l.kind, l.order, [&](auto tag) -> int {
        using L = decltype(tag);
        auto go = [&](auto ntlc) {
          constexpr int NTL = decltype(ntlc)::value;
          using T = TileShape<NTL>;
          if (l.prec)
            hipLaunchKernelGGL((k_persist_tensor<L, NTL, T::CX, T::CY, T::IPT, float>), dim3(l.grid),
                               dim3(2 * T::CX * T::CY / T::IPT), 0, st,
                               static_cast<const TensorArgs *>(descs), items, seg, ctr);
          else
            hipLaunchKernelGGL((k_persist_tensor<L, NTL, T::CX, T::CY, T::IPT, double>), dim3(l.grid),
                               dim3(2 * T::CX * T::CY / T::IPT), 0, st,
                               static_cast<const TensorArgs *>(descs), items, seg, ctr);
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
      case PO_COPY: hipLaunchKernelGGL(k_persist_point<PO_COPY>, dim3(l.grid), dim3(256), 0, st, pd, items, seg, ctr, l.pt); break;
      case PO_RESTRICT_VC: hipLaunchKernelGGL(k_persist_point<PO_RESTRICT_VC>, dim3(l.grid), dim3(256), 0, st, pd, items, seg, ctr, l.pt); break;
      case PO_RESTRICT_CC: hipLaunchKernelGGL(k_persist_point<PO_RESTRICT_CC>, dim3(l.grid), dim3(256), 0, st, pd, items, seg, ctr, l.pt); break;
      case PO_INTERP: hipLaunchKernelGGL(k_persist_point<PO_INTERP>, dim3(l.grid), dim3(256), 0, st, pd, items, seg, ctr, l.pt); break;
      case PO_TVD: hipLaunchKernelGGL(k_persist_point<PO_TVD>, dim3(l.grid), dim3(256), 0, st, pd, items, seg, ctr, l.pt); break;
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
