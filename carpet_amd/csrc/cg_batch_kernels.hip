// cg_batch_kernels.hip -- kernels of the batched level step.
//
// The transfers of one phase that share a kernel become one launch over a
// device-resident list of work items -- (descriptor, tile) pairs, transfers
// by decreasing size, each transfer's tiles in spatial order -- arranged in
// 8 segments, one per XCD, every transfer striped over all 8 in slabs of
// equal cost.  Two launch modes:
//  * persistent mode (default for tensor tiles): resident-capacity
//    workgroups take items from per-XCD counters, their own segment first,
//    then the others' (work stealing); items are taken in order, so the
//    tiles in flight on an XCD are neighbours that march through z together
//    and share halo rows in its L2; the loop ends for every workgroup once
//    all counters pass their segment ends, so the grid always drains;
//  * grid mode (default for pointwise runs): one workgroup per item;
//    workgroup b runs on XCD b % 8 (the hardware dispatches workgroups
//    round-robin over the XCDs) and takes slot b / 8 of that XCD's segment.
// Phases are separate launches in stream order (a kernel boundary makes one
// phase's results visible to the next across all XCD L2s), and the counters
// are cleared by a small kernel at the start of the step, so the whole step
// can be captured as one hipGraph.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <queue>
#include <vector>

#include "cg_batch_dev.hpp"

namespace cg {

// 256-point items per workgroup of pointwise classes of small regions in
// grid mode (k_grid_point_n)
#ifndef CG_POINT_NI
#define CG_POINT_NI 4
#endif

// destination stores of the pointwise kernels: nontemporal (1) or plain (0)
#ifndef CG_POINT_NT
#define CG_POINT_NT 1
#endif
__device__ __forceinline__ void point_store(double v, double *p) {
  if constexpr (CG_POINT_NT)
    __builtin_nontemporal_store(v, p);
  else
    *p = v;
}

// regions thin in x and long in y may take the 4x16-base tiles
#ifndef CG_SLIM_TILES
#define CG_SLIM_TILES 1
#endif

// tensor transfers of at most this many tiles run whole on one XCD
#ifndef CG_WHOLE_XCD_ITEMS
#define CG_WHOLE_XCD_ITEMS 16
#endif
constexpr int WHOLE_XCD_ITEMS = CG_WHOLE_XCD_ITEMS;

// default work items: pointwise = PT_DEFAULT consecutive destination points
// (x fastest); tensor = one x/y tile over ZCHUNK_DEFAULT coarse z bases
constexpr int PT_DEFAULT = 4096;
constexpr int ZCHUNK_DEFAULT = 16;


// points per thread and round of the pointwise kernel: items of 256, 512,
// 1024 points take one round of 1, 2, 4 points per thread, larger items
// rounds of 8 (a plain copy streams fastest with one point per thread and
// many small workgroups; stencils with several loads per point prefer more
// independent points per thread)
constexpr int pu_of(int pt) { return pt >= 2048 ? 8 : pt >= 1024 ? 4 : pt >= 512 ? 2 : 1; }

// coarse cells [tile*pt, tile*pt + pt) of a TVD region (fx/fxy divide by
// the parent-cell extents), each with all its children, NTL time levels
template <int NTL>
__device__ __forceinline__ void tvd_item(const PointArgs &a, unsigned tile, int pt) {
  int qx0, qy0, qz0, nqx, nqy, nqz;
  tvd_range(a, 0, &qx0, &nqx);
  tvd_range(a, 1, &qy0, &nqy);
  tvd_range(a, 2, &qz0, &nqz);
  const unsigned nxy = (unsigned)nqx * (unsigned)nqy;
  const unsigned tot = nxy * (unsigned)nqz; // < 2^31 (checked on the host)
  const unsigned e0 = tile * (unsigned)pt;
  const unsigned e1 = min(tot, e0 + (unsigned)pt);
  // TU cells per thread per round: every cell's loads before any store
  constexpr int TU = NTL == 1 ? 4 : 2;
  for (unsigned eb = e0; eb < e1; eb += 256u * TU) {
    TvdCell v[TU][NTL];
    int q[TU][3];
#pragma unroll
    for (int u = 0; u < TU; ++u) {
      const unsigned e = eb + u * 256u + threadIdx.x;
      q[u][0] = -1;
      if (e < e1) {
        const unsigned k = fastdiv(e, a.fxy);
        const unsigned r = e - k * nxy;
        const unsigned j = fastdiv(r, a.fx);
        const unsigned i = r - j * (unsigned)nqx;
        q[u][0] = qx0 + (int)i;
        q[u][1] = qy0 + (int)j;
        q[u][2] = qz0 + (int)k;
#pragma unroll
        for (int tl = 0; tl < NTL; ++tl)
          v[u][tl] = tvd_load(a, tl, q[u][0], q[u][1], q[u][2]);
      }
    }
#pragma unroll
    for (int u = 0; u < TU; ++u)
      if (q[u][0] >= 0)
        tvd_store<NTL>(a, q[u][0], q[u][1], q[u][2], v[u]);
  }
}

// points [tile*pt, tile*pt + pt) of a region, x fastest; for TVD the
// items count the region's coarse parent cells
template <int OP, int PU>
__device__ __forceinline__ void point_item(const PointArgs &a, unsigned tile, int pt) {
  if constexpr (OP == PO_TVD) {
    if (a.ntl == 1)
      tvd_item<1>(a, tile, pt);
    else if (a.ntl == 2)
      tvd_item<2>(a, tile, pt);
    else
      tvd_item<3>(a, tile, pt);
  } else {
    const unsigned nx = (unsigned)a.n[0];
    const unsigned nxy = nx * (unsigned)a.n[1];
    const unsigned tot = nxy * (unsigned)a.n[2]; // < 2^31 (checked on the host)
    const unsigned e0 = tile * (unsigned)pt;
    const unsigned e1 = min(tot, e0 + (unsigned)pt);
    // PU points per thread per round: all loads first, then all stores, so
    // that PU loads per thread are in flight at once
    for (unsigned eb = e0; eb < e1; eb += 256u * PU) {
      double v[PU];
      long long o[PU];
#pragma unroll
      for (int u = 0; u < PU; ++u) {
        const unsigned e = eb + u * 256u + threadIdx.x;
        o[u] = -1;
        if (e < e1) {
          const unsigned k = fastdiv(e, a.fxy);
          const unsigned r = e - k * nxy;
          const unsigned j = fastdiv(r, a.fx);
          const unsigned i = r - j * nx;
          v[u] = point_value<OP>(a, (int)i, (int)j, (int)k);
          o[u] = (long long)i + a.dj * j + a.dk * k;
        }
      }
#pragma unroll
      for (int u = 0; u < PU; ++u)
        if (o[u] >= 0)
          point_store(v[u], a.dst + o[u]);
    }
  }
}

template <int OP, int PU>
__global__ __launch_bounds__(256) void k_persist_point(
    const PointArgs *__restrict__ descs, const Item *__restrict__ items,
    const int *__restrict__ seg, int *counter, int pt) {
  __shared__ int s_item;
  for (;;) {
    const int it = dequeue(counter, seg, &s_item);
    if (it < 0)
      return;
    const Item w = items[it];
    point_item<OP, PU>(descs[w.desc], w.tile, pt);
  }
}

template <int OP, int PU>
__global__ __launch_bounds__(256) void k_grid_point(const PointArgs *__restrict__ descs,
                                                    const Item *__restrict__ items,
                                                    const int *__restrict__ seg, int pt) {
  const int it = grid_item(seg);
  if (it < 0)
    return;
  const Item w = items[it];
  point_item<OP, PU>(descs[w.desc], w.tile, pt);
}

// Grid mode for classes of 256-point items: workgroup b takes NI
// consecutive slots of its XCD's segment (slots (b/8)*NI ..), one point of
// each per thread, and issues the loads of all NI items before the first
// store.  Regions of a hierarchy's ghost shells are a few hundred points;
// with one item per workgroup each thread has a single load in flight and
// the launch is bound by workgroup latency.
template <int OP, int NI>
__global__ __launch_bounds__(256) void k_grid_point_n(const PointArgs *__restrict__ descs,
                                                      const Item *__restrict__ items,
                                                      const int *__restrict__ seg) {
  static_assert(OP != PO_TVD, "TVD items run per coarse cell");
  const int x = blockIdx.x & 7;
  const int i0 = seg[x] + (int)(blockIdx.x >> 3) * NI, i1 = seg[x + 1];
  double v[NI];
  double *q[NI]; // destination of each value, null outside the item
#pragma unroll
  for (int u = 0; u < NI; ++u) {
    q[u] = nullptr;
    const int it = i0 + u;
    if (it < i1) {
      const Item w = items[it];
      const PointArgs &a = descs[w.desc];
      const unsigned nx = (unsigned)a.n[0];
      const unsigned nxy = nx * (unsigned)a.n[1];
      const unsigned tot = nxy * (unsigned)a.n[2];
      const unsigned e = w.tile * 256u + threadIdx.x;
      if (e < tot) {
        const unsigned k = fastdiv(e, a.fxy);
        const unsigned r = e - k * nxy;
        const unsigned j = fastdiv(r, a.fx);
        const unsigned i = r - j * nx;
        v[u] = point_value<OP>(a, (int)i, (int)j, (int)k);
        q[u] = a.dst + ((long long)i + a.dj * j + a.dk * k);
      }
    }
  }
#pragma unroll
  for (int u = 0; u < NI; ++u)
    if (q[u])
      point_store(v[u], q[u]);
}

// TVD items run per coarse cell with their own unrolling: one instance
// points per thread and round instantiated for an operator: TVD always 8
// (per coarse cell), the staggered operators 1 (up to 64 taps per point and
// time level), the others as asked
template <int OP> constexpr int pu_inst(int pu) {
  return OP == PO_TVD ? 8 : (OP == PO_STAGGER || OP == PO_RESTRICT_STAGGER) ? 1 : pu;
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
  int small; // tensor regions routed to the 8x8-base (2) or 4x16-base (3) tiles, else 0
  int bigsh; // other tensor regions: the wide (0) or narrow (1) shape, per region
  std::vector<int> members; // transfer indices
  bool same(const Group &g) const {
    return phase == g.phase && is_tensor == g.is_tensor && kind == g.kind &&
           order == g.order && ntl == g.ntl && op == g.op && prec == g.prec && small == g.small &&
           bigsh == g.bigsh;
  }
};

// Lanes a region keeps busy in tiles of shape sh: tiles x (planes + stencil
// span) x (tile width) x (tile rows + span), the work of the window loads
// and of the x- and y-passes.
double tile_busy(const TensorArgs &a, int ntl, int prec, int span, int sh) {
  int cx, cy, th;
  tile_shape(ntl, prec, &cx, &cy, &th, sh);
  return (double)((a.nq[0] + cx - 1) / cx) * ((a.nq[1] + cy - 1) / cy) * (a.nq[2] + span) * cx *
         (cy + span);
}

int line_span(int kind, int order) {
  int span = 0;
  with_line(kind, order, [&](auto tag) -> int {
    using L = decltype(tag);
    span = L::TMAXALL - L::TMINALL;
    return 0;
  });
  return span;
}

// An FP64 region goes to the small tiles when they keep at most 1/1.3 of
// the lanes busy that the better of the two large shapes would: a small
// tile's lane costs more (more tiles per region, each with its own ramp-up
// planes and halo).  (Measured, both workloads in one process: the level
// step of a 3-level hierarchy, whose boundary prolongation regions are 3-
// and 6-point shells, 1.14 ms against 1.61 ms without small tiles; 2000
// regions of 8-96 points per axis 19.41 ms against 19.53 ms, and 19.83 ms
// when routing at the break-even point 1.0.)
// Of the two small shapes: 4x16 for regions at most 4 bases across in x
// (one tile wide) where it keeps fewer lanes busy than 8x8 -- ghost and
// buffer faces normal to x -- else 8x8.  (Routing every region by the
// lane count alone sent 2000 regions of 8-96 points per axis with
// 5-12 bases in x to the 4x16 tiles: 18.59 ms per step against 17.46.)
int small_tiles(const ProlongPlan &p, int ntl, int prec) {
  if (prec != CG_PREC_F64)
    return 0;
  const int span = line_span(p.kind, p.order);
  const double big = std::min(tile_busy(p.t, ntl, prec, span, 0), tile_busy(p.t, ntl, prec, span, 1));
  const double b2 = tile_busy(p.t, ntl, prec, span, 2), b3 = tile_busy(p.t, ntl, prec, span, 3);
  const int sh = CG_SLIM_TILES && p.t.nq[0] <= 4 && b3 < b2 ? 3 : 2;
  return (sh == 3 ? b3 : b2) * 1.3 <= big ? sh : 0;
}

// one transfer's descriptor in launch l: a tensor transfer resolved to one
// time level (its target time matches level 0 exactly) inside a class of
// more levels keeps nlev = 1 and reads level 0 for every level slot; a
// same-level interpolation resolved to a copy stays an interpolation of one
// level
TensorArgs tensor_desc(const BatchLaunch &l, const Transfer &t) {
  TensorArgs a = t.pr.t;
  for (int k = a.nlev; k < l.ntl; ++k)
    a.src[k] = a.src[0];
  TileCfg cfg;
  int th = 0;
  tile_shape(l.ntl, l.prec, &cfg.cx, &cfg.cy, &th, l.shape);
  cfg.zchunk = l.zchunk;
  (void)tensor_tiles(a, cfg);
  return a;
}

PointArgs point_desc(const BatchLaunch &l, const Transfer &t) {
  PointArgs a = t.is_prolong ? t.pr.p : t.pt.a;
  if (l.op == PO_INTERP && !t.is_prolong && t.pt.op == PO_COPY) {
    a.ntl = 1;
    a.s[0] = 1.0;
  }
  // work units: destination points, or TVD's coarse parent cells
  long long u[3];
  for (int d = 0; d < 3; ++d)
    u[d] = l.op == PO_TVD ? ((a.off[d] + a.n[d] - 1) >> 1) - (a.off[d] >> 1) + 1 : a.n[d];
  a.fx = make_fastdiv((unsigned)u[0]);
  a.fxy = make_fastdiv((unsigned)(u[0] * u[1]));
  return a;
}

template <int SH> const void *tensor_fn_sh(const BatchLaunch &l, int *threads) {
  switch (l.kind) {
  case LK_VLAG: return batch_tensor_fn_kind<LK_VLAG, SH>(l, threads);
  case LK_CLAG: return batch_tensor_fn_kind<LK_CLAG, SH>(l, threads);
  case LK_ENO: return batch_tensor_fn_kind<LK_ENO, SH>(l, threads);
  case LK_VENO: return batch_tensor_fn_kind<LK_VENO, SH>(l, threads);
  }
  return nullptr;
}

template <int SH>
int tensor_run_sh(const BatchLaunch &l, const TensorArgs *td, const Item *items, const int *seg, int *ctr,
                  hipStream_t st) {
  switch (l.kind) {
  case LK_VLAG: return batch_tensor_kind<LK_VLAG, SH>(l, td, items, seg, ctr, st);
  case LK_CLAG: return batch_tensor_kind<LK_CLAG, SH>(l, td, items, seg, ctr, st);
  case LK_ENO: return batch_tensor_kind<LK_ENO, SH>(l, td, items, seg, ctr, st);
  case LK_VENO: return batch_tensor_kind<LK_VENO, SH>(l, td, items, seg, ctr, st);
  }
  return -1;
}

template <int OP> const void *persist_point_pu(int pu) {
  switch (pu_inst<OP>(pu)) {
  case 1: return reinterpret_cast<const void *>(&k_persist_point<OP, 1>);
  case 2: return reinterpret_cast<const void *>(&k_persist_point<OP, 2>);
  case 4: return reinterpret_cast<const void *>(&k_persist_point<OP, 4>);
  }
  return reinterpret_cast<const void *>(&k_persist_point<OP, pu_inst<OP>(8)>);
}

const void *persist_point_fn(int op, int pt) {
  const int pu = pu_of(pt);
  switch (op) {
  case PO_COPY: return persist_point_pu<PO_COPY>(pu);
  case PO_RESTRICT_VC: return persist_point_pu<PO_RESTRICT_VC>(pu);
  case PO_RESTRICT_CC: return persist_point_pu<PO_RESTRICT_CC>(pu);
  case PO_INTERP: return persist_point_pu<PO_INTERP>(pu);
  case PO_TVD: return persist_point_pu<PO_TVD>(pu);
  case PO_RESTRICT_CC_O3: return persist_point_pu<PO_RESTRICT_CC_O3>(pu);
  case PO_RESTRICT_CC_O5: return persist_point_pu<PO_RESTRICT_CC_O5>(pu);
  case PO_MONO5: return persist_point_pu<PO_MONO5>(pu);
  case PO_STAGGER: return persist_point_pu<PO_STAGGER>(pu);
  case PO_RESTRICT_STAGGER: return persist_point_pu<PO_RESTRICT_STAGGER>(pu);
  case PO_RESTRICT_VC2: return persist_point_pu<PO_RESTRICT_VC2>(pu);
  }
  return nullptr;
}

const void *launch_fn(const BatchLaunch &l, int *threads) {
  if (!l.is_tensor) {
    *threads = 256;
    return persist_point_fn(l.op, l.pt);
  }
  switch (l.shape) {
  case 0: return tensor_fn_sh<0>(l, threads);
  case 1: return tensor_fn_sh<1>(l, threads);
  case 2: return tensor_fn_sh<2>(l, threads);
  case 3: return tensor_fn_sh<3>(l, threads);
  }
  return nullptr;
}

} // namespace

void batch_device_free(BatchDevice &dev) {
  for (hipStream_t s : dev.aux)
    (void)hipStreamSynchronize(s);
  for (hipStream_t s : dev.aux)
    (void)hipStreamDestroy(s);
  for (hipEvent_t e : dev.join)
    (void)hipEventDestroy(e);
  for (hipEvent_t e : dev.fork)
    (void)hipEventDestroy(e);
  dev.aux.clear();
  dev.join.clear();
  dev.fork.clear();
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

// Workgroups of fn that one CU holds at once, from the kernel's registers
// and LDS and the device's per-CU LDS: each SIMD has 512 VGPRs per lane
// (allocated in blocks of 8) and at most 8 waves; a workgroup's waves are
// spread over the 4 SIMDs.  The runtime's occupancy query is the fallback
// and the bound when it reports more.
static int resident_per_cu(const void *fn, int threads, int dev) {
  int api = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&api, fn, threads, 0) != hipSuccess)
    api = 1;
  hipFuncAttributes fa;
  int lds_cu = 0;
  if (hipFuncGetAttributes(&fa, fn) != hipSuccess ||
      hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) !=
          hipSuccess ||
      fa.numRegs <= 0)
    return std::max(1, api);
  const int waves = (threads + 63) / 64;
  const int per_simd = (waves + 3) / 4;
  const int vgprs = (fa.numRegs + 7) / 8 * 8;
  const int by_vgpr = std::min(8, 512 / vgprs) / per_simd;
  const int by_lds = fa.sharedSizeBytes > 0 ? lds_cu / (int)fa.sharedSizeBytes : 32;
  const int by_waves = 32 / waves;
  return std::max({1, api, std::min({by_vgpr, by_lds, by_waves})});
}

static int ilog2(long long v) {
  int r = 0;
  while (v > 1) {
    v >>= 1;
    ++r;
  }
  return r;
}

void batch_launch_descs(const BatchLaunch &l, const std::vector<Transfer> &tr, char *out) {
  for (size_t i = 0; i < l.members.size(); ++i) {
    const Transfer &t = tr[l.members[i]];
    if (l.is_tensor) {
      const TensorArgs a = tensor_desc(l, t);
      std::memcpy(out + i * sizeof(TensorArgs), &a, sizeof a);
    } else {
      const PointArgs a = point_desc(l, t);
      std::memcpy(out + i * sizeof(PointArgs), &a, sizeof a);
    }
  }
}

int batch_device_build(BatchDevice &dev, const std::vector<Transfer> &tr,
                       const std::vector<int> &phase, const std::vector<int> &prec,
                       const std::vector<int> &cls_ntl, int tile_cfg, const TuneMap *tuned,
                       int only) {
  batch_device_free(dev);
  // group transfers by (phase, kernel)
  std::vector<Group> groups;
  for (size_t i = 0; i < tr.size(); ++i) {
    Group g{};
    g.phase = phase[i];
    g.prec = prec[i];
    if (tr[i].is_prolong && tr[i].pr.kind != PK_TVD && tr[i].pr.kind != PK_MONO5 &&
        tr[i].pr.kind != PK_STAGGER) {
      g.is_tensor = 1;
      g.small = small_tiles(tr[i].pr, cls_ntl[i], prec[i]);
      // the large shape that keeps fewer lanes busy on this region -- tiles x
      // (planes + stencil span) x (tile width) x (tile rows + span), the work
      // of the window loads and of the x- and y-passes -- with a 25% margin
      // for the wide tile, which is faster per lane on large regions
      // (single 512^3 patch: narrow 0.81x the busy lanes of wide, yet 4%
      // slower).  Per region rather than per class: 0.6% faster on 2000
      // regions of 8-96 points per axis, 1.8% and 1.2% on the level steps
      // of the 3-level and the 2000-box hierarchies.
      g.bigsh = 0;
      if (!g.small) {
        const int span = line_span(tr[i].pr.kind, tr[i].pr.order);
        const double b0 = tile_busy(tr[i].pr.t, cls_ntl[i], prec[i], span, 0);
        const double b1 = tile_busy(tr[i].pr.t, cls_ntl[i], prec[i], span, 1);
        g.bigsh = b1 < 0.75 * b0 ? 1 : 0;
      }
      g.kind = tr[i].pr.kind;
      g.order = tr[i].pr.order;
      g.ntl = cls_ntl[i];
      g.op = -1;
    } else {
      g.is_tensor = 0;
      g.kind = g.order = -1;
      g.ntl = 1;
      g.op = !tr[i].is_prolong            ? tr[i].pt.op
             : tr[i].pr.kind == PK_MONO5   ? PO_MONO5
             : tr[i].pr.kind == PK_STAGGER ? PO_STAGGER
                                           : PO_TVD;
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
  for (size_t gi = 0; gi < groups.size(); ++gi) {
    const Group &g = groups[gi];
    if (only >= 0 && (int)gi != only)
      continue;
    if (g.prec != CG_PREC_F64 && (!g.is_tensor || fp64_only(g.kind)))
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
    long long median_pts = 0;
    {
      std::vector<long long> pts;
      for (int m : g.members)
        pts.push_back(tr[m].dst_points());
      std::nth_element(pts.begin(), pts.begin() + pts.size() / 2, pts.end());
      median_pts = pts[pts.size() / 2];
      char key[128];
      snprintf(key, sizeof key, "%s k%d o%d t%d p%d op%d h%d n%d",
               !g.is_tensor ? "point" : g.small == 3 ? "tensor4" : g.small ? "tensor8" : g.bigsh == 1 ? "tensorN" : "tensor",
               g.kind, g.order, g.ntl, g.prec, g.op, ilog2(median_pts),
               ilog2((long long)g.members.size()));
      l.key = key;
    }
    l.zchunk = tile_cfg < 0 ? ZCHUNK_DEFAULT : tile_cfg;
    l.pt = PT_DEFAULT;
    // pointwise classes of small regions (median under 2048 points, such as
    // the ghost and restriction slabs of a hierarchy's components): items of
    // 256 points, one per thread, so a region spreads over several
    // workgroups (the tuner's choice on such classes)
    l.ni = 1;
    if (!g.is_tensor && median_pts < 2048) {
      l.pt = 256;
      l.ni = CG_POINT_NI;
    }
    // TVD items count coarse parent cells (8 fine points each): 1024 cells
    // (measured on 8 variables of 192^3 -> 384^3: 0.98 ms against 1.05 ms
    // with 4096 and 1.11 ms with 256); cell-centred restriction (8 or more
    // children per point) likewise 1024 points (the 384^3 -> 192^3
    // restriction of the same set: 0.77 ms against 0.81 ms with 4096)
    else if (g.op == PO_TVD || g.op == PO_RESTRICT_CC || g.op == PO_RESTRICT_CC_O3 || g.op == PO_RESTRICT_CC_O5)
      l.pt = 1024;
    // copies: 2048 points (2000 regions of 8-96 points per axis: 19.03 ms
    // per step against 19.43 ms with 4096)
    else if (g.op == PO_COPY)
      l.pt = 2048;
    // tensor tiles of large regions (median 2^20 points and up): persistent
    // grid (its workgroups take the tiles in order, so neighbouring tiles
    // march through z together and share halo rows in L2); other tensor
    // classes and pointwise runs: one workgroup per item.  (Measured: on
    // 2000 regions of 8-96 points per axis both modes are within 0.2%; on a
    // hierarchy's shells one workgroup per item takes 0.45 ms less per
    // step for the small tiles and 0.09 ms less for the large ones.)
    l.wpc = g.is_tensor && !g.small && median_pts >= (1LL << 20) ? 0 : -1;
    // tile shape: the class's (its regions were routed by shape)
    l.shape = !g.is_tensor ? 0 : g.small ? g.small : g.bigsh;
    if (tuned) {
      auto it = tuned->find(l.key);
      if (it != tuned->end()) {
        l.zchunk = it->second.zchunk;
        l.pt = it->second.pt;
        l.wpc = it->second.wpc;
        // items per workgroup: 1, 2, 4, 8 or 16 (the largest of these not
        // above the cached value)
        l.ni = 1;
        while (l.ni < 16 && 2 * l.ni <= it->second.ni)
          l.ni *= 2;
        if (!g.is_tensor && l.wpc == -2) // static rounds: tensor tiles only
          l.wpc = -1;
        if (g.is_tensor && !g.small && it->second.shape >= 0 && it->second.shape < 2)
          l.shape = it->second.shape;
      }
    }
    // persistent grid: resident workgroups per CU x CUs, a multiple of the
    // 8 XCDs
    int threads = 0;
    const void *fn = launch_fn(l, &threads);
    if (!fn)
      return set_error(CG_ERR_ORDER, "no persistent kernel for this transfer class");
    const int per_cu = resident_per_cu(fn, threads, dev_id);
    const int full_grid = std::max(8, (per_cu * ncu) / 8 * 8);
    // transfers largest first; a transfer's items in spatial order (x tile
    // fastest, then y, then z chunk), so consecutive items are neighbours
    std::vector<int> members = g.members;
    std::stable_sort(members.begin(), members.end(), [&](int x, int y) {
      return tr[x].dst_points() > tr[y].dst_points();
    });
    std::vector<Item> items;
    std::vector<double> cost;
    std::vector<int> first; // first item of each transfer, plus the end
    if (g.is_tensor) {
      TileCfg cfg;
      int tthreads = 0;
      tile_shape(g.ntl, g.prec, &cfg.cx, &cfg.cy, &tthreads, l.shape);
      // z chunk: the one whose items finish soonest on the resident
      // workgroups, under greedy list scheduling in dispatch order (what the
      // work counters and the dispatcher do) with an item's cost its planes
      // plus the ramp-up of the stencil span; a chunk adds those ramp-up
      // planes, and whole-region items leave workgroups idle when regions
      // are few.  Near-ties go to the longer chunk (neighbouring tiles
      // then march through z together and share halo rows in L2).
      if (tile_cfg < 0 && !(tuned && tuned->count(l.key))) {
        const int span = line_span(g.kind, g.order);
        std::vector<double> mk;
        const int zcs[5] = {128, 64, 32, 16, 8};
        for (int zc : zcs) {
          std::priority_queue<double, std::vector<double>, std::greater<double>> q;
          for (int w = 0; w < full_grid; ++w)
            q.push(0.0);
          for (int m : members) {
            const TensorArgs &a = tr[m].pr.t;
            const long long nxy =
                (long long)((a.nq[0] + cfg.cx - 1) / cfg.cx) * ((a.nq[1] + cfg.cy - 1) / cfg.cy);
            for (int z0 = 0; z0 < a.nq[2]; z0 += zc)
              for (long long t = 0; t < nxy; ++t) {
                const double at = q.top();
                q.pop();
                q.push(at + std::min(zc, a.nq[2] - z0) + span + 1);
              }
          }
          double end = 0;
          for (; !q.empty(); q.pop())
            end = std::max(end, q.top());
          mk.push_back(end);
        }
        // ENO kinds (data-dependent stencils, several times the arithmetic
        // per plane) march slower and drift further apart: at most 64
        // planes (8 variables of 192^3 -> 384^3 with ENO<3>: 1.64 ms at 64
        // against 1.74 ms at 128)
        const int first = fp64_only(g.kind) ? 1 : 0;
        const double best = *std::min_element(mk.begin() + first, mk.end());
        for (int i = first; i < 5; ++i)
          if (mk[i] <= 1.03 * best) {
            l.zchunk = zcs[i];
            break;
          }
      }
      std::vector<TensorArgs> descs;
      for (int m : members) {
        const TensorArgs a = tensor_desc(l, tr[m]);
        const long long nt = (long long)a.ntiles[0] * a.ntiles[1] * a.ntiles[2];
        const unsigned d = (unsigned)descs.size();
        descs.push_back(a);
        first.push_back((int)items.size());
        for (long long t = 0; t < nt; ++t) {
          items.push_back(Item{d, (unsigned)t});
          // cost ~ coarse planes (with the halo) x tile area: a tile's
          // lanes are busy whether or not the region fills it
          const int tz = (int)(t / ((long long)a.ntiles[0] * a.ntiles[1]));
          const int nz = std::min(a.zchunk, a.nq[2] - tz * a.zchunk) + 5;
          cost.push_back((double)nz * cfg.cx * cfg.cy);
        }
      }
      l.desc_off = append(descs.data(), descs.size() * sizeof(TensorArgs));
      l.desc_bytes = descs.size() * sizeof(TensorArgs);
    } else {
      std::vector<PointArgs> descs;
      for (int m : members) {
        const PointArgs a = point_desc(l, tr[m]);
        // work units: destination points, or TVD's coarse parent cells
        long long u[3];
        for (int d = 0; d < 3; ++d)
          u[d] = g.op == PO_TVD ? ((a.off[d] + a.n[d] - 1) >> 1) - (a.off[d] >> 1) + 1 : a.n[d];
        const long long tot = u[0] * u[1] * u[2];
        if (tot >= (1LL << 31) - l.pt)
          return set_error(CG_ERR_ARG, "region too large for the batch");
        const unsigned d = (unsigned)descs.size();
        descs.push_back(a);
        first.push_back((int)items.size());
        const long long nt = (tot + l.pt - 1) / l.pt;
        for (long long t = 0; t < nt; ++t) {
          items.push_back(Item{d, (unsigned)t});
          cost.push_back((double)std::min<long long>(l.pt, tot - t * l.pt));
        }
      }
      l.desc_off = append(descs.data(), descs.size() * sizeof(PointArgs));
      l.desc_bytes = descs.size() * sizeof(PointArgs);
    }
    l.members = members;
    first.push_back((int)items.size());
    l.nitems = (int)items.size();
    // One segment per XCD.  Every transfer is striped over the XCDs: its
    // items are cut into 8 contiguous parts of equal cost (spatial slabs,
    // e.g. ranges of y tiles) and part x goes to XCD x, so all XCDs work on
    // the same transfer at the same time, as one launch per transfer would,
    // with each slab's halos shared in one L2.  A tensor transfer of at
    // most WHOLE_XCD_ITEMS tiles goes whole to the XCD with the least work
    // so far: its tiles' stencil halos overlap, and on one XCD the rows two
    // tiles share are fetched from HBM once, where striped over XCDs every
    // tile fetches them into its own L2.  A pointwise transfer with fewer
    // than 8 items gives each item to the least-loaded XCD.
    int segb[9];
    {
      std::vector<Item> part[8];
      double load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (size_t m = 0; m + 1 < first.size(); ++m) {
        const int i0 = first[m], i1 = first[m + 1];
        if (g.is_tensor && i1 - i0 <= WHOLE_XCD_ITEMS) {
          const int x = (int)(std::min_element(load, load + 8) - load);
          for (int i = i0; i < i1; ++i) {
            part[x].push_back(items[i]);
            load[x] += cost[i];
          }
          continue;
        }
        if (i1 - i0 < 8) {
          for (int i = i0; i < i1; ++i) {
            const int x = (int)(std::min_element(load, load + 8) - load);
            part[x].push_back(items[i]);
            load[x] += cost[i];
          }
          continue;
        }
        double total = 0;
        for (int i = i0; i < i1; ++i)
          total += cost[i];
        double acc = 0;
        int x = 0;
        for (int i = i0; i < i1; ++i) {
          // item i belongs to the part its cost midpoint falls in
          const double mid = acc + 0.5 * cost[i];
          while (x < 7 && mid >= total * (x + 1) / 8.0)
            ++x;
          part[x].push_back(items[i]);
          load[x] += cost[i];
          acc += cost[i];
        }
      }
      items.clear();
      segb[0] = 0;
      for (int x = 0; x < 8; ++x) {
        items.insert(items.end(), part[x].begin(), part[x].end());
        segb[x + 1] = (int)items.size();
      }
      l.seg_off = append(segb, sizeof segb);
    }
    l.item_off = append(items.data(), items.size() * sizeof(Item));
    l.counter = (int)dev.launches.size();
    if (l.wpc == -1) {
      // grid mode: 8 x the longest segment (in groups of ni items)
      long long longest = 1;
      for (int x = 0; x < 8; ++x)
        longest = std::max<long long>(longest, segb[x + 1] - segb[x]);
      const int ni = (!g.is_tensor && l.pt == 256 && g.op != PO_TVD) ? l.ni : 1;
      longest = (longest + ni - 1) / ni;
      if (8 * longest > 0x7fffffffLL)
        return set_error(CG_ERR_ARG, "too many work items for one launch");
      l.grid = (int)(8 * longest);
    } else {
      const long long all = std::max(8, (l.nitems + 7) / 8 * 8);
      // persistent (counters or static rounds): a multiple of 8 workgroups
      const long long grid =
          l.wpc > 0 ? std::max(8LL, (long long)l.wpc * ncu / 8 * 8) : full_grid;
      l.grid = (int)std::min(grid, all);
    }
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

// helper streams (default priority, non-blocking) and the fork/join events
// of the launches that share a phase
static int ensure_aux(BatchDevice &dev, size_t n) {
  if (dev.fork.empty()) {
    hipEvent_t e;
    hipError_t he = hipEventCreateWithFlags(&e, hipEventDisableTiming);
    if (he != hipSuccess)
      return hip_error(he, "hipEventCreate");
    dev.fork.push_back(e);
  }
  while (dev.aux.size() < n) {
    hipStream_t s;
    hipEvent_t e;
    hipError_t he = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (he != hipSuccess)
      return hip_error(he, "hipStreamCreate");
    dev.aux.push_back(s);
    he = hipEventCreateWithFlags(&e, hipEventDisableTiming);
    if (he != hipSuccess)
      return hip_error(he, "hipEventCreate");
    dev.join.push_back(e);
  }
  return CG_OK;
}

size_t batch_phase_end(const BatchDevice &dev, size_t i) {
  size_t j = i + 1;
  while (j < dev.launches.size() && dev.launches[j].phase == dev.launches[i].phase)
    ++j;
  return j;
}

// The launches of one phase write disjoint regions (dh's lists never send
// one point twice), so the classes of a phase may run side by side: the
// first on st, the others on helper streams forked from st by an event and
// joined back before the next phase.  Under stream capture the fork and
// join become branches of the graph.
static int run_phase(BatchDevice &dev, size_t i, size_t j, bool side, hipStream_t st) {
  const size_t extra = side ? j - i - 1 : 0;
  if (extra > 0) {
    if (int e = ensure_aux(dev, extra))
      return e;
    hipError_t he = hipEventRecord(dev.fork[0], st);
    for (size_t k = 0; k < extra && he == hipSuccess; ++k)
      he = hipStreamWaitEvent(dev.aux[k], dev.fork[0], 0);
    if (he != hipSuccess)
      return hip_error(he, "phase fork");
  }
  for (size_t k = i; k < j; ++k) {
    int e = run_launch(dev, dev.launches[k], (k == i || !side) ? st : dev.aux[k - i - 1]);
    if (e)
      return e;
  }
  for (size_t k = 0; k < extra; ++k) {
    hipError_t he = hipEventRecord(dev.join[k], dev.aux[k]);
    if (he == hipSuccess)
      he = hipStreamWaitEvent(st, dev.join[k], 0);
    if (he != hipSuccess)
      return hip_error(he, "phase join");
  }
  return CG_OK;
}

static bool phase_beside(const BatchDevice &dev, size_t i) {
  return dev.beside.size() != dev.launches.size() || dev.beside[i] != 0;
}

int batch_device_run(BatchDevice &dev, hipStream_t st) {
  if (dev.launches.empty())
    return CG_OK;
  // only persistent launches use the work counters
  bool persistent = false;
  for (const BatchLaunch &l : dev.launches)
    persistent = persistent || l.wpc >= 0;
  if (persistent)
    hipLaunchKernelGGL(k_zero_counters, dim3((8 * dev.ncounters + 255) / 256), dim3(256), 0, st,
                       dev.counters, 8 * dev.ncounters);
  for (size_t i = 0; i < dev.launches.size();) {
    const size_t j = batch_phase_end(dev, i);
    if (int e = run_phase(dev, i, j, phase_beside(dev, i), st))
      return e;
    i = j;
  }
  return CG_OK;
}

int batch_device_run_phase(BatchDevice &dev, size_t i, size_t j, bool side, hipStream_t st) {
  if (i >= j || j > dev.launches.size())
    return set_error(CG_ERR_ARG, "no launches [%zu, %zu)", i, j);
  hipLaunchKernelGGL(k_zero_counters, dim3((8 * dev.ncounters + 255) / 256), dim3(256), 0, st,
                     dev.counters, 8 * dev.ncounters);
  return run_phase(dev, i, j, side, st);
}

int batch_device_run_one(BatchDevice &dev, int idx, hipStream_t st) {
  if (idx < 0 || idx >= (int)dev.launches.size())
    return set_error(CG_ERR_ARG, "no launch %d", idx);
  hipLaunchKernelGGL(k_zero_counters, dim3((8 * dev.ncounters + 255) / 256), dim3(256), 0, st,
                     dev.counters, 8 * dev.ncounters);
  return run_launch(dev, dev.launches[idx], st);
}

static int run_launch(BatchDevice &dev, const BatchLaunch &l, hipStream_t st) {
  const void *descs = dev.blob + l.desc_off;
  const Item *items = reinterpret_cast<const Item *>(dev.blob + l.item_off);
  const int *seg = reinterpret_cast<const int *>(dev.blob + l.seg_off);
  int *ctr = dev.counters + 8 * l.counter;
  const bool grid_mode = l.wpc == -1;
  int rc = 0;
  if (l.is_tensor) {
    const TensorArgs *td = static_cast<const TensorArgs *>(descs);
    switch (l.shape) {
    case 0: rc = tensor_run_sh<0>(l, td, items, seg, ctr, st); break;
    case 1: rc = tensor_run_sh<1>(l, td, items, seg, ctr, st); break;
    case 2: rc = tensor_run_sh<2>(l, td, items, seg, ctr, st); break;
    case 3: rc = tensor_run_sh<3>(l, td, items, seg, ctr, st); break;
    default: rc = -1;
    }
  } else {
    const PointArgs *pd = static_cast<const PointArgs *>(descs);
    auto go_pu = [&](auto opc, auto puc) {
      constexpr int OP = decltype(opc)::value;
      constexpr int PU = decltype(puc)::value;
      if constexpr (OP != PO_TVD) {
        if (grid_mode && l.pt == 256 && l.ni > 1) {
          switch (l.ni) {
          case 2: hipLaunchKernelGGL((k_grid_point_n<OP, 2>), dim3(l.grid), dim3(256), 0, st, pd, items, seg); break;
          case 4: hipLaunchKernelGGL((k_grid_point_n<OP, 4>), dim3(l.grid), dim3(256), 0, st, pd, items, seg); break;
          case 8: hipLaunchKernelGGL((k_grid_point_n<OP, 8>), dim3(l.grid), dim3(256), 0, st, pd, items, seg); break;
          default: hipLaunchKernelGGL((k_grid_point_n<OP, 16>), dim3(l.grid), dim3(256), 0, st, pd, items, seg);
          }
          return 0;
        }
      }
      if (grid_mode)
        hipLaunchKernelGGL((k_grid_point<OP, PU>), dim3(l.grid), dim3(256), 0, st, pd, items, seg, l.pt);
      else
        hipLaunchKernelGGL((k_persist_point<OP, PU>), dim3(l.grid), dim3(256), 0, st, pd, items, seg,
                           ctr, l.pt);
      return 0;
    };
    auto go = [&](auto opc) {
      constexpr int OP = decltype(opc)::value;
      switch (pu_inst<OP>(pu_of(l.pt))) {
      case 1: return go_pu(opc, std::integral_constant<int, pu_inst<OP>(1)>{});
      case 2: return go_pu(opc, std::integral_constant<int, pu_inst<OP>(2)>{});
      case 4: return go_pu(opc, std::integral_constant<int, pu_inst<OP>(4)>{});
      }
      return go_pu(opc, std::integral_constant<int, pu_inst<OP>(8)>{});
    };
    switch (l.op) {
    case PO_COPY: rc = go(std::integral_constant<int, PO_COPY>{}); break;
    case PO_RESTRICT_VC: rc = go(std::integral_constant<int, PO_RESTRICT_VC>{}); break;
    case PO_RESTRICT_CC: rc = go(std::integral_constant<int, PO_RESTRICT_CC>{}); break;
    case PO_INTERP: rc = go(std::integral_constant<int, PO_INTERP>{}); break;
    case PO_TVD: rc = go(std::integral_constant<int, PO_TVD>{}); break;
    case PO_RESTRICT_CC_O3: rc = go(std::integral_constant<int, PO_RESTRICT_CC_O3>{}); break;
    case PO_RESTRICT_CC_O5: rc = go(std::integral_constant<int, PO_RESTRICT_CC_O5>{}); break;
    case PO_MONO5: rc = go(std::integral_constant<int, PO_MONO5>{}); break;
    case PO_STAGGER: rc = go(std::integral_constant<int, PO_STAGGER>{}); break;
    case PO_RESTRICT_STAGGER: rc = go(std::integral_constant<int, PO_RESTRICT_STAGGER>{}); break;
    case PO_RESTRICT_VC2: rc = go(std::integral_constant<int, PO_RESTRICT_VC2>{}); break;
    default: rc = -1;
    }
  }
  if (rc)
    return set_error(CG_ERR_ORDER, "no batch kernel for this transfer class");
  hipError_t he = hipGetLastError();
  if (he != hipSuccess)
    return hip_error(he, "batch launch");
  return CG_OK;
}

} // namespace cg
