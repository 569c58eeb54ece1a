// cg_batch.cpp -- batched level step (replaces the per-region walk of
// ggf::transfer_from_all, CarpetLib/src/ggf.cc:500-591).
//
// The caller adds every (region, variable) transfer of a level step once.
// Transfers are resolved and validated on the host exactly as
// cg_transfer_from would (same plans, same error codes).  "plain" runs them
// as one launch each, the shape of the reference's per-region operator
// calls.  "batched" runs each phase as one persistent launch over a
// device-resident descriptor list (cg_batch_kernels.hip).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "cg_batch.hpp"
#include "cg_internal.hpp"

using namespace cg;

struct cg_batch {
  std::vector<Transfer> transfers;
  std::vector<int> phase;
  std::vector<int> prec;
  cg::BatchDevice dev;
  bool finalized = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

extern "C" {

int cg_batch_create(cg_batch **out) {
  if (!out)
    return set_error(CG_ERR_ARG, "null out");
  *out = new (std::nothrow) cg_batch();
  return *out ? CG_OK : set_error(CG_ERR_NOMEM, "batch allocation failed");
}

int cg_batch_destroy(cg_batch *b) {
  if (!b)
    return set_error(CG_ERR_ARG, "null batch");
  if (b->exec)
    (void)hipGraphExecDestroy(b->exec);
  if (b->graph)
    (void)hipGraphDestroy(b->graph);
  batch_device_free(b->dev);
  delete b;
  return CG_OK;
}

int cg_batch_add(cg_batch *b, int phase, const cg_gdata *dst,
                 const cg_gdata *const *srcs, const double *times, int ntimes,
                 const int dstregbbox[9], const int srcregbbox[9],
                 const cg_islab *slab, double time, int transport_operator,
                 int centering, int order_space, int order_time, int prec) {
  if (!b)
    return set_error(CG_ERR_ARG, "null batch");
  if (b->finalized)
    return set_error(CG_ERR_STATE, "batch already finalized");
  if (phase < 0 || phase > 15)
    return set_error(CG_ERR_ARG, "phase must be in [0,15]");
  if (prec != CG_PREC_F64 && prec != CG_PREC_MIXED)
    return set_error(CG_ERR_ARG, "unknown precision %d", prec);
  Transfer tr;
  int e = resolve_transfer(dst, srcs, times, ntimes, dstregbbox, srcregbbox, slab,
                           time, transport_operator, centering, order_space,
                           order_time, tr);
  if (e)
    return e;
  if (prec == CG_PREC_MIXED && !(tr.is_prolong && (tr.pr.kind == PK_VLAG)))
    return set_error(CG_ERR_OPERATOR, "mixed precision applies to Lagrange prolongation only");
  b->transfers.push_back(tr);
  b->phase.push_back(phase);
  b->prec.push_back(prec);
  return (int)b->transfers.size() - 1;
}

int cg_batch_finalize(cg_batch *b, int tile_cfg) {
  if (!b)
    return set_error(CG_ERR_ARG, "null batch");
  if (b->finalized)
    return set_error(CG_ERR_STATE, "batch already finalized");
  int e = batch_device_build(b->dev, b->transfers, b->phase, b->prec, tile_cfg);
  if (e)
    return e;
  b->finalized = true;
  return CG_OK;
}

int cg_batch_run(cg_batch *b, cg_stream_t stream) {
  if (!b)
    return set_error(CG_ERR_ARG, "null batch");
  if (!b->finalized)
    return set_error(CG_ERR_STATE, "batch not finalized");
  return batch_device_run(b->dev, static_cast<hipStream_t>(stream));
}

int cg_batch_run_plain(cg_batch *b, cg_stream_t stream) {
  if (!b)
    return set_error(CG_ERR_ARG, "null batch");
  if (!b->finalized)
    return set_error(CG_ERR_STATE, "batch not finalized");
  hipStream_t st = static_cast<hipStream_t>(stream);
  // phase order, then insertion order
  std::vector<int> order(b->transfers.size());
  for (size_t i = 0; i < order.size(); ++i)
    order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int x, int y) { return b->phase[x] < b->phase[y]; });
  for (int i : order) {
    int e = run_transfer(b->transfers[i], b->prec[i], st);
    if (e)
      return e;
  }
  return CG_OK;
}

int cg_batch_capture_graph(cg_batch *b, cg_stream_t stream) {
  if (!b)
    return set_error(CG_ERR_ARG, "null batch");
  if (!b->finalized)
    return set_error(CG_ERR_STATE, "batch not finalized");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (!st)
    return set_error(CG_ERR_ARG, "graph capture needs a non-default stream");
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  if (b->graph) {
    (void)hipGraphDestroy(b->graph);
    b->graph = nullptr;
  }
  hipError_t he = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
  if (he != hipSuccess)
    return hip_error(he, "hipStreamBeginCapture");
  int e = batch_device_run(b->dev, st);
  hipGraph_t g = nullptr;
  he = hipStreamEndCapture(st, &g);
  if (e)
    return e;
  if (he != hipSuccess)
    return hip_error(he, "hipStreamEndCapture");
  b->graph = g;
  he = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
  if (he != hipSuccess)
    return hip_error(he, "hipGraphInstantiate");
  return CG_OK;
}

int cg_batch_launch_graph(cg_batch *b, cg_stream_t stream) {
  if (!b)
    return set_error(CG_ERR_ARG, "null batch");
  if (!b->exec)
    return set_error(CG_ERR_STATE, "no captured graph");
  hipError_t he = hipGraphLaunch(b->exec, static_cast<hipStream_t>(stream));
  return he == hipSuccess ? CG_OK : hip_error(he, "hipGraphLaunch");
}

int cg_batch_stats(const cg_batch *b, long long *ntransfers, long long *dst_points,
                   long long *alg_bytes) {
  if (!b)
    return set_error(CG_ERR_ARG, "null batch");
  long long n = 0, pts = 0, bytes = 0;
  for (size_t i = 0; i < b->transfers.size(); ++i) {
    const Transfer &t = b->transfers[i];
    const long long es = b->prec[i] == CG_PREC_MIXED ? 4 : 8;
    ++n;
    if (t.is_prolong) {
      pts += t.pr.dst_points;
      bytes += es * (t.pr.dst_points + t.pr.src_points * t.pr.ntl);
    } else {
      const int ntl = t.pt.op == PO_INTERP ? t.pt.a.ntl : 1;
      pts += t.pt.dst_points;
      bytes += 8 * (t.pt.dst_points + t.pt.src_points * ntl);
    }
  }
  if (ntransfers)
    *ntransfers = n;
  if (dst_points)
    *dst_points = pts;
  if (alg_bytes)
    *alg_bytes = bytes;
  return CG_OK;
}

} // extern "C"
