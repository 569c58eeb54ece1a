// cg_batch.hpp -- device side of the batched level step: descriptor lists
// and persistent launches (built in cg_batch_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <utility>
#include <vector>

#include "cg_internal.hpp"

namespace cg {

// One persistent launch: the transfers of one phase that share a kernel
// (tensor kind/order/time levels or one pointwise operator).
struct BatchLaunch {
  int phase;
  int is_tensor;
  int kind, order, ntl, op, prec;
  size_t desc_off; // byte offset of the descriptor array in the blob
  size_t item_off; // byte offset of the (descriptor, tile) item array
  size_t seg_off;  // byte offset of the 9 segment bounds (one segment per XCD)
  int nitems;
  int counter;     // index of this launch's block of 8 work counters
  int grid;        // persistent workgroups
  int zchunk;      // tensor work item: coarse z bases per chunk (0 = all)
  int pt;          // pointwise work item: destination points
  int wpc;         // workgroups per CU of a persistent grid taking items
                   // from counters (0 = resident capacity); -1 one
                   // workgroup per item; -2 resident capacity taking items
                   // in fixed rounds (tensor tiles)
  int shape;       // tensor tile shape (0: 32 x 8 bases, 1: 16 x 16)
  int ni;          // grid mode, 256-point pointwise items: items per workgroup
  std::string key; // tuner cache key (class + box-size histogram bucket)
  std::vector<int> members; // transfer indices in descriptor order
  size_t desc_bytes;        // bytes of the descriptor array
};

// one work item: a tile (tensor) or a run of points (pointwise) of one
// transfer's descriptor
struct Item {
  unsigned desc;
  unsigned tile;
};

// batched tensor launches per line kind (LK_*) and tile shape (0 32x8 wide,
// 1 16x16 narrow, 2 8x8 small and 3 4x16 slim, FP64 only); defined in
// cg_batch_tensor.hpp, instantiated in cg_batch_{vlag,clag,eno,veno}_{w,n,s,t}.hip
template <int KIND, int SH>
int batch_tensor_kind(const BatchLaunch &l, const TensorArgs *td, const Item *items,
                      const int *seg, int *ctr, hipStream_t st);
template <int KIND, int SH> const void *batch_tensor_fn_kind(const BatchLaunch &l, int *threads);

// tuned parameters of one launch class; wpc 0 = resident capacity;
// shape -1 = chosen from the region sizes
struct Tune {
  int zchunk, pt, wpc, shape = -1, ni = 1;
};
using TuneMap = std::map<std::string, Tune>;

struct BatchDevice {
  std::vector<BatchLaunch> launches;
  char *blob = nullptr;   // descriptors + items (device)
  size_t blob_bytes = 0;
  int *counters = nullptr; // one work counter per launch (device)
  int ncounters = 0;
  // launches of one phase after the first run on helper streams forked from
  // the caller's stream and joined back (graph branches under capture)
  std::vector<hipStream_t> aux;
  std::vector<hipEvent_t> fork, join;
  // per launch, read at the first launch of each phase: 1 = the phase's
  // launches run side by side, 0 = one after the other on the caller's
  // stream (empty: side by side)
  std::vector<char> beside;
};

// tile_cfg selects the z chunk of tensor tiles in coarse bases (0 = whole
// region, < 0 = automatic); cls_ntl: per transfer the time levels of its
// launch class (a tensor transfer resolved to fewer levels -- an exact time
// match -- runs in its class with nlev = 1); only >= 0 builds launch `only`
// of the full build alone (the tuner's trials)
int batch_device_build(BatchDevice &dev, const std::vector<Transfer> &tr,
                       const std::vector<int> &phase, const std::vector<int> &prec,
                       const std::vector<int> &cls_ntl, int tile_cfg,
                       const TuneMap *tuned = nullptr, int only = -1);
// the descriptor array of launch l from the current transfers, in the
// layout batch_device_build uploads (the per-step parameter path rewrites
// it in place: same items, new time weights and array pointers)
void batch_launch_descs(const BatchLaunch &l, const std::vector<Transfer> &tr, char *out);
int batch_device_run(BatchDevice &dev, hipStream_t st);
// end of the phase group that starts at launch i (launches of one phase are
// consecutive)
size_t batch_phase_end(const BatchDevice &dev, size_t i);
// runs the launches [i, j) of one phase, side by side or one after the
// other, their work counters cleared first (the tuner's phase trials)
int batch_device_run_phase(BatchDevice &dev, size_t i, size_t j, bool side, hipStream_t st);
// runs launch idx alone (its counter cleared first)
int batch_device_run_one(BatchDevice &dev, int idx, hipStream_t st);
void batch_device_free(BatchDevice &dev);

} // namespace cg
