// cg_batch.cpp -- batched level step (replaces the per-region walk of
// ggf::transfer_from_all, CarpetLib/src/ggf.cc:500-591).
//
// The caller adds every (region, variable) transfer of a level step once.
// Transfers are resolved and validated on the host exactly as
// cg_transfer_from would (same plans, same error codes).  "plain" runs them
// as one launch each, the shape of the reference's per-region operator
// calls.  "batched" runs each (phase, kernel class) as one launch over a
// device-resident descriptor list (cg_batch_kernels.hip).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <new>
#include <sstream>
#include <vector>

#include "cg_batch.hpp"
#include "cg_internal.hpp"

using namespace cg;

// the arguments of one cg_batch_add, kept so that a transfer can be
// resolved again with a new target time or new storage
struct BatchAddArgs {
  cg_gdata dst;
  std::vector<cg_gdata> srcs;
  std::vector<double> times;
  int dreg[9], sreg[9];
  bool has_slab;
  cg_islab slab;
  double time;
  int op, cent, order_space, order_time;
};

struct cg_batch {
  std::vector<Transfer> transfers;
  std::vector<int> phase;
  std::vector<int> prec;
  std::vector<int> cls_ntl; // time levels of each transfer's launch class
  std::vector<BatchAddArgs> args;
  cg::BatchDevice dev;
  cg::TuneMap tuned;
  // how the launches of a phase run (1 side by side, 0 one after the
  // other): set per phase by the caller, chosen per phase (key: its launch
  // classes) by the tuner; otherwise side by side exactly when every launch
  // of the phase runs one workgroup per item
  std::map<int, int> phase_mode;
  std::map<std::string, int> tuned_side;
  int tile_cfg = -1;
  bool finalized = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // per-step parameter path: launches whose descriptors changed since the
  // last upload, and a pinned staging image of the descriptor blob
  std::vector<int> launch_of; // transfer -> launch index
  std::vector<char> dirty;    // per launch
  char *stage = nullptr;
  size_t stage_bytes = 0;
  hipEvent_t stage_done = nullptr; // the last upload from stage has completed
  bool stage_used = false;
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
  if (b->stage_done) {
    (void)hipEventSynchronize(b->stage_done);
    (void)hipEventDestroy(b->stage_done);
  }
  if (b->stage)
    (void)hipHostFree(b->stage);
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
  if (prec != CG_PREC_F64 && prec != CG_PREC_MIXED && prec != CG_PREC_MIXED_DOT2)
    return set_error(CG_ERR_ARG, "unknown precision %d", prec);
  Transfer tr{};
  int e = resolve_transfer(dst, srcs, times, ntimes, dstregbbox, srcregbbox, slab,
                           time, transport_operator, centering, order_space,
                           order_time, tr, prec == CG_PREC_F64 ? 8 : 4);
  if (e)
    return e;
  if (tr.tmp_ntl)
    return set_error(CG_ERR_ORDER, "prolongation with %d time levels runs per call (cg_transfer_from), "
                                   "not in a batch", tr.tmp_ntl);
  BatchAddArgs a{};
  a.dst = *dst;
  a.srcs.reserve(ntimes);
  for (int l = 0; l < ntimes; ++l)
    a.srcs.push_back(*srcs[l]);
  a.times.assign(times, times + ntimes);
  std::memcpy(a.dreg, dstregbbox, sizeof a.dreg);
  std::memcpy(a.sreg, srcregbbox, sizeof a.sreg);
  a.has_slab = slab != nullptr;
  if (slab)
    a.slab = *slab;
  a.time = time;
  a.op = transport_operator;
  a.cent = centering;
  a.order_space = order_space;
  a.order_time = order_time;
  b->transfers.push_back(tr);
  b->phase.push_back(phase);
  b->prec.push_back(prec);
  b->cls_ntl.push_back(tr.is_prolong ? tr.pr.ntl : 1);
  b->args.push_back(std::move(a));
  return (int)b->transfers.size() - 1;
}

// the launch classes of the phase group [i, j)
static std::string group_key(const BatchDevice &dev, size_t i, size_t j) {
  std::string k;
  for (size_t m = i; m < j; ++m)
    k += (m > i ? " + " : "") + dev.launches[m].key;
  return k;
}

// A persistent launch sizes its grid to the resident capacity of the whole
// GPU; beside another launch its workgroups wait for CUs the other holds,
// and the phase takes longer than the two launches one after the other (the
// finest-level prolongation of the 2000-box hierarchy step: 7.16 ms side by
// side against 3.65 + 2.13 ms in turn).  Launches of one workgroup per item
// fill each other's tails.
static void apply_sides(cg_batch *b) {
  BatchDevice &dev = b->dev;
  dev.beside.assign(dev.launches.size(), 1);
  for (size_t i = 0; i < dev.launches.size();) {
    const size_t j = batch_phase_end(dev, i);
    if (j - i > 1) {
      int side = 1;
      for (size_t m = i; m < j; ++m)
        side = side && dev.launches[m].wpc == -1;
      const auto t = b->tuned_side.find(group_key(dev, i, j));
      if (t != b->tuned_side.end())
        side = t->second;
      const auto f = b->phase_mode.find(dev.launches[i].phase);
      if (f != b->phase_mode.end())
        side = f->second;
      dev.beside[i] = (char)side;
    }
    i = j;
  }
}

// transfer -> launch map after a (re)build; nothing is pending after it
static void index_launches(cg_batch *b) {
  b->launch_of.assign(b->transfers.size(), -1);
  for (size_t l = 0; l < b->dev.launches.size(); ++l)
    for (int m : b->dev.launches[l].members)
      b->launch_of[m] = (int)l;
  b->dirty.assign(b->dev.launches.size(), 0);
  apply_sides(b);
}

int cg_batch_finalize(cg_batch *b, int tile_cfg) {
  if (!b)
    return set_error(CG_ERR_ARG, "null batch");
  if (b->finalized)
    return set_error(CG_ERR_STATE, "batch already finalized");
  b->tile_cfg = tile_cfg;
  int e = batch_device_build(b->dev, b->transfers, b->phase, b->prec, b->cls_ntl, tile_cfg, &b->tuned);
  if (e)
    return e;
  index_launches(b);
  b->finalized = true;
  return CG_OK;
}

// ---- autotuner --------------------------------------------------------------
static void drop_graph(cg_batch *b) {
  if (b->exec)
    (void)hipGraphExecDestroy(b->exec);
  if (b->graph)
    (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

static void load_cache(const char *path, TuneMap &m, std::map<std::string, int> &sides) {
  if (!path)
    return;
  std::ifstream f(path);
  std::string line;
  while (std::getline(f, line)) {
    const size_t bar = line.rfind('|');
    if (line.empty() || line[0] == '#' || bar == std::string::npos)
      continue;
    if (line.compare(0, 6, "phase ") == 0) {
      int side = 1;
      if (sscanf(line.c_str() + bar + 1, "%d", &side) == 1)
        sides[line.substr(6, bar - 6)] = side ? 1 : 0;
      continue;
    }
    int z = 0, p = 0, w = 0, sh = -1, ni = 1;
    if (sscanf(line.c_str() + bar + 1, "%d %d %d %d %d", &z, &p, &w, &sh, &ni) >= 2)
      m[line.substr(0, bar)] = Tune{z, p, w, sh, std::max(1, ni)};
  }
}

static int save_cache(const char *path, const TuneMap &m, const std::map<std::string, int> &sides) {
  if (!path)
    return CG_OK;
  std::ofstream f(path);
  if (!f)
    return set_error(CG_ERR_ARG, "cannot write tuner cache %s", path);
  f << "# carpet_amd batch tuner cache: <class key>|<zchunk> <points per item> "
       "<workgroups per CU: 0 resident capacity, -1 one per item> "
       "<tile shape: 0 wide (32x8 coarse bases, 64x8 in the FP32-storage FP64 mode), 1 16x16, "
       "2 8x8 (FP64 regions thin in x and y or z, fixed), 3 4x16 (FP64 regions thin in x, fixed), "
       "-1 by region size> <256-point items per workgroup in grid mode>; "
       "phase <class key> + <class key> ...|<1 side by side, 0 one after the other>\n";
  for (const auto &kv : m)
    f << kv.first << "|" << kv.second.zchunk << " " << kv.second.pt << " " << kv.second.wpc << " "
      << kv.second.shape << " " << kv.second.ni << "\n";
  for (const auto &kv : sides)
    f << "phase " << kv.first << "|" << kv.second << "\n";
  return CG_OK;
}

static int time_launch(BatchDevice &dev, int idx, hipStream_t st, int reps, float *ms) {
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return hip_error(hipGetLastError(), "hipEventCreate");
  int e = batch_device_run_one(dev, idx, st); // warm-up
  float best = 1e30f;
  for (int r = 0; r < reps && !e; ++r) {
    (void)hipEventRecord(e0, st);
    e = batch_device_run_one(dev, idx, st);
    (void)hipEventRecord(e1, st);
    if (!e && hipEventSynchronize(e1) == hipSuccess) {
      float t = 0;
      (void)hipEventElapsedTime(&t, e0, e1);
      best = std::min(best, t);
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = best;
  return e;
}

static int time_phase(BatchDevice &dev, size_t i, size_t j, bool side, hipStream_t st, int reps,
                      float *ms) {
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return hip_error(hipGetLastError(), "hipEventCreate");
  int e = batch_device_run_phase(dev, i, j, side, st); // warm-up
  float best = 1e30f;
  for (int r = 0; r < reps && !e; ++r) {
    (void)hipEventRecord(e0, st);
    e = batch_device_run_phase(dev, i, j, side, st);
    (void)hipEventRecord(e1, st);
    if (!e && hipEventSynchronize(e1) == hipSuccess) {
      float t = 0;
      (void)hipEventElapsedTime(&t, e0, e1);
      best = std::min(best, t);
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = best;
  return e;
}

extern "C" int cg_batch_autotune(cg_batch *b, cg_stream_t stream, int reps,
                                 const char *cache_path) {
  if (!b)
    return set_error(CG_ERR_ARG, "null batch");
  if (!b->finalized)
    return set_error(CG_ERR_STATE, "batch not finalized");
  hipStream_t st = static_cast<hipStream_t>(stream);
  reps = std::max(1, reps);
  // classes named in the cache take its parameters, every other class is
  // searched (the launches run on the batch's own arrays)
  b->tuned.clear();
  b->tuned_side.clear();
  load_cache(cache_path, b->tuned, b->tuned_side);
  // CG_TUNE_LOG set: every trial's time on stderr
  const bool tune_log = std::getenv("CG_TUNE_LOG") != nullptr;
  drop_graph(b);
  // current choices for every class, then search the classes the cache lacks
  for (size_t i = 0; i < b->dev.launches.size(); ++i) {
    const BatchLaunch l = b->dev.launches[i];
    if (b->tuned.count(l.key))
      continue;
    // item size first (in the class's default mode), then the other modes
    std::vector<Tune> cands;
    if (l.is_tensor)
      for (int z : {8, 16, 32, 64, 128, 0})
        cands.push_back({z, l.pt, l.wpc, l.shape});
    else
      for (int p : {256, 512, 1024, 2048, 4096, 16384})
        cands.push_back({l.zchunk, p, l.wpc, -1, l.ni});
    float best = 1e30f;
    Tune bestc = cands[0];
    // each trial builds this launch alone into its own descriptor blob
    // (launch i of the full build; the other launches are not rebuilt)
    auto trial_run = [&](const Tune &c) -> int {
      TuneMap trial = b->tuned;
      trial[l.key] = c;
      BatchDevice one;
      int e = batch_device_build(one, b->transfers, b->phase, b->prec, b->cls_ntl, b->tile_cfg, &trial,
                                 (int)i);
      float ms = 0;
      if (!e && one.launches.size() != 1)
        e = set_error(CG_ERR_STATE, "tuner: launch %zu did not build alone", i);
      if (!e)
        e = time_launch(one, 0, st, reps, &ms);
      if (!e) {
        hipError_t he = hipStreamSynchronize(st);
        if (he != hipSuccess)
          e = hip_error(he, "tuner trial");
      }
      batch_device_free(one);
      if (e)
        return e;
      if (tune_log)
        fprintf(stderr, "tune [%s] zchunk %d pt %d wpc %d shape %d ni %d: %.4f ms\n", l.key.c_str(), c.zchunk,
                c.pt, c.wpc, c.shape, c.ni, ms);
      if (ms < best) {
        best = ms;
        bestc = c;
      }
      return CG_OK;
    };
    for (const Tune &c : cands)
      if (int e = trial_run(c))
        return e;
    const Tune base = bestc;
    for (int w : {-1, 0, 2, -2})
      if (w != base.wpc && (w != -2 || l.is_tensor))
        if (int e = trial_run(Tune{base.zchunk, base.pt, w, base.shape, base.ni}))
          return e;
    if (l.is_tensor && l.shape < 2) // the other large tile shape, in the best mode
      if (int e = trial_run(Tune{bestc.zchunk, bestc.pt, bestc.wpc, 1 - l.shape, 1}))
        return e;
    // 256-point items in grid mode: items per workgroup
    if (!l.is_tensor && l.op != PO_TVD && bestc.pt == 256 && bestc.wpc == -1) {
      const Tune b2 = bestc;
      for (int ni : {1, 2, 4, 8, 16})
        if (ni != b2.ni)
          if (int e = trial_run(Tune{b2.zchunk, b2.pt, b2.wpc, b2.shape, ni}))
            return e;
    }
    b->tuned[l.key] = bestc;
  }
  int e = batch_device_build(b->dev, b->transfers, b->phase, b->prec, b->cls_ntl, b->tile_cfg, &b->tuned);
  if (e)
    return e;
  index_launches(b);
  // phases of several launches: side by side or one after the other
  for (size_t i = 0; i < b->dev.launches.size();) {
    const size_t j = batch_phase_end(b->dev, i);
    const std::string k = group_key(b->dev, i, j);
    if (j - i > 1 && !b->tuned_side.count(k)) {
      float ms[2] = {0, 0};
      for (int side = 0; side < 2; ++side)
        if (int e2 = time_phase(b->dev, i, j, side != 0, st, reps, &ms[side]))
          return e2;
      hipError_t he = hipStreamSynchronize(st);
      if (he != hipSuccess)
        return hip_error(he, "tuner phase trial");
      b->tuned_side[k] = ms[1] <= ms[0] ? 1 : 0;
      if (tune_log)
        fprintf(stderr, "tune phase [%s]: in turn %.4f ms, side by side %.4f ms\n", k.c_str(), ms[0], ms[1]);
    }
    i = j;
  }
  apply_sides(b);
  return save_cache(cache_path, b->tuned, b->tuned_side);
}

extern "C" int cg_batch_describe(const cg_batch *b, char *buf, size_t len) {
  if (!b || !buf || !len)
    return set_error(CG_ERR_ARG, "null argument");
  std::ostringstream o;
  long long items = 0;
  for (const BatchLaunch &l : b->dev.launches)
    items += l.nitems;
  o << "transfers " << b->transfers.size() << " launches " << b->dev.launches.size()
    << " work items " << items << " device descriptor bytes " << b->dev.blob_bytes << "\n";
  for (size_t i = 0; i < b->dev.launches.size();) {
    const size_t j = batch_phase_end(b->dev, i);
    for (size_t m = i; m < j; ++m) {
      const BatchLaunch &l = b->dev.launches[m];
      o << "phase " << l.phase << " [" << l.key << "] items " << l.nitems << " grid " << l.grid
        << " zchunk " << l.zchunk << " pt " << l.pt << " wpc " << l.wpc << " shape " << l.shape
        << " ni " << l.ni;
      if (j - i > 1)
        o << (b->dev.beside.size() != b->dev.launches.size() || b->dev.beside[i] ? " side by side"
                                                                                 : " in turn");
      o << "\n";
    }
    i = j;
  }
  const std::string s = o.str();
  const size_t n = std::min(len - 1, s.size());
  std::memcpy(buf, s.data(), n);
  buf[n] = 0;
  return (int)s.size();
}

// ---- per-step parameters ----------------------------------------------------
// A transfer resolved again (new target time or storage) must run in the
// launch class it was finalized in: the same operator and stencil, and for
// tensor classes the class's time levels or one (an exact time match, run
// with nlev = 1); pointwise prolongations combine their levels at run time.
static int same_class(const Transfer &old, const Transfer &nu, int cls_ntl, int idx) {
  if (nu.tmp_ntl)
    return set_error(CG_ERR_STATE, "transfer %d would need time-level temporaries, which a batch does not run",
                     idx);
  if (old.is_prolong != nu.is_prolong)
    return set_error(CG_ERR_STATE, "transfer %d changes its operator; build a new batch", idx);
  if (old.is_prolong) {
    if (old.pr.kind != nu.pr.kind || old.pr.order != nu.pr.order)
      return set_error(CG_ERR_STATE, "transfer %d changes its stencil; build a new batch", idx);
    const bool tensor = nu.pr.kind != PK_TVD && nu.pr.kind != PK_MONO5 && nu.pr.kind != PK_STAGGER;
    if (tensor && nu.pr.ntl != cls_ntl && nu.pr.ntl != 1)
      return set_error(CG_ERR_STATE,
                       "transfer %d now combines %d time levels; its launch was built for %d (or an exact "
                       "time match); build a new batch",
                       idx, nu.pr.ntl, cls_ntl);
    return CG_OK;
  }
  if (old.pt.op == nu.pt.op || (old.pt.op == PO_INTERP && nu.pt.op == PO_COPY))
    return CG_OK;
  return set_error(CG_ERR_STATE,
                   "transfer %d changes from %s to %s; build a new batch", idx,
                   old.pt.op == PO_COPY ? "a copy" : "its operator",
                   nu.pt.op == PO_INTERP ? "a time interpolation" : "another operator");
}

// resolves the transfers idx[] with their (updated) arguments; all or nothing
static int reresolve(cg_batch *b, const std::vector<int> &idx, const std::vector<BatchAddArgs> &na) {
  std::vector<Transfer> nt(idx.size());
  for (size_t i = 0; i < idx.size(); ++i) {
    const BatchAddArgs &a = na[i];
    std::vector<const cg_gdata *> sp(a.srcs.size());
    for (size_t l = 0; l < a.srcs.size(); ++l)
      sp[l] = &a.srcs[l];
    const int t = idx[i];
    int e = resolve_transfer(&a.dst, sp.data(), a.times.data(), (int)a.times.size(), a.dreg, a.sreg,
                             a.has_slab ? &a.slab : nullptr, a.time, a.op, a.cent, a.order_space,
                             a.order_time, nt[i], b->prec[t] == CG_PREC_F64 ? 8 : 4);
    if (e)
      return e;
    if ((e = same_class(b->transfers[t], nt[i], b->cls_ntl[t], t)))
      return e;
  }
  for (size_t i = 0; i < idx.size(); ++i) {
    const int t = idx[i];
    b->transfers[t] = nt[i];
    b->args[t] = na[i];
    if (b->finalized && b->launch_of[t] >= 0)
      b->dirty[b->launch_of[t]] = 1;
  }
  return CG_OK;
}

// uploads the descriptors of the launches changed since the last upload, in
// stream order (a pinned staging image, reused once its last copy is done)
static int upload_pending(cg_batch *b, hipStream_t st) {
  bool any = false;
  for (char d : b->dirty)
    any = any || d;
  if (!any)
    return CG_OK;
  hipError_t he;
  if (b->stage_bytes < b->dev.blob_bytes) {
    if (b->stage) {
      (void)hipEventSynchronize(b->stage_done);
      (void)hipHostFree(b->stage);
      b->stage = nullptr;
      b->stage_bytes = 0;
    }
    he = hipHostMalloc(reinterpret_cast<void **>(&b->stage), b->dev.blob_bytes, hipHostMallocDefault);
    if (he != hipSuccess)
      return set_error(CG_ERR_NOMEM, "pinned staging of the batch descriptors: %s", hipGetErrorString(he));
    b->stage_bytes = b->dev.blob_bytes;
  }
  if (!b->stage_done) {
    he = hipEventCreateWithFlags(&b->stage_done, hipEventDisableTiming);
    if (he != hipSuccess)
      return hip_error(he, "hipEventCreate");
  }
  if (b->stage_used) {
    he = hipEventSynchronize(b->stage_done);
    if (he != hipSuccess)
      return hip_error(he, "hipEventSynchronize (descriptor staging)");
  }
  for (size_t l = 0; l < b->dev.launches.size(); ++l) {
    if (!b->dirty[l])
      continue;
    const BatchLaunch &L = b->dev.launches[l];
    batch_launch_descs(L, b->transfers, b->stage + L.desc_off);
    he = hipMemcpyAsync(b->dev.blob + L.desc_off, b->stage + L.desc_off, L.desc_bytes, hipMemcpyHostToDevice,
                        st);
    if (he != hipSuccess)
      return hip_error(he, "descriptor upload");
    b->dirty[l] = 0;
  }
  he = hipEventRecord(b->stage_done, st);
  if (he != hipSuccess)
    return hip_error(he, "hipEventRecord (descriptor staging)");
  b->stage_used = true;
  return CG_OK;
}

int cg_batch_set_phase_mode(cg_batch *b, int phase, int mode) {
  if (!b)
    return set_error(CG_ERR_ARG, "null batch");
  if (mode < -1 || mode > 1)
    return set_error(CG_ERR_ARG, "phase mode %d (1 side by side, 0 one after the other, -1 automatic)", mode);
  if (mode < 0)
    b->phase_mode.erase(phase);
  else
    b->phase_mode[phase] = mode;
  drop_graph(b);
  if (b->finalized)
    apply_sides(b);
  return CG_OK;
}

int cg_batch_set_time(cg_batch *b, int phase, const double *times, int ntimes, double time) {
  if (!b || !times || ntimes <= 0)
    return set_error(CG_ERR_ARG, "null argument");
  std::vector<int> idx;
  std::vector<BatchAddArgs> na;
  for (size_t t = 0; t < b->transfers.size(); ++t)
    if (b->phase[t] == phase && (int)b->args[t].times.size() == ntimes) {
      idx.push_back((int)t);
      na.push_back(b->args[t]);
      na.back().times.assign(times, times + ntimes);
      na.back().time = time;
    }
  int e = reresolve(b, idx, na);
  return e ? e : (int)idx.size();
}

int cg_batch_set_storage(cg_batch *b, int n, const cg_gdata *from, const cg_gdata *to) {
  if (!b || n < 0 || (n > 0 && (!from || !to)))
    return set_error(CG_ERR_ARG, "null argument");
  std::map<const void *, int> map;
  for (int i = 0; i < n; ++i) {
    if (std::memcmp(from[i].padext, to[i].padext, sizeof from[i].padext) ||
        std::memcmp(from[i].ext, to[i].ext, sizeof from[i].ext) ||
        std::memcmp(from[i].bbox, to[i].bbox, sizeof from[i].bbox))
      return set_error(CG_ERR_ARG, "storage pair %d: the arrays have different layouts", i);
    if (!map.emplace(from[i].storage, i).second)
      return set_error(CG_ERR_ARG, "storage pair %d: the same storage appears twice", i);
  }
  std::vector<int> idx;
  std::vector<BatchAddArgs> na;
  for (size_t t = 0; t < b->transfers.size(); ++t) {
    BatchAddArgs a = b->args[t];
    bool hit = false;
    auto swap = [&](cg_gdata &g) -> int {
      auto it = map.find(g.storage);
      if (it == map.end())
        return CG_OK;
      const cg_gdata &f = from[it->second];
      if (std::memcmp(g.padext, f.padext, sizeof g.padext) || std::memcmp(g.bbox, f.bbox, sizeof g.bbox))
        return set_error(CG_ERR_ARG, "storage pair %d: transfer %zu holds that storage with another layout",
                         it->second, t);
      g = to[it->second];
      hit = true;
      return CG_OK;
    };
    int e = swap(a.dst);
    for (size_t l = 0; l < a.srcs.size() && !e; ++l)
      e = swap(a.srcs[l]);
    if (e)
      return e;
    if (hit) {
      idx.push_back((int)t);
      na.push_back(std::move(a));
    }
  }
  int e = reresolve(b, idx, na);
  return e ? e : (int)idx.size();
}

int cg_batch_run(cg_batch *b, cg_stream_t stream) {
  if (!b)
    return set_error(CG_ERR_ARG, "null batch");
  if (!b->finalized)
    return set_error(CG_ERR_STATE, "batch not finalized");
  if (int e = upload_pending(b, static_cast<hipStream_t>(stream)))
    return e;
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
  // pending descriptor changes go up before (not into) the captured step
  if (int e = upload_pending(b, st))
    return e;
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
  if (int e = upload_pending(b, static_cast<hipStream_t>(stream)))
    return e;
  hipError_t he = hipGraphLaunch(b->exec, static_cast<hipStream_t>(stream));
  return he == hipSuccess ? CG_OK : hip_error(he, "hipGraphLaunch");
}

// destination points and algorithmic bytes of transfer i: element size x
// (destination points + unique source footprint per time level)
static void transfer_bytes(const cg_batch *b, size_t i, long long *pts, long long *bytes) {
  const Transfer &t = b->transfers[i];
  const long long es = b->prec[i] == CG_PREC_F64 ? 8 : 4;
  if (t.is_prolong) {
    *pts = t.pr.dst_points;
    *bytes = es * (t.pr.dst_points + t.pr.src_points * t.pr.ntl);
  } else {
    const int ntl = t.pt.op == PO_INTERP ? t.pt.a.ntl : 1;
    *pts = t.pt.dst_points;
    *bytes = 8 * (t.pt.dst_points + t.pt.src_points * ntl);
  }
}

int cg_batch_launch_count(const cg_batch *b) {
  if (!b)
    return set_error(CG_ERR_ARG, "null batch");
  if (!b->finalized)
    return set_error(CG_ERR_STATE, "batch not finalized");
  return (int)b->dev.launches.size();
}

int cg_batch_run_launch(cg_batch *b, int idx, cg_stream_t stream) {
  if (!b)
    return set_error(CG_ERR_ARG, "null batch");
  if (!b->finalized)
    return set_error(CG_ERR_STATE, "batch not finalized");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (int e = upload_pending(b, st))
    return e;
  return batch_device_run_one(b->dev, idx, st);
}

int cg_batch_launch_stats(const cg_batch *b, int idx, int *phase, long long *dst_points,
                          long long *alg_bytes) {
  if (!b)
    return set_error(CG_ERR_ARG, "null batch");
  if (!b->finalized)
    return set_error(CG_ERR_STATE, "batch not finalized");
  if (idx < 0 || idx >= (int)b->dev.launches.size())
    return set_error(CG_ERR_ARG, "no launch %d", idx);
  const BatchLaunch &l = b->dev.launches[idx];
  long long pts = 0, bytes = 0;
  for (int m : l.members) {
    long long p, q;
    transfer_bytes(b, (size_t)m, &p, &q);
    pts += p;
    bytes += q;
  }
  if (phase)
    *phase = l.phase;
  if (dst_points)
    *dst_points = pts;
  if (alg_bytes)
    *alg_bytes = bytes;
  return CG_OK;
}

int cg_batch_stats(const cg_batch *b, long long *ntransfers, long long *dst_points,
                   long long *alg_bytes) {
  if (!b)
    return set_error(CG_ERR_ARG, "null batch");
  long long n = 0, pts = 0, bytes = 0;
  for (size_t i = 0; i < b->transfers.size(); ++i) {
    long long p, q;
    transfer_bytes(b, i, &p, &q);
    ++n;
    pts += p;
    bytes += q;
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
