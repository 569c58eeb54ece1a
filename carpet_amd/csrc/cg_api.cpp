// cg_api.cpp -- extern "C" entry points of libcarpet_amd.so (per-call
// operators, transfer dispatch, device storage) and the transfer dispatch
// itself.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "cg_internal.hpp"

namespace cg {

const char *last_error_message();

static hipStream_t S(cg_stream_t s) { return static_cast<hipStream_t>(s); }

// staggered directions of op STAGGER011..111 (bit d = direction d), else 0
static int stagger_dirs(int op) {
  switch (op) {
  case CG_OP_STAGGER011: return CG_STAGGER011;
  case CG_OP_STAGGER101: return CG_STAGGER101;
  case CG_OP_STAGGER110: return CG_STAGGER110;
  case CG_OP_STAGGER111: return CG_STAGGER111;
  }
  return 0;
}

// region-sized, unpadded temporary over the transfer's destination region
static Arr tmp_arr(const Transfer &tr, const void *p) {
  Arr a;
  a.p = p;
  for (int k = 0; k < 3; ++k)
    a.pad[k] = a.ext[k] = tr.tmp_reg.sizes(k);
  a.box = tr.tmp_reg;
  return a;
}

// one time level of a transfer through temporaries, written to out
static int plan_level(const Transfer &tr, const void *src, const Arr &out, Transfer &one) {
  std::memset(&one, 0, sizeof one);
  Arr sa = tr.tmp_src;
  sa.p = src;
  switch (tr.tmp_step) {
  case 0: {
    const void *s1[1] = {src};
    const double w[1] = {1.0};
    one.is_prolong = true;
    return plan_prolong(tr.tmp_kind, tr.tmp_order, sa, s1, 1, w, out, tr.tmp_reg, one.pr, 8, nullptr,
                        tr.tmp_stag);
  }
  case 1:
    return plan_copy(sa, out, tr.tmp_sreg, tr.tmp_reg, tr.tmp_has_slab ? &tr.tmp_slab : nullptr, one.pt);
  case 2:
    if (tr.tmp_vc)
      return plan_restrict_vc(tr.tmp_vc & 7, sa, out, tr.tmp_reg, one.pt);
    return tr.tmp_stag ? plan_restrict_stagger(tr.tmp_stag, sa, out, tr.tmp_reg, one.pt)
                       : plan_restrict(tr.tmp_cell != 0, sa, out, tr.tmp_reg, one.pt, tr.tmp_rorder);
  }
  return set_error(CG_ERR_ARG, "unknown transfer step");
}

// completes a transfer through temporaries and validates it with a
// stand-in temporary (its address is never dereferenced)
static int setup_tmp(Transfer &tr, int ntl, const void *const *sp, const double *tw, double time,
                     const Arr &s0, const Arr &d, const Box &sreg, const Box &dreg) {
  tr.tmp_src = s0;
  tr.tmp_dst = d;
  tr.tmp_sreg = sreg;
  tr.tmp_reg = dreg;
  for (int l = 0; l < ntl; ++l) {
    tr.tmp_srcs[l] = sp[l];
    tr.tmp_times[l] = tw[l];
  }
  tr.tmp_time = time;
  const Arr stand_in = tmp_arr(tr, s0.p);
  Transfer one;
  int e;
  if ((e = plan_level(tr, sp[0], stand_in, one)))
    return e;
  PointPlan ip;
  if ((e = plan_interp(ntl, sp, tw, time, stand_in, d, dreg, ip, tr.tmp_eno)))
    return e;
  tr.tmp_ntl = ntl;
  tr.is_prolong = tr.tmp_step == 0;
  tr.pr.dst_points = tr.pt.dst_points = dreg.size();
  return CG_OK;
}

// data<T>::transfer_time -> transfer_p_r -> transfer_prolongate /
// transfer_restrict / copy_from_innerloop, with the time-level choice of
// gdata::find_source_timelevel (CarpetLib/src/data.cc:485-556, 572-966,
// 976-1171, 1182-1299; CarpetLib/src/gdata.cc:288-349)
int resolve_transfer(const cg_gdata *dst, const cg_gdata *const *srcs,
                     const double *times, int ntimes, const int *dstregbbox,
                     const int *srcregbbox, const cg_islab *slab, double time,
                     int op, int cent, int order_space, int order_time,
                     Transfer &tr, int es) {
  if (!dst || !srcs || !times || ntimes <= 0 || !dstregbbox || !srcregbbox)
    return set_error(CG_ERR_ARG, "null argument");
  for (int l = 0; l < ntimes; ++l)
    if (!srcs[l])
      return set_error(CG_ERR_ARG, "null source time level %d", l);
  if (op == CG_OP_ERROR || op == CG_OP_NONE)
    return set_error(CG_ERR_OPERATOR, "transport operator %d does not transfer", op);
  if (cent != CG_VERTEX_CENTERED && cent != CG_CELL_CENTERED)
    return set_error(CG_ERR_OPERATOR, "unknown centering %d", cent);
  std::memset(&tr, 0, sizeof tr);
  int tl0, ntl, e;
  if ((e = find_source_timelevel(times, ntimes, time, order_time, op, &tl0, &ntl)))
    return e;
  if (ntl > ntimes)
    return set_error(CG_ERR_ARG, "%d time levels needed, %d given", ntl, ntimes);
  const Arr d = arr_from(*dst);
  const Arr s0 = arr_from(*srcs[tl0]);
  for (int l = 0; l < ntl; ++l) {
    const cg_gdata &g = *srcs[tl0 + l];
    if (std::memcmp(g.bbox, srcs[tl0]->bbox, sizeof g.bbox) ||
        std::memcmp(g.padext, srcs[tl0]->padext, sizeof g.padext))
      return set_error(CG_ERR_ARG, "source time levels have different layouts");
  }
  const Box dreg = Box::from(dstregbbox);
  const Box sreg = Box::from(srcregbbox);
  const void *sp[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  for (int l = 0; l < ntl && l < 5; ++l)
    sp[l] = srcs[tl0 + l]->storage;
  // ENO, WENO, TVD and monotone groups interpolate 3 time levels with
  // interpolate_eno_3d_3tl (data.cc:1260-1292)
  const bool eno_group = op == CG_OP_ENO || op == CG_OP_WENO || op == CG_OP_TVD || op == CG_OP_LAGRANGE_MONOTONE;
  // 4 and 5 levels for Lagrange-type groups only: the ENO/WENO/TVD/
  // monotone branch of data<T>::time_interpolate asserts beyond order_time
  // 2 (data.cc:1260-1295)
  if (ntl > 5 || (ntl > 3 && eno_group))
    return set_error(CG_ERR_ORDER, "time interpolation order %d not supported", order_time);

  if (es != 8 && es != 4)
    return set_error(CG_ERR_ARG, "element size must be 8 or 4");
  bool same = true, coarser = true, finer = true;
  for (int k = 0; k < 3; ++k) {
    same &= s0.box.str[k] == d.box.str[k];
    coarser &= s0.box.str[k] > d.box.str[k];
    finer &= s0.box.str[k] < d.box.str[k];
  }
  if (es != 8 && !coarser)
    return set_error(CG_ERR_OPERATOR, "mixed precision applies to prolongation only; copies, "
                                      "restriction and time interpolation stay FP64");
  if (same) {
    if (ntl == 1) {
      tr.is_prolong = false;
      return plan_copy(s0, d, sreg, dreg, slab, tr.pt);
    }
    // copying each level into a temporary and interpolating in time equals
    // interpolating the sources directly (a copy is exact); with a slab
    // offset the copies go through temporaries as in the reference
    if (slab) {
      double tw[5];
      for (int l = 0; l < ntl; ++l)
        tw[l] = times[tl0 + l];
      tr.tmp_step = 1;
      tr.tmp_has_slab = true;
      tr.tmp_slab = *slab;
      tr.tmp_eno = ntl == 3 && eno_group;
      return setup_tmp(tr, ntl, sp, tw, time, s0, d, sreg, dreg);
    }
    // ENO and TVD groups interpolate 3 levels with interpolate_eno_3d_3tl
    // (data.cc:1262-1292)
    tr.is_prolong = false;
    return plan_interp(ntl, sp, times, time, s0, d, dreg, tr.pt,
                       eno_group);
  }
  if (coarser) {
    if (op == CG_OP_SYNC || op == CG_OP_RESTRICT)
      return set_error(CG_ERR_OPERATOR, "operator %d cannot prolongate", op);
    if (slab)
      return set_error(CG_ERR_ARG, "prolongation takes no slab");
    int kind = -1, order = order_space;
    switch (op) {
    case CG_OP_COPY:
    case CG_OP_LAGRANGE:
      if (cent == CG_VERTEX_CENTERED) {
        if (order_space < 0 || order_space > 11 || order_space % 2 == 0)
          return set_error(CG_ERR_ORDER,
                           "There is no vertex-centred stencil for op=\"LAGRANGE\" or "
                           "op=\"COPY\" with order_space not in {1, 3, 5, 7, 9, 11}");
        kind = PK_VLAG;
      } else {
        // use_dgfe selects prolongate_3d_dgfe_rf2<T,5> here (data.cc:617-627)
        if (dgfe_requested())
          return set_error(CG_ERR_OPERATOR, "use_dgfe: prolongate_3d_dgfe_rf2 evaluates HRSCCore's "
                                            "GLLElement<5>, which this library does not provide");
        if (order_space < 0 || order_space > 5)
          return set_error(CG_ERR_ORDER,
                           "There is no cell-centred stencil for op=\"LAGRANGE\" or "
                           "op=\"COPY\" with order_space not in {0, 1, 2, 3, 4, 5}");
        kind = PK_CLAG;
      }
      break;
    case CG_OP_ENO:
      if (cent == CG_VERTEX_CENTERED) {
        // prolongate_3d_eno for order_space 3 and 5 (data.cc:658-690)
        if (order_space == 1)
          return set_error(CG_ERR_ORDER, "There is no stencil for op=\"ENO\" with order_space=1");
        if (order_space != 3 && order_space != 5)
          return set_error(CG_ERR_ORDER, "There is no stencil for op=\"ENO\" with order_space!=3");
        kind = PK_VENO;
        order = 3;
        break;
      }
      if (order_space < 2 || order_space > 5)
        return set_error(CG_ERR_ORDER, "There is no cell-centred stencil for "
                                       "op=\"ENO\" with order_space not in {2,3,4,5}");
      kind = PK_ENO;
      order = order_space == 5 ? 3 : 2; // data.cc:697-704
      break;
    case CG_OP_WENO:
      // vertex-centred WENO runs prolongate_3d_eno for order_space 5
      // (data.cc:727-762)
      if (cent != CG_VERTEX_CENTERED)
        return set_error(CG_ERR_OPERATOR, "There are currently no cell-centred stencils for op=\"WENO\"");
      if (order_space == 1 || order_space == 3)
        return set_error(CG_ERR_ORDER, "There is no stencil for op=\"WENO\" with order_space=%d",
                         order_space);
      if (order_space != 5)
        return set_error(CG_ERR_ORDER, "There is no stencil for op=\"WENO\" with order_space!=5");
      kind = PK_VENO;
      order = 3;
      break;
    case CG_OP_TVD:
      // the reference's vertex-centred TVD (prolongate_3d_real8_tvd.F90)
      // aborts unconditionally
      if (cent == CG_VERTEX_CENTERED)
        return set_error(CG_ERR_OPERATOR, "vertex-centred TVD: the reference routine aborts (\"This "
                                          "routine has been resurrected from an older version of Carpet\")");
      if (order_space != 1)
        return set_error(CG_ERR_ORDER, "There is no stencil for op=\"TVD\" with order_space!=1");
      kind = PK_TVD;
      order = 1;
      break;
    case CG_OP_LAGRANGE_MONOTONE:
      // prolongate_3d_o5_monotone_rf2 for order_space 5 (data.cc:807-832)
      if (order_space == 1 || order_space == 3)
        return set_error(CG_ERR_ORDER, "There is no stencil for op=\"Lagrange_monotone\" with order_space=%d",
                         order_space);
      if (order_space != 5)
        return set_error(CG_ERR_ORDER, "There is no stencil for op=\"Lagrange_monotone\" with order_space!=5");
      if (cent != CG_VERTEX_CENTERED)
        return set_error(CG_ERR_OPERATOR, "op=\"Lagrange_monotone\" is provided for vertex centring only");
      kind = PK_MONO5;
      order = 5;
      break;
    case CG_OP_STAGGER011:
    case CG_OP_STAGGER101:
    case CG_OP_STAGGER110:
    case CG_OP_STAGGER111:
      // prolongate_3d_stagger*<T,2> for order_space 2, <T,3> for 3..5
      // (data.cc:834-920); other orders index past the reference's table
      if (order_space < 2 || order_space > 5)
        return set_error(CG_ERR_ORDER, "There is no staggered stencil with order_space=%d", order_space);
      kind = PK_STAGGER;
      order = order_space == 2 ? 2 : 3;
      break;
    default:
      return set_error(CG_ERR_OPERATOR, "prolongation operator %d is not provided", op);
    }
    if (ntl > 3) {
      // every level prolongated into a temporary, then combined by
      // interpolate_3d_4tl/5tl (data.cc:497-521, 1227-1256)
      if (es != 8)
        return set_error(CG_ERR_OPERATOR, "mixed precision with more than 3 time levels");
      double tw[5];
      for (int l = 0; l < ntl; ++l)
        tw[l] = times[tl0 + l];
      tr.tmp_step = 0;
      tr.tmp_kind = kind;
      tr.tmp_order = order;
      tr.tmp_stag = stagger_dirs(op);
      return setup_tmp(tr, ntl, sp, tw, time, s0, d, sreg, dreg);
    }
    double sf[3] = {1.0, 0.0, 0.0};
    EnoTime ew;
    const bool eno_time = ntl == 3 && eno_group;
    if (ntl > 1) {
      // each level prolongated, then combined in time (data.cc:497-521):
      // ENO/WENO/TVD/monotone groups with 3 levels by
      // interpolate_eno_3d_3tl, the others
      // by interpolate_3d_2tl/3tl (data.cc:1182-1299)
      if (eno_time) {
        if ((e = eno_time_factors(times, time, ew)))
          return e;
        for (int l = 0; l < 3; ++l)
          sf[l] = ew.s3[l];
      } else if ((e = time_factors(ntl, times, time, sf))) {
        return e;
      }
    }
    tr.is_prolong = true;
    return plan_prolong(kind, order, s0, sp, ntl, sf, d, dreg, tr.pr, es,
                        eno_time ? &ew : nullptr, stagger_dirs(op));
  }
  if (finer) {
    if (op == CG_OP_SYNC)
      return set_error(CG_ERR_OPERATOR, "op_sync cannot restrict");
    // the staggered groups restrict with their own operators whatever the
    // centring (data.cc:989-1027)
    const int stag = stagger_dirs(op);
    int cell = 0, rorder = 1, vc = 0;
    if (stag) {
      if (slab)
        return set_error(CG_ERR_ARG, "staggered restriction takes no slab");
    } else if (cent == CG_VERTEX_CENTERED) {
      if (slab)
        return set_error(CG_ERR_ARG, "vertex-centred restriction takes no slab");
    } else if (slab && !(slab->is_centered[0] == 1 && slab->is_centered[1] == 1 &&
                         slab->is_centered[2] == 1)) {
      // restrict_3d_vc_rf2 for a slab with one vertex-centred direction
      // (data.cc:1106-1127); other centrings reach the reference's assert(0)
      const int *c = slab->is_centered;
      for (int k = 0; k < 3; ++k)
        if (c[k] != 0 && c[k] != 1)
          return set_error(CG_ERR_ARG, "is_centered must hold 0 or 1");
      if (c[0] + c[1] + c[2] != 2)
        return set_error(CG_ERR_OPERATOR, "cell-centred restriction with is_centered (%d,%d,%d): the reference "
                                          "provides (1,1,1), (0,1,1), (1,0,1) and (1,1,0) only",
                         c[0], c[1], c[2]);
      vc = c[0] | (c[1] << 1) | (c[2] << 2);
    } else {
      // use_dgfe selects restrict_3d_dgfe_rf2<T,5> here (data.cc:1055-1062)
      if (dgfe_requested())
        return set_error(CG_ERR_OPERATOR, "use_dgfe: restrict_3d_dgfe_rf2 evaluates HRSCCore's "
                                          "GLLElement<5>, which this library does not provide");
      // data.cc:1064-1094: higher-order restriction when the parameter asks
      // for it, except for ENO and WENO groups
      int use_ho = 0, ro = 1;
      restriction_setting(&use_ho, &ro);
      cell = 1;
      rorder = use_ho && op != CG_OP_ENO && op != CG_OP_WENO ? ro : 1;
    }
    if (ntl > 1) {
      // every level restricted into a temporary, then interpolated in time
      // (data.cc:497-521)
      double tw[5];
      for (int l = 0; l < ntl; ++l)
        tw[l] = times[tl0 + l];
      tr.tmp_step = 2;
      tr.tmp_stag = stag;
      tr.tmp_vc = vc;
      tr.tmp_cell = cell;
      tr.tmp_rorder = rorder;
      tr.tmp_eno = ntl == 3 && eno_group;
      return setup_tmp(tr, ntl, sp, tw, time, s0, d, sreg, dreg);
    }
    tr.is_prolong = false;
    if (vc)
      return plan_restrict_vc(vc, s0, d, dreg, tr.pt);
    if (stag)
      return plan_restrict_stagger(stag, s0, d, dreg, tr.pt);
    return plan_restrict(cell != 0, s0, d, dreg, tr.pt, rorder);
  }
  return set_error(CG_ERR_STRIDE, "source and destination strides are not ordered");
}

// transfers through temporaries: one single-level transfer per time level
// into region-sized temporaries at tmp (tmp_ntl x region points), then the
// interpolation in time; every temporary point is written before it is read
static int run_transfer_tmp_into(const Transfer &tr, double *tmp, hipStream_t st) {
  const long long npts = tr.tmp_reg.size();
  int e = CG_OK;
  const void *tp[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  for (int l = 0; l < tr.tmp_ntl && !e; ++l) {
    tp[l] = tmp + l * npts;
    Transfer one;
    e = plan_level(tr, tr.tmp_srcs[l], tmp_arr(tr, tp[l]), one);
    if (!e)
      e = run_transfer(one, 0, st);
  }
  if (!e) {
    Transfer ip;
    std::memset(&ip, 0, sizeof ip);
    e = plan_interp(tr.tmp_ntl, tp, tr.tmp_times, tr.tmp_time, tmp_arr(tr, tmp), tr.tmp_dst, tr.tmp_reg, ip.pt,
                    tr.tmp_eno);
    if (!e)
      e = run_transfer(ip, 0, st);
  }
  return e;
}

// the temporaries from a stream-ordered allocation, released in stream order
static int run_transfer_tmp(const Transfer &tr, hipStream_t st) {
  double *tmp = nullptr;
  hipError_t he = hipMallocAsync(reinterpret_cast<void **>(&tmp),
                                 (size_t)tr.tmp_ntl * tr.tmp_reg.size() * sizeof(double), st);
  if (he != hipSuccess)
    return hip_error(he, "hipMallocAsync (time-level temporaries)");
  int e = run_transfer_tmp_into(tr, tmp, st);
  he = hipFreeAsync(tmp, st);
  if (!e && he != hipSuccess)
    e = hip_error(he, "hipFreeAsync (time-level temporaries)");
  return e;
}

// Temporaries of the allocation-free transfer: one library-owned device
// array per GPU (a static device symbol, so loading the library reserves it
// and no call allocates).  In host code the symbol's name is not a device
// address: its address on the current device comes from
// hipGetSymbolAddress, taken once per device on that device's first call and
// used from then on for every kernel argument.
static __device__ __align__(256) double scratch[CG_NOALLOC_SCRATCH_BYTES / sizeof(double)];
static std::mutex scratch_mu;
static std::vector<void *> scratch_at; // per device: address of scratch

static int scratch_base(double **out) {
  int dev = 0;
  hipError_t he = hipGetDevice(&dev);
  if (he != hipSuccess)
    return hip_error(he, "hipGetDevice");
  std::lock_guard<std::mutex> lock(scratch_mu);
  if ((size_t)dev >= scratch_at.size())
    scratch_at.resize((size_t)dev + 1, nullptr);
  if (!scratch_at[dev]) {
    void *p = nullptr;
    he = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (he != hipSuccess)
      return hip_error(he, "hipGetSymbolAddress (scratch)");
    if (!p)
      return set_error(CG_ERR_HIP, "hipGetSymbolAddress (scratch) gave no address on device %d", dev);
    scratch_at[dev] = p;
  }
  *out = static_cast<double *>(scratch_at[dev]);
  return CG_OK;
}

int run_transfer(const Transfer &tr, int prec, hipStream_t st) {
  if (tr.tmp_ntl)
    return run_transfer_tmp(tr, st);
  if (tr.is_prolong)
    return launch_prolong(tr.pr, tr.pr.es == 4 && prec == CG_PREC_F64 ? CG_PREC_MIXED : prec, st);
  const int r = launch_point(tr.pt.op, tr.pt.a, st);
  if (r == -1)
    return set_error(CG_ERR_ARG, "region too large for the pointwise kernel");
  if (r)
    return hip_error(hipGetLastError(), "kernel launch");
  return CG_OK;
}

} // namespace cg

using namespace cg;

extern "C" {

int cg_version(void) { return 1; }

const char *cg_error_string(int code) {
  switch (code) {
  case CG_OK: return "success";
  case CG_ERR_ARG: return "invalid argument";
  case CG_ERR_STRIDE: return "strides disagree";
  case CG_ERR_REFFACT: return "refinement factor is not 2";
  case CG_ERR_ALIGN: return "bbox not aligned";
  case CG_ERR_EMPTY: return "region extent is empty";
  case CG_ERR_CONTAIN: return "region extent is not contained in array extent";
  case CG_ERR_ORDER: return "no stencil for this order";
  case CG_ERR_TIME_SAME: return "arrays have same time";
  case CG_ERR_TIME_EXTRAP: return "extrapolation in time";
  case CG_ERR_SHAPE: return "array sizes don't agree with bounding boxes";
  case CG_ERR_OPERATOR: return "operator not supported";
  case CG_ERR_HIP: return "HIP runtime error";
  case CG_ERR_NOMEM: return "out of memory";
  case CG_ERR_STATE: return "batch used in the wrong state";
  case CG_ERR_DEVICE: return "device-side check failed";
  case CG_ERR_SCHEDULE: return "schedule consistency check failed";
  case CG_ERR_SCRATCH: return "allocation-free scratch too small";
  }
  return "unknown error";
}

const char *cg_last_error_message(void) { return cg::last_error_message(); }

int cg_set_device(int device) {
  hipError_t e = hipSetDevice(device);
  return e == hipSuccess ? CG_OK : hip_error(e, "hipSetDevice");
}

int cg_device_count(int *count) {
  if (!count)
    return set_error(CG_ERR_ARG, "null count");
  hipError_t e = hipGetDeviceCount(count);
  if (e != hipSuccess) {
    *count = 0;
    return hip_error(e, "hipGetDeviceCount");
  }
  return CG_OK;
}

// ---- storage --------------------------------------------------------------
static long long align_up(long long x, long long a) { return (x + a - 1) / a * a; }
static long long next_pow2(long long x) {
  long long p = 1;
  while (p < x)
    p <<= 1;
  return p;
}

// pad_shape(shape, offset), CarpetLib/src/cacheinfo.cc:40-209, with
// VECTORISE_ALIGNED_ARRAYS, VECTORISE_ALIGN_FOR_CACHE and one cache level
int cg_pad_shape(const int shape[3], const int offset[3], int vec_size,
                 int linesize, int align_interior, int padded_shape[3],
                 int padding_offset[3]) {
  if (!shape || !offset || !padded_shape || !padding_offset || vec_size < 1 ||
      linesize < 1 || (linesize & (linesize - 1)))
    return set_error(CG_ERR_ARG, "bad pad_shape arguments");
  for (int d = 0; d < 3; ++d)
    if (shape[d] < 0 || offset[d] < 0 || offset[d] > shape[d])
      return set_error(CG_ERR_ARG, "pad_shape: shape/offset out of range");
  if (shape[0] == 0 || shape[1] == 0 || shape[2] == 0) {
    for (int d = 0; d < 3; ++d) {
      padded_shape[d] = shape[d];
      padding_offset[d] = 0;
    }
    return CG_OK;
  }
  long long acc = 1;
  for (int d = 0; d < 3; ++d) {
    long long np = shape[d], bnd = offset[d];
    if (d == 0) {
      if (align_interior) {
        long long nb = align_up(bnd, vec_size);
        np += nb - bnd;
        bnd = nb;
      }
      np = align_up(np, vec_size);
    }
    if (np * acc < linesize) {
      np = next_pow2(np);
    } else {
      if (align_interior && d == 0) {
        long long nb = align_up(bnd, linesize);
        np += nb - bnd;
        bnd = nb;
      }
      long long tot = align_up(np * acc, linesize);
      if (tot % acc)
        return set_error(CG_ERR_ARG, "pad_shape: inconsistent line size");
      np = tot / acc;
    }
    if (np > 0x7fffffff)
      return set_error(CG_ERR_ARG, "pad_shape: extent overflows");
    padded_shape[d] = (int)np;
    padding_offset[d] = (int)(bnd - offset[d]);
    acc *= np;
  }
  return CG_OK;
}

int cg_gdata_alloc(cg_gdata *out, const int bbox[9], const int owned_bbox[9],
                   size_t elem_size) {
  if (!out || !bbox || (elem_size != 4 && elem_size != 8))
    return set_error(CG_ERR_ARG, "bad gdata_alloc arguments");
  const Box b = Box::from(bbox);
  for (int d = 0; d < 3; ++d)
    if (b.str[d] <= 0)
      return set_error(CG_ERR_ARG, "non-positive stride");
  if (b.empty())
    return set_error(CG_ERR_EMPTY, "empty extent");
  int shape[3], offset[3] = {0, 0, 0}, pad[3], poff[3];
  for (int d = 0; d < 3; ++d)
    shape[d] = b.sizes(d);
  if (owned_bbox) {
    const Box o = Box::from(owned_bbox);
    if (!o.contained_in(b) || o.empty())
      return set_error(CG_ERR_CONTAIN, "owned box not inside the extent");
    for (int d = 0; d < 3; ++d)
      offset[d] = (o.lo[d] - b.lo[d]) / b.str[d];
  }
  // 128-byte x-rows: a wavefront's 64 lanes then read whole cache lines
  const int vec = (int)(128 / elem_size);
  int e = cg_pad_shape(shape, offset, vec, vec, owned_bbox ? 1 : 0, pad, poff);
  if (e)
    return e;
  // the padding offset shifts the array inside the allocation so the owned
  // interior starts on a line; the storage pointer points at element (0,0,0)
  const size_t n = (size_t)pad[0] * pad[1] * pad[2] + (size_t)poff[0];
  void *p = nullptr;
  hipError_t he = hipMalloc(&p, n * elem_size);
  if (he != hipSuccess)
    return set_error(CG_ERR_NOMEM, "hipMalloc(%zu): %s", n * elem_size, hipGetErrorString(he));
  out->storage = static_cast<char *>(p) + (size_t)poff[0] * elem_size;
  for (int d = 0; d < 3; ++d) {
    out->padext[d] = pad[d];
    out->ext[d] = shape[d];
  }
  std::memcpy(out->bbox, bbox, sizeof out->bbox);
  // cg_gdata_free recovers the allocation from the storage pointer: the
  // offset is under one 128-byte row and hipMalloc aligns to 256 bytes
  return CG_OK;
}

int cg_gdata_free(cg_gdata *g) {
  if (!g || !g->storage)
    return set_error(CG_ERR_ARG, "null gdata");
  // the allocation base is the storage rounded down to the allocation
  // granularity: hipMalloc returns 256-byte aligned blocks and the offset is
  // below one 128-byte row
  void *base = reinterpret_cast<void *>(reinterpret_cast<uintptr_t>(g->storage) & ~uintptr_t(255));
  hipError_t e = hipFree(base);
  g->storage = nullptr;
  return e == hipSuccess ? CG_OK : hip_error(e, "hipFree");
}

int cg_dev_malloc(void **ptr, size_t bytes) {
  if (!ptr)
    return set_error(CG_ERR_ARG, "null ptr");
  hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
  if (e != hipSuccess) {
    *ptr = nullptr;
    return set_error(CG_ERR_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  }
  return CG_OK;
}
int cg_dev_free(void *ptr) {
  hipError_t e = hipFree(ptr);
  return e == hipSuccess ? CG_OK : hip_error(e, "hipFree");
}
int cg_memcpy_htod(void *dst, const void *src, size_t bytes) {
  hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
  return e == hipSuccess ? CG_OK : hip_error(e, "hipMemcpy H2D");
}
int cg_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
  hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
  return e == hipSuccess ? CG_OK : hip_error(e, "hipMemcpy D2H");
}
int cg_memcpy_dtod(void *dst, const void *src, size_t bytes, cg_stream_t s) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, S(s));
  return e == hipSuccess ? CG_OK : hip_error(e, "hipMemcpyAsync D2D");
}
int cg_memset_d8(void *dst, int value, size_t bytes, cg_stream_t s) {
  hipError_t e = hipMemsetAsync(dst, value, bytes, S(s));
  return e == hipSuccess ? CG_OK : hip_error(e, "hipMemsetAsync");
}
int cg_fill_f64(double *dst, double value, size_t n, cg_stream_t s) {
  if (!dst && n)
    return set_error(CG_ERR_ARG, "null dst");
  return launch_fill(dst, value, n, S(s)) ? hip_error(hipGetLastError(), "fill") : CG_OK;
}
int cg_stream_copy_f64(double *dst, const double *src, size_t n, cg_stream_t s) {
  if ((!dst || !src) && n)
    return set_error(CG_ERR_ARG, "null pointer");
  if ((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) & 7)
    return set_error(CG_ERR_ARG, "stream copy needs 8-byte aligned arrays");
  if (n / 512 >= 0x7fffffffull)
    return set_error(CG_ERR_ARG, "stream copy of %zu doubles exceeds one launch", n);
  return launch_stream_copy(dst, src, n, S(s)) ? hip_error(hipGetLastError(), "stream copy") : CG_OK;
}
int cg_fill_hash_f64(double *dst, const int padext[3], const int ext[3],
                     unsigned long long seed, double lo, double hi, cg_stream_t s) {
  if (!dst || !padext || !ext)
    return set_error(CG_ERR_ARG, "null argument");
  for (int d = 0; d < 3; ++d)
    if (ext[d] < 0 || padext[d] < ext[d])
      return set_error(CG_ERR_SHAPE, "extent does not fit the padding");
  return launch_fill_hash(dst, padext[0], (long long)padext[0] * padext[1], ext, seed, lo, hi,
                          S(s))
             ? hip_error(hipGetLastError(), "fill_hash")
             : CG_OK;
}
int cg_fill_hash_many_f64(int n, const cg_hash_fill *fills, cg_stream_t s) {
  if (n < 0 || (n > 0 && !fills))
    return set_error(CG_ERR_ARG, "null argument");
  if (n == 0)
    return CG_OK;
  std::vector<HashFill> f((size_t)n);
  std::vector<long long> first((size_t)n + 1, 0);
  for (int a = 0; a < n; ++a) {
    const cg_hash_fill &h = fills[a];
    if (!h.dst)
      return set_error(CG_ERR_ARG, "fill %d: null array", a);
    for (int d = 0; d < 3; ++d)
      if (h.ext[d] < 0 || h.padext[d] < h.ext[d])
        return set_error(CG_ERR_SHAPE, "fill %d: extent does not fit the padding", a);
    f[a] = HashFill{h.dst, h.padext[0], (long long)h.padext[0] * h.padext[1], h.ext[0], h.ext[1], h.seed, h.lo,
                    h.hi};
    first[a + 1] = first[a] + (long long)h.ext[0] * h.ext[1] * h.ext[2];
  }
  const size_t fb = (size_t)n * sizeof(HashFill), pb = ((size_t)n + 1) * sizeof(long long);
  char *buf = nullptr;
  hipError_t he = hipMalloc(reinterpret_cast<void **>(&buf), fb + pb);
  if (he != hipSuccess)
    return set_error(CG_ERR_NOMEM, "fill descriptors: %s", hipGetErrorString(he));
  int e = CG_OK;
  if ((he = hipMemcpy(buf, f.data(), fb, hipMemcpyHostToDevice)) != hipSuccess ||
      (he = hipMemcpy(buf + fb, first.data(), pb, hipMemcpyHostToDevice)) != hipSuccess)
    e = hip_error(he, "fill descriptor upload");
  if (!e && launch_fill_hash_many(reinterpret_cast<const HashFill *>(buf),
                                  reinterpret_cast<const long long *>(buf + fb), n, first[n], S(s)))
    e = hip_error(hipGetLastError(), "fill_hash_many");
  // the descriptors are released once the fill has run
  if ((he = hipStreamSynchronize(S(s))) != hipSuccess && !e)
    e = hip_error(he, "hipStreamSynchronize (fill_hash_many)");
  (void)hipFree(buf);
  return e;
}

int cg_fill_fourier_f64(const cg_gdata *g, double scale, double amp, int nmodes,
                        const int *kvec, const double *phase, cg_stream_t s) {
  if (!g || !g->storage || (nmodes > 0 && (!kvec || !phase)))
    return set_error(CG_ERR_ARG, "null argument");
  if (nmodes < 0 || nmodes > 8)
    return set_error(CG_ERR_ARG, "at most 8 modes");
  const int r = launch_fill_fourier(static_cast<double *>(g->storage), g->padext[0],
                                    (long long)g->padext[0] * g->padext[1], g->ext, g->bbox,
                                    g->bbox + 6, scale, amp, nmodes, kvec, phase, S(s));
  return r ? hip_error(hipGetLastError(), "fill_fourier") : CG_OK;
}
int cg_stream_create(cg_stream_t *s) {
  hipStream_t st;
  hipError_t e = hipStreamCreate(&st);
  if (e != hipSuccess)
    return hip_error(e, "hipStreamCreate");
  *s = st;
  return CG_OK;
}
int cg_stream_destroy(cg_stream_t s) {
  hipError_t e = hipStreamDestroy(S(s));
  return e == hipSuccess ? CG_OK : hip_error(e, "hipStreamDestroy");
}
int cg_stream_synchronize(cg_stream_t s) {
  hipError_t e = hipStreamSynchronize(S(s));
  return e == hipSuccess ? CG_OK : hip_error(e, "hipStreamSynchronize");
}
int cg_device_synchronize(void) {
  hipError_t e = hipDeviceSynchronize();
  return e == hipSuccess ? CG_OK : hip_error(e, "hipDeviceSynchronize");
}
int cg_event_create(void **ev) {
  hipEvent_t e;
  hipError_t r = hipEventCreate(&e);
  if (r != hipSuccess)
    return hip_error(r, "hipEventCreate");
  *ev = e;
  return CG_OK;
}
int cg_event_destroy(void *ev) {
  hipError_t r = hipEventDestroy(static_cast<hipEvent_t>(ev));
  return r == hipSuccess ? CG_OK : hip_error(r, "hipEventDestroy");
}
int cg_event_record(void *ev, cg_stream_t s) {
  hipError_t r = hipEventRecord(static_cast<hipEvent_t>(ev), S(s));
  return r == hipSuccess ? CG_OK : hip_error(r, "hipEventRecord");
}
int cg_event_elapsed_ms(float *ms, void *start, void *stop) {
  hipError_t r = hipEventSynchronize(static_cast<hipEvent_t>(stop));
  if (r != hipSuccess)
    return hip_error(r, "hipEventSynchronize");
  r = hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop));
  return r == hipSuccess ? CG_OK : hip_error(r, "hipEventElapsedTime");
}

// ---- per-call operators ---------------------------------------------------
int cg_copy_3d_f64(const double *src, const int srcpadext[3], const int srcext[3],
                   double *dst, const int dstpadext[3], const int dstext[3],
                   const int srcbbox[9], const int dstbbox[9],
                   const int srcregbbox[9], const int dstregbbox[9],
                   const cg_islab *slab, cg_stream_t stream) {
  if (!srcpadext || !srcext || !dstpadext || !dstext || !srcbbox || !dstbbox ||
      !srcregbbox || !dstregbbox)
    return set_error(CG_ERR_ARG, "null argument");
  PointPlan pl;
  int e = plan_copy(arr_from(src, srcpadext, srcext, srcbbox),
                    arr_from(dst, dstpadext, dstext, dstbbox),
                    Box::from(srcregbbox), Box::from(dstregbbox), slab, pl);
  if (e)
    return e;
  Transfer tr{};
  tr.is_prolong = false;
  tr.pt = pl;
  return run_transfer(tr, 0, S(stream));
}

static int prolong_call(int kind, int order, const double *src,
                        const int *srcpadext, const int *srcext, double *dst,
                        const int *dstpadext, const int *dstext,
                        const int *srcbbox, const int *dstbbox,
                        const int *dstregbbox, cg_stream_t stream) {
  if (!srcpadext || !srcext || !dstpadext || !dstext || !srcbbox || !dstbbox || !dstregbbox)
    return set_error(CG_ERR_ARG, "null argument");
  ProlongPlan pl;
  const void *sp[1] = {src};
  const double s1[1] = {1.0};
  int e = plan_prolong(kind, order, arr_from(src, srcpadext, srcext, srcbbox), sp, 1,
                       s1, arr_from(dst, dstpadext, dstext, dstbbox),
                       Box::from(dstregbbox), pl);
  if (e)
    return e;
  return launch_prolong(pl, 0, S(stream));
}

int cg_prolongate_3d_rf2_f64(int order, const double *src, const int srcpadext[3],
                             const int srcext[3], double *dst,
                             const int dstpadext[3], const int dstext[3],
                             const int srcbbox[9], const int dstbbox[9],
                             const int srcregbbox[9], const int dstregbbox[9],
                             cg_stream_t stream) {
  (void)srcregbbox;
  return prolong_call(PK_VLAG, order, src, srcpadext, srcext, dst, dstpadext,
                      dstext, srcbbox, dstbbox, dstregbbox, stream);
}

int cg_prolongate_3d_cc_rf2_f64(int order, const double *src, const int srcpadext[3],
                                const int srcext[3], double *dst,
                                const int dstpadext[3], const int dstext[3],
                                const int srcbbox[9], const int dstbbox[9],
                                const int srcregbbox[9], const int dstregbbox[9],
                                cg_stream_t stream) {
  (void)srcregbbox;
  return prolong_call(PK_CLAG, order, src, srcpadext, srcext, dst, dstpadext,
                      dstext, srcbbox, dstbbox, dstregbbox, stream);
}

int cg_prolongate_3d_cc_eno_rf2_f64(int order, const double *src,
                                    const int srcpadext[3], const int srcext[3],
                                    double *dst, const int dstpadext[3],
                                    const int dstext[3], const int srcbbox[9],
                                    const int dstbbox[9], const int srcregbbox[9],
                                    const int dstregbbox[9], cg_stream_t stream) {
  (void)srcregbbox;
  return prolong_call(PK_ENO, order, src, srcpadext, srcext, dst, dstpadext,
                      dstext, srcbbox, dstbbox, dstregbbox, stream);
}

int cg_prolongate_3d_eno_f64(const double *src, const int srcpadext[3], const int srcext[3],
                             double *dst, const int dstpadext[3], const int dstext[3],
                             const int srcbbox[9], const int dstbbox[9], const int srcregbbox[9],
                             const int dstregbbox[9], cg_stream_t stream) {
  (void)srcregbbox;
  return prolong_call(PK_VENO, 3, src, srcpadext, srcext, dst, dstpadext, dstext, srcbbox, dstbbox,
                      dstregbbox, stream);
}

int cg_prolongate_3d_o5_monotone_rf2_f64(const double *src, const int srcpadext[3], const int srcext[3],
                                         double *dst, const int dstpadext[3], const int dstext[3],
                                         const int srcbbox[9], const int dstbbox[9],
                                         const int srcregbbox[9], const int dstregbbox[9],
                                         cg_stream_t stream) {
  (void)srcregbbox;
  return prolong_call(PK_MONO5, 5, src, srcpadext, srcext, dst, dstpadext, dstext, srcbbox, dstbbox,
                      dstregbbox, stream);
}

int cg_prolongate_3d_stagger_f64(int stagger, int order, const double *src, const int srcpadext[3],
                                 const int srcext[3], double *dst, const int dstpadext[3],
                                 const int dstext[3], const int srcbbox[9], const int dstbbox[9],
                                 const int srcregbbox[9], const int dstregbbox[9], cg_stream_t stream) {
  (void)srcregbbox;
  if (!srcpadext || !srcext || !dstpadext || !dstext || !srcbbox || !dstbbox || !dstregbbox)
    return set_error(CG_ERR_ARG, "null argument");
  ProlongPlan pl;
  const void *sp[1] = {src};
  const double s1[1] = {1.0};
  int e = plan_prolong(PK_STAGGER, order, arr_from(src, srcpadext, srcext, srcbbox), sp, 1, s1,
                       arr_from(dst, dstpadext, dstext, dstbbox), Box::from(dstregbbox), pl, 8, nullptr,
                       stagger);
  if (e)
    return e;
  return launch_prolong(pl, 0, S(stream));
}

int cg_restrict_3d_stagger_f64(int stagger, const double *src, const int srcpadext[3],
                               const int srcext[3], double *dst, const int dstpadext[3],
                               const int dstext[3], const int srcbbox[9], const int dstbbox[9],
                               const int srcregbbox[9], const int dstregbbox[9], cg_stream_t stream) {
  (void)srcregbbox;
  if (!srcpadext || !srcext || !dstpadext || !dstext || !srcbbox || !dstbbox || !dstregbbox)
    return set_error(CG_ERR_ARG, "null argument");
  Transfer tr{};
  tr.is_prolong = false;
  int e = plan_restrict_stagger(stagger, arr_from(src, srcpadext, srcext, srcbbox),
                                arr_from(dst, dstpadext, dstext, dstbbox), Box::from(dstregbbox), tr.pt);
  if (e)
    return e;
  return run_transfer(tr, 0, S(stream));
}

int cg_prolongate_3d_cc_tvd_f64(const double *src, const int srcpadext[3],
                                const int srcext[3], double *dst,
                                const int dstpadext[3], const int dstext[3],
                                const int srcbbox[9], const int dstbbox[9],
                                const int srcregbbox[9], const int dstregbbox[9],
                                cg_stream_t stream) {
  (void)srcregbbox;
  return prolong_call(PK_TVD, 1, src, srcpadext, srcext, dst, dstpadext, dstext,
                      srcbbox, dstbbox, dstregbbox, stream);
}

static int restrict_call(bool cell, const double *src, const int *srcpadext,
                         const int *srcext, double *dst, const int *dstpadext,
                         const int *dstext, const int *srcbbox,
                         const int *dstbbox, const int *dstregbbox,
                         cg_stream_t stream, int order = 1) {
  if (!srcpadext || !srcext || !dstpadext || !dstext || !srcbbox || !dstbbox || !dstregbbox)
    return set_error(CG_ERR_ARG, "null argument");
  Transfer tr{};
  tr.is_prolong = false;
  int e = plan_restrict(cell, arr_from(src, srcpadext, srcext, srcbbox),
                        arr_from(dst, dstpadext, dstext, dstbbox),
                        Box::from(dstregbbox), tr.pt, order);
  if (e)
    return e;
  return run_transfer(tr, 0, S(stream));
}

int cg_restrict_3d_rf2_f64(const double *src, const int srcpadext[3],
                           const int srcext[3], double *dst, const int dstpadext[3],
                           const int dstext[3], const int srcbbox[9],
                           const int dstbbox[9], const int srcregbbox[9],
                           const int dstregbbox[9], cg_stream_t stream) {
  (void)srcregbbox;
  return restrict_call(false, src, srcpadext, srcext, dst, dstpadext, dstext,
                       srcbbox, dstbbox, dstregbbox, stream);
}

int cg_restrict_3d_cc_rf2_f64(const double *src, const int srcpadext[3],
                              const int srcext[3], double *dst,
                              const int dstpadext[3], const int dstext[3],
                              const int srcbbox[9], const int dstbbox[9],
                              const int srcregbbox[9], const int dstregbbox[9],
                              cg_stream_t stream) {
  (void)srcregbbox;
  return restrict_call(true, src, srcpadext, srcext, dst, dstpadext, dstext,
                       srcbbox, dstbbox, dstregbbox, stream);
}

int cg_restrict_3d_vc_rf2_f64(int centi, int centj, int centk, const double *src,
                              const int srcpadext[3], const int srcext[3], double *dst,
                              const int dstpadext[3], const int dstext[3], const int srcbbox[9],
                              const int dstbbox[9], const int srcregbbox[9], const int dstregbbox[9],
                              cg_stream_t stream) {
  (void)srcregbbox;
  if (!src || !dst || !srcpadext || !srcext || !dstpadext || !dstext || !srcbbox || !dstbbox || !dstregbbox)
    return set_error(CG_ERR_ARG, "null argument");
  if ((centi != 0 && centi != 1) || (centj != 0 && centj != 1) || (centk != 0 && centk != 1))
    return set_error(CG_ERR_ARG, "centi, centj, centk must be 0 or 1");
  Transfer tr{};
  int e = plan_restrict_vc(centi | (centj << 1) | (centk << 2), arr_from(src, srcpadext, srcext, srcbbox),
                           arr_from(dst, dstpadext, dstext, dstbbox), Box::from(dstregbbox), tr.pt);
  if (e)
    return e;
  return run_transfer(tr, 0, S(stream));
}

int cg_restrict_3d_cc_o3_rf2_f64(const double *src, const int srcpadext[3],
                                 const int srcext[3], double *dst,
                                 const int dstpadext[3], const int dstext[3],
                                 const int srcbbox[9], const int dstbbox[9],
                                 const int srcregbbox[9], const int dstregbbox[9],
                                 cg_stream_t stream) {
  (void)srcregbbox;
  return restrict_call(true, src, srcpadext, srcext, dst, dstpadext, dstext,
                       srcbbox, dstbbox, dstregbbox, stream, 3);
}

int cg_restrict_3d_cc_o5_rf2_f64(const double *src, const int srcpadext[3],
                                 const int srcext[3], double *dst,
                                 const int dstpadext[3], const int dstext[3],
                                 const int srcbbox[9], const int dstbbox[9],
                                 const int srcregbbox[9], const int dstregbbox[9],
                                 cg_stream_t stream) {
  (void)srcregbbox;
  return restrict_call(true, src, srcpadext, srcext, dst, dstpadext, dstext,
                       srcbbox, dstbbox, dstregbbox, stream, 5);
}

static int interp_call(int ntl, const double *const *srcs, const double *times,
                       const int *srcpadext, const int *srcext, double *dst,
                       double t, const int *dstpadext, const int *dstext,
                       const int *srcbbox, const int *dstbbox,
                       const int *dstregbbox, cg_stream_t stream, bool eno = false) {
  if (!srcpadext || !srcext || !dstpadext || !dstext || !srcbbox || !dstbbox || !dstregbbox)
    return set_error(CG_ERR_ARG, "null argument");
  Transfer tr{};
  tr.is_prolong = false;
  const void *sp[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  for (int l = 0; l < ntl && l < 5; ++l)
    sp[l] = srcs[l];
  int e = plan_interp(ntl, sp, times, t, arr_from(srcs[0], srcpadext, srcext, srcbbox),
                      arr_from(dst, dstpadext, dstext, dstbbox),
                      Box::from(dstregbbox), tr.pt, eno);
  if (e)
    return e;
  return run_transfer(tr, 0, S(stream));
}

int cg_interpolate_3d_2tl_f64(const double *src1, double t1, const double *src2,
                              double t2, const int srcpadext[3], const int srcext[3],
                              double *dst, double t, const int dstpadext[3],
                              const int dstext[3], const int srcbbox[9],
                              const int dstbbox[9], const int srcregbbox[9],
                              const int dstregbbox[9], cg_stream_t stream) {
  (void)srcregbbox;
  const double *srcs[2] = {src1, src2};
  const double tt[2] = {t1, t2};
  return interp_call(2, srcs, tt, srcpadext, srcext, dst, t, dstpadext, dstext,
                     srcbbox, dstbbox, dstregbbox, stream);
}

int cg_interpolate_3d_3tl_f64(const double *src1, double t1, const double *src2,
                              double t2, const double *src3, double t3,
                              const int srcpadext[3], const int srcext[3],
                              double *dst, double t, const int dstpadext[3],
                              const int dstext[3], const int srcbbox[9],
                              const int dstbbox[9], const int srcregbbox[9],
                              const int dstregbbox[9], cg_stream_t stream) {
  (void)srcregbbox;
  const double *srcs[3] = {src1, src2, src3};
  const double tt[3] = {t1, t2, t3};
  return interp_call(3, srcs, tt, srcpadext, srcext, dst, t, dstpadext, dstext,
                     srcbbox, dstbbox, dstregbbox, stream);
}

int cg_interpolate_3d_4tl_f64(const double *src1, double t1, const double *src2, double t2,
                              const double *src3, double t3, const double *src4, double t4,
                              const int srcpadext[3], const int srcext[3], double *dst, double t,
                              const int dstpadext[3], const int dstext[3], const int srcbbox[9],
                              const int dstbbox[9], const int srcregbbox[9], const int dstregbbox[9],
                              cg_stream_t stream) {
  (void)srcregbbox;
  const double *srcs[4] = {src1, src2, src3, src4};
  const double tt[4] = {t1, t2, t3, t4};
  return interp_call(4, srcs, tt, srcpadext, srcext, dst, t, dstpadext, dstext, srcbbox, dstbbox, dstregbbox,
                     stream);
}

int cg_interpolate_3d_5tl_f64(const double *src1, double t1, const double *src2, double t2,
                              const double *src3, double t3, const double *src4, double t4,
                              const double *src5, double t5, const int srcpadext[3], const int srcext[3],
                              double *dst, double t, const int dstpadext[3], const int dstext[3],
                              const int srcbbox[9], const int dstbbox[9], const int srcregbbox[9],
                              const int dstregbbox[9], cg_stream_t stream) {
  (void)srcregbbox;
  const double *srcs[5] = {src1, src2, src3, src4, src5};
  const double tt[5] = {t1, t2, t3, t4, t5};
  return interp_call(5, srcs, tt, srcpadext, srcext, dst, t, dstpadext, dstext, srcbbox, dstbbox, dstregbbox,
                     stream);
}

int cg_interpolate_eno_3d_3tl_f64(const double *src1, double t1, const double *src2,
                                  double t2, const double *src3, double t3,
                                  const int srcpadext[3], const int srcext[3],
                                  double *dst, double t, const int dstpadext[3],
                                  const int dstext[3], const int srcbbox[9],
                                  const int dstbbox[9], const int srcregbbox[9],
                                  const int dstregbbox[9], cg_stream_t stream) {
  (void)srcregbbox;
  const double *srcs[3] = {src1, src2, src3};
  const double tt[3] = {t1, t2, t3};
  return interp_call(3, srcs, tt, srcpadext, srcext, dst, t, dstpadext, dstext,
                     srcbbox, dstbbox, dstregbbox, stream, true);
}

static int prolong_mixed_call(int kind, int order, const float *src, const int *srcpadext,
                              const int *srcext, float *dst, const int *dstpadext,
                              const int *dstext, const int *srcbbox, const int *dstbbox,
                              const int *dstregbbox, cg_stream_t stream) {
  if (!srcpadext || !srcext || !dstpadext || !dstext || !srcbbox || !dstbbox || !dstregbbox)
    return set_error(CG_ERR_ARG, "null argument");
  ProlongPlan pl;
  const void *sp[1] = {src};
  const double s1[1] = {1.0};
  int e = plan_prolong(kind, order, arr_from(src, srcpadext, srcext, srcbbox), sp, 1, s1,
                       arr_from(dst, dstpadext, dstext, dstbbox), Box::from(dstregbbox), pl, 4);
  if (e)
    return e;
  return launch_prolong(pl, 1, S(stream));
}

int cg_prolongate_3d_rf2_mixed(int order, const float *src, const int srcpadext[3],
                               const int srcext[3], float *dst, const int dstpadext[3],
                               const int dstext[3], const int srcbbox[9], const int dstbbox[9],
                               const int srcregbbox[9], const int dstregbbox[9],
                               cg_stream_t stream) {
  (void)srcregbbox;
  return prolong_mixed_call(PK_VLAG, order, src, srcpadext, srcext, dst, dstpadext, dstext,
                            srcbbox, dstbbox, dstregbbox, stream);
}

int cg_prolongate_3d_cc_rf2_mixed(int order, const float *src, const int srcpadext[3],
                                  const int srcext[3], float *dst, const int dstpadext[3],
                                  const int dstext[3], const int srcbbox[9], const int dstbbox[9],
                                  const int srcregbbox[9], const int dstregbbox[9],
                                  cg_stream_t stream) {
  (void)srcregbbox;
  return prolong_mixed_call(PK_CLAG, order, src, srcpadext, srcext, dst, dstpadext, dstext,
                            srcbbox, dstbbox, dstregbbox, stream);
}

int cg_transfer_from_mixed(const cg_gdata *dst, const cg_gdata *const *srcs,
                           const double *times, int ntimes, const int dstregbbox[9],
                           const int srcregbbox[9], double time, int transport_operator,
                           int centering, int order_space, int order_time, cg_stream_t stream) {
  Transfer tr{};
  int e = resolve_transfer(dst, srcs, times, ntimes, dstregbbox, srcregbbox, nullptr, time,
                           transport_operator, centering, order_space, order_time, tr, 4);
  if (e)
    return e;
  return run_transfer(tr, 1, S(stream));
}

int cg_transfer_from(const cg_gdata *dst, const cg_gdata *const *srcs,
                     const double *times, int ntimes, const int dstregbbox[9],
                     const int srcregbbox[9], const cg_islab *slab, double time,
                     int transport_operator, int centering, int order_space,
                     int order_time, cg_stream_t stream) {
  Transfer tr{};
  int e = resolve_transfer(dst, srcs, times, ntimes, dstregbbox, srcregbbox, slab,
                           time, transport_operator, centering, order_space,
                           order_time, tr);
  if (e)
    return e;
  return run_transfer(tr, 0, S(stream));
}

int cg_transfer_from_noalloc(const cg_gdata *dst, const cg_gdata *const *srcs,
                             const double *times, int ntimes, const int dstregbbox[9],
                             const int srcregbbox[9], const cg_islab *slab, double time,
                             int transport_operator, int centering, int order_space,
                             int order_time, cg_stream_t stream) {
  Transfer tr{};
  int e = resolve_transfer(dst, srcs, times, ntimes, dstregbbox, srcregbbox, slab,
                           time, transport_operator, centering, order_space,
                           order_time, tr);
  if (e)
    return e;
  if (!tr.tmp_ntl) // a single kernel on the caller's arrays: nothing to stage
    return run_transfer(tr, 0, S(stream));
  const unsigned long long need =
      (unsigned long long)tr.tmp_ntl * (unsigned long long)tr.tmp_reg.size() * sizeof(double);
  if (need > CG_NOALLOC_SCRATCH_BYTES)
    return set_error(CG_ERR_SCRATCH,
                     "%d time-level temporaries of %lld points need %llu bytes of scratch; the "
                     "allocation-free transfer holds at most %llu",
                     tr.tmp_ntl, tr.tmp_reg.size(), need, (unsigned long long)CG_NOALLOC_SCRATCH_BYTES);
  double *tmp = nullptr;
  if ((e = scratch_base(&tmp)))
    return e;
  return run_transfer_tmp_into(tr, tmp, S(stream));
}

} // extern "C"
