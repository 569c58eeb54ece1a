// cg_internal.hpp -- host-side helpers of libcarpet_amd: boxes, error
// reporting and the validated "plans" that turn an operator call into
// kernel arguments.  Shared by the per-call entry points, the transfer
// dispatch and the batch path.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/carpet_amd.h"
#include "cg_kernels.hpp"

namespace cg {

// bbox<int,3> semantics (CarpetLib/src/bbox.cc:64-126)
struct Box {
  int lo[3], up[3], str[3];
  static Box from(const int *b) {
    Box r;
    for (int d = 0; d < 3; ++d) {
      r.lo[d] = b[d];
      r.up[d] = b[3 + d];
      r.str[d] = b[6 + d];
    }
    return r;
  }
  bool empty() const {
    return lo[0] > up[0] || lo[1] > up[1] || lo[2] > up[2];
  }
  int shape(int d) const { return empty() ? 0 : up[d] - lo[d] + str[d]; }
  int sizes(int d) const { return shape(d) / str[d]; }
  long long size() const {
    return empty() ? 0 : (long long)sizes(0) * sizes(1) * sizes(2);
  }
  bool contained_in(const Box &b) const {
    if (empty())
      return true;
    for (int d = 0; d < 3; ++d)
      if (lo[d] < b.lo[d] || up[d] > b.up[d])
        return false;
    return true;
  }
};

// Array view: storage + padded extents + extents + box
struct Arr {
  const void *p;
  int pad[3];
  int ext[3];
  Box box;
  long long sj() const { return pad[0]; }
  long long sk() const { return (long long)pad[0] * pad[1]; }
  long long at(int i, int j, int k) const { return i + sj() * j + sk() * k; }
};

int set_error(int code, const char *fmt, ...);
int hip_error(hipError_t e, const char *where);

// ext must equal the box sizes and fit in the padding (the reference's
// data<T> guarantees this; F90 TVD checks it explicitly, F90:140-147)
int check_array(const Arr &a, const char *name);

// kinds of tensor prolongation (cg::LineKind) plus TVD
// the tensor kinds equal the line kinds (LK_*) of the kernels
enum ProlongKind { PK_VLAG = 0, PK_CLAG = 1, PK_ENO = 2, PK_TVD = 3, PK_VENO = 4, PK_MONO5 = 5,
                   PK_STAGGER = 6 };

// A validated prolongation: kernel arguments for one (region, time levels)
// plus the statistics the batch path reports.
struct ProlongPlan {
  int kind, order, ntl;
  TensorArgs t;   // tensor kinds
  PointArgs p;    // TVD
  long long dst_points;
  long long src_points; // unique source footprint per time level
  int es;               // element size (8 FP64, 4 mixed)
};

// validates and builds a prolongation plan (srcs[0..ntl-1] share src's
// layout); s[] are the time weights
// es: element size of the arrays (8 = FP64 drop-in, 4 = mixed precision)
// interpolate_eno_3d_3tl's weights (interpolate_eno_3d_3tl.cc:104-137): the
// 3-level weights, the fallback pair's weights and its first level
struct EnoTime {
  double s3[3];
  double s2[2];
  int l0;
};
int eno_time_factors(const double *tt, double t, EnoTime &w);

// eno: the time levels combine with interpolate_eno_3d_3tl's rule (3 levels)
// stag: PK_STAGGER's directions (bit d set = staggered in direction d)
int plan_prolong(int kind, int order, const Arr &src, const void *const *srcs,
                 int ntl, const double *s, const Arr &dst, const Box &reg,
                 ProlongPlan &plan, int es = 8, const EnoTime *eno = nullptr, int stag = 0);
int launch_prolong(const ProlongPlan &plan, int prec, hipStream_t st);

struct PointPlan {
  int op;
  PointArgs a;
  long long dst_points;
  long long src_points;
};
int plan_copy(const Arr &src, const Arr &dst, const Box &srcreg1,
              const Box &dreg, const cg_islab *slab, PointPlan &plan);
// order: 1 (restrict_3d_rf2 / restrict_3d_cc_rf2), 3 or 5 (cell-centred
// restrict_3d_cc_o3_rf2 / restrict_3d_cc_o5_rf2)
int plan_restrict(bool cell, const Arr &src, const Arr &dst, const Box &reg,
                  PointPlan &plan, int order = 1);
// restrict_3d_stagger{011,101,110,111} (bit d of stag set = staggered in
// direction d)
int plan_restrict_stagger(int stag, const Arr &src, const Arr &dst, const Box &reg, PointPlan &plan);
// restrict_3d_vc_rf2 (bit d of cmask set = direction d cell-centred)
int plan_restrict_vc(int cmask, const Arr &src, const Arr &dst, const Box &reg, PointPlan &plan);
// CarpetLib's use_dgfe parameter (process-wide, default off): the DGFE
// operators it selects are not provided, so the transfers that would use
// them are refused
bool dgfe_requested();
// CarpetLib's support_staggered_operators parameter (process-wide, default
// off): the staggered operators refuse to run without it
bool staggered_supported();
// CarpetLib's use_higher_order_restriction / restriction_order_space
// parameters, as seen by the dispatch (process-wide, default off / 1)
void restriction_setting(int *use_higher_order, int *order_space);
int plan_interp(int ntl, const void *const *srcs, const double *times,
                double t, const Arr &src, const Arr &dst, const Box &reg,
                PointPlan &plan, bool eno = false);

// time weights of interpolate_3d_2tl/3tl (interpolate_3d_2tl.cc:97-107,
// interpolate_3d_3tl.cc:98-115)
int time_factors(int ntl, const double *tt, double t, double *s);
// gdata::find_source_timelevel (CarpetLib/src/gdata.cc:288-349)
int find_source_timelevel(const double *times, int ntimes, double time,
                          int order_time, int op, int *tl0, int *ntl);

// A fully resolved transfer (what cg_transfer_from executes): either a
// pointwise plan or a prolongation plan.
struct Transfer {
  bool is_prolong;
  PointPlan pt;
  ProlongPlan pr;
  // transfers through temporaries, as data<T>::transfer_time does them
  // (data.cc:497-521): each time level is prolongated (4 or 5 levels),
  // copied through a slab or restricted into a region-sized temporary
  // (stream-ordered allocation), then the temporaries are interpolated in
  // time; per call only (the batch refuses them)
  int tmp_ntl;  // 0: not such a transfer
  int tmp_step; // 0 prolongation, 1 slabbed copy, 2 restriction
  int tmp_kind, tmp_order, tmp_stag; // prolongation kind; staggered directions
  int tmp_cell, tmp_rorder;          // restriction: cell-centred, order
  int tmp_vc;                        // restriction: restrict_3d_vc_rf2 centring mask (0: none)
  bool tmp_has_slab;
  cg_islab tmp_slab;
  bool tmp_eno; // interpolate_eno_3d_3tl for the 3 levels
  Arr tmp_src, tmp_dst;
  Box tmp_sreg, tmp_reg;
  const void *tmp_srcs[5];
  double tmp_times[5], tmp_time;
  long long dst_points() const { return is_prolong ? pr.dst_points : pt.dst_points; }
};
int resolve_transfer(const cg_gdata *dst, const cg_gdata *const *srcs,
                     const double *times, int ntimes, const int *dstregbbox,
                     const int *srcregbbox, const cg_islab *slab, double time,
                     int op, int cent, int order_space, int order_time,
                     Transfer &tr, int es = 8);
int run_transfer(const Transfer &tr, int prec, hipStream_t st);

Arr arr_from(const void *p, const int *pad, const int *ext, const int *bbox);
Arr arr_from(const cg_gdata &g);

} // namespace cg
