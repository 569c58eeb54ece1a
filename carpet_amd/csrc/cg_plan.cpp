// cg_plan.cpp -- host-side validation of operator calls and construction of
// kernel arguments.  Every check restates a precondition of the reference
// operator (cited per function) and returns the matching CG_ERR_* code
// where the reference aborts.  Where the reference's own containment check
// is looser than the stencil it evaluates (prolongate_3d_rf2.cc:399-404 for
// ORDER >= 5 and the cell-centred checks), the exact stencil footprint is
// required to lie in the source array as well, so that no kernel reads
// outside the arrays it was given.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>

#include "cg_internal.hpp"
#include "cg_tiles.hpp"
#include "cg_stencil.hpp"

namespace cg {

static thread_local char g_errmsg[512];

int set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_errmsg, sizeof g_errmsg, fmt, ap);
  va_end(ap);
  return code;
}

const char *last_error_message() { return g_errmsg; }

int hip_error(hipError_t e, const char *where) {
  return set_error(CG_ERR_HIP, "%s: %s", where, hipGetErrorString(e));
}

Arr arr_from(const void *p, const int *pad, const int *ext, const int *bbox) {
  Arr a;
  a.p = p;
  for (int d = 0; d < 3; ++d) {
    a.pad[d] = pad[d];
    a.ext[d] = ext[d];
  }
  a.box = Box::from(bbox);
  return a;
}

Arr arr_from(const cg_gdata &g) { return arr_from(g.storage, g.padext, g.ext, g.bbox); }

int check_array(const Arr &a, const char *name) {
  if (!a.p)
    return set_error(CG_ERR_ARG, "%s: null storage", name);
  for (int d = 0; d < 3; ++d) {
    if (a.box.str[d] <= 0)
      return set_error(CG_ERR_ARG, "%s: stride %d is not positive", name, a.box.str[d]);
    if (a.ext[d] != a.box.sizes(d) || a.pad[d] < a.ext[d])
      return set_error(CG_ERR_SHAPE,
                       "%s: array sizes don't agree with bounding box (dir %d: "
                       "ext=%d padext=%d box size=%d)",
                       name, d, a.ext[d], a.pad[d], a.box.sizes(d));
  }
  return CG_OK;
}

static int check_region_stride(const Box &r) {
  for (int d = 0; d < 3; ++d)
    if (r.str[d] <= 0)
      return set_error(CG_ERR_ARG, "region stride %d is not positive", r.str[d]);
  return CG_OK;
}

static LineInfo line_info_rt(int kind, int order, bool &ok) {
  ok = true;
  switch (kind) {
  case PK_VLAG:
    switch (order) {
    case 1: return line_info<LK_VLAG, 1>();
    case 3: return line_info<LK_VLAG, 3>();
    case 5: return line_info<LK_VLAG, 5>();
    case 7: return line_info<LK_VLAG, 7>();
    case 9: return line_info<LK_VLAG, 9>();
    case 11: return line_info<LK_VLAG, 11>();
    }
    break;
  case PK_CLAG:
    switch (order) {
    case 0: return line_info<LK_CLAG, 0>();
    case 1: return line_info<LK_CLAG, 1>();
    case 2: return line_info<LK_CLAG, 2>();
    case 3: return line_info<LK_CLAG, 3>();
    case 4: return line_info<LK_CLAG, 4>();
    case 5: return line_info<LK_CLAG, 5>();
    }
    break;
  case PK_ENO:
    switch (order) {
    case 2: return line_info<LK_ENO, 2>();
    case 3: return line_info<LK_ENO, 3>();
    }
    break;
  case PK_VENO:
    if (order == 3)
      return line_info<LK_VENO, 3>();
    break;
  }
  ok = false;
  return LineInfo{};
}

// prolongate_3d_rf2.cc:373-417 (vertex), prolongate_3d_cc_rf2.cc:253-321
// (cell Lagrange), prolongate_3d_cc_eno_rf2.cc:962-1030 (cell ENO),
// prolongate_3d_cc_real8_tvd.F90:94-179 (cell TVD)
int plan_prolong(int kind, int order, const Arr &src, const void *const *srcs,
                 int ntl, const double *s, const Arr &dst, const Box &reg,
                 ProlongPlan &pl, int es, const EnoTime *eno, int stag) {
  if (es != 8 && !(es == 4 && (kind == PK_VLAG || kind == PK_CLAG)))
    return set_error(CG_ERR_OPERATOR, "mixed precision applies to Lagrange prolongation only");
  int e;
  if ((e = check_array(src, "source")) || (e = check_array(dst, "destination")))
    return e;
  if ((e = check_region_stride(reg)))
    return e;
  if (ntl < 1 || ntl > 3)
    return set_error(CG_ERR_ORDER, "unsupported number of time levels %d", ntl);
  for (int l = 0; l < ntl; ++l)
    if (!srcs[l])
      return set_error(CG_ERR_ARG, "null source time level %d", l);
  const Box &sb = src.box, &db = dst.box;
  std::memset(&pl, 0, sizeof pl);
  pl.kind = kind;
  pl.order = order;
  pl.ntl = ntl;
  pl.es = es;

  if (kind == PK_MONO5 || kind == PK_STAGGER) {
    // prolongate_3d_o5_monotone_rf2.cc:56-97 and
    // prolongate_3d_stagger011.cc:30-90 (the other three alike): vertex
    // geometry; the exact footprint must lie in the source array (monotone:
    // b-2..b+3 where a point lies between coarse points; staggered: the tap
    // sets of the direction's kind, b-2..b+2 at most)
    if (kind == PK_STAGGER) {
      if (order != 2 && order != 3)
        return set_error(CG_ERR_ORDER, "There is no staggered stencil with order %d", order);
      if (stag < 1 || stag > 7)
        return set_error(CG_ERR_ARG, "staggered directions %d not in 1..7", stag);
      if (!staggered_supported())
        return set_error(CG_ERR_OPERATOR, "When using staggered restriction and prolongation operators "
                                          "you need to set CarpetLib::support_staggered_operators");
    }
    for (int d = 0; d < 3; ++d) {
      if (sb.str[d] <= reg.str[d] || db.str[d] != reg.str[d])
        return set_error(CG_ERR_STRIDE, "Internal error: strides disagree");
      if (sb.str[d] != 2 * db.str[d])
        return set_error(CG_ERR_REFFACT, "Internal error: source strides are not twice the destination strides");
      if (sb.lo[d] % sb.str[d] || db.lo[d] % db.str[d] || reg.lo[d] % reg.str[d])
        return set_error(CG_ERR_ALIGN, "Internal error: bboxes are not aligned with vertices");
    }
    if (reg.empty())
      return set_error(CG_ERR_EMPTY, "Internal error: region extent is empty");
    if (!reg.contained_in(db))
      return set_error(CG_ERR_CONTAIN, "Internal error: region extent is not contained in array extent");
    PointArgs &p = pl.p;
    long long srcpts = 1;
    int doff[3];
    for (int d = 0; d < 3; ++d) {
      if ((reg.lo[d] - sb.lo[d]) % reg.str[d] || (reg.lo[d] - db.lo[d]) % reg.str[d])
        return set_error(CG_ERR_ALIGN, "region is not aligned with the source/destination grids (dir %d)", d);
      p.off[d] = (reg.lo[d] - sb.lo[d]) / reg.str[d];
      doff[d] = (reg.lo[d] - db.lo[d]) / reg.str[d];
      p.n[d] = reg.sizes(d);
      int lo = 1 << 30, hi = -(1 << 30);
      for (int f : {0, 1, p.n[d] - 2, p.n[d] - 1}) {
        if (f < 0 || f >= p.n[d])
          continue;
        const int g = p.off[d] + f, b = g >> 1;
        if (kind == PK_MONO5) {
          lo = std::min(lo, (g & 1) ? b - 2 : b);
          hi = std::max(hi, (g & 1) ? b + 3 : b);
        } else {
          // tap sets: staggered f=0 b-1..b+1 (order 2) / b-2..b+1 (order 3),
          // staggered f=1 and unstaggered f=1 b-1..b+order-1, unstaggered
          // f=0 b (order 2) / b-1..b+1 (order 3)
          const bool st = (stag >> d) & 1, odd = g & 1;
          const int l = st && !odd && order == 3 ? b - 2 : (!st && !odd && order == 2) ? b : b - 1;
          const int h = odd ? b + order - 1 : st ? b + 1 : (order == 2 ? b : b + 1);
          lo = std::min(lo, l);
          hi = std::max(hi, h);
        }
      }
      if (p.off[d] < 0 || lo < 0 || hi >= src.ext[d])
        return set_error(CG_ERR_CONTAIN,
                         "Internal error: region extent is not contained in array extent (stencil "
                         "footprint [%d,%d] vs source [0,%d) in dir %d)", lo, hi, src.ext[d], d);
      srcpts *= (long long)(hi - lo + 1);
    }
    p.sj = src.sj();
    p.sk = src.sk();
    p.dst = static_cast<double *>(const_cast<void *>(dst.p)) + dst.at(doff[0], doff[1], doff[2]);
    p.dj = dst.sj();
    p.dk = dst.sk();
    p.ntl = ntl;
    for (int l = 0; l < ntl; ++l) {
      p.src[l] = static_cast<const double *>(srcs[l]);
      p.s[l] = s[l];
    }
    p.stag = stag;
    p.sorder = order;
    if (eno && ntl == 3) { // monotone groups: interpolate_eno_3d_3tl (data.cc:1260-1292)
      p.teno = 1;
      p.l0 = eno->l0;
      p.s2[0] = eno->s2[0];
      p.s2[1] = eno->s2[1];
    }
    pl.dst_points = (long long)p.n[0] * p.n[1] * p.n[2];
    pl.src_points = srcpts;
    return CG_OK;
  }

  if (kind == PK_TVD) {
    for (int d = 0; d < 3; ++d) {
      if (sb.str[d] <= reg.str[d] || db.str[d] != reg.str[d])
        return set_error(CG_ERR_STRIDE, "Internal error: strides disagree");
      if (sb.str[d] % db.str[d])
        return set_error(CG_ERR_REFFACT, "Internal error: destination strides "
                                         "are not integer multiples of the source strides");
      if (reg.lo[d] > reg.up[d])
        return set_error(CG_ERR_EMPTY, "Internal error: region extent is empty");
    }
    for (int d = 0; d < 3; ++d) {
      if (sb.str[d] / db.str[d] != 2)
        return set_error(CG_ERR_REFFACT, "Internal error: refinement factor is not 2");
      if (reg.str[d] % 2)
        return set_error(CG_ERR_ALIGN, "Internal error: region stride is not a multiple of 2");
    }
    int off[3], doff[3], n[3];
    long long srcpts = 1;
    for (int d = 0; d < 3; ++d) {
      n[d] = (reg.up[d] - reg.lo[d]) / reg.str[d] + 1;
      // every axis adds the x stride's half (F90:165-167)
      off[d] = (reg.lo[d] - sb.lo[d] + reg.str[0] / 2) / db.str[d];
      doff[d] = (reg.lo[d] - db.lo[d]) / db.str[d];
      if (off[d] < 0)
        return set_error(CG_ERR_CONTAIN, "Internal error: source array offset is negative");
      if (doff[d] < 0 || doff[d] + n[d] > dst.ext[d] || reg.lo[d] < db.lo[d])
        return set_error(CG_ERR_CONTAIN, "Internal error: region extent is not "
                                         "contained in array extent");
      const int plo = off[d] / 2, phi = (off[d] + n[d] - 1) / 2;
      if (plo - 1 < 0 || phi + 1 >= src.ext[d])
        return set_error(CG_ERR_CONTAIN, "TVD stencil footprint [%d,%d] outside "
                                         "source array [0,%d) in dir %d",
                         plo - 1, phi + 1, src.ext[d], d);
      srcpts *= (phi + 1) - (plo - 1) + 1;
    }
    PointArgs &p = pl.p;
    p.src[0] = static_cast<const double *>(srcs[0]);
    p.sj = src.sj();
    p.sk = src.sk();
    p.dst = static_cast<double *>(const_cast<void *>(dst.p)) + dst.at(doff[0], doff[1], doff[2]);
    p.dj = dst.sj();
    p.dk = dst.sk();
    for (int d = 0; d < 3; ++d) {
      p.n[d] = n[d];
      p.off[d] = off[d];
    }
    // time levels: each prolongated, then combined (data.cc:497-521)
    p.ntl = ntl;
    for (int l = 0; l < ntl; ++l) {
      p.src[l] = static_cast<const double *>(srcs[l]);
      p.s[l] = s[l];
    }
    if (eno && ntl == 3) {
      p.teno = 1;
      p.l0 = eno->l0;
      p.s2[0] = eno->s2[0];
      p.s2[1] = eno->s2[1];
    }
    pl.dst_points = (long long)n[0] * n[1] * n[2];
    pl.src_points = srcpts;
    return CG_OK;
  }

  bool ok;
  const LineInfo li = line_info_rt(kind, order, ok);
  if (!ok)
    return set_error(CG_ERR_ORDER, "There is no %s stencil with order_space=%d",
                     kind == PK_VLAG ? "vertex-centred Lagrange"
                     : kind == PK_CLAG ? "cell-centred Lagrange"
                     : kind == PK_VENO ? "vertex-centred ENO"
                                       : "cell-centred ENO",
                     order);
  const bool cell = kind == PK_CLAG || kind == PK_ENO;
  for (int d = 0; d < 3; ++d)
    if (sb.str[d] <= reg.str[d] || db.str[d] != reg.str[d])
      return set_error(CG_ERR_STRIDE, "Internal error: strides disagree");
  for (int d = 0; d < 3; ++d)
    if (sb.str[d] != 2 * db.str[d])
      return set_error(CG_ERR_REFFACT, "Internal error: source strides are not "
                                       "twice the destination strides");
  if (!cell) {
    for (int d = 0; d < 3; ++d) {
      if (sb.lo[d] % sb.str[d])
        return set_error(CG_ERR_ALIGN, "Internal error: source bbox is not aligned with vertices");
      if (db.lo[d] % db.str[d])
        return set_error(CG_ERR_ALIGN, "Internal error: destination bbox is not aligned with vertices");
      if (reg.lo[d] % reg.str[d])
        return set_error(CG_ERR_ALIGN, "Internal error: prolongation region bbox is not aligned with vertices");
    }
  } else {
    for (int d = 0; d < 3; ++d)
      if (db.str[d] % 2)
        return set_error(CG_ERR_ALIGN, "Internal error: destination strides are not even");
  }
  if (reg.empty())
    return set_error(CG_ERR_EMPTY, "Internal error: region extent is empty");
  if (!reg.contained_in(db))
    return set_error(CG_ERR_CONTAIN, "Internal error: region extent is not contained in array extent");

  TensorArgs &t = pl.t;
  long long srcpts = 1;
  int doff[3];
  for (int d = 0; d < 3; ++d) {
    const int num = reg.lo[d] - sb.lo[d] + (cell ? reg.str[d] / 2 : 0);
    if (num % reg.str[d] || (reg.lo[d] - db.lo[d]) % reg.str[d])
      return set_error(CG_ERR_ALIGN, "region is not aligned with the source/destination grids (dir %d)", d);
    t.off[d] = num / reg.str[d];
    doff[d] = (reg.lo[d] - db.lo[d]) / reg.str[d];
    t.n[d] = reg.sizes(d);
    int lo = 1 << 30, hi = -(1 << 30);
    const int fs[4] = {0, 1, t.n[d] - 2, t.n[d] - 1};
    for (int f : fs) {
      if (f < 0 || f >= t.n[d])
        continue;
      const int g = t.off[d] + f, b = g >> 1, p = g & 1;
      lo = std::min(lo, b + li.tmin[p]);
      hi = std::max(hi, b + li.tmin[p] + li.ntap[p] - 1);
    }
    if (lo < 0 || hi >= src.ext[d])
      return set_error(CG_ERR_CONTAIN,
                       "Internal error: region extent is not contained in array "
                       "extent (stencil footprint [%d,%d] vs source [0,%d) in dir %d)",
                       lo, hi, src.ext[d], d);
    t.flo[d] = lo;
    t.fhi[d] = hi;
    t.qlo[d] = t.off[d] >> 1;
    t.nq[d] = ((t.off[d] + t.n[d] - 1) >> 1) - t.qlo[d] + 1;
    srcpts *= (long long)(hi - lo + 1);
  }
  for (int l = 0; l < 3; ++l)
    t.src[l] = l < ntl ? srcs[l] : nullptr;
  t.sj = src.sj();
  t.sk = src.sk();
  t.dst = static_cast<char *>(const_cast<void *>(dst.p)) + (long long)es * dst.at(doff[0], doff[1], doff[2]);
  t.dj = dst.sj();
  t.dk = dst.sk();
  for (int l = 0; l < 3; ++l)
    t.s[l] = l < ntl ? s[l] : 0.0;
  t.nlev = ntl;
  if (eno && ntl == 3) {
    t.teno = 1;
    t.l0 = eno->l0;
    t.s2[0] = eno->s2[0];
    t.s2[1] = eno->s2[1];
  }
  pl.dst_points = (long long)t.n[0] * t.n[1] * t.n[2];
  pl.src_points = srcpts;
  pl.es = es;
  return CG_OK;
}

int launch_prolong(const ProlongPlan &pl, int prec, hipStream_t st) {
  int r;
  if (pl.kind == PK_TVD) {
    r = launch_point(PO_TVD, pl.p, st);
  } else if (pl.kind == PK_MONO5) {
    r = launch_point(PO_MONO5, pl.p, st);
  } else if (pl.kind == PK_STAGGER) {
    r = launch_point(PO_STAGGER, pl.p, st);
  } else {
    TileCfg cfg{0, 0, 0};
    // z chunks: at most 128 planes (tiles that march much further apart in
    // z stop sharing their halo rows through L2), at least 2^18 threads (a
    // full round of resident workgroups), at least 16 planes per chunk
    TensorArgs a = pl.t;
    int threads = 0;
    tile_shape(pl.ntl, prec, &cfg.cx, &cfg.cy, &threads);
    const long long cols = (long long)((a.nq[0] + cfg.cx - 1) / cfg.cx) *
                           ((a.nq[1] + cfg.cy - 1) / cfg.cy);
    int zc = std::min(a.nq[2], 128);
    while (zc > 16 && cols * ((a.nq[2] + zc - 1) / zc) * threads < (1LL << 18))
      zc = (zc + 1) / 2;
    cfg.zchunk = zc;
    r = launch_tensor(pl.kind, pl.order, pl.ntl, prec, a, cfg, st);
  }
  if (r == -1)
    return set_error(CG_ERR_ORDER, "no kernel for kind %d order %d ntl %d prec %d",
                     pl.kind, pl.order, pl.ntl, prec);
  if (r)
    return hip_error(hipGetLastError(), "kernel launch");
  return CG_OK;
}

// copy_3d.cc:36-82
int plan_copy(const Arr &src, const Arr &dst, const Box &srcreg1,
              const Box &dreg, const cg_islab *slab, PointPlan &pl) {
  int e;
  if ((e = check_array(src, "source")) || (e = check_array(dst, "destination")))
    return e;
  if ((e = check_region_stride(dreg)))
    return e;
  const Box &sb = src.box, &db = dst.box;
  for (int d = 0; d < 3; ++d)
    if (sb.str[d] != dreg.str[d] || db.str[d] != dreg.str[d] ||
        srcreg1.str[d] != dreg.str[d])
      return set_error(CG_ERR_STRIDE, "Internal error: strides disagree");
  if (dreg.empty())
    return set_error(CG_ERR_EMPTY, "Internal error: region extent is empty");
  Box sreg = dreg;
  if (slab)
    for (int d = 0; d < 3; ++d) {
      sreg.lo[d] += slab->offset[d] * dreg.str[d];
      sreg.up[d] += slab->offset[d] * dreg.str[d];
    }
  if (!sreg.contained_in(srcreg1))
    return set_error(CG_ERR_CONTAIN, "shifted source region is not contained in srcregbbox");
  if (!sreg.contained_in(sb) || !dreg.contained_in(db))
    return set_error(CG_ERR_CONTAIN, "Internal error: region extent is not contained in array extent");
  int so[3], dof[3];
  std::memset(&pl, 0, sizeof pl);
  for (int d = 0; d < 3; ++d) {
    if ((sreg.lo[d] - sb.lo[d]) % sb.str[d] || (dreg.lo[d] - db.lo[d]) % db.str[d])
      return set_error(CG_ERR_ALIGN, "region is not aligned with the arrays (dir %d)", d);
    so[d] = (sreg.lo[d] - sb.lo[d]) / sb.str[d];
    dof[d] = (dreg.lo[d] - db.lo[d]) / db.str[d];
    pl.a.n[d] = dreg.sizes(d);
  }
  pl.op = PO_COPY;
  pl.a.src[0] = static_cast<const double *>(src.p) + src.at(so[0], so[1], so[2]);
  pl.a.sj = src.sj();
  pl.a.sk = src.sk();
  pl.a.dst = static_cast<double *>(const_cast<void *>(dst.p)) + dst.at(dof[0], dof[1], dof[2]);
  pl.a.dj = dst.sj();
  pl.a.dk = dst.sk();
  pl.dst_points = dreg.size();
  pl.src_points = dreg.size();
  return CG_OK;
}

// restrict_3d_rf2.cc:38-68 (vertex), restrict_3d_cc_rf2.cc:38-73 (cell)
int plan_restrict(bool cell, const Arr &src, const Arr &dst, const Box &reg,
                  PointPlan &pl, int order) {
  int e;
  if (order != 1 && !(cell && (order == 3 || order == 5)))
    return set_error(CG_ERR_ORDER, "There is no restriction stencil with restriction_order_space==%d",
                     order);
  // fine cells beyond the children that the stencil reads, per side
  const int halo = order == 3 ? 1 : order == 5 ? 2 : 0;
  if ((e = check_array(src, "source")) || (e = check_array(dst, "destination")))
    return e;
  if ((e = check_region_stride(reg)))
    return e;
  const Box &sb = src.box, &db = dst.box;
  for (int d = 0; d < 3; ++d)
    if (sb.str[d] >= reg.str[d] || db.str[d] != reg.str[d])
      return set_error(CG_ERR_STRIDE, "Internal error: strides disagree");
  for (int d = 0; d < 3; ++d)
    if (2 * sb.str[d] != db.str[d])
      return set_error(CG_ERR_REFFACT, "Internal error: destination strides are "
                                       "not twice the source strides");
  if (reg.empty())
    return set_error(CG_ERR_EMPTY, "Internal error: region extent is empty");
  std::memset(&pl, 0, sizeof pl);
  int so[3], dof[3];
  for (int d = 0; d < 3; ++d) {
    pl.a.n[d] = reg.shape(d) / reg.str[d];
    if (cell) {
      if (sb.str[d] % 2)
        return set_error(CG_ERR_ALIGN, "source stride is not even");
      const int num = reg.lo[d] - sb.lo[d] - sb.str[d] / 2;
      if (num % sb.str[d])
        return set_error(CG_ERR_ALIGN, "region is not aligned with the source cells (dir %d)", d);
      so[d] = num / sb.str[d];
      if (so[d] - halo < 0 || so[d] + 2 * pl.a.n[d] + halo > src.ext[d])
        return set_error(CG_ERR_CONTAIN, "Internal error: region extent is not contained in array "
                                         "extent (restriction footprint, dir %d)", d);
    } else {
      if ((reg.lo[d] - sb.lo[d]) % sb.str[d])
        return set_error(CG_ERR_ALIGN, "region is not aligned with the source (dir %d)", d);
      so[d] = (reg.lo[d] - sb.lo[d]) / sb.str[d];
    }
    if ((reg.lo[d] - db.lo[d]) % db.str[d])
      return set_error(CG_ERR_ALIGN, "region is not aligned with the destination (dir %d)", d);
    dof[d] = (reg.lo[d] - db.lo[d]) / db.str[d];
  }
  if ((!cell && !reg.contained_in(sb)) || !reg.contained_in(db))
    return set_error(CG_ERR_CONTAIN, "Internal error: region extent is not contained in array extent");
  pl.op = !cell ? PO_RESTRICT_VC
          : order == 3 ? PO_RESTRICT_CC_O3
          : order == 5 ? PO_RESTRICT_CC_O5
                       : PO_RESTRICT_CC;
  pl.a.src[0] = static_cast<const double *>(src.p) + src.at(so[0], so[1], so[2]);
  pl.a.sj = src.sj();
  pl.a.sk = src.sk();
  pl.a.dst = static_cast<double *>(const_cast<void *>(dst.p)) + dst.at(dof[0], dof[1], dof[2]);
  pl.a.dj = dst.sj();
  pl.a.dk = dst.sk();
  pl.dst_points = (long long)pl.a.n[0] * pl.a.n[1] * pl.a.n[2];
  pl.src_points = pl.dst_points * (cell ? 8 : 1);
  return CG_OK;
}

// restrict_3d_vc_rf2<T, centi, centj, centk> (restrict_3d_vc_rf2.cc:
// 107-200): a cell-centred group's restriction whose slab has vertex-centred
// directions (cmask bit d set = direction d cell-centred); the children of
// coarse point i start at fine index srcoff + 2i, srcoff = (reg.lo - src.lo
// - src.str/2) / src.str, and every fine point read (srcoff + 2i, and
// srcoff + 2i + 1 in cell-centred directions) must lie in the source array
int plan_restrict_vc(int cmask, const Arr &src, const Arr &dst, const Box &reg, PointPlan &pl) {
  int e;
  if (cmask < 0 || cmask > 7)
    return set_error(CG_ERR_ARG, "centring mask %d not in 0..7", cmask);
  if ((e = check_array(src, "source")) || (e = check_array(dst, "destination")))
    return e;
  if ((e = check_region_stride(reg)))
    return e;
  const Box &sb = src.box, &db = dst.box;
  for (int d = 0; d < 3; ++d)
    if (sb.str[d] >= reg.str[d] || db.str[d] != reg.str[d])
      return set_error(CG_ERR_STRIDE, "Internal error: strides disagree");
  for (int d = 0; d < 3; ++d)
    if (2 * sb.str[d] != db.str[d])
      return set_error(CG_ERR_REFFACT, "Internal error: destination strides are "
                                       "not twice the source strides");
  if (reg.empty())
    return set_error(CG_ERR_EMPTY, "Internal error: region extent is empty");
  std::memset(&pl, 0, sizeof pl);
  int so[3], dof[3];
  for (int d = 0; d < 3; ++d) {
    pl.a.n[d] = reg.sizes(d);
    if (sb.str[d] % 2)
      return set_error(CG_ERR_ALIGN, "source stride is not even");
    const int num = reg.lo[d] - sb.lo[d] - sb.str[d] / 2;
    if (num % sb.str[d] || (reg.lo[d] - db.lo[d]) % db.str[d])
      return set_error(CG_ERR_ALIGN, "region is not aligned with the source/destination (dir %d)", d);
    so[d] = num / sb.str[d];
    dof[d] = (reg.lo[d] - db.lo[d]) / db.str[d];
    const int hi = so[d] + 2 * (pl.a.n[d] - 1) + ((cmask >> d) & 1);
    if (so[d] < 0 || hi >= src.ext[d])
      return set_error(CG_ERR_CONTAIN, "Internal error: region extent is not contained in array extent "
                                       "(footprint [%d,%d] vs source [0,%d) in dir %d)",
                       so[d], hi, src.ext[d], d);
  }
  if (!reg.contained_in(db))
    return set_error(CG_ERR_CONTAIN, "Internal error: region extent is not contained in array extent");
  pl.op = PO_RESTRICT_VC2;
  pl.a.stag = cmask;
  pl.a.src[0] = static_cast<const double *>(src.p) + src.at(so[0], so[1], so[2]);
  pl.a.sj = src.sj();
  pl.a.sk = src.sk();
  pl.a.dst = static_cast<double *>(const_cast<void *>(dst.p)) + dst.at(dof[0], dof[1], dof[2]);
  pl.a.dj = dst.sj();
  pl.a.dk = dst.sk();
  pl.dst_points = (long long)pl.a.n[0] * pl.a.n[1] * pl.a.n[2];
  pl.src_points = pl.dst_points * (1 + (cmask & 1)) * (1 + ((cmask >> 1) & 1)) * (1 + ((cmask >> 2) & 1));
  return CG_OK;
}

// restrict_3d_stagger011.cc:26-100 (the other three alike): vertex
// geometry; every fine point read (2i-1 .. 2i+2 in staggered directions,
// 2i-1 .. 2i+1 in the others) must lie in the source array
int plan_restrict_stagger(int stag, const Arr &src, const Arr &dst, const Box &reg, PointPlan &pl) {
  int e;
  if (stag < 1 || stag > 7)
    return set_error(CG_ERR_ARG, "staggered directions %d not in 1..7", stag);
  if ((e = check_array(src, "source")) || (e = check_array(dst, "destination")))
    return e;
  if ((e = check_region_stride(reg)))
    return e;
  const Box &sb = src.box, &db = dst.box;
  for (int d = 0; d < 3; ++d)
    if (sb.str[d] >= reg.str[d] || db.str[d] != reg.str[d])
      return set_error(CG_ERR_STRIDE, "Internal error: strides disagree");
  for (int d = 0; d < 3; ++d)
    if (2 * sb.str[d] != db.str[d])
      return set_error(CG_ERR_REFFACT, "Internal error: destination strides are "
                                       "not twice the source strides");
  if (reg.empty())
    return set_error(CG_ERR_EMPTY, "Internal error: region extent is empty");
  if (!reg.contained_in(db))
    return set_error(CG_ERR_CONTAIN, "Internal error: region extent is not contained in array extent");
  if (!staggered_supported())
    return set_error(CG_ERR_OPERATOR, "When using staggered restriction and prolongation operators "
                                      "you need to set CarpetLib::support_staggered_operators");
  std::memset(&pl, 0, sizeof pl);
  int so[3], dof[3];
  for (int d = 0; d < 3; ++d) {
    pl.a.n[d] = reg.sizes(d);
    if ((reg.lo[d] - sb.lo[d]) % sb.str[d] || (reg.lo[d] - db.lo[d]) % db.str[d])
      return set_error(CG_ERR_ALIGN, "region is not aligned with the source/destination (dir %d)", d);
    so[d] = (reg.lo[d] - sb.lo[d]) / sb.str[d];
    dof[d] = (reg.lo[d] - db.lo[d]) / db.str[d];
    const int hi = so[d] + 2 * (pl.a.n[d] - 1) + (((stag >> d) & 1) ? 2 : 1);
    if (so[d] - 1 < 0 || hi >= src.ext[d])
      return set_error(CG_ERR_CONTAIN, "Internal error: region extent is not contained in array extent "
                                       "(stencil footprint [%d,%d] vs source [0,%d) in dir %d)",
                       so[d] - 1, hi, src.ext[d], d);
  }
  pl.op = PO_RESTRICT_STAGGER;
  pl.a.stag = stag;
  pl.a.src[0] = static_cast<const double *>(src.p) + src.at(so[0], so[1], so[2]);
  pl.a.sj = src.sj();
  pl.a.sk = src.sk();
  pl.a.dst = static_cast<double *>(const_cast<void *>(dst.p)) + dst.at(dof[0], dof[1], dof[2]);
  pl.a.dj = dst.sj();
  pl.a.dk = dst.sk();
  pl.dst_points = (long long)pl.a.n[0] * pl.a.n[1] * pl.a.n[2];
  long long sp = 1;
  for (int d = 0; d < 3; ++d)
    sp *= 2 * (long long)pl.a.n[d] + (((stag >> d) & 1) ? 2 : 1);
  pl.src_points = sp;
  return CG_OK;
}

static std::atomic<int> g_staggered{0};

bool staggered_supported() { return g_staggered.load(std::memory_order_relaxed) != 0; }

extern "C" int cg_set_staggered_operators(int support_staggered_operators) {
  g_staggered.store(support_staggered_operators ? 1 : 0, std::memory_order_relaxed);
  return CG_OK;
}

static std::atomic<int> g_dgfe{0};

bool dgfe_requested() { return g_dgfe.load(std::memory_order_relaxed) != 0; }

extern "C" int cg_set_dgfe(int use_dgfe) {
  g_dgfe.store(use_dgfe ? 1 : 0, std::memory_order_relaxed);
  return CG_OK;
}

static std::atomic<int> g_ho_restrict{0}, g_restrict_order{3};

void restriction_setting(int *use_higher_order, int *order_space) {
  *use_higher_order = g_ho_restrict.load(std::memory_order_relaxed);
  *order_space = g_restrict_order.load(std::memory_order_relaxed);
}

extern "C" int cg_set_restriction(int use_higher_order_restriction, int restriction_order_space) {
  if (restriction_order_space != 1 && restriction_order_space != 3 && restriction_order_space != 5)
    return set_error(CG_ERR_ORDER, "There is no restriction stencil with restriction_order_space==%d",
                     restriction_order_space);
  g_ho_restrict.store(use_higher_order_restriction ? 1 : 0, std::memory_order_relaxed);
  g_restrict_order.store(restriction_order_space, std::memory_order_relaxed);
  return CG_OK;
}

int time_factors(int ntl, const double *tt, double t, double *s) {
  const double eps = 1.0e-12;
  if (ntl == 1) {
    s[0] = 1.0;
    return CG_OK;
  }
  if (ntl == 2) {
    const double t1 = tt[0], t2 = tt[1];
    if (std::fabs(t1 - t2) < eps)
      return set_error(CG_ERR_TIME_SAME, "Internal error: arrays have same time");
    if (t < std::min(t1, t2) - eps || t > std::max(t1, t2) + eps)
      return set_error(CG_ERR_TIME_EXTRAP, "Internal error: extrapolation in time");
    s[0] = (t - t2) / (t1 - t2);
    s[1] = (t - t1) / (t2 - t1);
    return CG_OK;
  }
  if (ntl == 3) {
    const double t1 = tt[0], t2 = tt[1], t3 = tt[2];
    if (std::fabs(t1 - t2) < eps || std::fabs(t1 - t3) < eps || std::fabs(t2 - t3) < eps)
      return set_error(CG_ERR_TIME_SAME,
                       "Internal error: arrays have same time (t1=%.17g t2=%.17g t3=%.17g)",
                       t1, t2, t3);
    if (t < std::fmin(std::fmin(t1, t2), t3) - eps || t > std::fmax(std::fmax(t1, t2), t3) + eps)
      return set_error(CG_ERR_TIME_EXTRAP, "Internal error: extrapolation in time");
    s[0] = (t - t2) * (t - t3) / ((t1 - t2) * (t1 - t3));
    s[1] = (t - t1) * (t - t3) / ((t2 - t1) * (t2 - t3));
    s[2] = (t - t1) * (t - t2) / ((t3 - t1) * (t3 - t2));
    return CG_OK;
  }
  if (ntl == 4 || ntl == 5) {
    // interpolate_3d_4tl.cc:98-119, interpolate_3d_5tl.cc:101-128: per level
    // the product of (t - t_m) over the other levels in level order over the
    // product of (t_k - t_m) in the same order
    double tmin = tt[0], tmax = tt[0];
    for (int a = 0; a < ntl; ++a) {
      for (int b = a + 1; b < ntl; ++b)
        if (std::fabs(tt[a] - tt[b]) < eps)
          return set_error(CG_ERR_TIME_SAME, "Internal error: arrays have same time");
      tmin = std::fmin(tmin, tt[a]);
      tmax = std::fmax(tmax, tt[a]);
    }
    if (t < tmin - eps || t > tmax + eps)
      return set_error(CG_ERR_TIME_EXTRAP, "Internal error: extrapolation in time");
    for (int k = 0; k < ntl; ++k) {
      double num = 0.0, den = 0.0;
      bool first = true;
      for (int m = 0; m < ntl; ++m) {
        if (m == k)
          continue;
        num = first ? (t - tt[m]) : num * (t - tt[m]);
        den = first ? (tt[k] - tt[m]) : den * (tt[k] - tt[m]);
        first = false;
      }
      s[k] = num / den;
    }
    return CG_OK;
  }
  return set_error(CG_ERR_ORDER, "no time interpolation with %d time levels", ntl);
}

// interpolate_3d_2tl.cc:40-107, interpolate_3d_3tl.cc:41-115
int eno_time_factors(const double *tt, double t, EnoTime &w) {
  const double t1 = tt[0], t2 = tt[1], t3 = tt[2];
  const double tmin = std::fmin(std::fmin(t1, t2), t3);
  const double tmax = std::fmax(std::fmax(t1, t2), t3);
  const double eps = 1.0e-12 * (tmax - tmin);
  if (std::fabs(t1 - t2) < eps || std::fabs(t1 - t3) < eps || std::fabs(t2 - t3) < eps)
    return set_error(CG_ERR_TIME_SAME, "Internal error: arrays have same time");
  if (t < tmin - eps || t > tmax + eps)
    return set_error(CG_ERR_TIME_EXTRAP, "Internal error: extrapolation in time");
  w.s3[0] = (t - t2) * (t - t3) / ((t1 - t2) * (t1 - t3));
  w.s3[1] = (t - t1) * (t - t3) / ((t2 - t1) * (t2 - t3));
  w.s3[2] = (t - t1) * (t - t2) / ((t3 - t1) * (t3 - t2));
  const bool use_12 = t >= std::min(t1, t2) - eps && t <= std::max(t1, t2) + eps;
  const bool use_23 = t >= std::min(t2, t3) - eps && t <= std::max(t2, t3) + eps;
  if (!use_12 && !use_23)
    return set_error(CG_ERR_TIME_EXTRAP, "Internal error: no time interval holds t");
  if (use_12) {
    w.l0 = 0;
    w.s2[0] = (t - t2) / (t1 - t2);
    w.s2[1] = (t - t1) / (t2 - t1);
  } else {
    w.l0 = 1;
    w.s2[0] = (t - t3) / (t2 - t3);
    w.s2[1] = (t - t2) / (t3 - t2);
  }
  return CG_OK;
}

int plan_interp(int ntl, const void *const *srcs, const double *times, double t,
                const Arr &src, const Arr &dst, const Box &reg, PointPlan &pl, bool eno) {
  int e;
  if ((e = check_array(src, "source")) || (e = check_array(dst, "destination")))
    return e;
  if ((e = check_region_stride(reg)))
    return e;
  for (int l = 0; l < ntl; ++l)
    if (!srcs[l])
      return set_error(CG_ERR_ARG, "null source time level %d", l);
  const Box &sb = src.box, &db = dst.box;
  for (int d = 0; d < 3; ++d)
    if (sb.str[d] != reg.str[d] || db.str[d] != reg.str[d])
      return set_error(CG_ERR_STRIDE, "Internal error: strides disagree");
  if (reg.empty())
    return set_error(CG_ERR_EMPTY, "Internal error: region extent is empty");
  if (!reg.contained_in(sb) || !reg.contained_in(db))
    return set_error(CG_ERR_CONTAIN, "Internal error: region extent is not contained in array extent");
  std::memset(&pl, 0, sizeof pl);
  if (eno && ntl == 3) {
    EnoTime w;
    if ((e = eno_time_factors(times, t, w)))
      return e;
    for (int l = 0; l < 3; ++l)
      pl.a.s[l] = w.s3[l];
    pl.a.s2[0] = w.s2[0];
    pl.a.s2[1] = w.s2[1];
    pl.a.l0 = w.l0;
    pl.a.teno = 1;
  } else if ((e = time_factors(ntl, times, t, pl.a.s))) {
    return e;
  }
  if (ntl < 2 || ntl > 5)
    return set_error(CG_ERR_ORDER, "time interpolation needs 2 to 5 time levels");
  int so[3], dof[3];
  for (int d = 0; d < 3; ++d) {
    if ((reg.lo[d] - sb.lo[d]) % sb.str[d] || (reg.lo[d] - db.lo[d]) % db.str[d])
      return set_error(CG_ERR_ALIGN, "region is not aligned with the arrays (dir %d)", d);
    so[d] = (reg.lo[d] - sb.lo[d]) / sb.str[d];
    dof[d] = (reg.lo[d] - db.lo[d]) / db.str[d];
    pl.a.n[d] = reg.sizes(d);
  }
  pl.op = PO_INTERP;
  pl.a.ntl = ntl;
  for (int l = 0; l < ntl; ++l)
    pl.a.src[l] = static_cast<const double *>(srcs[l]) + src.at(so[0], so[1], so[2]);
  pl.a.sj = src.sj();
  pl.a.sk = src.sk();
  pl.a.dst = static_cast<double *>(const_cast<void *>(dst.p)) + dst.at(dof[0], dof[1], dof[2]);
  pl.a.dj = dst.sj();
  pl.a.dk = dst.sk();
  pl.dst_points = reg.size();
  pl.src_points = reg.size();
  return CG_OK;
}

int find_source_timelevel(const double *times, int ntimes, double time,
                          int order_time, int op, int *tl0, int *ntl) {
  if (!times || ntimes <= 0 || order_time < 0)
    return set_error(CG_ERR_ARG, "bad time levels");
  const double eps = 1.0e-12;
  double mn = times[0], mx = times[0];
  for (int i = 1; i < ntimes; ++i) {
    mn = std::min(mn, times[i]);
    mx = std::max(mx, times[i]);
  }
  const double some_time = std::fabs(mn) + std::fabs(mx) + 1.0;
  if (op != CG_OP_COPY && (time < mn - eps * some_time || time > mx + eps * some_time))
    return set_error(CG_ERR_TIME_EXTRAP, "Internal error: extrapolation in time. time=%.17g", time);
  int tl = -1;
  if (ntimes == 1)
    tl = 0;
  if (tl == -1 && op == CG_OP_COPY)
    tl = 0;
  if (tl == -1)
    for (int i = 0; i < ntimes; ++i)
      if (std::fabs(times[i] - time) < eps * some_time) {
        tl = i;
        break;
      }
  if (tl >= 0) {
    *tl0 = tl;
    *ntl = 1;
  } else {
    *tl0 = 0;
    *ntl = order_time + 1;
  }
  return CG_OK;
}

} // namespace cg
