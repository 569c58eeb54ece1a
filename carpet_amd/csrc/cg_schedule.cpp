// cg_schedule.cpp -- the communication schedule of a grid hierarchy: the
// region lists dh::regrid builds (CarpetLib/src/dh.cc:211-1000) for one
// multigrid level, as (send box, source component, receive box, destination
// component) entries:
//   refinement prolongation  (dh.cc:766-830, fast_ref_prol_sendrecv)
//   synchronisation          (dh.cc:834-890, fast_sync_sendrecv)
//   boundary prolongation    (dh.cc:894-975, fast_ref_bnd_prol_sendrecv)
//   refinement restriction   (dh.cc:981-1080, fast_ref_rest_sendrecv)
// together with each component's region description (exterior, owned,
// interior, communicated, dh.cc:283-470) and the level's buffer / overlap /
// active split (dh.cc:472-618).  Host code; every entry is one transfer for
// cg_transfer_from or cg_batch_add.
//
// Box sets are lists of disjoint boxes on one lattice.  The set operations
// have pointwise meanings (bboxset2.hh:1141-1200): expand is a dilation by
// a box of lattice steps, expanded_for maps every point to the target
// lattice points that bracket it, contracted_for keeps the target points
// whose open neighbourhood of one source stride lies inside the set.  A set
// may be split into boxes differently from the reference's normal form;
// the union of the boxes of each (list, source, destination) is the same,
// and so is every transferred value.
#include <algorithm>
#include <cstring>
#include <vector>

#include "cg_internal.hpp"

namespace cg {
namespace {

struct IB {
  int lo[3], up[3], st[3];
  bool empty() const { return lo[0] > up[0] || lo[1] > up[1] || lo[2] > up[2]; }
  static IB none(const int *st) {
    IB b;
    for (int d = 0; d < 3; ++d) {
      b.lo[d] = 0;
      b.up[d] = -st[d];
      b.st[d] = st[d];
    }
    return b;
  }
  static IB from(const int *b) {
    IB r;
    for (int d = 0; d < 3; ++d) {
      r.lo[d] = b[d];
      r.up[d] = b[3 + d];
      r.st[d] = b[6 + d];
    }
    return r;
  }
  void to(int *b) const {
    for (int d = 0; d < 3; ++d) {
      b[d] = lo[d];
      b[3 + d] = up[d];
      b[6 + d] = st[d];
    }
  }
  bool contained_in(const IB &b) const {
    if (empty())
      return true;
    for (int d = 0; d < 3; ++d)
      if (lo[d] < b.lo[d] || up[d] > b.up[d])
        return false;
    return true;
  }
  // bbox::expand (bbox.cc:182-193): by lattice steps, negative shrinks
  IB expand(const int *l, const int *h) const {
    IB r = *this;
    for (int d = 0; d < 3; ++d) {
      r.lo[d] = lo[d] - l[d] * st[d];
      r.up[d] = up[d] + h[d] * st[d];
    }
    return r;
  }
  IB expand(int l, int h) const {
    const int L[3] = {l, l, l}, H[3] = {h, h, h};
    return expand(L, H);
  }
};

int imod(int a, int b) { return ((a % b) + b) % b; }

// bbox::expanded_for / contracted_for / anti_contracted_for (bbox.cc:209-242)
IB expanded_for(const IB &a, const IB &t) {
  if (a.empty())
    return IB::none(t.st);
  IB r;
  for (int d = 0; d < 3; ++d) {
    const int s = t.st[d];
    r.lo[d] = a.lo[d] - imod(a.lo[d] - t.lo[d], s);
    r.up[d] = a.up[d] + (s - imod(a.up[d] - t.lo[d], s)) % s;
    r.st[d] = s;
  }
  return r;
}
IB contracted_for(const IB &a, const IB &t) {
  if (a.empty())
    return IB::none(t.st);
  IB r;
  for (int d = 0; d < 3; ++d) {
    const int s = t.st[d];
    r.lo[d] = a.lo[d] + (s - imod(a.lo[d] - t.lo[d], s)) % s;
    r.up[d] = a.up[d] - imod(a.up[d] - t.lo[d], s);
    r.st[d] = s;
  }
  return r;
}
IB anti_contracted_for(const IB &a, const IB &t) {
  if (a.empty())
    return IB::none(t.st);
  return expanded_for(a.expand(1, 1), t).expand(-1, -1);
}

// intersection of two boxes on one lattice (empty when the lattices differ)
IB meet(const IB &a, const IB &b) {
  IB r = a;
  for (int d = 0; d < 3; ++d) {
    if (a.st[d] != b.st[d] || imod(a.lo[d] - b.lo[d], a.st[d]) != 0)
      return IB::none(a.st);
    r.lo[d] = std::max(a.lo[d], b.lo[d]);
    r.up[d] = std::min(a.up[d], b.up[d]);
  }
  return r.empty() ? IB::none(a.st) : r;
}

// a minus b as at most 6 disjoint boxes
void subtract(const IB &a, const IB &b, std::vector<IB> &out) {
  const IB i = meet(a, b);
  if (i.empty()) {
    out.push_back(a);
    return;
  }
  IB cur = a;
  for (int d = 0; d < 3; ++d) {
    if (cur.lo[d] < i.lo[d]) {
      IB p = cur;
      p.up[d] = i.lo[d] - cur.st[d];
      out.push_back(p);
      cur.lo[d] = i.lo[d];
    }
    if (cur.up[d] > i.up[d]) {
      IB p = cur;
      p.lo[d] = i.up[d] + cur.st[d];
      out.push_back(p);
      cur.up[d] = i.up[d];
    }
  }
}

struct Set {
  std::vector<IB> b; // disjoint, non-empty

  bool empty() const { return b.empty(); }
  long long size() const {
    long long n = 0;
    for (const IB &x : b)
      n += (long long)((x.up[0] - x.lo[0]) / x.st[0] + 1) * ((x.up[1] - x.lo[1]) / x.st[1] + 1) *
           ((x.up[2] - x.lo[2]) / x.st[2] + 1);
    return n;
  }
  static Set of(const IB &x) {
    Set s;
    if (!x.empty())
      s.b.push_back(x);
    return s;
  }
  void add(const IB &x) {
    if (x.empty())
      return;
    std::vector<IB> pieces{x}, next;
    for (const IB &e : b) {
      next.clear();
      for (const IB &p : pieces)
        subtract(p, e, next);
      pieces.swap(next);
      if (pieces.empty())
        return;
    }
    b.insert(b.end(), pieces.begin(), pieces.end());
  }
  Set operator|(const Set &o) const {
    Set r = *this;
    for (const IB &x : o.b)
      r.add(x);
    r.coalesce();
    return r;
  }
  Set operator&(const Set &o) const {
    Set r;
    for (const IB &x : b)
      for (const IB &y : o.b) {
        const IB i = meet(x, y);
        if (!i.empty())
          r.b.push_back(i);
      }
    r.coalesce();
    return r;
  }
  Set operator&(const IB &y) const { return *this & Set::of(y); }
  Set operator-(const Set &o) const {
    Set r;
    for (const IB &x : b) {
      std::vector<IB> pieces{x}, next;
      for (const IB &y : o.b) {
        next.clear();
        for (const IB &p : pieces)
          subtract(p, y, next);
        pieces.swap(next);
        if (pieces.empty())
          break;
      }
      r.b.insert(r.b.end(), pieces.begin(), pieces.end());
    }
    r.coalesce();
    return r;
  }
  // dilation by lo/hi lattice steps (bboxset2.hh:1141-1158)
  Set expand(const int *l, const int *h) const {
    Set r;
    for (const IB &x : b)
      r.add(x.expand(l, h));
    r.coalesce();
    return r;
  }
  Set expanded_for(const IB &t) const {
    Set r;
    for (const IB &x : b)
      r.add(cg::expanded_for(x, t));
    r.coalesce();
    return r;
  }
  IB container() const {
    IB c = b.front();
    for (const IB &x : b)
      for (int d = 0; d < 3; ++d) {
        c.lo[d] = std::min(c.lo[d], x.lo[d]);
        c.up[d] = std::max(c.up[d], x.up[d]);
      }
    return c;
  }
  // target points whose open neighbourhood of one source stride lies in
  // the set: the complement construction of bboxset2.hh:1186-1199
  Set contracted_for(const IB &t) const {
    if (empty())
      return Set();
    const IB cont = container();
    const IB world = cont.expand(2, 2);
    const Set outside = Set::of(world) - *this;
    Set touched;
    for (const IB &x : outside.b)
      touched.add(anti_contracted_for(x, t));
    Set r = Set::of(cg::contracted_for(cont, t)) - touched;
    r.coalesce();
    return r;
  }
  // merge neighbours that share a face and agree on the other two axes
  void coalesce() {
    bool merged = true;
    while (merged) {
      merged = false;
      for (size_t i = 0; i < b.size() && !merged; ++i)
        for (size_t j = i + 1; j < b.size() && !merged; ++j)
          for (int d = 0; d < 3 && !merged; ++d) {
            const int e = (d + 1) % 3, f = (d + 2) % 3;
            const IB &x = b[i], &y = b[j];
            if (x.lo[e] != y.lo[e] || x.up[e] != y.up[e] || x.lo[f] != y.lo[f] || x.up[f] != y.up[f])
              continue;
            if (x.up[d] + x.st[d] == y.lo[d] || y.up[d] + y.st[d] == x.lo[d]) {
              IB m = x;
              m.lo[d] = std::min(x.lo[d], y.lo[d]);
              m.up[d] = std::max(x.up[d], y.up[d]);
              b[i] = m;
              b.erase(b.begin() + (long)j);
              merged = true;
            }
          }
    }
    std::sort(b.begin(), b.end(), [](const IB &x, const IB &y) {
      for (int d = 2; d >= 0; --d)
        if (x.lo[d] != y.lo[d])
          return x.lo[d] < y.lo[d];
      for (int d = 2; d >= 0; --d)
        if (x.up[d] != y.up[d])
          return x.up[d] < y.up[d];
      return false;
    });
  }
  bool contained_in(const Set &o) const { return (*this - o).empty(); }
  bool contained_in(const IB &o) const { return (*this - Set::of(o)).empty(); }
};

struct SR {
  IB send;
  int src;
  IB recv;
  int dst;
};

struct Comp {
  IB exterior, interior, owned, communicated;
  Set ghosts, boundaries, outer_boundaries, buffers, overlaps, active, sync, bndref;
  bool outer[2][3];
};

struct Level {
  IB domain;
  int reffact[3];
  int gw[2][3], bw[2][3], ow[2][3];
  int order;
  std::vector<Comp> comps;
  Set allbuffers, alloverlaps, allactive;
  std::vector<SR> lists[CG_SCHED_NLISTS];
};

} // namespace
} // namespace cg

struct cg_schedule {
  int centering;
  std::vector<cg::Level> levels;
};

namespace cg {
namespace {

#define SCHED_CHECK(cond, ...)                                                                     \
  do {                                                                                             \
    if (!(cond))                                                                                   \
      return set_error(CG_ERR_SCHEDULE, __VA_ARGS__);                                             \
  } while (0)

// regions of one level (dh.cc:283-470) and its buffer/overlap/active split
// (dh.cc:472-618)
int level_regions(Level &L, const int bnd[2][3]) {
  bool trivial = true;
  for (int f = 0; f < 2; ++f)
    for (int d = 0; d < 3; ++d)
      trivial = trivial && bnd[f][d] == 0 && L.gw[f][d] == 0 && L.bw[f][d] == 0 && L.ow[f][d] == 0;
  const IB &dom = L.domain;
  IB act = dom; // domain_active = domain_exterior.expand(-boundary_width)
  for (int d = 0; d < 3; ++d) {
    act.lo[d] = dom.lo[d] + bnd[0][d] * dom.st[d];
    act.up[d] = dom.up[d] - bnd[1][d] * dom.st[d];
  }
  SCHED_CHECK(act.contained_in(dom), "The active part of the domain must be contained in the exterior");
  for (size_t c = 0; c < L.comps.size(); ++c) {
    Comp &k = L.comps[c];
    const IB &in = k.interior;
    SCHED_CHECK(in.contained_in(dom), "The interior must be contained in the domain (component %zu)", c);
    int glo[3], ghi[3], blo[3], bhi[3];
    for (int d = 0; d < 3; ++d) {
      k.outer[0][d] = in.lo[d] == dom.lo[d];
      k.outer[1][d] = in.up[d] == dom.up[d];
      glo[d] = k.outer[0][d] ? 0 : L.gw[0][d];
      ghi[d] = k.outer[1][d] ? 0 : L.gw[1][d];
      blo[d] = k.outer[0][d] ? -bnd[0][d] : 0;
      bhi[d] = k.outer[1][d] ? -bnd[1][d] : 0;
    }
    k.exterior = trivial ? in : in.expand(glo, ghi);
    SCHED_CHECK(k.exterior.contained_in(dom), "The exterior must be contained in the domain (component %zu)", c);
    k.ghosts = trivial ? Set() : Set::of(k.exterior) - Set::of(in);
    k.communicated = trivial ? k.exterior : k.exterior.expand(blo, bhi);
    SCHED_CHECK(k.communicated.contained_in(act),
                "The communicated region must be contained in the active part of the domain "
                "(component %zu)", c);
    k.outer_boundaries = trivial ? Set() : Set::of(k.exterior) - Set::of(k.communicated);
    k.owned = trivial ? in : in.expand(blo, bhi);
    SCHED_CHECK(k.owned.contained_in(act),
                "The owned region must be contained in the active part of the domain (component %zu)", c);
    k.boundaries = trivial ? Set() : Set::of(k.communicated) - Set::of(k.owned);
    SCHED_CHECK(k.boundaries.contained_in(act),
                "The boundary must be contained in the active part of the domain (component %zu)", c);
  }
  // all owned regions; they must be disjoint
  Set allowned;
  long long nowned = 0;
  for (const Comp &k : L.comps) {
    allowned.add(k.owned);
    nowned += Set::of(k.owned).size();
  }
  allowned.coalesce();
  SCHED_CHECK(allowned.size() == nowned, "All owned regions must be disjoint");
  SCHED_CHECK(allowned.contained_in(act), "The owned regions must be contained in the active part of the domain");
  const Set notowned = Set::of(act) - allowned;
  int nal[3], nah[3];
  for (int d = 0; d < 3; ++d) {
    nal[d] = L.bw[0][d] + L.ow[0][d];
    nah[d] = L.bw[1][d] + L.ow[1][d];
  }
  const Set notactive = notowned.expand(nal, nah);
  L.allbuffers = allowned & notowned.expand(L.bw[0], L.bw[1]);
  L.alloverlaps = (allowned & notactive) - L.allbuffers;
  L.allactive = allowned - notactive;
  for (Comp &k : L.comps) {
    if (trivial) {
      k.buffers = Set();
      k.overlaps = Set();
      k.active = Set::of(k.owned);
    } else {
      k.buffers = L.allbuffers & k.owned;
      k.overlaps = L.alloverlaps & k.owned;
      k.active = L.allactive & k.owned;
    }
  }
  return CG_OK;
}

// the coarse active points of obox as seen on the fine lattice of box
// (dh.cc:795-800, "expanded_oactive")
Set expanded_oactive(const Comp &obox, const Comp &box, const int *reffact, int centering) {
  int e[3];
  for (int d = 0; d < 3; ++d)
    e[d] = centering == CG_VERTEX_CENTERED ? reffact[d] : reffact[d] - 1;
  return obox.active.expanded_for(box.interior).expand(e, e);
}

// prolongation from level orl into needrecv of component c (dh.cc:776-826
// and 926-970)
int prolong_list(cg_schedule &S, int rl, int c, Set needrecv, std::vector<SR> &out, Set *filled,
                 const char *what, Set *left = nullptr) {
  Level &L = S.levels[rl];
  const Level &O = S.levels[rl - 1];
  const Comp &box = L.comps[c];
  const int sten[3] = {L.order / 2, L.order / 2, L.order / 2};
  for (size_t cc = 0; cc < O.comps.size() && !needrecv.empty(); ++cc) {
    const Comp &obox = O.comps[cc];
    const Set ovlp = needrecv & expanded_oactive(obox, box, L.reffact, S.centering);
    for (const IB &recv : ovlp.b) {
      const IB send = expanded_for(recv, obox.interior).expand(sten, sten);
      SCHED_CHECK(send.contained_in(obox.exterior),
                  "%s: Send region must be contained in exterior (level %d, component %d from %zu)", what,
                  rl, c, cc);
      out.push_back(SR{send, (int)cc, recv, c});
    }
    needrecv = needrecv - ovlp;
    if (filled)
      *filled = *filled | ovlp;
  }
  if (left) {
    *left = needrecv;
    return CG_OK;
  }
  SCHED_CHECK(needrecv.empty(), "%s: All points must have been received (level %d, component %d)", what, rl, c);
  return CG_OK;
}

int build(cg_schedule &S, const int bnd[2][3], int ho_restrict, int restrict_order) {
  const int nl = (int)S.levels.size();
  for (int rl = 0; rl < nl; ++rl)
    if (int e = level_regions(S.levels[rl], bnd))
      return e;
  for (int rl = 0; rl < nl; ++rl) {
    Level &L = S.levels[rl];
    for (size_t c = 0; c < L.comps.size(); ++c) {
      Comp &box = L.comps[c];
      // refinement prolongation: active and overlap points (dh.cc:766-830)
      if (rl > 0)
        if (int e = prolong_list(S, rl, (int)c, box.active | box.overlaps, L.lists[CG_SCHED_REF_PROL], nullptr,
                                 "Refinement prolongation"))
          return e;
      // synchronisation: the ghost zones (outer boundary included) from
      // the other components' interiors (dh.cc:834-890)
      {
        Set needrecv = box.ghosts;
        for (size_t cc = 0; cc < L.comps.size() && !needrecv.empty(); ++cc) {
          const Set ovlp = needrecv & L.comps[cc].interior;
          SCHED_CHECK(cc != c || ovlp.empty(), "A region may not synchronise from itself (component %zu)", c);
          for (const IB &recv : ovlp.b)
            L.lists[CG_SCHED_SYNC].push_back(SR{recv, (int)cc, recv, (int)c});
          needrecv = needrecv - ovlp;
          box.sync = box.sync | ovlp;
        }
      }
      // boundary prolongation: ghost points not synchronised (outer
      // boundary excluded) and the buffer zones (dh.cc:894-975)
      if (rl > 0) {
        const Set needrecv = ((box.ghosts - box.sync) & box.communicated) | box.buffers;
        if (int e = prolong_list(S, rl, (int)c, needrecv, L.lists[CG_SCHED_REF_BND_PROL], &box.bndref,
                                 "Boundary prolongation"))
          return e;
      }
    }
    // refinement restriction into level rl-1 (dh.cc:981-1080)
    if (rl > 0) {
      Level &O = S.levels[rl - 1];
      const IB &odom = O.domain;
      Set allrestricted;
      if (S.centering == CG_VERTEX_CENTERED) {
        allrestricted = L.allactive.contracted_for(odom);
      } else if (!L.allactive.empty()) {
        // coarse cells at least one coarse cell inside the region the fine
        // active points cover (dh.cc:998-1011)
        const IB all_source = L.allactive.container().expand(10, 10);
        const IB all_target = contracted_for(all_source, odom);
        const Set tmp1 = L.allactive.expanded_for(odom);
        const Set tmp2 = Set::of(all_target) - tmp1;
        const int one[3] = {1, 1, 1};
        const Set tmp3 = tmp2.expand(one, one);
        allrestricted = Set::of(all_target) - tmp3;
      }
      // with higher-order restriction or the staggered operators, restrict
      // into interiors only, from sources shrunk by order/2 (cell centring)
      // or by 2 (staggered) (dh.cc:1019-1050)
      const bool stag = staggered_supported();
      const int shrink = (ho_restrict && S.centering == CG_CELL_CENTERED) ? restrict_order / 2 : stag ? 2 : 0;
      const int sh[3] = {shrink, shrink, shrink}, msh[3] = {-shrink, -shrink, -shrink};
      for (size_t oc = 0; oc < O.comps.size(); ++oc) {
        const Comp &obox = O.comps[oc];
        Set needrecv = ho_restrict || stag ? allrestricted & obox.interior : allrestricted & obox.exterior;
        SCHED_CHECK((allrestricted & obox.buffers).empty(),
                    "Refinement restriction: cannot restrict into buffer zones (level %d, component %zu)",
                    rl - 1, oc);
        for (size_t c = 0; c < L.comps.size() && !needrecv.empty(); ++c) {
          const Comp &box = L.comps[c];
          const IB cext = contracted_for(box.exterior.expand(msh, msh), odom);
          const Set ovlp = needrecv & cext;
          for (const IB &recv : ovlp.b) {
            const IB send = expanded_for(recv, box.exterior).expand(sh, sh);
            SCHED_CHECK(send.contained_in(box.exterior),
                        "Refinement restriction: Send region must be contained in exterior (level %d, "
                        "component %zu)", rl, c);
            O.lists[CG_SCHED_REF_REST].push_back(SR{send, (int)c, recv, (int)oc});
          }
          needrecv = needrecv - ovlp;
        }
        SCHED_CHECK(needrecv.empty(),
                    "Refinement restriction: All points must have been received (level %d, component %zu)",
                    rl - 1, oc);
      }
    }
  }
  return CG_OK;
}

const Set *set_of(const Comp &k, const Level &L, int which) {
  switch (which) {
  case CG_SCHED_GHOSTS: return &k.ghosts;
  case CG_SCHED_BUFFERS: return &k.buffers;
  case CG_SCHED_OVERLAPS: return &k.overlaps;
  case CG_SCHED_ACTIVE: return &k.active;
  case CG_SCHED_SYNCED: return &k.sync;
  case CG_SCHED_BNDREF: return &k.bndref;
  case CG_SCHED_BOUNDARIES: return &k.boundaries;
  }
  (void)L;
  return nullptr;
}

} // namespace
} // namespace cg

using namespace cg;

extern "C" int cg_schedule_build(int centering, const int boundary_width[6], int nlevels,
                                 const cg_level_spec *levels, int use_higher_order_restriction,
                                 int restriction_order_space, cg_schedule **out) {
  if (!out || !levels || !boundary_width || nlevels <= 0)
    return set_error(CG_ERR_ARG, "null argument");
  *out = nullptr;
  if (centering != CG_VERTEX_CENTERED && centering != CG_CELL_CENTERED)
    return set_error(CG_ERR_ARG, "centering must be vertex or cell");
  if (use_higher_order_restriction && restriction_order_space != 1 && restriction_order_space != 3 &&
      restriction_order_space != 5)
    return set_error(CG_ERR_ORDER, "There is no restriction stencil with restriction_order_space==%d",
                     restriction_order_space);
  int bnd[2][3];
  for (int f = 0; f < 2; ++f)
    for (int d = 0; d < 3; ++d) {
      bnd[f][d] = boundary_width[3 * f + d];
      if (bnd[f][d] < 0)
        return set_error(CG_ERR_ARG, "The gh boundary widths must not be negative");
    }
  auto S = new cg_schedule;
  S->centering = centering;
  S->levels.resize(nlevels);
  for (int rl = 0; rl < nlevels; ++rl) {
    const cg_level_spec &sp = levels[rl];
    Level &L = S->levels[rl];
    if (sp.ncomp < 0 || (sp.ncomp > 0 && !sp.interior)) {
      delete S;
      return set_error(CG_ERR_ARG, "level %d: bad component list", rl);
    }
    L.domain = IB::from(sp.baseextent);
    for (int d = 0; d < 3; ++d) {
      if (L.domain.st[d] <= 0 || sp.reffact[d] <= 0) {
        delete S;
        return set_error(CG_ERR_ARG, "level %d: strides and refinement factors must be positive", rl);
      }
      for (int f = 0; f < 2; ++f) {
        L.gw[f][d] = sp.ghost_width[3 * f + d];
        L.bw[f][d] = sp.buffer_width[3 * f + d];
        L.ow[f][d] = sp.overlap_width[3 * f + d];
        if (L.gw[f][d] < 0 || L.bw[f][d] < 0 || L.ow[f][d] < 0) {
          delete S;
          return set_error(CG_ERR_ARG, "level %d: widths must not be negative", rl);
        }
      }
      L.reffact[d] = 1;
    }
    if (rl > 0)
      for (int d = 0; d < 3; ++d) {
        const int r = sp.reffact[d], ro = levels[rl - 1].reffact[d];
        if (r % ro != 0) {
          delete S;
          return set_error(CG_ERR_REFFACT, "Refinement factors must be integer multiples of each other");
        }
        L.reffact[d] = r / ro;
      }
    L.order = sp.prolongation_order;
    if (L.order < 0) {
      delete S;
      return set_error(CG_ERR_ORDER, "level %d: negative prolongation order", rl);
    }
    L.comps.resize(sp.ncomp);
    for (int c = 0; c < sp.ncomp; ++c) {
      L.comps[c].interior = IB::from(sp.interior + 9 * c);
      for (int d = 0; d < 3; ++d)
        if (L.comps[c].interior.st[d] != L.domain.st[d]) {
          delete S;
          return set_error(CG_ERR_STRIDE, "level %d component %d: stride differs from the domain's", rl, c);
        }
    }
  }
  if (int e = build(*S, bnd, use_higher_order_restriction, restriction_order_space)) {
    delete S;
    return e;
  }
  *out = S;
  return CG_OK;
}

extern "C" void cg_schedule_free(cg_schedule *s) { delete s; }

extern "C" int cg_schedule_count(const cg_schedule *s, int rl, int list) {
  if (!s || rl < 0 || rl >= (int)s->levels.size() || list < 0 || list >= CG_SCHED_NLISTS)
    return set_error(CG_ERR_ARG, "no such level or list");
  return (int)s->levels[rl].lists[list].size();
}

extern "C" int cg_schedule_region(const cg_schedule *s, int rl, int list, int i, int *src_comp, int send[9],
                                  int *dst_comp, int recv[9]) {
  if (!s || rl < 0 || rl >= (int)s->levels.size() || list < 0 || list >= CG_SCHED_NLISTS)
    return set_error(CG_ERR_ARG, "no such level or list");
  const auto &v = s->levels[rl].lists[list];
  if (i < 0 || i >= (int)v.size() || !src_comp || !send || !dst_comp || !recv)
    return set_error(CG_ERR_ARG, "no such entry");
  const SR &r = v[i];
  *src_comp = r.src;
  *dst_comp = r.dst;
  r.send.to(send);
  r.recv.to(recv);
  return CG_OK;
}

extern "C" int cg_schedule_component(const cg_schedule *s, int rl, int c, int which, int bbox[9]) {
  if (!s || rl < 0 || rl >= (int)s->levels.size() || !bbox)
    return set_error(CG_ERR_ARG, "no such level");
  const Level &L = s->levels[rl];
  if (c < 0 || c >= (int)L.comps.size())
    return set_error(CG_ERR_ARG, "no such component");
  const Comp &k = L.comps[c];
  switch (which) {
  case CG_SCHED_EXTERIOR: k.exterior.to(bbox); return CG_OK;
  case CG_SCHED_INTERIOR: k.interior.to(bbox); return CG_OK;
  case CG_SCHED_OWNED: k.owned.to(bbox); return CG_OK;
  case CG_SCHED_COMMUNICATED: k.communicated.to(bbox); return CG_OK;
  }
  return set_error(CG_ERR_ARG, "no such box kind");
}

extern "C" int cg_schedule_set(const cg_schedule *s, int rl, int c, int which, int cap, int *bboxes) {
  if (!s || rl < 0 || rl >= (int)s->levels.size())
    return set_error(CG_ERR_ARG, "no such level");
  const Level &L = s->levels[rl];
  if (c < 0 || c >= (int)L.comps.size())
    return set_error(CG_ERR_ARG, "no such component");
  const Set *st = set_of(L.comps[c], L, which);
  if (!st)
    return set_error(CG_ERR_ARG, "no such set kind");
  const int n = (int)st->b.size();
  if (bboxes)
    for (int i = 0; i < std::min(n, cap); ++i)
      st->b[i].to(bboxes + 9 * i);
  return n;
}

// Regridding schedule (dh.cc:1712-1850): the active and overlap points of
// every new component, first from the owned regions of the same level's old
// components (fast_old2new_sync_sendrecv), the rest by prolongation from
// the new coarser level (fast_old2new_ref_prol_sendrecv).
extern "C" int cg_schedule_regrid(cg_schedule *s, const cg_schedule *old) {
  if (!s)
    return set_error(CG_ERR_ARG, "null schedule");
  if (old && old->centering != s->centering)
    return set_error(CG_ERR_ARG, "old and new hierarchies differ in centering");
  const bool have_old = old && !old->levels.empty();
  if (old)
    for (size_t rl = 0; rl < std::min(s->levels.size(), old->levels.size()); ++rl)
      for (int d = 0; d < 3; ++d)
        if (old->levels[rl].domain.st[d] != s->levels[rl].domain.st[d])
          return set_error(CG_ERR_STRIDE, "old and new hierarchies differ in the stride of level %zu", rl);
  for (size_t rl = 0; rl < s->levels.size(); ++rl) {
    Level &L = s->levels[rl];
    L.lists[CG_SCHED_OLD2NEW_SYNC].clear();
    L.lists[CG_SCHED_OLD2NEW_PROL].clear();
  }
  for (size_t rl = 0; rl < s->levels.size(); ++rl) {
    Level &L = s->levels[rl];
    for (size_t c = 0; c < L.comps.size(); ++c) {
      const Comp &box = L.comps[c];
      Set needrecv = box.active | box.overlaps;
      if (old && rl < old->levels.size()) {
        const Level &OL = old->levels[rl];
        for (size_t cc = 0; cc < OL.comps.size() && !needrecv.empty(); ++cc) {
          const Set ovlp = needrecv & OL.comps[cc].owned;
          for (const IB &recv : ovlp.b)
            L.lists[CG_SCHED_OLD2NEW_SYNC].push_back(SR{recv, (int)cc, recv, (int)c});
          needrecv = needrecv - ovlp;
        }
      }
      if (rl > 0 && !needrecv.empty()) {
        Set left;
        if (int e = prolong_list(*s, (int)rl, (int)c, needrecv, L.lists[CG_SCHED_OLD2NEW_PROL], nullptr,
                                 "Regridding prolongation", &left))
          return e;
        needrecv = left;
      }
      if (have_old)
        SCHED_CHECK(needrecv.empty(),
                    "Regridding prolongation: All points must have been received (level %zu, component %zu)",
                    rl, c);
    }
  }
  return CG_OK;
}
