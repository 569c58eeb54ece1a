"""Oracle for the communication schedule: dh::regrid's region description
and communication lists (CarpetLib/src/dh.cc:211-1080), restated on boolean
masks over each level's lattice.

Test infrastructure only: tests/ use it to check the library's box-algebra
implementation (carpet_amd/csrc/cg_schedule.cpp); nothing in the product
imports it.

Every region set is a boolean array over the lattice points of one level
(stride s, points o + s*k), so each set operation is written in its
pointwise meaning, independent of any decomposition into boxes:
  expand(lo, hi)      dilation by [-lo, hi] lattice steps (bboxset2.hh:1141)
  expanded_for(t)     every point maps to the target lattice points that
                      bracket it per axis (bbox.cc:209-220, applied per box,
                      bboxset2.hh:1172-1183)
  contracted_for(t)   target points whose source points within an open
                      distance of one source stride, per axis, all lie in
                      the set (the complement construction of
                      bboxset2.hh:1186-1199)
The list construction follows dh.cc line by line: refinement prolongation
(766-830), synchronisation (834-890), boundary prolongation (894-975) and
refinement restriction (981-1080), including the order in which source
components are tried (the first one that covers a point sends it).
"""
from dataclasses import dataclass, field

import numpy as np

VERTEX, CELL = 1, 2
MARGIN = 16  # lattice points beyond the domain on each side


class ScheduleError(Exception):
    pass


def _check(cond, msg):
    if not cond:
        raise ScheduleError(msg)


@dataclass
class Grid:
    """Lattice points o + s*k, k in [k0, k0 + n) per axis."""
    s: tuple
    k0: tuple
    n: tuple

    @staticmethod
    def around(box, margin=MARGIN):
        lo, up, st = box
        k0 = tuple(lo[d] // st[d] - margin for d in range(3))
        n = tuple((up[d] - lo[d]) // st[d] + 1 + 2 * margin for d in range(3))
        off = tuple(lo[d] % st[d] for d in range(3))
        return Grid(tuple(st), k0, n), off

    def coords(self, off, d):
        return off[d] + self.s[d] * (self.k0[d] + np.arange(self.n[d]))


@dataclass
class Lat:
    """A level's lattice: grid plus the offset o (0 <= o < s)."""
    g: Grid
    off: tuple

    def empty(self):
        return np.zeros(self.g.n, dtype=bool)

    def x(self, d):
        return self.g.coords(self.off, d)

    def box(self, b):
        """Mask of the lattice points of box b = (lo, up, stride)."""
        lo, up, st = b
        m = self.empty()
        if any(lo[d] > up[d] for d in range(3)):
            return m
        _check(tuple(st) == self.g.s, f"box stride {st} is not the lattice stride {self.g.s}")
        sel = []
        for d in range(3):
            xd = self.x(d)
            _check((lo[d] - self.off[d]) % st[d] == 0, "box not aligned with the lattice")
            _check(xd[0] <= lo[d] and up[d] <= xd[-1], "box outside the oracle's grid; widen MARGIN")
            sel.append((xd >= lo[d]) & (xd <= up[d]))
        m[np.ix_(*sel)] = True
        return m

    def expand(self, m, lo, hi):
        """Dilation by lo/hi lattice steps per axis (negative: erosion)."""
        r = m
        for d in range(3):
            a, b = -lo[d], hi[d]
            if a <= b:
                acc = np.zeros_like(r)
                for t in range(a, b + 1):
                    acc |= _shift(r, d, t)
                r = acc
            else:  # erosion: all shifts must hold
                acc = np.ones_like(r)
                for t in range(b, a + 1):
                    acc &= _shift(r, d, t, fill=False)
                r = acc
        return r

    def expanded_for(self, m, tl):
        """Target-lattice points bracketing the set's points."""
        r = tl.empty()
        idx = np.nonzero(m)
        if idx[0].size == 0:
            return r
        t = []
        for d in range(3):
            x = self.x(d)[idx[d]]
            s, o = tl.g.s[d], tl.off[d]
            lo = x - np.mod(x - o, s)
            hi = x + np.mod(s - np.mod(x - o, s), s)
            klo = (lo - o) // s - tl.g.k0[d]
            khi = (hi - o) // s - tl.g.k0[d]
            _check(klo.min() >= 0 and khi.max() < tl.g.n[d], "expanded_for leaves the oracle's grid")
            t.append((klo, khi))
        for cx in (0, 1):
            for cy in (0, 1):
                for cz in (0, 1):
                    r[t[0][cx], t[1][cy], t[2][cz]] = True
        return r

    def contracted_for(self, m, tl):
        """Target points whose open neighbourhood of one source stride, per
        axis, lies in the set."""
        axes = []
        for d in range(3):
            q = tl.x(d)
            s, o = self.g.s[d], self.off[d]
            # source points x with |x - q| < s: x0 = largest x <= q, and x0 + s
            # when q is strictly between
            x0 = q - np.mod(q - o, s)
            x1 = np.where(np.mod(q - o, s) == 0, x0, x0 + s)
            k0 = (x0 - o) // s - self.g.k0[d]
            k1 = (x1 - o) // s - self.g.k0[d]
            axes.append((k0, k1))
        r = np.ones(tl.g.n, dtype=bool)
        for cx in (0, 1):
            for cy in (0, 1):
                for cz in (0, 1):
                    ks = [axes[0][cx], axes[1][cy], axes[2][cz]]
                    ok = [(k >= 0) & (k < self.g.n[d]) for d, k in enumerate(ks)]
                    kc = [np.clip(k, 0, self.g.n[d] - 1) for d, k in enumerate(ks)]
                    inside = ok[0][:, None, None] & ok[1][None, :, None] & ok[2][None, None, :]
                    v = m[np.ix_(*kc)] & inside
                    r &= v
        return r


def _shift(m, d, t, fill=False):
    """out[k] = m[k - t] along axis d (points move by +t)."""
    out = np.full_like(m, fill)
    n = m.shape[d]
    if abs(t) >= n:
        return out
    src = [slice(None)] * 3
    dst = [slice(None)] * 3
    if t >= 0:
        src[d] = slice(0, n - t)
        dst[d] = slice(t, n)
    else:
        src[d] = slice(-t, n)
        dst[d] = slice(0, n + t)
    out[tuple(dst)] = m[tuple(src)]
    return out


def box_expand(b, lo, hi):
    l, u, s = b
    return (tuple(l[d] - lo[d] * s[d] for d in range(3)), tuple(u[d] + hi[d] * s[d] for d in range(3)), tuple(s))


@dataclass
class Comp:
    interior: tuple
    exterior: tuple = None
    owned: tuple = None
    communicated: tuple = None
    ghosts: np.ndarray = None
    buffers: np.ndarray = None
    overlaps: np.ndarray = None
    active: np.ndarray = None
    sync: np.ndarray = None
    bndref: np.ndarray = None


@dataclass
class Level:
    domain: tuple
    interiors: list
    reffact: tuple
    gw: tuple  # ((lo x,y,z), (hi x,y,z))
    bw: tuple
    ow: tuple
    order: int
    lat: Lat = None
    comps: list = field(default_factory=list)
    allactive: np.ndarray = None
    # lists: name -> {(src, dst): recv mask}
    lists: dict = field(default_factory=dict)


def _add(lst, key, m):
    if key in lst:
        lst[key] = lst[key] | m
    else:
        lst[key] = m.copy()


def build(centering, boundary, levels, ho_restrict=False, restrict_order=1, staggered=False):
    """levels: list of Level (domain, interiors, reffact, widths, order).
    Returns the levels with regions and lists filled (masks)."""
    bnd = boundary
    for L in levels:
        g, off = Grid.around(L.domain)
        L.lat = Lat(g, off)
    rel = [tuple(1 for _ in range(3))]
    for rl in range(1, len(levels)):
        _check(all(levels[rl].reffact[d] % levels[rl - 1].reffact[d] == 0 for d in range(3)),
               "Refinement factors must be integer multiples of each other")
        rel.append(tuple(levels[rl].reffact[d] // levels[rl - 1].reffact[d] for d in range(3)))
    # regions (dh.cc:283-618)
    for L in levels:
        lat = L.lat
        trivial = all(v == 0 for w in (bnd, L.gw, L.bw, L.ow) for f in w for v in f)
        dom = L.domain
        act = box_expand(dom, [-v for v in bnd[0]], [-v for v in bnd[1]])
        dom_m, act_m = lat.box(dom), lat.box(act)
        _check(not (act_m & ~dom_m).any(), "The active part of the domain must be contained in the exterior")
        L.comps = []
        for b in L.interiors:
            k = Comp(interior=b)
            intr_m = lat.box(b)
            _check(not (intr_m & ~dom_m).any(), "The interior must be contained in the domain")
            outer = [[b[0][d] == dom[0][d] for d in range(3)], [b[1][d] == dom[1][d] for d in range(3)]]
            glo = [0 if outer[0][d] else L.gw[0][d] for d in range(3)]
            ghi = [0 if outer[1][d] else L.gw[1][d] for d in range(3)]
            blo = [-bnd[0][d] if outer[0][d] else 0 for d in range(3)]
            bhi = [-bnd[1][d] if outer[1][d] else 0 for d in range(3)]
            k.exterior = b if trivial else box_expand(b, glo, ghi)
            ext_m = lat.box(k.exterior)
            _check(not (ext_m & ~dom_m).any(), "The exterior must be contained in the domain")
            k.ghosts = lat.empty() if trivial else ext_m & ~intr_m
            k.communicated = k.exterior if trivial else box_expand(k.exterior, blo, bhi)
            comm_m = lat.box(k.communicated)
            _check(not (comm_m & ~act_m).any(), "The communicated region must be contained in the active part")
            k.owned = b if trivial else box_expand(b, blo, bhi)
            own_m = lat.box(k.owned)
            _check(not (own_m & ~act_m).any(), "The owned region must be contained in the active part")
            bound = lat.empty() if trivial else comm_m & ~own_m
            _check(not (bound & ~act_m).any(), "The boundary must be contained in the active part")
            L.comps.append(k)
        allowned = lat.empty()
        for k in L.comps:
            om = lat.box(k.owned)
            _check(not (allowned & om).any(), "All owned regions must be disjoint")
            allowned |= om
        notowned = act_m & ~allowned
        nal = [L.bw[0][d] + L.ow[0][d] for d in range(3)]
        nah = [L.bw[1][d] + L.ow[1][d] for d in range(3)]
        notactive = lat.expand(notowned, nal, nah)
        allbuffers = allowned & lat.expand(notowned, L.bw[0], L.bw[1])
        alloverlaps = (allowned & notactive) & ~allbuffers
        L.allactive = allowned & ~notactive
        for k in L.comps:
            om = lat.box(k.owned)
            if trivial:
                k.buffers, k.overlaps, k.active = lat.empty(), lat.empty(), om
            else:
                k.buffers, k.overlaps, k.active = om & allbuffers, om & alloverlaps, om & L.allactive
            k.sync = lat.empty()
            k.bndref = lat.empty()
        L.lists = {"ref_prol": {}, "sync": {}, "ref_bnd_prol": {}, "ref_rest": {}, "old2new_sync": {},
                   "old2new_prol": {}}

    def prolong(rl, c, needrecv, name, filled):
        L, O = levels[rl], levels[rl - 1]
        box = L.comps[c]
        e = [rel[rl][d] if centering == VERTEX else rel[rl][d] - 1 for d in range(3)]
        sten = (L.order // 2,) * 3
        for cc, obox in enumerate(O.comps):
            if not needrecv.any():
                break
            exp = L.lat.expand(O.lat.expanded_for(obox.active, L.lat), e, e)
            ovlp = needrecv & exp
            if ovlp.any():
                # every receive point's stencil inside the source exterior
                # (the per-box assertion of dh.cc:810-813, pointwise)
                foot = O.lat.expand(L.lat.expanded_for(ovlp, O.lat), sten, sten)
                _check(not (foot & ~O.lat.box(obox.exterior)).any(),
                       f"{name}: Send region must be contained in exterior")
                _add(L.lists[name], (cc, c), ovlp)
            needrecv = needrecv & ~ovlp
            if filled is not None:
                filled |= ovlp
        _check(not needrecv.any(), f"{name}: All points must have been received")

    for rl, L in enumerate(levels):
        lat = L.lat
        for c, box in enumerate(L.comps):
            if rl > 0:
                prolong(rl, c, box.active | box.overlaps, "ref_prol", None)
            needrecv = box.ghosts.copy()
            for cc, obox in enumerate(L.comps):
                if not needrecv.any():
                    break
                ovlp = needrecv & lat.box(obox.interior)
                _check(cc != c or not ovlp.any(), "A region may not synchronise from itself")
                if ovlp.any():
                    _add(L.lists["sync"], (cc, c), ovlp)
                needrecv &= ~ovlp
                box.sync |= ovlp
            if rl > 0:
                nr = ((box.ghosts & ~box.sync) & lat.box(box.communicated)) | box.buffers
                prolong(rl, c, nr, "ref_bnd_prol", box.bndref)
        if rl > 0:
            O = levels[rl - 1]
            olat = O.lat
            if centering == VERTEX:
                allrestricted = lat.contracted_for(L.allactive, olat)
            else:
                all_target = olat.box(olat_box_all(olat))
                tmp1 = lat.expanded_for(L.allactive, olat)
                tmp2 = all_target & ~tmp1
                tmp3 = olat.expand(tmp2, (1, 1, 1), (1, 1, 1))
                allrestricted = all_target & ~tmp3
            # dh.cc:1019-1050: support_staggered_operators also restricts into
            # interiors only, from sources shrunk by 2
            shrink = restrict_order // 2 if (ho_restrict and centering == CELL) else 2 if staggered else 0
            for oc, obox in enumerate(O.comps):
                tgt = obox.interior if (ho_restrict or staggered) else obox.exterior
                needrecv = allrestricted & olat.box(tgt)
                _check(not (allrestricted & obox.buffers).any(),
                       "Refinement restriction: cannot restrict into buffer zones")
                for c, box in enumerate(L.comps):
                    if not needrecv.any():
                        break
                    # a box rule here (bbox::contracted_for), dh.cc:1047-1050
                    cext = olat.box(_contract_box(box_expand(box.exterior, [-shrink] * 3, [-shrink] * 3),
                                                  O.domain))
                    ovlp = needrecv & cext
                    if ovlp.any():
                        kids = lat.expand(olat.expanded_for(ovlp, lat), (shrink,) * 3, (shrink,) * 3)
                        _check(not (kids & ~lat.box(box.exterior)).any(),
                               "Refinement restriction: Send region must be contained in exterior")
                        _add(O.lists["ref_rest"], (c, oc), ovlp)
                    needrecv &= ~ovlp
                _check(not needrecv.any(), "Refinement restriction: All points must have been received")
    return levels


def olat_box_all(lat):
    """The whole grid of a lattice as a box."""
    lo = tuple(int(lat.x(d)[0]) for d in range(3))
    up = tuple(int(lat.x(d)[-1]) for d in range(3))
    return (lo, up, lat.g.s)


def _contract_box(b, t):
    """Coarse lattice points within the extent of box b (the box rule of
    bbox.cc:222-233 for one box, restated as a range)."""
    lo, up, _ = b
    tlo, _, ts = t
    rlo, rup = [], []
    for d in range(3):
        s = ts[d]
        a = lo[d] + (-(lo[d] - tlo[d])) % s
        z = up[d] - (up[d] - tlo[d]) % s
        rlo.append(a)
        rup.append(z)
    return (tuple(rlo), tuple(rup), tuple(ts))


def regrid(centering, new_levels, old_levels):
    """Regridding schedule (dh.cc:1712-1850) of new_levels (built by build())
    after a regrid from old_levels (built, or None): per new component the
    active and overlap points, first from the owned regions of the old
    components of the same level ("old2new_sync", keys (old component, new
    component)), the rest by prolongation from the new coarser level
    ("old2new_prol")."""
    have_old = old_levels is not None and len(old_levels) > 0
    rel = [None] + [tuple(new_levels[rl].reffact[d] // new_levels[rl - 1].reffact[d] for d in range(3))
                    for rl in range(1, len(new_levels))]
    for rl, L in enumerate(new_levels):
        lat = L.lat
        L.lists["old2new_sync"] = {}
        L.lists["old2new_prol"] = {}
        for c, box in enumerate(L.comps):
            needrecv = box.active | box.overlaps
            if have_old and rl < len(old_levels):
                for cc, obox in enumerate(old_levels[rl].comps):
                    if not needrecv.any():
                        break
                    ovlp = needrecv & lat.box(obox.owned)
                    if ovlp.any():
                        _add(L.lists["old2new_sync"], (cc, c), ovlp)
                    needrecv &= ~ovlp
            if rl > 0 and needrecv.any():
                O = new_levels[rl - 1]
                e = [rel[rl][d] if centering == VERTEX else rel[rl][d] - 1 for d in range(3)]
                sten = (L.order // 2,) * 3
                for cc, obox in enumerate(O.comps):
                    if not needrecv.any():
                        break
                    exp = lat.expand(O.lat.expanded_for(obox.active, lat), e, e)
                    ovlp = needrecv & exp
                    if ovlp.any():
                        foot = O.lat.expand(lat.expanded_for(ovlp, O.lat), sten, sten)
                        _check(not (foot & ~O.lat.box(obox.exterior)).any(),
                               "Regridding prolongation: Send region must be contained in exterior")
                        _add(L.lists["old2new_prol"], (cc, c), ovlp)
                    needrecv &= ~ovlp
            if have_old:
                _check(not needrecv.any(), "Regridding prolongation: All points must have been received")
    return new_levels
