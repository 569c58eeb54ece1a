"""Communication schedule (dh::regrid's region lists, CarpetLib/src/dh.cc:
211-1080): the library's box-algebra implementation (cg_schedule_build)
against the mask oracle (oracle/schedule_oracle.py) on random nested
hierarchies, and the oracle pinned by hand-computed hierarchies and by the
single-box rules of bbox.cc.  Host code only: no GPU."""
import numpy as np
import pytest

import schedule_oracle as so

cg = pytest.importorskip("carpet_amd")
from carpet_amd import schedule as S  # noqa: E402

V, CC = so.VERTEX, so.CELL
Z3 = (0, 0, 0)


def W(n):
    return ((n, n, n), (n, n, n))


# ---------------------------------------------------------------- oracle pins
def test_oracle_sync_two_components():
    """Two components splitting a 32^3 domain in x, ghost width 3: each
    receives the 3-plane slab next to it from the other (hand computed)."""
    lv = so.Level(((0, 0, 0), (31, 31, 31), (1, 1, 1)),
                  [((0, 0, 0), (15, 31, 31), (1, 1, 1)), ((16, 0, 0), (31, 31, 31), (1, 1, 1))],
                  (1, 1, 1), W(3), W(0), W(0), 1)
    so.build(V, (Z3, Z3), [lv])
    lat = lv.lat
    sync = lv.lists["sync"]
    assert set(sync) == {(1, 0), (0, 1)}
    assert np.array_equal(sync[(1, 0)], lat.box(((16, 0, 0), (18, 31, 31), (1, 1, 1))))
    assert np.array_equal(sync[(0, 1)], lat.box(((13, 0, 0), (15, 31, 31), (1, 1, 1))))
    # outer faces: no ghosts, so the exteriors grow only inwards
    assert lv.comps[0].exterior == ((0, 0, 0), (18, 31, 31), (1, 1, 1))


@pytest.mark.parametrize("cent", [V, CC])
def test_oracle_nested_closed_forms(cent):
    """One fine component in the middle of a coarse one: boundary
    prolongation fills the ghost shell plus the buffer zones, restriction
    fills the coarse points under the fine active region; sizes in closed
    form."""
    if cent == V:
        cdom = ((0, 0, 0), (64, 64, 64), (2, 2, 2))
        fdom = ((0, 0, 0), (64, 64, 64), (1, 1, 1))
        fin = ((16, 16, 16), (48, 48, 48), (1, 1, 1))  # 33 points per axis
    else:  # cell centres: fine at odd, coarse at 2 mod 4
        cdom = ((2, 2, 2), (62, 62, 62), (4, 4, 4))
        fdom = ((1, 1, 1), (63, 63, 63), (2, 2, 2))
        fin = ((17, 17, 17), (47, 47, 47), (2, 2, 2))  # 16 cells per axis
    L0 = so.Level(cdom, [cdom], (1, 1, 1), W(3), W(0), W(0), 5)
    L1 = so.Level(fdom, [fin], (2, 2, 2), W(3), W(6), W(0), 5)
    so.build(cent, (Z3, Z3), [L0, L1])
    n = 33 if cent == V else 16
    # owned = interior (no outer boundary); active = owned minus 6 buffer layers
    assert L1.comps[0].active.sum() == (n - 12) ** 3
    assert L1.comps[0].buffers.sum() == n ** 3 - (n - 12) ** 3
    bnd = L1.lists["ref_bnd_prol"][(0, 0)]
    assert bnd.sum() == (n + 6) ** 3 - (n - 12) ** 3
    rest = L0.lists["ref_rest"][(0, 0)]
    if cent == V:  # coarse points (even) inside the fine active box [22, 42]
        assert rest.sum() == 11 ** 3
    else:  # the coarse cells whose two children per axis are both active:
        # fine cells 29..35 are coarse cells 30 and 34
        assert rest.sum() == 2 ** 3
    # refinement prolongation fills the active points
    assert L1.lists["ref_prol"][(0, 0)].sum() == (n - 12) ** 3


def test_oracle_cell_restriction_interior():
    """Cell centring: the restricted coarse cells are exactly those whose
    children are all fine active cells (the bracketing expanded_for adds one
    coarse layer, which dh.cc:1000-1011 removes again)."""
    cdom = ((2, 2, 2), (126, 126, 126), (4, 4, 4))
    fdom = ((1, 1, 1), (127, 127, 127), (2, 2, 2))
    fin = ((17, 17, 17), (111, 111, 111), (2, 2, 2))  # fine cells 17..111: 48 cells
    L0 = so.Level(cdom, [cdom], (1, 1, 1), W(3), W(0), W(0), 5)
    L1 = so.Level(fdom, [fin], (2, 2, 2), W(3), W(6), W(0), 5)
    so.build(CC, (Z3, Z3), [L0, L1])
    # fine active: 48 - 12 = 36 cells = 18 coarse cells per axis
    assert L0.lists["ref_rest"][(0, 0)].sum() == 18 ** 3


@pytest.mark.parametrize("seed", range(6))
def test_oracle_set_rules_match_box_rules(seed):
    """For one box the pointwise set rules reduce to bbox.cc's box rules:
    expanded_for = the smallest target box around it, contracted_for = the
    target points within its extent (both centrings)."""
    rng = np.random.default_rng(seed)
    for cent in (V, CC):
        if cent == V:
            fine, coarse = ((0, 0, 0), (60, 60, 60), (1, 1, 1)), ((0, 0, 0), (60, 60, 60), (2, 2, 2))
        else:
            fine, coarse = ((1, 1, 1), (59, 59, 59), (2, 2, 2)), ((2, 2, 2), (58, 58, 58), (4, 4, 4))
        gf, of = so.Grid.around(fine)
        gc, oc = so.Grid.around(coarse)
        lf, lc = so.Lat(gf, of), so.Lat(gc, oc)
        for _ in range(5):
            lo = [int(v) for v in rng.integers(0, 20, 3)]
            n = [int(v) for v in rng.integers(1, 12, 3)]
            s = fine[2][0]
            lo = [fine[0][d] + s * lo[d] for d in range(3)]
            b = (tuple(lo), tuple(lo[d] + s * n[d] for d in range(3)), fine[2])
            m = lf.box(b)
            got = lf.contracted_for(m, lc)
            exp = lc.box(so._contract_box(b, coarse))
            assert np.array_equal(got, exp)
            up = lf.expanded_for(m, lc)
            # smallest coarse box around b: every coarse point whose bracket
            # meets b, found by brute force along each axis
            lo2, up2 = [], []
            for d in range(3):
                xs = lc.x(d)
                ok = [q for q in xs if any(abs(q - x) < coarse[2][d] for x in range(b[0][d], b[1][d] + 1, s))]
                lo2.append(min(ok))
                up2.append(max(ok))
            assert np.array_equal(up, lc.box((tuple(lo2), tuple(up2), coarse[2])))


# ------------------------------------------------ library against the oracle
def _random_hierarchy(rng, cent, nlev):
    """Nested components on nlev levels (refinement factor 2 per level),
    each level's components a tiling of a box well inside the previous
    level's, so that every dh assertion holds."""
    ns = 2 ** (nlev - 1)
    levels = []
    # two levels: any order and buffer width; three levels: small margins
    # (the oracle's masks grow with the finest level's size)
    order = int(rng.choice([1, 3, 5] if nlev == 2 else [1, 3]))
    gw = int(rng.integers(order // 2 + 1, 4))
    bw = int(rng.choice([0, gw, 2 * gw] if nlev == 2 else [0, gw]))
    ow = int(rng.choice([0, 1]))
    margin = 2 * (gw + bw + ow + order // 2) + 4
    ncoarse = margin + int(rng.integers(8, 12)) + (margin // 2 + 6 if nlev == 3 else 0)
    prev = None
    for rl in range(nlev):
        s = ns // 2 ** rl
        if cent == V:
            dom = (Z3, (ncoarse * ns,) * 3, (s,) * 3)
            off = 0
        else:
            dom = ((s // 2,) * 3, (ncoarse * ns - s // 2,) * 3, (s,) * 3)
            off = s // 2
        npts = (dom[1][0] - dom[0][0]) // s + 1
        if rl == 0:
            blo, bhi = [0, 0, 0], [npts - 1] * 3
        else:
            # inside the previous region, margin of 6 coarse points
            plo, phi = prev
            m = 2 * (gw + bw + ow + order // 2) + 4
            blo, bhi = [], []
            for d in range(3):
                a0, a1 = 2 * plo[d] + m, 2 * phi[d] - m
                if cent == CC:
                    a1 += 1
                if a1 - a0 < 12:
                    return None
                if rl < nlev - 1:  # room for the next level inside
                    lo = int(rng.integers(a0, a0 + (a1 - a0) // 4 + 1))
                    hi = int(rng.integers(a1 - (a1 - a0) // 4, a1 + 1))
                else:
                    lo = int(rng.integers(a0, a0 + (a1 - a0) // 3 + 1))
                    hi = int(rng.integers(lo + 5, a1 + 1))
                blo.append(lo)
                bhi.append(hi)
        # tile the box into 1-3 components along a random axis
        ax = int(rng.integers(0, 3))
        nt = int(rng.integers(1, 4))
        cuts = sorted(set(int(v) for v in rng.integers(blo[ax] + 2, bhi[ax] - 1, nt - 1))) if nt > 1 else []
        edges = [blo[ax]] + [c for c in cuts] + [bhi[ax] + 1]
        comps = []
        for a, b in zip(edges[:-1], edges[1:]):
            if b - a < 1:
                continue
            lo = list(blo)
            hi = list(bhi)
            lo[ax], hi[ax] = a, b - 1
            comps.append((tuple(off + s * v for v in lo), tuple(off + s * v for v in hi), (s,) * 3))
        prev = (blo, bhi)
        levels.append(dict(domain=dom, comps=comps, reffact=(2 ** rl,) * 3, gw=W(gw), bw=W(bw if rl else 0),
                           ow=W(ow if rl else 0), order=order))
    return levels


def _masks(lat, boxes):
    m = lat.empty()
    for b in boxes:
        bm = lat.box(b)
        assert not (m & bm).any(), "boxes of one set overlap"
        m |= bm
    return m


def _compare(levels, cent, bnd, ho, rorder, staggered=False):
    descs = [S.LevelDesc(L["domain"], L["comps"], L["reffact"], L["gw"], L["bw"], L["ow"], L["order"])
             for L in levels]
    olv = [so.Level(L["domain"], L["comps"], L["reffact"], L["gw"], L["bw"], L["ow"], L["order"]) for L in levels]
    try:
        so.build(cent, bnd, olv, ho, rorder, staggered)
        oerr = None
    except so.ScheduleError as e:
        oerr = e
    cg.set_staggered_operators(staggered)
    try:
        sch = S.Schedule(cent, bnd, descs, ho, rorder)
        perr = None
    except cg.CarpetError as e:
        perr = e
    finally:
        cg.set_staggered_operators(False)
    assert (oerr is None) == (perr is None), f"oracle: {oerr}; library: {perr}"
    if oerr is not None:
        return False
    for rl, L in enumerate(olv):
        lat = L.lat
        for c, k in enumerate(L.comps):
            assert sch.component(rl, c, S.EXTERIOR) == k.exterior
            assert sch.component(rl, c, S.OWNED) == k.owned
            assert sch.component(rl, c, S.COMMUNICATED) == k.communicated
            for which, m in ((S.GHOSTS, k.ghosts), (S.BUFFERS, k.buffers), (S.OVERLAPS, k.overlaps),
                             (S.ACTIVE, k.active), (S.SYNCED, k.sync), (S.BNDREF, k.bndref)):
                assert np.array_equal(_masks(lat, sch.region_set(rl, c, which)), m), (rl, c, which)
        for kind, name in S.LIST_NAMES.items():
            got = {}
            for r in sch.lists[rl][kind]:
                key = (r.src, r.dst)
                m = lat.box(r.recv)
                if key in got:
                    assert not (got[key] & m).any(), "receive boxes overlap"
                    got[key] |= m
                else:
                    got[key] = m
                # send boxes: sync sends what it receives; prolongation and
                # restriction send the receive box's footprint, inside the
                # source's exterior
                if kind == S.SYNC:
                    assert r.send == r.recv
                elif kind in (S.REF_PROL, S.REF_BND_PROL):
                    O = olv[rl - 1]
                    ext = O.lat.box(O.comps[r.src].exterior)
                    send = O.lat.box(r.send)
                    assert not (send & ~ext).any()
                    foot = O.lat.expand(lat.expanded_for(m, O.lat), (L.order // 2,) * 3, (L.order // 2,) * 3)
                    assert not (foot & ~send).any(), "prolongation send box misses stencil points"
                else:
                    F = olv[rl + 1]
                    ext = F.lat.box(F.comps[r.src].exterior)
                    send = F.lat.box(r.send)
                    assert not (send & ~ext).any()
                    kids = lat.expanded_for(m, F.lat)
                    assert not (kids & ~send).any(), "restriction send box misses children"
            exp = L.lists[name]
            assert set(got) == set(exp), (rl, name, set(got), set(exp))
            for key in exp:
                assert np.array_equal(got[key], exp[key]), (rl, name, key)
    sch.close()
    return True


@pytest.mark.parametrize("cent", [V, CC])
@pytest.mark.parametrize("seed", range(16))
def test_schedule_matches_oracle(cent, seed):
    rng = np.random.default_rng(100 * cent + seed)
    nlev = 3 if seed % 4 == 3 else 2
    levels = None
    for _ in range(500):
        levels = _random_hierarchy(rng, cent, nlev)
        if levels is not None:
            break
    assert levels is not None, "no valid random hierarchy drawn" 
    bnd = (Z3, Z3) if seed % 2 == 0 else (tuple(int(v) for v in rng.integers(0, 3, 3)), Z3)
    ho = cent == CC and seed % 3 == 1
    rorder = 3 if ho else 1
    print(f"{'vertex' if cent == V else 'cell'} seed {seed}: {nlev} levels, "
          f"{[len(L['comps']) for L in levels]} components, boundary {bnd}, ho {ho}", flush=True)
    _compare(levels, cent, bnd, ho, rorder)


@pytest.mark.parametrize("seed", range(8))
def test_schedule_staggered_matches_oracle(seed):
    """support_staggered_operators: restriction into interiors only, from
    sources shrunk by 2 (dh.cc:1019-1050), library against the oracle."""
    rng = np.random.default_rng(7000 + seed)
    nlev = 3 if seed % 4 == 3 else 2
    levels = None
    for _ in range(500):
        levels = _random_hierarchy(rng, V, nlev)
        if levels is not None:
            break
    assert levels is not None, "no valid random hierarchy drawn"
    bnd = (Z3, Z3) if seed % 2 == 0 else (tuple(int(v) for v in rng.integers(0, 3, 3)), Z3)
    print(f"staggered seed {seed}: {nlev} levels, {[len(L['comps']) for L in levels]} components", flush=True)
    _compare(levels, V, bnd, False, 1, staggered=True)


def test_schedule_rejects_what_dh_asserts():
    """A fine component whose ghost zones reach outside the coarse active
    region cannot be boundary-prolongated: both report the failed check."""
    cdom = (Z3, (32, 32, 32), (2, 2, 2))
    fdom = (Z3, (32, 32, 32), (1, 1, 1))
    levels = [dict(domain=cdom, comps=[((0, 0, 0), (12, 32, 32), (2, 2, 2))], reffact=(1, 1, 1), gw=W(3),
                   bw=W(0), ow=W(0), order=5),
              dict(domain=fdom, comps=[((4, 4, 4), (20, 20, 20), (1, 1, 1))], reffact=(2, 2, 2), gw=W(3),
                   bw=W(0), ow=W(0), order=5)]
    assert _compare(levels, V, (Z3, Z3), False, 1) is False
    descs = [S.LevelDesc(L["domain"], L["comps"], L["reffact"], L["gw"], L["bw"], L["ow"], L["order"])
             for L in levels]
    with pytest.raises(cg.CarpetError) as e:
        S.Schedule(V, (Z3, Z3), descs)
    assert e.value.code == -16


@pytest.mark.parametrize("cent", [V, CC])
def test_hierarchy_step_transfers_are_consistent(cent):
    """The level step built from a schedule (workloads.hierarchy_step): every
    destination and source region lies in its array, and within one phase
    no two transfers write the same point of one array (dh's lists are
    disjoint per destination)."""
    from carpet_amd.workloads import hierarchy_step, nested_levels
    lv = nested_levels(ncoarse=32, nlevels=3, split=(2, 1, 1), centering=cent, buffer=6, order=5 if cent == V else 3)
    w = hierarchy_step(lv, cent, n_gf=1)
    written = {}
    for t in w.steps:
        dbox = w.arrays[t.dst].box
        assert all(dl <= rl and ru <= du for dl, rl, ru, du in zip(dbox.lower, t.dstreg.lower, t.dstreg.upper, dbox.upper))
        sbox = w.arrays[t.srcs[0]].box
        assert all(sl <= rl and ru <= su for sl, rl, ru, su in zip(sbox.lower, t.srcreg.lower, t.srcreg.upper, sbox.upper))
        key = (t.phase, t.dst)
        pts = set(itertools_product(t.dstreg))
        assert not (written.get(key, set()) & pts), f"overlapping writes in phase {t.phase} to {t.dst}"
        written.setdefault(key, set()).update(pts)
    assert {t.phase for t in w.steps} >= {0, 2, 4, 5, 7, 8}


def itertools_product(b):
    import itertools
    rng = [range(lo, up + 1, st) for lo, up, st in zip(b.lower, b.upper, b.stride)]
    return itertools.product(*rng)


def _olv(descs):
    return [so.Level(L.domain, L.components, L.reffact, L.ghost_width, L.buffer_width, L.overlap_width,
                     L.prolongation_order) for L in descs]


@pytest.mark.parametrize("cent", [V, CC])
@pytest.mark.parametrize("seed", range(8))
def test_regrid_schedule_matches_oracle(cent, seed):
    """Regridding lists (dh.cc:1712-1850) from an old to a new nested
    hierarchy: copies from the old components of each level, prolongation
    from the new coarser level for the rest; library against the oracle."""
    from carpet_amd.workloads import nested_levels
    rng = np.random.default_rng(500 + 10 * cent + seed)
    nold = int(rng.integers(1, 4))
    nnew = int(rng.integers(2, 4))
    splits = [(2, 1, 1), (1, 2, 1), (1, 1, 2), (2, 2, 1), (1, 1, 1)]
    order = 5 if cent == V else 3
    kw = dict(ncoarse=int(rng.integers(28, 36)), centering=cent, buffer=int(rng.choice([3, 6])), order=order,
              maxlevels=3)
    old = nested_levels(nlevels=nold, split=splits[int(rng.integers(0, 5))],
                        fraction=float(rng.choice([0.45, 0.5, 0.6])), **kw)
    new = nested_levels(nlevels=nnew, split=splits[int(rng.integers(0, 5))],
                        fraction=float(rng.choice([0.5, 0.6, 0.7])), **kw)
    print(f"old {nold} levels, new {nnew} levels", flush=True)
    bnd = (Z3, Z3)
    so_old = so.build(cent, bnd, _olv(old))
    so_new = so.build(cent, bnd, _olv(new))
    s_old = S.Schedule(cent, bnd, old)
    s_new = S.Schedule(cent, bnd, new)
    try:
        so.regrid(cent, so_new, so_old)
        oerr = None
    except so.ScheduleError as e:
        oerr = e
    try:
        s_new.regrid(s_old)
        perr = None
    except cg.CarpetError as e:
        perr = e
    assert (oerr is None) == (perr is None), f"oracle: {oerr}; library: {perr}"
    if oerr is not None:
        return
    for rl, L in enumerate(so_new):
        lat = L.lat
        for kind, name in ((S.OLD2NEW_SYNC, "old2new_sync"), (S.OLD2NEW_PROL, "old2new_prol")):
            got = {}
            for r in s_new.lists[rl][kind]:
                m = lat.box(r.recv)
                key = (r.src, r.dst)
                assert not (got.get(key, lat.empty()) & m).any()
                got[key] = got.get(key, lat.empty()) | m
                if kind == S.OLD2NEW_SYNC:
                    assert r.send == r.recv
                else:
                    O = so_new[rl - 1]
                    send = O.lat.box(r.send)
                    assert not (send & ~O.lat.box(O.comps[r.src].exterior)).any()
                    foot = O.lat.expand(lat.expanded_for(m, O.lat), (L.order // 2,) * 3, (L.order // 2,) * 3)
                    assert not (foot & ~send).any()
            exp = L.lists[name]
            assert set(got) == set(exp), (rl, name)
            for key in exp:
                assert np.array_equal(got[key], exp[key]), (rl, name, key)
    # every active / overlap point of the new hierarchy is filled exactly once
    for rl, L in enumerate(so_new):
        filled = L.lat.empty()
        for name in ("old2new_sync", "old2new_prol"):
            for m in L.lists[name].values():
                assert not (filled & m).any()
                filled |= m
        want = L.lat.empty()
        for k in L.comps:
            want |= k.active | k.overlaps
        assert np.array_equal(filled, want)


def test_tiled_levels_make_a_valid_hierarchy():
    """The 2000-box hierarchy of the level-step benchmark: 1995 boxes on 5
    levels, every edge in [8,96] points, each level tiling its region
    without gaps or overlaps, nested so that dh's consistency checks pass
    (cg_schedule_build raises on a violation); every component has sync,
    every finer level boundary prolongation and restriction lists."""
    from carpet_amd.workloads import tiled_levels

    lv = tiled_levels()
    assert [len(L.components) for L in lv] == [64, 125, 294, 512, 1000]
    for L in lv:
        pts = 0
        lo = [min(c[0][d] for c in L.components) for d in range(3)]
        hi = [max(c[1][d] for c in L.components) for d in range(3)]
        s = L.components[0][2][0]
        for c in L.components:
            for d in range(3):
                e = (c[1][d] - c[0][d]) // s + 1
                assert 8 <= e <= 96
            pts += np.prod([(c[1][d] - c[0][d]) // s + 1 for d in range(3)])
        assert pts == np.prod([(hi[d] - lo[d]) // s + 1 for d in range(3)])  # a tiling of the bounding box
    sch = S.Schedule(1, ((0, 0, 0), (0, 0, 0)), lv)
    for rl in range(5):
        assert len(sch.lists[rl][S.SYNC]) > 0
        if rl > 0:
            assert len(sch.lists[rl][S.REF_BND_PROL]) > 0
        if rl < 4:
            assert len(sch.lists[rl][S.REF_REST]) > 0
    sch.close()
