"""A level step of a real nested hierarchy, generated from its
communication schedule (dh's lists: restriction, boundary prolongation with
3-level time interpolation, sync; workloads.hierarchy_step), executed by the
batch (one launch per phase and kernel class) and by one cg_transfer_from
per region, both equal to the CPU oracle's replay bit for bit."""
import pytest

from helpers import first_diff, same
from workload_check import run_oracle

cg = pytest.importorskip("carpet_amd")
from carpet_amd.workloads import hierarchy_step, nested_levels  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cent,order,op", [(cg.VERTEX_CENTERED, 5, cg.OP_LAGRANGE), (cg.CELL_CENTERED, 3, cg.OP_LAGRANGE),
                                           (cg.VERTEX_CENTERED, 5, cg.OP_ENO), (cg.CELL_CENTERED, 3, cg.OP_ENO),
                                           (cg.CELL_CENTERED, 5, cg.OP_ENO),
                                           (cg.VERTEX_CENTERED, 5, cg.OP_LAGRANGE_MONOTONE)],
                         ids=["vertex-o5", "cell-o3", "vertex-eno", "cell-eno2", "cell-eno3", "vertex-monotone"])
def test_hierarchy_step_matches_oracle(cent, order, op):
    """Every tensor kind's boundary prolongation shells go through the small
    (8x8- and 4x16-base) tiles as well as the large ones."""
    lv = nested_levels(ncoarse=32, nlevels=3, split=(2, 1, 1), centering=cent, buffer=6, order=order)
    w = hierarchy_step(lv, cent, n_gf=2, op=op)
    print(f"{w.name}: {len(w.steps)} transfers", flush=True)
    ref = run_oracle(w)
    st = cg.Stream()
    dev = w.allocate()
    b = w.batch(dev)
    b.finalize()
    print(b.describe(), flush=True)
    if op != cg.OP_LAGRANGE_MONOTONE:
        # the shells' small tiles: 8x8 bases, and 4x16 for faces thin in x
        assert "[tensor8 " in b.describe() or "[tensor4 " in b.describe()
    b.run(st)
    st.synchronize()
    for k in ref:
        got = dev[k].download()
        if not same(got, ref[k]):
            raise AssertionError(f"batched: {k} differs {first_diff(got, ref[k])}")
    b.close()
    w.free(dev)
    dev = w.allocate()
    w.run_steps(dev)
    cg.synchronize()
    for k in ref:
        got = dev[k].download()
        if not same(got, ref[k]):
            raise AssertionError(f"per region: {k} differs {first_diff(got, ref[k])}")
    w.free(dev)


@pytest.mark.parametrize("mode", [0, 1], ids=["in-turn", "side-by-side"])
def test_phase_modes_match_oracle(mode, tmp_path):
    """The launches of a phase (8x8-base and large tiles of one boundary
    prolongation) run one after the other or side by side, direct and from a
    captured graph; then the tuner chooses per phase (its choices cached)."""
    import re
    lv = nested_levels(ncoarse=32, nlevels=3, split=(2, 1, 1), centering=cg.VERTEX_CENTERED, buffer=6, order=5)
    w = hierarchy_step(lv, cg.VERTEX_CENTERED, n_gf=2)
    ref = run_oracle(w)
    st = cg.Stream()
    dev = w.allocate()
    b = w.batch(dev)
    b.finalize()
    phases = [int(p) for p in re.findall(r"^phase (\d+) ", b.describe(), re.M)]
    multi = sorted({p for p in phases if phases.count(p) > 1})
    assert multi, b.describe()
    for p in multi:
        b.set_phase_mode(p, mode)
    print(b.describe(), flush=True)
    assert ("side by side" if mode else "in turn") in b.describe()
    assert ("in turn" if mode else "side by side") not in b.describe()
    for how in ("run", "graph", "tuned"):
        for k in ref:  # back to the initial state
            dev[k].fill(float("nan"))
            dev[k].fill_hash(w.arrays[k].seed)
        if how == "run":
            b.run(st)
        elif how == "graph":
            b.capture_graph(st)
            b.launch_graph(st)
        else:
            for p in multi:
                b.set_phase_mode(p, -1)
            cache = tmp_path / "tune.txt"
            b.autotune(st, reps=1, cache_path=str(cache))
            assert "phase " in cache.read_text()
            print(b.describe(), flush=True)
            for k in ref:  # the tuner ran the launches on these arrays
                dev[k].fill(float("nan"))
                dev[k].fill_hash(w.arrays[k].seed)
            b.run(st)
        st.synchronize()
        for k in ref:
            got = dev[k].download()
            if not same(got, ref[k]):
                raise AssertionError(f"{how}: {k} differs {first_diff(got, ref[k])}")
    b.close()
    w.free(dev)


@pytest.mark.parametrize("cent,order", [(cg.VERTEX_CENTERED, 5), (cg.CELL_CENTERED, 3)], ids=["vertex-o5", "cell-o3"])
def test_recompose_fill_matches_oracle(cent, order):
    """Regrid from a 2-level to a 3-level hierarchy with regrown fine
    regions: copies from the old storage and prolongation (with time
    interpolation) from the new coarser level, batched and per region,
    equal to the oracle's replay."""
    from carpet_amd.workloads import recompose_step
    kw = dict(ncoarse=32, centering=cent, buffer=6, order=order, maxlevels=3)
    old = nested_levels(nlevels=2, split=(2, 1, 1), fraction=0.35, **kw)
    new = nested_levels(nlevels=3, split=(1, 2, 1), fraction=0.6, **kw)
    w = recompose_step(old, new, cent, n_gf=2)
    print(f"{w.name}: {len(w.steps)} transfers", flush=True)
    ref = run_oracle(w)
    for mode in ("batched", "per region"):
        dev = w.allocate()
        if mode == "batched":
            b = w.batch(dev)
            b.finalize()
            b.run()
            cg.synchronize()
            b.close()
        else:
            w.run_steps(dev)
            cg.synchronize()
        for k in ref:
            got = dev[k].download()
            if not same(got, ref[k]):
                raise AssertionError(f"{mode}: {k} differs {first_diff(got, ref[k])}")
        w.free(dev)


@pytest.mark.parametrize("cent,order,op", [(cg.VERTEX_CENTERED, 5, cg.OP_LAGRANGE), (cg.CELL_CENTERED, 5, cg.OP_ENO)],
                         ids=["vertex-o5", "cell-eno3"])
def test_captured_step_advances_in_time(cent, order, op):
    """One captured level step replayed as a run advances: the boundary
    prolongation at the intermediate fine time (-1/2, interpolating), at the
    coarse-aligned time (0, an exact match: one level), then after every
    level's time levels rotate (ggf::cycle_all) at 1/2 and 1 against new
    level times -- new target times and storage go in through
    cg_batch_set_time / cg_batch_set_storage, the graph is never rebuilt.
    After each replay every array equals the oracle's sequential replay of
    the same steps, bit for bit."""
    from dataclasses import replace

    from workload_check import host_state

    lv = nested_levels(ncoarse=32, nlevels=3, split=(2, 1, 1), centering=cent, buffer=6, order=order)
    w = hierarchy_step(lv, cent, n_gf=2, op=op)
    host = host_state(w)
    st = cg.Stream()
    dev = w.allocate()
    b = w.batch(dev)
    b.finalize()
    b.capture_graph(st)
    prol = sorted({t.phase for t in w.steps if len(t.srcs) == 3})
    rot = {"t0": "t2", "t1": "t0", "t2": "t1"}  # storage of level tl after one cycle: old tl-1

    def ren(name):
        return name[:-2] + rot[name[-2:]]

    steps = list(w.steps)
    plan = [("interpolating -1/2", None, [0.0, -1.0, -2.0], -0.5), ("exact 0", None, [0.0, -1.0, -2.0], 0.0),
            ("rotated, interpolating 1/2", "cycle", [1.0, 0.0, -1.0], 0.5),
            ("rotated, exact 1", None, [1.0, 0.0, -1.0], 1.0)]
    for k, (what, cycle, times, t) in enumerate(plan):
        print(f"step {k}: {what}", flush=True)
        if cycle:
            n = b.set_storage([(dev[a], dev[ren(a)]) for a in dev])
            assert n == len(w.steps)
            steps = [replace(s, dst=ren(s.dst), srcs=[ren(x) for x in s.srcs]) for s in steps]
        if k:
            for ph in prol:
                assert b.set_time(ph, times, t) > 0
        steps = [replace(s, times=list(times), time=t) if len(s.srcs) == 3 else s for s in steps]
        b.launch_graph(st)
        st.synchronize()
        run_oracle(w, host, steps)
        for name, h in host.items():
            got = dev[name].download()
            if not same(got, h.data):
                raise AssertionError(f"step {k} ({what}): {name} differs {first_diff(got, h.data)}")
    # a target time outside the levels is refused and changes nothing
    with pytest.raises(cg.CarpetError):
        b.set_time(prol[0], [1.0, 0.0, -1.0], 5.0)
    b.close()
    w.free(dev)


def test_tiled_hierarchy_step_matches_oracle():
    """A hierarchy whose levels are tiled by boxes of random edges (the
    generator of the 2000-box level step, here 3 levels of 27, 27 and 64
    boxes of 16-40 points): the captured batched step equals the oracle's
    replay bit for bit; arrays no transfer touches are not allocated."""
    from carpet_amd.workloads import tiled_levels

    lv = tiled_levels(counts=((3, 3, 3), (3, 3, 3), (4, 4, 4)), edge=(16, 40), seed=11)
    w = hierarchy_step(lv, cg.VERTEX_CENTERED, n_gf=1, only_used=True)
    print(f"{w.name}: {len(w.steps)} transfers, {len(w.arrays)} arrays", flush=True)
    ref = run_oracle(w)
    st = cg.Stream()
    dev = w.allocate(pool=True)
    b = w.batch(dev)
    b.finalize()
    print(b.describe(), flush=True)
    b.capture_graph(st)
    b.launch_graph(st)
    st.synchronize()
    for k in ref:
        got = dev[k].download()
        if not same(got, ref[k]):
            raise AssertionError(f"{k} differs {first_diff(got, ref[k])}")
    b.close()
    w.free(dev)
