"""Batched level step on the GPU: batched launches over the descriptor list
(persistent grids with per-XCD work counters, and one workgroup per item),
autotuned launches, the plain per-region launches and hipGraph replays all
equal the CPU oracle's replay of the same transfers, bit for bit."""
import re

import numpy as np
import pytest

from helpers import first_diff, same
from workload_check import run_oracle

cg = pytest.importorskip("carpet_amd")
from carpet_amd.workloads import config1, config4, hierarchy_step, nested_levels  # noqa: E402

pytestmark = pytest.mark.gpu


def _outputs(w, dev):
    return {k: dev[k].download() for k in w.outputs}


def _compare(got, ref, what):
    for k in ref:
        if not same(got[k], ref[k]):
            raise AssertionError(f"{what}: {k} differs {first_diff(got[k], ref[k])}")


def _poison_outputs(w, dev, stream):
    for k in w.outputs:
        spec = w.arrays[k]
        a = dev[k]
        a.fill(np.nan, stream=stream)
        if spec.seed is not None:
            a.fill_hash(spec.seed, stream=stream)
        elif spec.init is not None:
            stream.synchronize()
            a.upload(spec.init, fill=np.nan)


def _thin_faces(cent, op, order, ntl=3):
    """Prolongation (ntl time levels, target between them) onto regions a
    few fine points thick in x and long in y and z -- the shell faces normal
    to x that go to the 4x16-base tiles -- at both parities, plus one thin
    in y (large tiles) in the same phase."""
    from carpet_amd.workloads import ArraySpec, Box, Transfer, Workload
    w = Workload(f"thin-faces c{cent} op{op} o{order}")
    # fine points at f (vertex, stride 1) or cells centred at 1 + 2f (cell,
    # stride 2); coarse points at 2c or cells centred at 2 + 4c
    vc = cent == cg.VERTEX_CENTERED
    fs, cs, fo, co = (1, 2, 0, 0) if vc else (2, 4, 1, 2)
    cbox = Box((co - 6 * cs,) * 3, (co + 56 * cs,) * 3, (cs, cs, cs))
    for tl in range(ntl):
        w.arrays[f"c{tl}"] = ArraySpec(cbox, seed=900 + tl)
    times = [0.0, -1.0, -2.0][:ntl]
    target = -0.5 if ntl > 1 else 0.0
    regions = [((10, 4, 6), (12, 83, 71)), ((21, 5, 3), (26, 60, 44)), ((40, 2, 9), (42, 41, 88)),
               ((7, 30, 2), (80, 32, 50))]
    for r, (lo, up) in enumerate(regions):
        reg = Box(tuple(fo + fs * l for l in lo), tuple(fo + fs * u for u in up), (fs, fs, fs))
        fbox = reg.expand((3, 3, 3), (3, 3, 3))
        w.arrays[f"f{r}"] = ArraySpec(fbox, owned=reg, seed=950 + r)
        w.steps.append(Transfer(f"prolong f{r}", f"f{r}", [f"c{tl}" for tl in range(ntl)], times, reg, reg,
                                target, op, cent, order, ntl - 1, phase=0))
        w.outputs.append(f"f{r}")
    return w


@pytest.mark.parametrize("cent,op,order,ntl", [(cg.VERTEX_CENTERED, cg.OP_LAGRANGE, 5, 3),
                                               (cg.VERTEX_CENTERED, cg.OP_LAGRANGE, 5, 2),
                                               (cg.VERTEX_CENTERED, cg.OP_LAGRANGE, 3, 1),
                                               (cg.CELL_CENTERED, cg.OP_LAGRANGE, 3, 3),
                                               (cg.CELL_CENTERED, cg.OP_ENO, 5, 3),
                                               (cg.VERTEX_CENTERED, cg.OP_ENO, 5, 2)],
                         ids=["vertex-o5-3tl", "vertex-o5-2tl", "vertex-o3-1tl", "cell-o3-3tl", "cell-eno3-3tl",
                              "vertex-eno-2tl"])
def test_slim_tiles_match_oracle(cent, op, order, ntl):
    """Regions thin in x run on the 4x16-base tiles (class tensor4), direct
    and from a captured graph, equal to the oracle bit for bit."""
    w = _thin_faces(cent, op, order, ntl)
    ref = run_oracle(w)
    st = cg.Stream()
    dev = w.allocate()
    b = w.batch(dev)
    b.finalize()
    print(b.describe(), flush=True)
    assert "[tensor4 " in b.describe()
    b.run(st)
    st.synchronize()
    _compare(_outputs(w, dev), ref, "batched")
    _poison_outputs(w, dev, st)
    b.capture_graph(st)
    b.launch_graph(st)
    st.synchronize()
    _compare(_outputs(w, dev), ref, "graph")
    b.close()
    w.free(dev)


def test_persistent_phase_runs_in_turn():
    """A phase whose large-region class runs persistent (the default for
    regions of 2^20 points and more) and whose thin regions run one
    workgroup per item runs its launches one after the other by default,
    and equals the oracle."""
    from carpet_amd.workloads import ArraySpec, Box, Transfer
    w = _thin_faces(cg.VERTEX_CENTERED, cg.OP_LAGRANGE, 5)
    big = Box((2, 2, 2), (105, 105, 98), (1, 1, 1))  # 104 x 104 x 97 > 2^20 points
    w.arrays["fbig"] = ArraySpec(big.expand((3, 3, 3), (3, 3, 3)), owned=big, seed=990)
    w.steps.append(Transfer("prolong fbig", "fbig", ["c0", "c1", "c2"], [0.0, -1.0, -2.0], big, big, -0.5,
                            cg.OP_LAGRANGE, cg.VERTEX_CENTERED, 5, 2, phase=0))
    w.outputs.append("fbig")
    ref = run_oracle(w)
    st = cg.Stream()
    dev = w.allocate()
    b = w.batch(dev)
    b.finalize()
    d = b.describe()
    print(d, flush=True)
    assert re.search(r"\[tensor k0 o5 t3 [^\]]*\] .* wpc 0 .* in turn", d), d
    assert "side by side" not in d
    b.run(st)
    st.synchronize()
    _compare(_outputs(w, dev), ref, "in turn")
    b.set_phase_mode(0, 1)
    assert "side by side" in b.describe()
    _poison_outputs(w, dev, st)
    b.run(st)
    st.synchronize()
    _compare(_outputs(w, dev), ref, "side by side")
    b.close()
    w.free(dev)


def _shells():
    # a 2-level vertex hierarchy's step: its boundary prolongation regions are
    # 3- and 6-point shells, which go to the 8x8-base tiles
    lv = nested_levels(ncoarse=32, nlevels=2, split=(2, 1, 1), centering=cg.VERTEX_CENTERED, buffer=6, order=5)
    return hierarchy_step(lv, cg.VERTEX_CENTERED, n_gf=1)


@pytest.mark.parametrize("make", [lambda: config1(), lambda: config4(n_regions=45, n_gf=2, edge=(4, 40)), _shells],
                         ids=["config1", "config4-small", "hierarchy-shells"])
def test_batch_modes_match_oracle(make, tmp_path):
    w = make()
    print(f"workload {w.name}: {len(w.steps)} transfers", flush=True)
    ref = run_oracle(w)
    st = cg.Stream()
    dev = w.allocate(pool=True)
    b = w.batch(dev)
    b.finalize()
    print("stats", b.stats(), flush=True)
    print(b.describe(), flush=True)
    if make is _shells:
        # shell edges and corners on 8x8-base tiles, faces thin in x on 4x16
        assert "[tensor8 " in b.describe() or "[tensor4 " in b.describe()
    # batched launches, one workgroup per item
    b.run(st)
    st.synchronize()
    _compare(_outputs(w, dev), ref, "batched")
    if w.name != "config1":
        # every launch mode (a tuner cache naming every class with one
        # workgroup per item, wpc -1, persistent grids taking items from
        # counters, wpc 0 and 2, and in fixed rounds, wpc -2, which point
        # classes run as wpc -1),
        # then the tuner's own search; each time from poisoned outputs
        # 256-point items also with 2, 4, 8 and 16 items per workgroup (ni;
        # a cached 3 runs as 2)
        keys = re.findall(r"\[([^\]]+)\]", b.describe())
        for wpc, zc, pt, ni in ((-1, 16, 256, 1), (-1, 16, 256, 2), (-1, 16, 256, 3), (-1, 16, 256, 4),
                                (-1, 16, 256, 8), (-1, 16, 256, 16), (0, 8, 512, 1), (2, 0, 65536, 1),
                                (-2, 32, 1024, 1)):
            cache = tmp_path / f"tune_{wpc}_{ni}.txt"
            cache.write_text("".join(f"{k}|{zc} {pt} {wpc} -1 {ni}\n" for k in keys))
            b.autotune(st, reps=1, cache_path=str(cache))
            assert f"wpc {wpc}" in b.describe()
            if ni > 1:
                assert f"pt 256 wpc -1 shape 0 ni {2 if ni == 3 else ni}" in b.describe()
            _poison_outputs(w, dev, st)
            b.run(st)
            st.synchronize()
            _compare(_outputs(w, dev), ref, f"persistent wpc {wpc} ni {ni}")
        cache = tmp_path / "tune_search.txt"
        b.autotune(st, reps=1, cache_path=str(cache))
        print("tuned:\n" + cache.read_text() + b.describe(), flush=True)
        _poison_outputs(w, dev, st)
        b.run(st)
        st.synchronize()
        _compare(_outputs(w, dev), ref, "autotuned")
    # plain: one launch per transfer
    _poison_outputs(w, dev, st)
    if w.name == "config1":
        w.free(dev)
        dev = w.allocate(pool=True)
        b = w.batch(dev)
        b.finalize()
    b.run_plain(st)
    st.synchronize()
    _compare(_outputs(w, dev), ref, "plain")
    # hipGraph: capture once, replay several times with eager work between
    if w.name == "config1":
        w.free(dev)
        dev = w.allocate(pool=True)
        b = w.batch(dev)
        b.finalize()
    b.capture_graph(st)
    for rep in range(3):
        # every replay from the seeded state: a hierarchy step is not
        # idempotent (its restriction feeds its own prolongation)
        if w.name != "config1":
            _poison_outputs(w, dev, st)
            st.synchronize()
        b.launch_graph(st)
        st.synchronize()
        print(f"graph replay {rep}", flush=True)
        if w.name != "config1" or rep == 0:
            _compare(_outputs(w, dev), ref, f"graph replay {rep}")
    b.close()
    w.free(dev)


def test_batch_rejects_wrong_state():
    b = cg.Batch()
    with pytest.raises(cg.CarpetError):
        b.run()
    b.finalize()
    with pytest.raises(cg.CarpetError):
        b.finalize()
    b.close()
