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
    (8x8-base) tiles as well as the large ones."""
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
        assert "[tensor8 " in b.describe()
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
