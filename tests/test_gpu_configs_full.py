"""The benchmark configurations at their stated sizes against the oracle.

- The batched uneven level step at 2000 regions with edges uniform in
  [8,96] (one grid function; the whole step runs on the GPU, a fixed sample
  of regions is replayed on the oracle and compared bit for bit -- every
  region has its own arrays, so a sample replays independently).
- The cell-centred set through the batched path the benchmark times
  (default launch settings: 1024-cell TVD items, ENO z chunks capped at 64
  planes, 1024-point restriction items), sampled sub-boxes bit for bit.
- Mixed precision on the 8 x 193^3 -> 384^3 shape: max and RMS relative
  error against the FP64 oracle on sampled sub-boxes, within 1e-6.
"""
import numpy as np
import pytest

import pyoracle as orc
from helpers import first_diff, same
from workload_check import _host, run_oracle

cg = pytest.importorskip("carpet_amd")
from carpet_amd.workloads import Config3, Config5, config4, hash_field  # noqa: E402

pytestmark = pytest.mark.gpu


def test_config4_full_size_sample_matches_oracle():
    w = config4(n_regions=2000, n_gf=1)
    print(f"{w.name}: {len(w.steps)} transfers, {len(w.arrays)} arrays", flush=True)
    st = cg.Stream()
    dev = w.allocate(pool=True)
    b = w.batch(dev)
    b.finalize()
    print(b.describe(), flush=True)
    b.capture_graph(st)
    b.launch_graph(st)
    st.synchronize()
    # every 8th region of each kind, up to 250 regions
    pick = [t for i, t in enumerate(w.steps) if i % 8 == 0][:250]
    names = sorted({t.dst for t in pick} | {s for t in pick for s in t.srcs})
    host = {}
    for k in names:
        spec = w.arrays[k]
        e = spec.box.sizes()
        host[k] = _host(spec, hash_field(spec.seed, e))
    run_oracle(w, host, pick)
    print(f"oracle replayed {len(pick)} regions", flush=True)
    for t in pick:
        got = dev[t.dst].download()
        if not same(got, host[t.dst].data):
            raise AssertionError(f"{t.name}: {first_diff(got, host[t.dst].data)}")
    b.close()
    w.free(dev)


def _sub_boxes(rng, reg, n, edge, zmin=None):
    """n random sub-boxes of reg (a Box) with `edge` points per
    axis; the first reaches z planes above zmin points from the region start
    when given."""
    lo, up, st = reg.lower, reg.upper, reg.stride
    sizes = reg.sizes()
    out = []
    for c in range(n):
        o = [int(rng.integers(0, sizes[d] - edge + 1)) for d in range(3)]
        if c == 0 and zmin is not None:
            o[2] = max(o[2], min(sizes[2] - edge, zmin))
        blo = tuple(lo[d] + o[d] * st[d] for d in range(3))
        bup = tuple(blo[d] + (edge - 1) * st[d] for d in range(3))
        out.append(cg.Box(blo, bup, st))
    return out


def _part(arr_box, a, sub):
    """the sub-box of array data a ((z, y, x) over arr_box)"""
    o = [(sub.lower[d] - arr_box.lower[d]) // arr_box.stride[d] for d in range(3)]
    e = sub.sizes()
    return a[o[2]:o[2] + e[2], o[1]:o[1] + e[1], o[0]:o[0] + e[0]]


@pytest.mark.parametrize("op,order_space", [(cg.OP_TVD, 1), (cg.OP_ENO, 5), (cg.OP_ENO, 3)],
                         ids=["tvd", "eno3", "eno2"])
def test_config3_batched_matches_oracle(op, order_space):
    """The cell-centred set (4 variables of 192^3 -> 384^3) through the
    batched path the benchmark times, default launch settings; sampled
    24^3 sub-boxes of the prolongation (one deeper than 64 coarse planes,
    past the ENO z-chunk cap) and of the restriction back equal the
    oracle bit for bit."""
    cfg = Config3(n_var=4)
    cbox, owned, fbox, reg = cfg.boxes()
    st = cg.Stream()
    coarse, fine, back = cfg.allocate(stream=st)
    bp = cfg.prolong_batch(coarse, fine, op, order_space)
    bp.finalize()
    br = cfg.restrict_batch(fine, back)
    br.finalize()
    print(bp.describe(), br.describe(), flush=True)
    bp.run(st)
    br.run(st)
    st.synchronize()
    rng = np.random.default_rng(30 + order_space + op)
    kind, order = (orc.KIND_TVD, 1) if op == cg.OP_TVD else (orc.KIND_ENO, 3 if order_space == 5 else 2)
    for v in range(cfg.n_var):
        cdata = cfg.coarse_data(v)
        src = orc.HostArr((cbox.lower, cbox.upper, cbox.stride), np.ascontiguousarray(cdata))
        got_f = fine[v].download()
        for sub in _sub_boxes(rng, reg, 2, 24, zmin=140):
            exp = orc.HostArr((sub.lower, sub.upper, sub.stride), np.full((24, 24, 24), np.nan))
            assert orc.prolongate_time(kind, order, [src], [0.0], exp, 0.0,
                                       (sub.lower, sub.upper, sub.stride), 0) == 0
            g = _part(fbox, got_f, sub)
            if not same(g, exp.data):
                raise AssertionError(f"prolongation var {v} {sub}: {first_diff(g, exp.data)}")
        # restriction of the GPU's fine data back onto sampled coarse sub-boxes
        fh = orc.HostArr((fbox.lower, fbox.upper, fbox.stride), np.ascontiguousarray(got_f))
        got_c = back[v].download()
        for sub in _sub_boxes(rng, owned, 2, 16):
            exp = orc.HostArr((sub.lower, sub.upper, sub.stride), np.full((16, 16, 16), np.nan))
            assert orc.restrict_3d_cc_rf2(fh, exp, (sub.lower, sub.upper, sub.stride)) == 0
            g = _part(cbox, got_c, sub)
            if not same(g, exp.data):
                raise AssertionError(f"restriction var {v} {sub}: {first_diff(g, exp.data)}")
    for b in (bp, br):
        b.close()
    for a in coarse + fine + back:
        a.free()


def test_config5_mixed_on_config3_shape_error():
    """Mixed precision (FP32 storage, FP64 arithmetic; and compensated FP32
    sums) on 8 variables of 193^3 -> 384^3, order 5 with 3 time levels at
    t = 1/3, batched: sampled 20^3 sub-boxes against the FP64 oracle on the
    original FP64 sources.  Max and RMS relative error (normwise, over the
    samples) are printed and held to 1e-6."""
    cfg = Config5(shape="cfg3")
    cbox, owned, fbox, reg = cfg.boxes()
    st = cg.Stream()
    src64, src32, src64r, dst64, dst32 = cfg.allocate(stream=st)
    rng = np.random.default_rng(55)
    subs = _sub_boxes(rng, reg, 3, 20, zmin=300)
    hosts = [[orc.HostArr((cbox.lower, cbox.upper, cbox.stride), np.ascontiguousarray(a.download()))
              for a in src64[v]] for v in range(cfg.n_var)]
    exps = []
    for v in range(cfg.n_var):
        for sub in subs:
            exp = orc.HostArr((sub.lower, sub.upper, sub.stride), np.full((20, 20, 20), np.nan))
            assert orc.prolongate_time(orc.KIND_VLAG, 5, hosts[v], list(cfg.times), exp, cfg.target,
                                       (sub.lower, sub.upper, sub.stride), 2) == 0
            exps.append(exp.data)
    ref = np.concatenate([e.ravel() for e in exps])
    for name, prec in (("mixed (FP64 arithmetic)", cg.PREC_MIXED), ("mixed dot2", cg.PREC_MIXED_DOT2)):
        b = cfg.batch(src32, dst32, prec)
        b.finalize()
        b.run(st)
        st.synchronize()
        got = []
        for v in range(cfg.n_var):
            d = dst32[v].download().astype(np.float64)
            got.extend(_part(fbox, d, sub).ravel() for sub in subs)
        got = np.concatenate(got)
        mx, rms = cg.relative_errors(got, ref)
        print(f"{name}: max rel {mx:.3e}, rms rel {rms:.3e} over {ref.size} points", flush=True)
        assert mx <= 1e-6 and rms <= 1e-6, (name, mx, rms)
        b.close()
    for a in [x for v in src64 + src32 + src64r for x in v] + dst64 + dst32:
        a.free()
