"""Full-size configurations on the GPU, checked where the CPU oracle cannot
afford the whole volume: bit-exact parity on sampled sub-regions (the
operators are local, so a sub-region computed by the oracle from the same
source must equal the GPU's values there) plus size-independent
properties (conservation, boundedness, region-only writes)."""
import numpy as np
import pytest

import pyoracle as orc
from helpers import same

cg = pytest.importorskip("carpet_amd")
from carpet_amd.workloads import Config2, Config3  # noqa: E402

pytestmark = pytest.mark.gpu


def _host(box, data):
    return orc.HostArr((box.lower, box.upper, box.stride), np.ascontiguousarray(data))


def _sub_regions(reg: cg.Box, rng, n=4, size=24):
    lo, up, st = reg.lower, reg.upper, reg.stride
    ext = reg.sizes()
    out = [reg.expand((0, 0, 0), tuple(-(e - min(e, size)) for e in ext))]  # lower corner
    out.append(cg.Box(tuple(u - (min(e, size) - 1) * s for u, e, s in zip(up, ext, st)), up, st))  # upper corner
    for _ in range(n):
        o = [int(rng.integers(0, e - min(e, size) + 1)) for e in ext]
        l2 = tuple(l + oo * s for l, oo, s in zip(lo, o, st))
        out.append(cg.Box(l2, tuple(l + (min(e, size) - 1) * s for l, e, s in zip(l2, ext, st)), st))
    return out


def _oracle_region(kind, order, src_h, times, dbox, sub, time, ntl):
    e = sub.sizes()
    h = orc.HostArr((sub.lower, sub.upper, sub.stride), np.full((e[2], e[1], e[0]), np.nan))
    rb = (sub.lower, sub.upper, sub.stride)
    assert orc.prolongate_time(kind, order, src_h, times, h, time, rb, ntl - 1) == 0
    return h.data


def _slice(full, dbox, sub):
    o = [(l - dl) // s for l, dl, s in zip(sub.lower, dbox.lower, dbox.stride)]
    e = sub.sizes()
    return full[o[2]:o[2] + e[2], o[1]:o[1] + e[1], o[0]:o[0] + e[0]]


def test_config2_full_size_sampled_parity():
    """One grid function of the 24 at full size (257^3 coarse, 512^3 fine,
    O=5 + 3 time levels at t=1/3): sampled sub-boxes bit-exact against the
    oracle; the whole region written, nothing outside it."""
    cfg = Config2(n_gf=1)
    st = cg.Stream()
    src, dst = cfg.allocate(stream=st)
    cbox, fbox, reg = cfg.boxes()
    b = cfg.batch(src, dst)
    b.finalize()
    b.run(st)
    st.synchronize()
    got = dst[(0, 0)].download()
    src_h = [_host(cbox, src[(0, tl)].download()) for tl in range(3)]
    inner = _slice(got, fbox, reg)
    assert not np.isnan(inner).any()
    assert np.isnan(got).sum() == got.size - inner.size
    rng = np.random.default_rng(2)
    for sub in _sub_regions(reg, rng):
        ref = _oracle_region(orc.KIND_VLAG, 5, src_h, list(cfg.times), fbox, sub, cfg.targets[0], 3)
        assert same(_slice(got, fbox, sub), ref), sub
    b.close()


@pytest.mark.parametrize("n_var", [2])
def test_config3_full_size(n_var):
    """Cell-centred set at full size (192^3 -> 384^3): TVD and ENO<3>/ENO<2>
    prolongation sampled bit-exact against the oracle; TVD followed by
    restrict_3d_cc_rf2 conserves every coarse cell to 1e-14 relative; ENO
    values stay within their 2x2x2 nearest coarse cells."""
    cfg = Config3(n_var=n_var)
    cbox, owned, fbox, reg = cfg.boxes()
    st = cg.Stream()
    rng = np.random.default_rng(3)
    for var in range(n_var):
        data = cfg.coarse_data(var)
        c = cg.DeviceArray(cbox, owned=owned)
        c.upload(data)
        src_h = [_host(cbox, data)]
        for op, order_space, kind, order in ((cg.OP_TVD, 1, orc.KIND_TVD, 1), (cg.OP_ENO, 5, orc.KIND_ENO, 3),
                                             (cg.OP_ENO, 3, orc.KIND_ENO, 2)):
            f = cg.DeviceArray(fbox, owned=reg)
            f.fill(np.nan, st)
            cg.transfer_from(f, [c], [0.0], reg, reg, 0.0, op, cg.CELL_CENTERED, order_space, 0, stream=st)
            st.synchronize()
            got = f.download()
            inner = _slice(got, fbox, reg)
            assert not np.isnan(inner).any()
            for sub in _sub_regions(reg, rng, n=3, size=16):
                ref = _oracle_region(kind, order, src_h, [0.0], fbox, sub, 0.0, 1)
                assert same(_slice(got, fbox, sub), ref), (var, op, sub)
            if op == cg.OP_TVD:
                back = cg.DeviceArray(cbox, owned=owned)
                back.fill(np.nan, st)
                cg.transfer_from(back, [f], [0.0], owned, owned, 0.0, op, cg.CELL_CENTERED, 1, 0, stream=st)
                st.synchronize()
                r = _slice(back.download(), cbox, owned)
                orig = _slice(data, cbox, owned)
                rel = np.max(np.abs(r - orig)) / np.max(np.abs(orig))
                print(f"var {var}: TVD->restrict_cc conservation max rel {rel:.3e}", flush=True)
                assert rel <= 1e-14
                back.free()
            else:
                # boundedness by the 8 nearest coarse cells: fine cell f has
                # nearest coarse cells floor((f-1)/2) and +1 per axis
                g = cfg.ghosts
                cz = _slice(data, cbox, cbox)
                fl = np.arange(inner.shape[0])
                lo_idx = (fl - 1) // 2 + g
                lo_idx = np.clip(lo_idx, 0, cz.shape[0] - 2)
                # per-axis min/max over the two nearest -> separable bound
                mn = np.minimum(cz[:-1], cz[1:])
                mx = np.maximum(cz[:-1], cz[1:])
                mn = np.minimum(mn[:, :-1], mn[:, 1:]); mx = np.maximum(mx[:, :-1], mx[:, 1:])
                mn = np.minimum(mn[:, :, :-1], mn[:, :, 1:]); mx = np.maximum(mx[:, :, :-1], mx[:, :, 1:])
                ix = np.ix_(lo_idx, lo_idx, lo_idx)
                assert np.all(inner >= mn[ix] - 1e-12) and np.all(inner <= mx[ix] + 1e-12)
            f.free()
        c.free()
