"""Mixed precision (FP32 storage; FP64 arithmetic, or compensated FP32 sums;
FP64 time interpolation): the error against the FP64 oracle on the same
(FP32-representable) inputs stays below 1e-6 normwise and is reported per
operator; inputs are rounded to FP32 first so the comparison isolates the
arithmetic and the final rounding."""
import numpy as np
import pytest

import pyoracle as orc
from helpers import field, host_array, pad_of, prolong_case, sizes

cg = pytest.importorskip("carpet_amd")
pytestmark = pytest.mark.gpu

TOL_MAX = 1e-6


def _dev32(h, box=None):
    lo, up, st = h.box
    a = cg.DeviceArray(cg.Box(lo, up, st), dtype=np.float32)
    a.upload(h.view().astype(np.float32), fill=np.nan)
    return a


@pytest.mark.parametrize("kind,order", [("vlag", 5), ("vlag", 3), ("clag", 5), ("clag", 3)])
def test_mixed_prolongation_error(kind, order):
    rng = np.random.default_rng(50 + order)
    worst = 0.0
    for c in range(4):
        sb, db, rb = prolong_case(rng, kind, order, n=(64, 64, 64) if c == 0 else None)
        src = host_array(sb, pad_of(sizes(sb), rng))
        src.view()[...] = field(rng, sizes(sb), "smooth").astype(np.float32).astype(np.float64)
        dst = host_array(db, sizes(db))
        f = orc.prolongate_3d_rf2 if kind == "vlag" else orc.prolongate_3d_cc_rf2
        assert f(order, src, dst, rb) == 0
        dsrc = _dev32(src)
        ddst = cg.DeviceArray(cg.Box(*db), dtype=np.float32)
        ddst.upload(np.full(dst.view().shape, np.nan, np.float32))
        g = cg.prolongate_3d_rf2_mixed if kind == "vlag" else cg.prolongate_3d_cc_rf2_mixed
        g(order, dsrc, ddst, cg.Box(*rb))
        got = ddst.download().astype(np.float64)
        m = ~np.isnan(dst.view())
        assert np.array_equal(m, ~np.isnan(got)), "mixed kernel wrote outside the region"
        emax, erms = cg.relative_errors(got[m], dst.view()[m])
        print(f"{kind} o{order} case {c}: max rel {emax:.3e} rms rel {erms:.3e}", flush=True)
        worst = max(worst, emax)
    assert worst <= TOL_MAX


@pytest.mark.parametrize("kind,order", [("vlag", 5), ("clag", 5)])
def test_mixed_transfer_with_time_interpolation(kind, order):
    rng = np.random.default_rng(7)
    sb, db, rb = prolong_case(rng, kind, order, n=(48, 40, 36))
    pad = sizes(sb)
    srcs = []
    for _ in range(3):
        h = host_array(sb, pad)
        h.view()[...] = field(rng, sizes(sb), "smooth").astype(np.float32).astype(np.float64)
        srcs.append(h)
    times = [1.0, 0.0, -1.0]
    exp = host_array(db, sizes(db))
    okind = orc.KIND_VLAG if kind == "vlag" else orc.KIND_CLAG
    assert orc.prolongate_time(okind, order, srcs, times, exp, 1 / 3, rb, 2) == 0
    ds = [_dev32(h) for h in srcs]
    dd = cg.DeviceArray(cg.Box(*db), dtype=np.float32)
    dd.upload(np.full(exp.view().shape, np.nan, np.float32))
    cent = cg.VERTEX_CENTERED if kind == "vlag" else cg.CELL_CENTERED
    cg.transfer_from_mixed(dd, ds, times, cg.Box(*rb), cg.Box(*rb), 1 / 3, cg.OP_LAGRANGE, cent, order, 2)
    got = dd.download().astype(np.float64)
    m = ~np.isnan(exp.view())
    emax, erms = cg.relative_errors(got[m], exp.view()[m])
    print(f"{kind} o{order} + 3tl: max rel {emax:.3e} rms rel {erms:.3e}", flush=True)
    assert emax <= TOL_MAX


def _ulp_distance(a, b):
    """FP32 units in the last place between a and b (same shape)."""
    def ordered(x):
        i = np.ascontiguousarray(x, dtype=np.float32).view(np.int32).astype(np.int64)
        return np.where(i < 0, -(i & 0x7FFFFFFF), i)
    return np.abs(ordered(a) - ordered(b))


@pytest.mark.parametrize("kind,order", [("vlag", 5), ("vlag", 3), ("clag", 5), ("clag", 3)])
def test_mixed_batch_modes(kind, order):
    """Both mixed modes through the batch: CG_PREC_MIXED (FP64 arithmetic on
    FP32 storage, the three time levels combined before the spatial stencil)
    is within one FP32 ulp of the FP64 oracle on the same FP32-representable
    inputs rounded once to FP32 (the two FP64 results differ only by
    reassociation, ~1e-16 relative, so they round to the same FP32 value
    except next to a rounding boundary); CG_PREC_MIXED_DOT2 (compensated FP32
    sums) stays within 1e-6 normwise."""
    rng = np.random.default_rng(11 + order)
    sb, db, rb = prolong_case(rng, kind, order, n=(40, 36, 34))
    pad = sizes(sb)
    srcs = []
    for _ in range(3):
        h = host_array(sb, pad)
        h.view()[...] = field(rng, sizes(sb), "smooth").astype(np.float32).astype(np.float64)
        srcs.append(h)
    times = [1.0, 0.0, -1.0]
    exp = host_array(db, sizes(db))
    okind = orc.KIND_VLAG if kind == "vlag" else orc.KIND_CLAG
    assert orc.prolongate_time(okind, order, srcs, times, exp, 1 / 3, rb, 2) == 0
    m = ~np.isnan(exp.view())
    ds = [_dev32(h) for h in srcs]
    cent = cg.VERTEX_CENTERED if kind == "vlag" else cg.CELL_CENTERED
    out = {}
    for prec in (cg.PREC_MIXED, cg.PREC_MIXED_DOT2):
        dd = cg.DeviceArray(cg.Box(*db), dtype=np.float32)
        dd.upload(np.full(exp.view().shape, np.nan, np.float32))
        b = cg.Batch()
        b.add(0, dd, ds, times, cg.Box(*rb), cg.Box(*rb), 1 / 3, cg.OP_LAGRANGE, cent, order, 2, prec=prec)
        b.finalize()
        b.run()
        cg.synchronize()
        out[prec] = dd.download()
        assert np.array_equal(m, ~np.isnan(out[prec])), "mixed kernel wrote outside the region"
        b.close()
    exp32 = exp.view()[m].astype(np.float32)
    ulps = _ulp_distance(out[cg.PREC_MIXED][m], exp32)
    print(f"{kind} o{order} + 3tl FP64 arithmetic: {np.count_nonzero(ulps)} of {ulps.size} points "
          f"one ulp from the rounded oracle", flush=True)
    assert ulps.max() <= 1, "FP64-arithmetic mixed mode is not the rounded oracle to one ulp"
    emax, erms = cg.relative_errors(out[cg.PREC_MIXED_DOT2][m].astype(np.float64), exp.view()[m])
    print(f"{kind} o{order} + 3tl Dot2: max rel {emax:.3e} rms rel {erms:.3e}", flush=True)
    assert emax <= TOL_MAX


def test_mixed_rejects_fp64_only_operators():
    box = cg.Box((0, 0, 0), (20, 20, 20), (2, 2, 2))
    fbox = cg.Box((0, 0, 0), (20, 20, 20), (1, 1, 1))
    a = cg.DeviceArray(box, dtype=np.float32)
    b = cg.DeviceArray(fbox, dtype=np.float32)
    reg = cg.Box((4, 4, 4), (10, 10, 10), (1, 1, 1))
    with pytest.raises(cg.CarpetError):
        cg.transfer_from_mixed(b, [a], [0.0], reg, reg, 0.0, cg.OP_TVD, cg.CELL_CENTERED, 1, 0)
    with pytest.raises(cg.CarpetError):
        cg.transfer_from_mixed(b, [b], [0.0], reg, reg, 0.0, cg.OP_LAGRANGE, cg.VERTEX_CENTERED, 5, 0)
