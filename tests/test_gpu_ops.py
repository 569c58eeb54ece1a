"""GPU parity of every per-call operator against the CPU oracle (bit-exact).

Each case builds seeded random geometry (region offsets of both parities,
random padding, regions crossing several tiles and z-chunks), poisons the
destination and the source padding with NaN, runs the HIP operator through
the C ABI and the oracle on the same inputs, and compares the whole padded
destination: the region must match bit for bit and everything outside it
must be left untouched.
"""
import numpy as np
import pytest

import pyoracle as orc
from helpers import field, first_diff, host_array, pad_of, prolong_case, same, sizes

cg = pytest.importorskip("carpet_amd")

pytestmark = pytest.mark.gpu


def _dev(h: orc.HostArr):
    lo, up, st = h.box
    a = cg.DeviceArray(cg.Box(lo, up, st), padext=h.padext)
    a.upload(h.data)
    return a


def _check(expected: orc.HostArr, dev: cg.DeviceArray, what: str):
    got = dev.download_padded()
    if not same(got, expected.data):
        raise AssertionError(f"{what}: mismatch {first_diff(got, expected.data)}")


def _src_host(rng, box, kind="mixed"):
    ext = sizes(box)
    h = host_array(box, pad_of(ext, rng))
    h.view()[...] = field(rng, ext, kind)
    return h


CASES = (
    [("vlag", o) for o in (1, 3, 5, 7, 9, 11)]
    + [("clag", o) for o in (0, 1, 2, 3, 4, 5)]
    + [("eno", 2), ("eno", 3), ("tvd", 1), ("veno", 3), ("mono", 5)]
)


def _run_prolong(kind, order, src, dst, reg, stream=None):
    box = cg.Box(*reg)
    if kind == "vlag":
        cg.prolongate_3d_rf2(order, src, dst, box, stream=stream)
    elif kind == "clag":
        cg.prolongate_3d_cc_rf2(order, src, dst, box, stream=stream)
    elif kind == "eno":
        cg.prolongate_3d_cc_eno_rf2(order, src, dst, box, stream=stream)
    elif kind == "veno":
        cg.prolongate_3d_eno(src, dst, box, stream=stream)
    elif kind == "mono":
        cg.prolongate_3d_o5_monotone_rf2(src, dst, box, stream=stream)
    else:
        cg.prolongate_3d_cc_tvd(src, dst, box, stream=stream)


def _orc_prolong(kind, order, src, dst, reg):
    if kind == "vlag":
        return orc.prolongate_3d_rf2(order, src, dst, reg)
    if kind == "clag":
        return orc.prolongate_3d_cc_rf2(order, src, dst, reg)
    if kind == "eno":
        return orc.prolongate_3d_cc_eno_rf2(order, src, dst, reg)
    if kind == "veno":
        return orc.prolongate_3d_eno(src, dst, reg)
    if kind == "mono":
        return orc.prolongate_3d_o5_monotone_rf2(src, dst, reg)
    return orc.prolongate_3d_cc_tvd(src, dst, reg)


@pytest.mark.parametrize("kind,order", CASES)
def test_prolongate_parity(kind, order):
    rng = np.random.default_rng(1000 + 17 * order + len(kind))
    ncases = 6
    for c in range(ncases):
        print(f"case {kind} o{order} #{c}", flush=True)
        if c == 0:
            n = (1, 1, 1)
        elif c == 1:
            n = (150, 3, 70)  # several x tiles and z chunks
        else:
            n = None
        sb, db, rb = prolong_case(rng, kind, order, n=n, scale=1 + (c % 2))
        data_kind = ("mixed", "smooth", "random")[c % 3]
        src = _src_host(rng, sb, data_kind)
        dst = host_array(db, pad_of(sizes(db), rng))
        exp = orc.HostArr(db, dst.data.copy())
        assert _orc_prolong(kind, order, src, exp, rb) == 0
        dsrc, ddst = _dev(src), _dev(dst)
        _run_prolong(kind, order, dsrc, ddst, rb)
        _check(exp, ddst, f"{kind} o{order} case {c} reg={rb}")


@pytest.mark.parametrize("order", [2, 3])
@pytest.mark.parametrize("shape", ["sum", "product"])
def test_eno_kinked_field_parity(order, shape):
    """The kinked fields that pin the oracle's ENO stencil choice
    (test_oracle_eno_pins): g(x)+h(y)+k(z) and g(x)*h(y)*k(z), each axis a
    polynomial with kinks so that neighbouring fine points select left,
    centred and right stencils.  Random padding and region offsets; the HIP
    operator must equal the oracle bit for bit."""
    from fractions import Fraction as F

    from test_oracle_eno_pins import kinked_axis

    rng = np.random.default_rng(70 + order + 3 * (shape == "product"))
    n = 40
    axes = [kinked_axis(n, order, [(F(2 * int(rng.integers(6, n - 6)) + 1, 2) + F(1, 2), +1),
                                   (F(2 * int(rng.integers(6, n - 6)) + 1, 2) - F(1, 2), -1)])
            for _ in range(3)]
    gx, gy, gz = (np.array([float(v) for v in a]) for a in axes)
    gx, gy, gz = gx / gx.max(), gy / gy.max(), gz / gz.max()
    if shape == "sum":
        vol = gz[:, None, None] + gy[None, :, None] + gx[None, None, :]
    else:
        vol = gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
    sb = ((2, 2, 2), (2 + 4 * (n - 1),) * 3, (4, 4, 4))
    src = host_array(sb, pad_of(sizes(sb), rng))
    src.view()[...] = vol
    lo = [int(rng.integers(6, 9)) for _ in range(3)]
    hi = [2 * n - int(rng.integers(7, 10)) for _ in range(3)]
    rb = (tuple(1 + 2 * v for v in lo), tuple(1 + 2 * v for v in hi), (2, 2, 2))
    db = (tuple(1 + 2 * (v - 1) for v in lo), tuple(1 + 2 * (v + 1) for v in hi), (2, 2, 2))
    dst = host_array(db, pad_of(sizes(db), rng))
    exp = orc.HostArr(db, dst.data.copy())
    assert orc.prolongate_3d_cc_eno_rf2(order, src, exp, rb) == 0
    dsrc, ddst = _dev(src), _dev(dst)
    cg.prolongate_3d_cc_eno_rf2(order, dsrc, ddst, cg.Box(*rb))
    _check(exp, ddst, f"eno{order} kinked {shape} reg={rb}")


@pytest.mark.parametrize("cell", [False, True])
def test_restrict_parity(cell):
    rng = np.random.default_rng(7 + cell)
    for c in range(6):
        print(f"restrict cell={cell} #{c}", flush=True)
        sf = 2 if cell else 1
        sc = 2 * sf
        n = tuple(int(v) for v in rng.integers(1, 40, size=3)) if c else (1, 1, 1)
        rlo, rup, slo, sup, dlo, dup = [], [], [], [], [], []
        for d in range(3):
            C0 = int(rng.integers(-4, 5))
            if cell:
                rlo.append(sc // 2 + C0 * sc)
                rup.append(sc // 2 + (C0 + n[d] - 1) * sc)
                F0 = 2 * C0 - int(rng.integers(0, 3))  # first fine cell of the source
                slo.append(sf // 2 + F0 * sf)
                sup.append(sf // 2 + (2 * (C0 + n[d]) + int(rng.integers(0, 3)) - 1) * sf)
            else:
                rlo.append(C0 * sc)
                rup.append((C0 + n[d] - 1) * sc)
                slo.append(rlo[-1] - int(rng.integers(0, 3)) * sf)
                sup.append(rup[-1] + int(rng.integers(0, 3)) * sf)
            dlo.append(rlo[-1] - int(rng.integers(0, 2)) * sc)
            dup.append(rup[-1] + int(rng.integers(0, 2)) * sc)
        sb = (tuple(slo), tuple(sup), (sf,) * 3)
        db = (tuple(dlo), tuple(dup), (sc,) * 3)
        rb = (tuple(rlo), tuple(rup), (sc,) * 3)
        src = _src_host(rng, sb, "random")
        dst = host_array(db, pad_of(sizes(db), rng))
        exp = orc.HostArr(db, dst.data.copy())
        f = orc.restrict_3d_cc_rf2 if cell else orc.restrict_3d_rf2
        assert f(src, exp, rb) == 0
        dsrc, ddst = _dev(src), _dev(dst)
        (cg.restrict_3d_cc_rf2 if cell else cg.restrict_3d_rf2)(dsrc, ddst, cg.Box(*rb))
        _check(exp, ddst, f"restrict cell={cell} case {c}")


def test_copy_parity():
    rng = np.random.default_rng(99)
    for c in range(8):
        print(f"copy #{c}", flush=True)
        s = int(rng.integers(1, 4))
        n = tuple(int(v) for v in rng.integers(1, 50, size=3))
        if c == 1:
            n = (3, 64, 64)  # a ghost slab
        rlo = tuple(int(v) * s for v in rng.integers(-10, 10, size=3))
        rup = tuple(l + (k - 1) * s for l, k in zip(rlo, n))
        off = tuple(int(v) for v in rng.integers(-3, 4, size=3)) if c % 2 else None
        srlo = tuple(l + (o or 0) * s for l, o in zip(rlo, off or (0, 0, 0)))
        srup = tuple(u + (o or 0) * s for u, o in zip(rup, off or (0, 0, 0)))
        sb = (tuple(l - int(rng.integers(0, 3)) * s for l in srlo),
              tuple(u + int(rng.integers(0, 3)) * s for u in srup), (s,) * 3)
        db = (tuple(l - int(rng.integers(0, 3)) * s for l in rlo),
              tuple(u + int(rng.integers(0, 3)) * s for u in rup), (s,) * 3)
        rb = (rlo, rup, (s,) * 3)
        srb = (srlo, srup, (s,) * 3)
        src = _src_host(rng, sb, "random")
        dst = host_array(db, pad_of(sizes(db), rng))
        exp = orc.HostArr(db, dst.data.copy())
        assert orc.copy_3d(src, exp, srb, rb, off) == 0
        dsrc, ddst = _dev(src), _dev(dst)
        cg.copy_3d(dsrc, ddst, cg.Box(*srb), cg.Box(*rb), slab_offset=off)
        _check(exp, ddst, f"copy case {c}")


@pytest.mark.parametrize("n", [1, 2, 1023, 2048, 2049, 100001])
def test_stream_copy(n):
    """The flat stream copy (the benchmark's copy ceiling) copies exactly n
    doubles, odd n included, and nothing beyond them."""
    import ctypes as C
    from carpet_amd._lib import check, lib
    m = n + 64
    box = cg.Box((0, 0, 0), (m - 1, 0, 0), (1, 1, 1))
    a, b = cg.DeviceArray(box, padext=(m, 1, 1)), cg.DeviceArray(box, padext=(m, 1, 1))
    src = np.random.default_rng(n).standard_normal(m)
    a.upload(src.reshape(1, 1, m))
    b.fill(-7.0)
    check(lib.cg_stream_copy_f64(C.c_void_p(b.ptr), C.c_void_p(a.ptr), n, None), "cg_stream_copy_f64")
    got = b.download().reshape(-1)
    assert np.array_equal(got[:n], src[:n])
    assert np.all(got[n:] == -7.0)
    a.free()
    b.free()


@pytest.mark.parametrize("ntl", [2, 3])
def test_interpolate_parity(ntl):
    rng = np.random.default_rng(5 + ntl)
    for c in range(5):
        print(f"interp ntl={ntl} #{c}", flush=True)
        s = 1
        n = tuple(int(v) for v in rng.integers(1, 40, size=3))
        rlo = tuple(int(v) for v in rng.integers(-5, 5, size=3))
        rup = tuple(l + k - 1 for l, k in zip(rlo, n))
        sb = (tuple(l - 1 for l in rlo), tuple(u + 2 for u in rup), (s,) * 3)
        db = (tuple(l - 2 for l in rlo), tuple(u + 1 for u in rup), (s,) * 3)
        rb = (rlo, rup, (s,) * 3)
        pad = pad_of(sizes(sb), rng)
        srcs = []
        for _ in range(ntl):
            h = host_array(sb, pad)
            h.view()[...] = field(rng, sizes(sb), "random")
            srcs.append(h)
        times = [1.0, 0.0, -1.0][:ntl]
        t = float(rng.uniform(min(times), max(times)))
        dst = host_array(db, pad_of(sizes(db), rng))
        exp = orc.HostArr(db, dst.data.copy())
        assert orc.interpolate_3d(srcs, times, exp, t, rb) == 0
        ds = [_dev(h) for h in srcs]
        ddst = _dev(dst)
        if ntl == 2:
            cg.interpolate_3d_2tl(ds[0], times[0], ds[1], times[1], ddst, t, cg.Box(*rb))
        else:
            cg.interpolate_3d_3tl(ds[0], times[0], ds[1], times[1], ds[2], times[2], ddst, t, cg.Box(*rb))
        _check(exp, ddst, f"interp ntl={ntl} case {c}")


KIND_IDS = {"vlag": (orc.KIND_VLAG, cg.OP_LAGRANGE, cg.VERTEX_CENTERED),
            "clag": (orc.KIND_CLAG, cg.OP_LAGRANGE, cg.CELL_CENTERED),
            "eno": (orc.KIND_ENO, cg.OP_ENO, cg.CELL_CENTERED),
            "tvd": (orc.KIND_TVD, cg.OP_TVD, cg.CELL_CENTERED),
            "veno": (orc.KIND_VENO, cg.OP_ENO, cg.VERTEX_CENTERED),
            "mono": (orc.KIND_MONO5, cg.OP_LAGRANGE_MONOTONE, cg.VERTEX_CENTERED),
            "weno": (orc.KIND_VENO, cg.OP_WENO, cg.VERTEX_CENTERED)}


def test_interpolate_eno_3d_3tl():
    """interpolate_eno_3d_3tl against the oracle, bit for bit, on independent
    random levels (the quadratic overshoots often, so both branches run) and
    on times in either order."""
    rng = np.random.default_rng(77)
    for c, (times, frac) in enumerate((([0.5, 0.25, 0.0], 0.3), ([0.0, 0.25, 0.5], 0.8),
                                       ([0.5, 0.25, 0.0], 0.9), ([0.2, 0.0, 0.1], 0.55))):
        print(f"interp_eno case {c}", flush=True)
        n = tuple(int(v) for v in rng.integers(5, 60, 3))
        box = ((0, 0, 0), tuple(v - 1 for v in n), (1, 1, 1))
        sz = sizes(box)
        rb = (tuple(int(rng.integers(0, 3)) for _ in range(3)),
              tuple(v - 1 - int(rng.integers(0, 3)) for v in sz), (1, 1, 1))
        pad = pad_of(sz, rng)
        srcs = []
        for _ in range(3):
            h = host_array(box, pad)
            h.view()[...] = field(rng, sz, "random")
            srcs.append(h)
        t = min(times) + frac * (max(times) - min(times))
        dst = host_array(box, pad_of(sz, rng))
        exp = orc.HostArr(box, dst.data.copy())
        assert orc.interpolate_eno_3d_3tl(srcs, times, exp, t, rb) == 0
        ds = [_dev(h) for h in srcs]
        ddst = _dev(dst)
        cg.interpolate_eno_3d_3tl(ds[0], times[0], ds[1], times[1], ds[2], times[2], ddst, t, cg.Box(*rb))
        _check(exp, ddst, f"interp_eno case {c}")


@pytest.mark.parametrize("kind,order,ntl", [("mono", 5, 3), ("mono", 5, 2), ("mono", 5, 1), ("weno", 3, 3),
                                            ("veno", 3, 3), ("veno", 3, 2), ("veno", 3, 1),
                                            ("vlag", 5, 3), ("vlag", 3, 2), ("vlag", 1, 3),
                                            ("clag", 3, 3), ("eno", 3, 2), ("eno", 2, 2),
                                            ("eno", 3, 3), ("eno", 2, 3), ("tvd", 1, 2), ("tvd", 1, 3)])
def test_transfer_time_interpolation(kind, order, ntl):
    """cg_transfer_from with time interpolation (one fused kernel) against the
    reference's shape (prolongate every level into a temporary, then
    interpolate_3d_2tl/3tl, or interpolate_eno_3d_3tl for ENO/WENO/TVD/monotone groups with
    3 levels), data.cc:497-521, 1182-1299; also through a batch."""
    rng = np.random.default_rng(31 * order + ntl)
    okind, op, cent = KIND_IDS[kind]
    # ENO order_space 5 -> ENO<3>, 2..4 -> ENO<2> (data.cc:697-704)
    order_space = {("eno", 3): 5, ("eno", 2): 3, ("veno", 3): 5, ("weno", 3): 5}.get((kind, order), order)
    for c in range(4):
        print(f"transfer {kind} o{order} ntl{ntl} #{c}", flush=True)
        sb, db, rb = prolong_case(rng, kind, order, n=(70, 20, 40) if c == 0 else None)
        pad = pad_of(sizes(sb), rng)
        srcs = []
        for _ in range(ntl):
            h = host_array(sb, pad)
            h.view()[...] = field(rng, sizes(sb), "mixed")
            srcs.append(h)
        times = [0.5, 0.25, 0.0][:ntl]
        t = min(times) + [0.4, 0.3, 0.1, 0.125][c] * (max(times) - min(times)) / 0.5
        dst = host_array(db, pad_of(sizes(db), rng))
        exp = orc.HostArr(db, dst.data.copy())
        assert orc.prolongate_time(okind, order, srcs, times, exp, t, rb, ntl - 1) == 0
        ds = [_dev(h) for h in srcs]
        ddst = _dev(dst)
        cg.transfer_from(ddst, ds, times, cg.Box(*rb), cg.Box(*rb), t, op, cent, order_space, ntl - 1)
        _check(exp, ddst, f"transfer {kind} case {c}")
        if c == 1:
            bdst = _dev(dst)
            b = cg.Batch()
            b.add(0, bdst, ds, times, cg.Box(*rb), cg.Box(*rb), t, op, cent, order_space, ntl - 1)
            b.finalize()
            b.run()
            cg.synchronize()
            _check(exp, bdst, f"batched transfer {kind} case {c}")
            b.close()


def _cc_restrict_case(rng, halo, n):
    """Cell-centred restriction geometry: the source holds the region's fine
    footprint plus `halo` cells per side (and 0..2 more)."""
    sf, sc = 2, 4
    rlo, rup, slo, sup, dlo, dup = [], [], [], [], [], []
    for d in range(3):
        C0 = int(rng.integers(-4, 5))
        rlo.append(sc // 2 + C0 * sc)
        rup.append(sc // 2 + (C0 + n[d] - 1) * sc)
        F0 = 2 * C0 - halo - int(rng.integers(0, 3))
        slo.append(sf // 2 + F0 * sf)
        sup.append(sf // 2 + (2 * (C0 + n[d]) - 1 + halo + int(rng.integers(0, 3))) * sf)
        dlo.append(rlo[-1] - int(rng.integers(0, 2)) * sc)
        dup.append(rup[-1] + int(rng.integers(0, 2)) * sc)
    return ((tuple(slo), tuple(sup), (sf,) * 3), (tuple(dlo), tuple(dup), (sc,) * 3),
            (tuple(rlo), tuple(rup), (sc,) * 3))


@pytest.mark.parametrize("order", [3, 5])
def test_restrict_cc_ho_parity(order):
    """restrict_3d_cc_o3_rf2 / restrict_3d_cc_o5_rf2 (bit-exact)."""
    rng = np.random.default_rng(70 + order)
    halo = (order - 1) // 2
    fn = cg.restrict_3d_cc_o3_rf2 if order == 3 else cg.restrict_3d_cc_o5_rf2
    for c in range(6):
        print(f"restrict_cc_o{order} #{c}", flush=True)
        n = (1, 1, 1) if c == 0 else (70, 3, 41) if c == 1 else tuple(
            int(v) for v in rng.integers(1, 40, size=3))
        sb, db, rb = _cc_restrict_case(rng, halo, n)
        src = _src_host(rng, sb, ("random", "smooth", "mixed")[c % 3])
        dst = host_array(db, pad_of(sizes(db), rng))
        exp = orc.HostArr(db, dst.data.copy())
        assert orc.restrict_3d_cc_ho_rf2(order, src, exp, rb) == 0
        dsrc, ddst = _dev(src), _dev(dst)
        fn(dsrc, ddst, cg.Box(*rb))
        _check(exp, ddst, f"restrict_cc_o{order} case {c} reg={rb}")


@pytest.mark.parametrize("order", [3, 5])
def test_restrict_cc_ho_needs_footprint(order):
    """A source without the stencil's extra fine cells is refused before any
    launch and the destination is left as it was."""
    rng = np.random.default_rng(80 + order)
    fn = cg.restrict_3d_cc_o3_rf2 if order == 3 else cg.restrict_3d_cc_o5_rf2
    sb, db, rb = _cc_restrict_case(rng, 0, (5, 4, 3))
    lo, up, st = sb
    # trim the source to exactly the order-1 footprint on the lower x side
    sb = ((rb[0][0] - 1, lo[1], lo[2]), up, st)
    src = _src_host(rng, sb, "random")
    dst = host_array(db, pad_of(sizes(db), rng))
    dsrc, ddst = _dev(src), _dev(dst)
    with pytest.raises(cg.CarpetError) as e:
        fn(dsrc, ddst, cg.Box(*rb))
    assert e.value.code == -6
    _check(orc.HostArr(db, dst.data.copy()), ddst, "untouched destination")


@pytest.mark.parametrize("order", [3, 5])
def test_transfer_higher_order_restriction(order):
    """cg_transfer_from on a finer cell-centred source follows
    use_higher_order_restriction / restriction_order_space for Lagrange groups
    and keeps the conservative stencil for ENO groups (data.cc transfer_restrict);
    also through a batch."""
    rng = np.random.default_rng(90 + order)
    _, op_lag, cent = KIND_IDS["clag"]
    op_eno = KIND_IDS["eno"][1]
    halo = (order - 1) // 2
    cg.set_restriction(True, order)
    try:
        for c, op in enumerate((op_lag, op_eno, op_lag)):
            print(f"transfer restrict o{order} op{op} #{c}", flush=True)
            sb, db, rb = _cc_restrict_case(rng, halo, tuple(int(v) for v in rng.integers(2, 30, size=3)))
            src = _src_host(rng, sb, "random")
            dst = host_array(db, pad_of(sizes(db), rng))
            exp = orc.HostArr(db, dst.data.copy())
            if op == op_lag:
                assert orc.restrict_3d_cc_ho_rf2(order, src, exp, rb) == 0
            else:
                assert orc.restrict_3d_cc_rf2(src, exp, rb) == 0
            dsrc, ddst = _dev(src), _dev(dst)
            cg.transfer_from(ddst, [dsrc], [0.0], cg.Box(*rb), cg.Box(*rb), 0.0, op, cent, 3, 0)
            _check(exp, ddst, f"transfer restrict o{order} op{op} case {c}")
            if c == 2:
                bdst = _dev(dst)
                b = cg.Batch()
                b.add(0, bdst, [dsrc], [0.0], cg.Box(*rb), cg.Box(*rb), 0.0, op, cent, 3, 0)
                b.finalize()
                b.run()
                cg.synchronize()
                _check(exp, bdst, f"batched transfer restrict o{order}")
                b.close()
    finally:
        cg.set_restriction(False, 3)
    with pytest.raises(cg.CarpetError):
        cg.set_restriction(True, 4)


def test_vertex_eno_dispatch():
    """Vertex-centred op ENO (order_space 3, 5) and op WENO (5) run
    prolongate_3d_eno (data.cc:654-762); the other vertex-centred ENO/WENO
    orders, cell-centred WENO and vertex-centred TVD (whose reference routine
    aborts) are refused before any launch."""
    rng = np.random.default_rng(4242)
    sb, db, rb = prolong_case(rng, "veno", 3, n=(21, 14, 9))
    src = _src_host(rng, sb, "mixed")
    dst = host_array(db, pad_of(sizes(db), rng))
    exp = orc.HostArr(db, dst.data.copy())
    assert orc.prolongate_3d_eno(src, exp, rb) == 0
    dsrc = _dev(src)
    for op, order in ((cg.OP_ENO, 3), (cg.OP_ENO, 5), (cg.OP_WENO, 5)):
        ddst = _dev(dst)
        cg.transfer_from(ddst, [dsrc], [0.0], cg.Box(*rb), cg.Box(*rb), 0.0, op, cg.VERTEX_CENTERED, order, 0)
        _check(exp, ddst, f"vertex op {op} order {order}")
    ddst = _dev(dst)
    for op, order, cent in ((cg.OP_ENO, 1, cg.VERTEX_CENTERED), (cg.OP_WENO, 3, cg.VERTEX_CENTERED),
                            (cg.OP_WENO, 5, cg.CELL_CENTERED), (cg.OP_TVD, 1, cg.VERTEX_CENTERED)):
        with pytest.raises(cg.CarpetError):
            cg.transfer_from(ddst, [dsrc], [0.0], cg.Box(*rb), cg.Box(*rb), 0.0, op, cent, order, 0)
    _check(orc.HostArr(db, dst.data.copy()), ddst, "refused transfers leave the destination untouched")


def test_monotone_dispatch():
    """op Lagrange_monotone with order_space 5 runs
    prolongate_3d_o5_monotone_rf2 (data.cc:807-832); orders 1 and 3 (and
    any other) are refused, as is cell centring, before any launch."""
    rng = np.random.default_rng(4343)
    sb, db, rb = prolong_case(rng, "mono", 5, n=(19, 12, 11))
    src = _src_host(rng, sb, "random")
    dst = host_array(db, pad_of(sizes(db), rng))
    exp = orc.HostArr(db, dst.data.copy())
    assert orc.prolongate_3d_o5_monotone_rf2(src, exp, rb) == 0
    dsrc, ddst = _dev(src), _dev(dst)
    cg.transfer_from(ddst, [dsrc], [0.0], cg.Box(*rb), cg.Box(*rb), 0.0, cg.OP_LAGRANGE_MONOTONE,
                     cg.VERTEX_CENTERED, 5, 0)
    _check(exp, ddst, "Lagrange_monotone order 5")
    ddst = _dev(dst)
    for order, cent in ((1, cg.VERTEX_CENTERED), (3, cg.VERTEX_CENTERED), (7, cg.VERTEX_CENTERED),
                        (5, cg.CELL_CENTERED)):
        with pytest.raises(cg.CarpetError):
            cg.transfer_from(ddst, [dsrc], [0.0], cg.Box(*rb), cg.Box(*rb), 0.0, cg.OP_LAGRANGE_MONOTONE, cent,
                             order, 0)
    _check(orc.HostArr(db, dst.data.copy()), ddst, "refused transfers leave the destination untouched")


# ---- staggered operators ---------------------------------------------------
STAGGERS = [(cg.STAGGER011, cg.OP_STAGGER011), (cg.STAGGER101, cg.OP_STAGGER101),
            (cg.STAGGER110, cg.OP_STAGGER110), (cg.STAGGER111, cg.OP_STAGGER111)]


@pytest.fixture
def staggered_on():
    cg.set_staggered_operators(True)
    yield
    cg.set_staggered_operators(False)


@pytest.mark.parametrize("order", [2, 3])
@pytest.mark.parametrize("stag", [s for s, _ in STAGGERS], ids=["011", "101", "110", "111"])
def test_stagger_prolong_parity(stag, order, staggered_on):
    """prolongate_3d_stagger{011,101,110,111}<ORDER> against the oracle, bit
    for bit, on regions of both parities (single points, thin slabs, blocks)."""
    rng = np.random.default_rng(900 + 10 * stag + order)
    for c, n in enumerate(((1, 1, 1), (2, 1, 3), (31, 3, 17), None, None)):
        print(f"stagger {stag} o{order} #{c}", flush=True)
        sb, db, rb = prolong_case(rng, "stag", order, n=n, scale=1 + (c % 2))
        src = _src_host(rng, sb, ("mixed", "smooth", "random")[c % 3])
        dst = host_array(db, pad_of(sizes(db), rng))
        exp = orc.HostArr(db, dst.data.copy())
        assert orc.prolongate_3d_stagger(stag, order, src, exp, rb) == 0
        dsrc, ddst = _dev(src), _dev(dst)
        cg.prolongate_3d_stagger(stag, order, dsrc, ddst, cg.Box(*rb))
        _check(exp, ddst, f"stagger {stag} o{order} case {c} reg={rb}")


def _stagger_restrict_case(rng, n, sf=1):
    """Vertex geometry; the source holds one fine point below the region's
    children and two above (the staggered directions' footprint), and 0..2
    more."""
    sc = 2 * sf
    rlo, rup, slo, sup, dlo, dup = [], [], [], [], [], []
    for d in range(3):
        C0 = int(rng.integers(-4, 5))
        rlo.append(C0 * sc)
        rup.append((C0 + n[d] - 1) * sc)
        slo.append(rlo[-1] - (1 + int(rng.integers(0, 3))) * sf)
        sup.append(rup[-1] + (2 + int(rng.integers(0, 3))) * sf)
        dlo.append(rlo[-1] - int(rng.integers(0, 2)) * sc)
        dup.append(rup[-1] + int(rng.integers(0, 2)) * sc)
    return ((tuple(slo), tuple(sup), (sf,) * 3), (tuple(dlo), tuple(dup), (sc,) * 3),
            (tuple(rlo), tuple(rup), (sc,) * 3))


@pytest.mark.parametrize("stag", [s for s, _ in STAGGERS], ids=["011", "101", "110", "111"])
def test_stagger_restrict_parity(stag, staggered_on):
    """restrict_3d_stagger{011,101,110,111} against the oracle, bit for bit."""
    rng = np.random.default_rng(950 + stag)
    for c in range(5):
        n = (1, 1, 1) if c == 0 else tuple(int(v) for v in rng.integers(1, 30, size=3))
        print(f"stagger restrict {stag} #{c} n={n}", flush=True)
        sb, db, rb = _stagger_restrict_case(rng, n, sf=1 + c % 2)
        src = _src_host(rng, sb, "random")
        dst = host_array(db, pad_of(sizes(db), rng))
        exp = orc.HostArr(db, dst.data.copy())
        assert orc.restrict_3d_stagger(stag, src, exp, rb) == 0
        dsrc, ddst = _dev(src), _dev(dst)
        cg.restrict_3d_stagger(stag, dsrc, ddst, cg.Box(*rb))
        _check(exp, ddst, f"stagger restrict {stag} case {c}")


def test_stagger_requires_parameter():
    """Without support_staggered_operators the staggered operators refuse
    to run (prolongate_3d_stagger011.cc:60-63), per call and through the
    dispatch, and leave the destination untouched."""
    cg.set_staggered_operators(False)
    rng = np.random.default_rng(4711)
    sb, db, rb = prolong_case(rng, "stag", 3, n=(7, 5, 6))
    src = _src_host(rng, sb, "random")
    dst = host_array(db, pad_of(sizes(db), rng))
    dsrc, ddst = _dev(src), _dev(dst)
    with pytest.raises(cg.CarpetError) as ei:
        cg.prolongate_3d_stagger(cg.STAGGER011, 3, dsrc, ddst, cg.Box(*rb))
    assert ei.value.code == -11  # CG_ERR_OPERATOR
    with pytest.raises(cg.CarpetError):
        cg.transfer_from(ddst, [dsrc], [0.0], cg.Box(*rb), cg.Box(*rb), 0.0, cg.OP_STAGGER111,
                         cg.VERTEX_CENTERED, 3, 0)
    rsb, rdb, rrb = _stagger_restrict_case(rng, (4, 3, 5))
    rsrc = _src_host(rng, rsb, "random")
    rdst = host_array(rdb, pad_of(sizes(rdb), rng))
    with pytest.raises(cg.CarpetError):
        cg.restrict_3d_stagger(cg.STAGGER101, _dev(rsrc), _dev(rdst), cg.Box(*rrb))
    _check(orc.HostArr(db, dst.data.copy()), ddst, "refused transfers leave the destination untouched")


@pytest.mark.parametrize("stag,op", STAGGERS, ids=["011", "101", "110", "111"])
def test_stagger_dispatch(stag, op, staggered_on):
    """op STAGGERxyz through cg_transfer_from: prolongation with order_space
    2 -> ORDER 2 and 3..5 -> ORDER 3 (data.cc:834-920), with 1-3 time levels
    (each prolongated, then Lagrange in time) and through a batch; other
    orders refused; restriction whatever the centring (data.cc:989-1027)."""
    rng = np.random.default_rng(990 + stag)
    orc.set_stagger(stag)
    for c, (order_space, ntl) in enumerate(((2, 1), (3, 3), (5, 2), (4, 1))):
        order = 2 if order_space == 2 else 3
        print(f"stagger dispatch {stag} order_space {order_space} ntl {ntl}", flush=True)
        sb, db, rb = prolong_case(rng, "stag", order)
        pad = pad_of(sizes(sb), rng)
        srcs = []
        for _ in range(ntl):
            h = host_array(sb, pad)
            h.view()[...] = field(rng, sizes(sb), "mixed")
            srcs.append(h)
        times = [0.5, 0.25, 0.0][:ntl]
        t = 0.3 if ntl > 1 else 0.5
        dst = host_array(db, pad_of(sizes(db), rng))
        exp = orc.HostArr(db, dst.data.copy())
        assert orc.prolongate_time(orc.KIND_STAGGER, order, srcs, times, exp, t, rb, ntl - 1) == 0
        ds = [_dev(h) for h in srcs]
        ddst = _dev(dst)
        cg.transfer_from(ddst, ds, times, cg.Box(*rb), cg.Box(*rb), t, op, cg.VERTEX_CENTERED, order_space, ntl - 1)
        _check(exp, ddst, f"stagger dispatch {stag} case {c}")
        if c == 1:
            bdst = _dev(dst)
            b = cg.Batch()
            b.add(0, bdst, ds, times, cg.Box(*rb), cg.Box(*rb), t, op, cg.VERTEX_CENTERED, order_space, ntl - 1)
            b.finalize()
            b.run()
            cg.synchronize()
            _check(exp, bdst, f"batched stagger {stag}")
            b.close()
    sb, db, rb = prolong_case(rng, "stag", 3, n=(6, 5, 4))
    dsrc, ddst = _dev(_src_host(rng, sb, "random")), _dev(host_array(db, pad_of(sizes(db), rng)))
    for order_space in (1, 6):
        with pytest.raises(cg.CarpetError):
            cg.transfer_from(ddst, [dsrc], [0.0], cg.Box(*rb), cg.Box(*rb), 0.0, op, cg.VERTEX_CENTERED,
                             order_space, 0)
    for cent in (cg.VERTEX_CENTERED, cg.CELL_CENTERED):
        rsb, rdb, rrb = _stagger_restrict_case(rng, (9, 4, 7))
        rsrc = _src_host(rng, rsb, "random")
        rdst = host_array(rdb, pad_of(sizes(rdb), rng))
        rexp = orc.HostArr(rdb, rdst.data.copy())
        assert orc.restrict_3d_stagger(stag, rsrc, rexp, rrb) == 0
        drdst = _dev(rdst)
        cg.transfer_from(drdst, [_dev(rsrc)], [0.0], cg.Box(*rrb), cg.Box(*rrb), 0.0, op, cent, 3, 0)
        _check(rexp, drdst, f"stagger restriction dispatch {stag} cent {cent}")


# ---- 4 and 5 time levels ---------------------------------------------------
@pytest.mark.parametrize("ntl", [4, 5])
def test_interpolate_4tl_5tl_parity(ntl):
    """interpolate_3d_4tl / interpolate_3d_5tl against the oracle, bit for
    bit, per call and through cg_transfer_from (same level, order_time
    ntl-1)."""
    rng = np.random.default_rng(40 + ntl)
    times = [0.5, 0.25, 0.0, -0.25, -0.5][:ntl]
    for c in range(3):
        print(f"interp ntl={ntl} #{c}", flush=True)
        n = tuple(int(v) for v in rng.integers(1, 40, size=3))
        rlo = tuple(int(v) for v in rng.integers(-5, 5, size=3))
        rup = tuple(l + k - 1 for l, k in zip(rlo, n))
        sb = (tuple(l - 1 for l in rlo), tuple(u + 2 for u in rup), (1,) * 3)
        db = (tuple(l - 2 for l in rlo), tuple(u + 1 for u in rup), (1,) * 3)
        rb = (rlo, rup, (1,) * 3)
        pad = pad_of(sizes(sb), rng)
        srcs = []
        for _ in range(ntl):
            h = host_array(sb, pad)
            h.view()[...] = field(rng, sizes(sb), "random")
            srcs.append(h)
        t = float(rng.uniform(min(times), max(times)))
        dst = host_array(db, pad_of(sizes(db), rng))
        exp = orc.HostArr(db, dst.data.copy())
        assert orc.interpolate_3d(srcs, times, exp, t, rb) == 0
        ds = [_dev(h) for h in srcs]
        ddst = _dev(dst)
        cg.interpolate_3d_ntl(ds, times, ddst, t, cg.Box(*rb))
        _check(exp, ddst, f"interp ntl={ntl} case {c}")
        ddst = _dev(dst)
        cg.transfer_from(ddst, ds, times, cg.Box(*rb), cg.Box(*rb), t, cg.OP_LAGRANGE, cg.VERTEX_CENTERED, 3,
                         ntl - 1)
        _check(exp, ddst, f"transfer interp ntl={ntl} case {c}")


@pytest.mark.parametrize("kind,order,ntl", [("vlag", 5, 4), ("vlag", 3, 5), ("clag", 3, 4), ("clag", 2, 5),
                                            ("stag", 3, 4), ("stag", 2, 5)])
def test_transfer_4tl_5tl_prolongation(kind, order, ntl):
    """Prolongation with 4 or 5 time levels (order_time 3, 4): each level
    into a temporary, then interpolate_3d_4tl/5tl, as data.cc:497-521 does;
    equal to the oracle's replay of that shape bit for bit.  The batch
    refuses it, and ENO-type groups refuse more than 3 levels."""
    rng = np.random.default_rng(60 * ntl + order + len(kind))
    okind, op, cent = {"vlag": (orc.KIND_VLAG, cg.OP_LAGRANGE, cg.VERTEX_CENTERED),
                       "clag": (orc.KIND_CLAG, cg.OP_LAGRANGE, cg.CELL_CENTERED),
                       "stag": (orc.KIND_STAGGER, cg.OP_STAGGER111, cg.VERTEX_CENTERED)}[kind]
    order_space = {("stag", 3): 4}.get((kind, order), order)
    if kind == "stag":
        cg.set_staggered_operators(True)
        orc.set_stagger(cg.STAGGER111)
    try:
        times = [0.5, 0.25, 0.0, -0.25, -0.5][:ntl]
        for c in range(3):
            print(f"transfer {kind} o{order} ntl{ntl} #{c}", flush=True)
            sb, db, rb = prolong_case(rng, kind, order, n=(40, 9, 13) if c == 0 else None)
            pad = pad_of(sizes(sb), rng)
            srcs = []
            for _ in range(ntl):
                h = host_array(sb, pad)
                h.view()[...] = field(rng, sizes(sb), "mixed")
                srcs.append(h)
            t = [0.3, 0.1, -0.2][c]
            dst = host_array(db, pad_of(sizes(db), rng))
            exp = orc.HostArr(db, dst.data.copy())
            assert orc.prolongate_time(okind, order, srcs, times, exp, t, rb, ntl - 1) == 0
            ds = [_dev(h) for h in srcs]
            ddst = _dev(dst)
            cg.transfer_from(ddst, ds, times, cg.Box(*rb), cg.Box(*rb), t, op, cent, order_space, ntl - 1)
            _check(exp, ddst, f"transfer {kind} ntl {ntl} case {c}")
        b = cg.Batch()
        with pytest.raises(cg.CarpetError):
            b.add(0, ddst, ds, times, cg.Box(*rb), cg.Box(*rb), t, op, cent, order_space, ntl - 1)
        b.close()
    finally:
        cg.set_staggered_operators(False)
    if kind != "stag":
        with pytest.raises(cg.CarpetError):
            cg.transfer_from(ddst, ds, times, cg.Box(*rb), cg.Box(*rb), t, cg.OP_TVD if kind == "clag" else cg.OP_ENO,
                             cent, 1 if kind == "clag" else 5, ntl - 1)


# ---- restriction and slabbed copies with time interpolation ----------------
def _tmp_interp(srcs, times, t, rb, per_level, eno):
    """The reference's transfer_time (data.cc:497-521): each level into a
    region-sized temporary, then interpolated in time."""
    tmps = []
    for h in srcs:
        tmp = host_array(rb, sizes(rb))
        assert per_level(h, tmp) == 0
        tmps.append(tmp)

    def finish(exp):
        if eno:
            return orc.interpolate_eno_3d_3tl(tmps, times, exp, t, rb)
        return orc.interpolate_3d(tmps, times, exp, t, rb)
    return finish


@pytest.mark.parametrize("what,op,ntl", [("vertex", cg.OP_LAGRANGE, 2), ("vertex", cg.OP_LAGRANGE, 3),
                                         ("cell", cg.OP_LAGRANGE, 3), ("cell", cg.OP_ENO, 3),
                                         ("cell", cg.OP_LAGRANGE, 5), ("stagger", cg.OP_STAGGER011, 3)])
def test_restriction_with_time_interpolation(what, op, ntl):
    """Restriction between levels at a time between the source levels:
    every level restricted into a temporary, then interpolated in time
    (interpolate_eno_3d_3tl for ENO groups), bit for bit against the oracle
    replay of that shape; the batch refuses it."""
    rng = np.random.default_rng(500 + ntl + 7 * len(what) + op)
    times = [0.5, 0.25, 0.0, -0.25, -0.5][:ntl]
    cent = cg.CELL_CENTERED if what == "cell" else cg.VERTEX_CENTERED
    if what == "stagger":
        cg.set_staggered_operators(True)
    try:
        for c in range(3):
            n = (1, 1, 1) if c == 0 else tuple(int(v) for v in rng.integers(2, 20, size=3))
            print(f"restrict+time {what} op {op} ntl {ntl} #{c} n={n}", flush=True)
            if what == "cell":
                sb, db, rb = _cc_restrict_case(rng, 0, n)
                per = orc.restrict_3d_cc_rf2
            else:
                sb, db, rb = _stagger_restrict_case(rng, n)
                per = (lambda s, d, r: orc.restrict_3d_stagger(cg.STAGGER011, s, d, r)) if what == "stagger" \
                    else orc.restrict_3d_rf2
            pad = pad_of(sizes(sb), rng)
            srcs = []
            for _ in range(ntl):
                h = host_array(sb, pad)
                h.view()[...] = field(rng, sizes(sb), "random")
                srcs.append(h)
            t = min(times) + [0.3, 0.7, 0.1][c] * (max(times) - min(times))
            dst = host_array(db, pad_of(sizes(db), rng))
            exp = orc.HostArr(db, dst.data.copy())
            finish = _tmp_interp(srcs, times, t, rb, lambda h, tmp: per(h, tmp, rb), ntl == 3 and op == cg.OP_ENO)
            assert finish(exp) == 0
            ds = [_dev(h) for h in srcs]
            ddst = _dev(dst)
            cg.transfer_from(ddst, ds, times, cg.Box(*rb), cg.Box(*rb), t, op, cent, 3, ntl - 1)
            _check(exp, ddst, f"restrict+time {what} case {c}")
        b = cg.Batch()
        with pytest.raises(cg.CarpetError):
            b.add(0, ddst, ds, times, cg.Box(*rb), cg.Box(*rb), t, op, cent, 3, ntl - 1)
        b.close()
    finally:
        cg.set_staggered_operators(False)


@pytest.mark.parametrize("ntl", [2, 3])
def test_slabbed_copy_with_time_interpolation(ntl):
    """A copy with a slab offset at a time between the source levels: each
    level copied through the slab into a temporary, then interpolated."""
    rng = np.random.default_rng(620 + ntl)
    times = [1.0, 0.0, -1.0][:ntl]
    for c in range(3):
        print(f"slab copy+time ntl {ntl} #{c}", flush=True)
        n = tuple(int(v) for v in rng.integers(1, 30, size=3))
        rlo = tuple(int(v) for v in rng.integers(-10, 10, size=3))
        rup = tuple(l + k - 1 for l, k in zip(rlo, n))
        off = tuple(int(v) for v in rng.integers(-3, 4, size=3))
        srlo = tuple(l + o for l, o in zip(rlo, off))
        srup = tuple(u + o for u, o in zip(rup, off))
        sb = (tuple(l - int(rng.integers(0, 3)) for l in srlo), tuple(u + int(rng.integers(0, 3)) for u in srup),
              (1,) * 3)
        db = (tuple(l - int(rng.integers(0, 3)) for l in rlo), tuple(u + int(rng.integers(0, 3)) for u in rup),
              (1,) * 3)
        rb, srb = (rlo, rup, (1,) * 3), (srlo, srup, (1,) * 3)
        pad = pad_of(sizes(sb), rng)
        srcs = []
        for _ in range(ntl):
            h = host_array(sb, pad)
            h.view()[...] = field(rng, sizes(sb), "random")
            srcs.append(h)
        t = min(times) + [0.4, 0.15, 0.9][c] * (max(times) - min(times))
        dst = host_array(db, pad_of(sizes(db), rng))
        exp = orc.HostArr(db, dst.data.copy())
        finish = _tmp_interp(srcs, times, t, rb, lambda h, tmp: orc.copy_3d(h, tmp, srb, rb, off), False)
        assert finish(exp) == 0
        ds = [_dev(h) for h in srcs]
        ddst = _dev(dst)
        cg.transfer_from(ddst, ds, times, cg.Box(*rb), cg.Box(*srb), t, cg.OP_LAGRANGE, cg.VERTEX_CENTERED, 1,
                         ntl - 1, slab_offset=off)
        _check(exp, ddst, f"slab copy+time case {c}")


# ---- restrict_3d_vc_rf2 -------------------------------------------------------
VC_CENTS = [(0, 1, 1), (1, 0, 1), (1, 1, 0)]


@pytest.mark.parametrize("cent", VC_CENTS, ids=["011", "101", "110"])
def test_restrict_vc_parity(cent):
    """restrict_3d_vc_rf2 against the oracle, bit for bit: per call, through
    cg_transfer_from with a slab of that is_centered (data.cc:1106-1127),
    with 2-level time interpolation (levels restricted into temporaries),
    and in a batch."""
    from test_oracle_pins import _vc_case
    rng = np.random.default_rng(1200 + cent[0] + 2 * cent[1])
    for c in range(4):
        n = (1, 1, 1) if c == 0 else tuple(int(v) for v in rng.integers(2, 40, size=3))
        print(f"restrict_vc {cent} #{c} n={n}", flush=True)
        sb, db, rb = _vc_case(rng, cent, n)
        src = _src_host(rng, sb, "random")
        dst = host_array(db, pad_of(sizes(db), rng))
        exp = orc.HostArr(db, dst.data.copy())
        assert orc.restrict_3d_vc_rf2(cent, src, exp, rb) == 0
        dsrc, ddst = _dev(src), _dev(dst)
        cg.restrict_3d_vc_rf2(cent, dsrc, ddst, cg.Box(*rb))
        _check(exp, ddst, f"restrict_vc {cent} case {c}")
        ddst = _dev(dst)
        cg.transfer_from(ddst, [dsrc], [0.0], cg.Box(*rb), cg.Box(*rb), 0.0, cg.OP_LAGRANGE, cg.CELL_CENTERED, 3, 0,
                         is_centered=cent)
        _check(exp, ddst, f"transfer restrict_vc {cent} case {c}")
        ddst = _dev(dst)
        b = cg.Batch()
        b.add(0, ddst, [dsrc], [0.0], cg.Box(*rb), cg.Box(*rb), 0.0, cg.OP_LAGRANGE, cg.CELL_CENTERED, 3, 0,
              is_centered=cent)
        b.finalize()
        b.run()
        cg.synchronize()
        _check(exp, ddst, f"batched restrict_vc {cent} case {c}")
        b.close()
        # two time levels: each restricted into a temporary, then interpolated
        src2 = orc.HostArr(sb, np.full_like(src.data, np.nan))  # same padding as level 1
        src2.view()[...] = field(rng, sizes(sb), "random")
        texp = orc.HostArr(db, dst.data.copy())
        finish = _tmp_interp([src, src2], [1.0, 0.0], 0.3, rb, lambda h, tmp: orc.restrict_3d_vc_rf2(cent, h, tmp, rb),
                             False)
        assert finish(texp) == 0
        ddst = _dev(dst)
        cg.transfer_from(ddst, [dsrc, _dev(src2)], [1.0, 0.0], cg.Box(*rb), cg.Box(*rb), 0.3, cg.OP_LAGRANGE,
                         cg.CELL_CENTERED, 3, 1, is_centered=cent)
        _check(texp, ddst, f"restrict_vc {cent} with time case {c}")
