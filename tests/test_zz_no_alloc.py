"""cg_transfer_from_noalloc against cg_transfer_from, bit for bit.

The allocation-free entry point stages the time-level temporaries of a
transfer in the library's per-GPU scratch array instead of a stream-ordered
allocation.  Each case runs two calls in a row through the C ABI -- the
second on a smaller region, so it reuses (and must fully overwrite) what the
first left in the scratch array -- and compares the whole padded
destination with what cg_transfer_from writes from the same inputs:
prolongation with 4 and 5 time levels, restriction and a slabbed copy with
time interpolation, and a single-kernel transfer that needs no scratch."""
import numpy as np
import pytest

from helpers import field, first_diff, host_array, pad_of, prolong_case, same, sizes

cg = pytest.importorskip("carpet_amd")

pytestmark = pytest.mark.gpu


def _dev(h):
    lo, up, st = h.box
    a = cg.DeviceArray(cg.Box(lo, up, st), padext=h.padext)
    a.upload(h.data)
    return a


def _both(srcs, dst, *args, **kw):
    """(cg_transfer_from result, cg_transfer_from_noalloc result), padded."""
    ds = [_dev(h) for h in srcs]
    d_ref, d_new = _dev(dst), _dev(dst)
    cg.transfer_from(d_ref, ds, *args, **kw)
    cg.transfer_from_noalloc(d_new, ds, *args, **kw)
    out = d_ref.download_padded(), d_new.download_padded()
    for a in ds + [d_ref, d_new]:
        a.free()
    return out


def _check(ref, got, what):
    assert not np.isnan(ref).all(), what
    if not same(got, ref):
        raise AssertionError(f"{what}: {first_diff(got, ref)}")


@pytest.mark.parametrize("ntl,order", [(4, 5), (5, 3)])
def test_noalloc_prolongation_levels(ntl, order):
    rng = np.random.default_rng(900 + ntl)
    times = [0.5, 0.25, 0.0, -0.25, -0.5][:ntl]
    for c, n in enumerate([(70, 41, 37), (9, 5, 12)]):  # second call smaller
        print(f"noalloc prolongation ntl={ntl} o{order} call {c} n={n}", flush=True)
        sb, db, rb = prolong_case(rng, "vlag", order, n=n)
        pad = pad_of(sizes(sb), rng)
        srcs = []
        for _ in range(ntl):
            h = host_array(sb, pad)
            h.view()[...] = field(rng, sizes(sb), "mixed")
            srcs.append(h)
        dst = host_array(db, pad_of(sizes(db), rng))
        t = [0.3, -0.1][c]
        ref, got = _both(srcs, dst, times, cg.Box(*rb), cg.Box(*rb), t, cg.OP_LAGRANGE, cg.VERTEX_CENTERED,
                         order, ntl - 1)
        _check(ref, got, f"prolongation ntl={ntl} call {c}")


def test_noalloc_restriction_and_slab_copy_with_time():
    rng = np.random.default_rng(911)
    times = [1.0, 0.0, -1.0]
    for c, n in enumerate([(40, 33, 21), (3, 7, 2)]):
        print(f"noalloc restriction+time call {c} n={n}", flush=True)
        # vertex restriction: fine stride 1 source, coarse stride 2 region
        rlo = tuple(2 * int(v) for v in rng.integers(-4, 4, size=3))
        rup = tuple(l + 2 * (k - 1) for l, k in zip(rlo, n))
        sb = (tuple(l - 2 for l in rlo), tuple(u + 3 for u in rup), (1,) * 3)
        db = (tuple(l - 4 for l in rlo), tuple(u + 2 for u in rup), (2,) * 3)
        rb = (rlo, rup, (2,) * 3)
        pad = pad_of(sizes(sb), rng)
        srcs = []
        for _ in range(3):
            h = host_array(sb, pad)
            h.view()[...] = field(rng, sizes(sb), "random")
            srcs.append(h)
        dst = host_array(db, pad_of(sizes(db), rng))
        ref, got = _both(srcs, dst, times, cg.Box(*rb), cg.Box(*rb), [0.4, -0.7][c], cg.OP_LAGRANGE,
                         cg.VERTEX_CENTERED, 3, 2)
        _check(ref, got, f"restriction+time call {c}")
    for c, n in enumerate([(60, 20, 30), (4, 4, 9)]):
        print(f"noalloc slab copy+time call {c} n={n}", flush=True)
        rlo = tuple(int(v) for v in rng.integers(-10, 10, size=3))
        rup = tuple(l + k - 1 for l, k in zip(rlo, n))
        off = tuple(int(v) for v in rng.integers(-3, 4, size=3))
        srlo = tuple(l + o for l, o in zip(rlo, off))
        srup = tuple(u + o for u, o in zip(rup, off))
        sb = (tuple(l - 1 for l in srlo), tuple(u + 2 for u in srup), (1,) * 3)
        db = (tuple(l - 2 for l in rlo), tuple(u + 1 for u in rup), (1,) * 3)
        pad = pad_of(sizes(sb), rng)
        srcs = []
        for _ in range(2):
            h = host_array(sb, pad)
            h.view()[...] = field(rng, sizes(sb), "random")
            srcs.append(h)
        dst = host_array(db, pad_of(sizes(db), rng))
        ref, got = _both(srcs, dst, [1.0, 0.0], cg.Box(rlo, rup, (1,) * 3), cg.Box(srlo, srup, (1,) * 3),
                         [0.35, 0.8][c], cg.OP_LAGRANGE, cg.VERTEX_CENTERED, 1, 1, slab_offset=off)
        _check(ref, got, f"slab copy+time call {c}")


def test_noalloc_single_kernel_transfer():
    """A fused 3-level prolongation needs no scratch: same as transfer_from."""
    rng = np.random.default_rng(923)
    for c, n in enumerate([(50, 30, 20), (6, 6, 6)]):
        sb, db, rb = prolong_case(rng, "vlag", 5, n=n)
        pad = pad_of(sizes(sb), rng)
        srcs = []
        for _ in range(3):
            h = host_array(sb, pad)
            h.view()[...] = field(rng, sizes(sb), "smooth")
            srcs.append(h)
        dst = host_array(db, pad_of(sizes(db), rng))
        ref, got = _both(srcs, dst, [1.0, 0.0, -1.0], cg.Box(*rb), cg.Box(*rb), 1 / 3, cg.OP_LAGRANGE,
                         cg.VERTEX_CENTERED, 5, 2)
        _check(ref, got, f"fused prolongation call {c}")
