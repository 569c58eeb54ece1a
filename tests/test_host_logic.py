"""Host-side logic of libcarpet_amd.so that needs no GPU: the C ABI exports
every declared symbol, pad_shape follows CarpetLib's rules, and malformed
operator calls are rejected with the reference's error class before
anything touches the GPU -- with the same codes the oracle returns."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import pyoracle as orc
from helpers import host_array, sizes

cg = pytest.importorskip("carpet_amd")
from carpet_amd._lib import I3, lib  # noqa: E402

HEADER = Path(__file__).resolve().parent.parent / "include" / "carpet_amd.h"


def test_library_exports_every_declared_symbol():
    names = set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(cg_\w+)\s*\(", HEADER.read_text(), re.M))
    assert len(names) > 30
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_error_strings():
    for code in range(0, -18, -1):
        assert lib.cg_error_string(code).decode() != "unknown error"


def test_pad_shape_rules():
    # x padded to a multiple of the vector size and to whole cache lines
    ps, po = cg.pad_shape((33, 33, 33), (0, 0, 0), 4, 8, False)
    assert ps[0] == 40 and ps[1] * ps[0] % 8 == 0 and po == (0, 0, 0)
    # extents below one line are rounded to a power of two
    ps, _ = cg.pad_shape((3, 5, 7), (0, 0, 0), 1, 16, False)
    assert ps[0] == 4
    # align_interior moves the owned interior to an aligned x position
    ps, po = cg.pad_shape((70, 10, 10), (3, 3, 3), 16, 16, True)
    assert (po[0] + 3) % 16 == 0 and po[0] + 70 <= ps[0]
    # empty arrays are not padded
    assert cg.pad_shape((0, 4, 4), (0, 0, 0), 16, 16, True)[0] == (0, 4, 4)
    # the reference's safety bound prod(padded) <= 2*prod(shape)+1000 holds
    for shape in [(64, 64, 64), (9, 17, 33), (100, 3, 3)]:
        ps, _ = cg.pad_shape(shape, (0, 0, 0), 16, 16, False)
        assert np.prod(ps) <= 2 * np.prod(shape) + 1000 or shape[0] < 16


FAKE = 0x10000  # never dereferenced: every call below fails validation first


def _call_prolong(order, sb, db, rb, sext=None, spad=None, cell=False):
    sext = sext or sizes(sb)
    spad = spad or sext
    dext = sizes(db)
    f = lib.cg_prolongate_3d_cc_rf2_f64 if cell else lib.cg_prolongate_3d_rf2_f64
    return f(order, FAKE, I3(*spad), I3(*sext), FAKE, I3(*dext), I3(*dext),
             cg.Box(*sb).as_c(), cg.Box(*db).as_c(), cg.Box(*rb).as_c(), cg.Box(*rb).as_c(), None)


def _orc_prolong(order, sb, db, rb, cell=False):
    src = host_array(sb, sizes(sb), 0.0)
    dst = host_array(db, sizes(db), 0.0)
    f = orc.prolongate_3d_cc_rf2 if cell else orc.prolongate_3d_rf2
    return f(order, src, dst, rb)


S = ((0, 0, 0), (40, 40, 40), (2, 2, 2))
D = ((0, 0, 0), (40, 40, 40), (1, 1, 1))
BAD = {
    "refinement factor": (S[:2] + ((4, 4, 4),), D, ((8, 8, 8), (12, 12, 12), (1, 1, 1)), -3),
    "strides disagree": (S, D, ((8, 8, 8), (12, 12, 12), (2, 2, 2)), -2),
    "unaligned source": (((1, 0, 0), (41, 40, 40), (2, 2, 2)), D, ((8, 8, 8), (12, 12, 12), (1, 1, 1)), -4),
    "empty region": (S, D, ((8, 8, 8), (7, 12, 12), (1, 1, 1)), -5),
    "outside destination": (S, D, ((8, 8, 8), (44, 12, 12), (1, 1, 1)), -6),
    "footprint outside source": (S, D, ((1, 8, 8), (12, 12, 12), (1, 1, 1)), -6),
}


@pytest.mark.parametrize("case", sorted(BAD))
def test_prolongate_validation_matches_oracle(case):
    sb, db, rb, code = BAD[case]
    assert _call_prolong(5, sb, db, rb) == code
    assert _orc_prolong(5, sb, db, rb) == code
    assert lib.cg_last_error_message().decode()


def test_unknown_orders_rejected():
    rb = ((8, 8, 8), (12, 12, 12), (1, 1, 1))
    for order in (0, 2, 4, 13):
        assert _call_prolong(order, S, D, rb) == -7
    for order in (-1, 6):
        assert _call_prolong(order, S, D, rb, cell=True) == -7
    assert lib.cg_prolongate_3d_cc_eno_rf2_f64(4, FAKE, I3(21, 21, 21), I3(21, 21, 21), FAKE, I3(41, 41, 41),
                                               I3(41, 41, 41), cg.Box(*S).as_c(), cg.Box(*D).as_c(),
                                               cg.Box(*rb).as_c(), cg.Box(*rb).as_c(), None) == -7


def test_shape_and_padding_checks():
    rb = ((8, 8, 8), (12, 12, 12), (1, 1, 1))
    # ext must equal the box sizes; padext >= ext
    assert _call_prolong(5, S, D, rb, sext=(20, 21, 21)) == -10
    assert _call_prolong(5, S, D, rb, spad=(20, 21, 21)) == -10


def test_time_interpolation_validation():
    box = cg.Box((0, 0, 0), (9, 9, 9), (1, 1, 1)).as_c()
    e = I3(10, 10, 10)
    # same time
    assert lib.cg_interpolate_3d_2tl_f64(FAKE, 1.0, FAKE, 1.0 + 1e-13, e, e, FAKE, 1.0, e, e, box, box, box,
                                         box, None) == -8
    # extrapolation
    assert lib.cg_interpolate_3d_3tl_f64(FAKE, 1.0, FAKE, 0.0, FAKE, -1.0, e, e, FAKE, 1.5, e, e, box, box, box,
                                         box, None) == -9
    # 4 and 5 levels (interpolate_3d_4tl.cc:98-108, interpolate_3d_5tl.cc:101-113)
    assert lib.cg_interpolate_3d_4tl_f64(FAKE, 0.5, FAKE, 0.25, FAKE, 0.25 + 1e-13, FAKE, -0.25, e, e, FAKE, 0.3,
                                         e, e, box, box, box, box, None) == -8
    assert lib.cg_interpolate_3d_5tl_f64(FAKE, 0.5, FAKE, 0.25, FAKE, 0.0, FAKE, -0.25, FAKE, -0.5, e, e, FAKE,
                                         0.6, e, e, box, box, box, box, None) == -9


def test_copy_validation():
    box = cg.Box((0, 0, 0), (9, 9, 9), (1, 1, 1)).as_c()
    e = I3(10, 10, 10)
    reg = cg.Box((2, 2, 2), (5, 5, 5), (1, 1, 1)).as_c()
    far = cg.Box((2, 2, 2), (12, 5, 5), (1, 1, 1)).as_c()
    assert lib.cg_copy_3d_f64(FAKE, e, e, FAKE, e, e, box, box, far, far, None, None) == -6
    s2 = cg.Box((0, 0, 0), (18, 18, 18), (2, 2, 2)).as_c()
    assert lib.cg_copy_3d_f64(FAKE, e, e, FAKE, e, e, s2, box, reg, reg, None, None) == -2
    empty = cg.Box((2, 2, 2), (1, 5, 5), (1, 1, 1)).as_c()
    assert lib.cg_copy_3d_f64(FAKE, e, e, FAKE, e, e, box, box, empty, empty, None, None) == -5


def test_stream_copy_validation():
    """The flat stream copy takes 8-byte aligned arrays only; null arrays
    are an error unless n is 0."""
    assert lib.cg_stream_copy_f64(FAKE + 4, FAKE, 10, None) == -1
    assert "8-byte" in lib.cg_last_error_message().decode()
    assert lib.cg_stream_copy_f64(FAKE, FAKE + 12, 10, None) == -1
    assert lib.cg_stream_copy_f64(None, FAKE, 10, None) == -1


def test_transfer_dispatch_validation():
    """cg_transfer_from rejects operator/centering/order combinations the
    reference dispatch rejects (data.cc:584-800) before launching."""
    from carpet_amd._lib import GData
    c = GData()
    c.storage = FAKE
    c.padext[:] = [21, 21, 21]
    c.ext[:] = [21, 21, 21]
    c.bbox[:] = [0, 0, 0, 40, 40, 40, 2, 2, 2]
    f = GData()
    f.storage = FAKE
    f.padext[:] = [41, 41, 41]
    f.ext[:] = [41, 41, 41]
    f.bbox[:] = [0, 0, 0, 40, 40, 40, 1, 1, 1]
    srcs = (C.POINTER(GData) * 1)(C.pointer(c))
    reg = cg.Box((8, 8, 8), (12, 12, 12), (1, 1, 1)).as_c()
    tt = (C.c_double * 1)(0.0)

    def call(op, cent, order, time=0.0, nt=1, ot=0):
        return lib.cg_transfer_from(C.byref(f), srcs, tt, nt, reg, reg, None, time, op, cent, order, ot, None)

    assert call(cg.OP_LAGRANGE, cg.VERTEX_CENTERED, 4) == -7  # even vertex order
    assert call(cg.OP_ENO, cg.VERTEX_CENTERED, 1) == -7  # vertex ENO: order_space 3 or 5
    assert call(cg.OP_WENO, cg.VERTEX_CENTERED, 3) == -7  # vertex WENO: order_space 5
    assert call(cg.OP_WENO, cg.CELL_CENTERED, 5) == -11  # no cell-centred WENO
    assert call(cg.OP_TVD, cg.VERTEX_CENTERED, 1) == -11  # the reference's vertex TVD aborts
    assert call(cg.OP_TVD, cg.CELL_CENTERED, 3) == -7  # TVD order 1 only
    assert call(cg.OP_SYNC, cg.VERTEX_CENTERED, 3) == -11  # op_sync cannot prolongate
    assert call(cg.OP_NONE, cg.VERTEX_CENTERED, 3) == -11
    assert call(cg.OP_LAGRANGE, cg.VERTEX_CENTERED, 3, time=5.0) == -9  # extrapolation in time
    # staggered operators: order_space 2..5 only, and only with
    # support_staggered_operators (prolongate_3d_stagger011.cc:60-63)
    cg.set_staggered_operators(False)
    assert call(cg.OP_STAGGER011, cg.VERTEX_CENTERED, 3) == -11
    cg.set_staggered_operators(True)
    try:
        assert call(cg.OP_STAGGER101, cg.VERTEX_CENTERED, 1) == -7
        assert call(cg.OP_STAGGER110, cg.VERTEX_CENTERED, 6) == -7
    finally:
        cg.set_staggered_operators(False)
    # restriction of a staggered group (fine -> coarse) without the parameter
    csrc = (C.POINTER(GData) * 1)(C.pointer(f))
    creg = cg.Box((8, 8, 8), (12, 12, 12), (2, 2, 2)).as_c()
    assert lib.cg_transfer_from(C.byref(c), csrc, tt, 1, creg, creg, None, 0.0, cg.OP_STAGGER111,
                                cg.CELL_CENTERED, 3, 0, None) == -11


def test_noalloc_rejects_temporaries_beyond_its_scratch():
    """cg_transfer_from_noalloc stages time-level temporaries in a fixed
    per-GPU array: a prolongation with 5 time levels onto a 120^3 region
    needs 5 x 120^3 x 8 bytes, more than CG_NOALLOC_SCRATCH_BYTES (64 MiB),
    and is refused before anything touches the GPU; so is every transfer
    cg_transfer_from refuses."""
    from carpet_amd._lib import GData
    n = 120
    c = GData()
    c.storage = FAKE
    nc = n // 2 + 8
    c.padext[:] = [nc] * 3
    c.ext[:] = [nc] * 3
    c.bbox[:] = [-8, -8, -8, -8 + 2 * (nc - 1), -8 + 2 * (nc - 1), -8 + 2 * (nc - 1), 2, 2, 2]
    f = GData()
    f.storage = FAKE
    f.padext[:] = [n] * 3
    f.ext[:] = [n] * 3
    f.bbox[:] = [0, 0, 0, n - 1, n - 1, n - 1, 1, 1, 1]
    srcs = (C.POINTER(GData) * 5)(*[C.pointer(c)] * 5)
    tt = (C.c_double * 5)(0.5, 0.25, 0.0, -0.25, -0.5)
    reg = cg.Box((0, 0, 0), (n - 1,) * 3, (1, 1, 1)).as_c()
    assert 5 * n**3 * 8 > 64 << 20
    assert lib.cg_transfer_from_noalloc(C.byref(f), srcs, tt, 5, reg, reg, None, 0.1, cg.OP_LAGRANGE,
                                        cg.VERTEX_CENTERED, 5, 4, None) == -17
    assert "scratch" in lib.cg_last_error_message().decode()
    assert lib.cg_error_string(-17).decode() != "unknown error"
    # the same validation as cg_transfer_from comes first
    assert lib.cg_transfer_from_noalloc(C.byref(f), srcs, tt, 5, reg, reg, None, 5.0, cg.OP_LAGRANGE,
                                        cg.VERTEX_CENTERED, 5, 4, None) == -9
    assert lib.cg_transfer_from_noalloc(C.byref(f), srcs, tt, 5, reg, reg, None, 0.1, cg.OP_LAGRANGE,
                                        cg.VERTEX_CENTERED, 4, 4, None) == -7


def test_batch_set_time_keeps_launch_classes():
    """cg_batch_set_time resolves the phase's transfers again; a transfer
    added at an exact time match (one level) cannot start to interpolate
    (its launch class has one level): CG_ERR_STATE, and the batch keeps its
    state; a target outside the levels is the reference's extrapolation
    error.  Host logic only (the batch is never finalized)."""
    from carpet_amd._lib import GData
    c = GData()
    c.storage = FAKE
    c.padext[:] = [21, 21, 21]
    c.ext[:] = [21, 21, 21]
    c.bbox[:] = [0, 0, 0, 40, 40, 40, 2, 2, 2]
    f = GData()
    f.storage = FAKE
    f.padext[:] = [41, 41, 41]
    f.ext[:] = [41, 41, 41]
    f.bbox[:] = [0, 0, 0, 40, 40, 40, 1, 1, 1]
    srcs = (C.POINTER(GData) * 3)(C.pointer(c), C.pointer(c), C.pointer(c))
    reg = cg.Box((8, 8, 8), (12, 12, 12), (1, 1, 1)).as_c()
    tt = (C.c_double * 3)(0.0, -1.0, -2.0)
    h = C.c_void_p()
    assert lib.cg_batch_create(C.byref(h)) == 0
    try:
        assert lib.cg_batch_add(h, 1, C.byref(f), srcs, tt, 3, reg, reg, None, 0.0, cg.OP_LAGRANGE,
                                cg.VERTEX_CENTERED, 5, 2, 0) == 0
        assert lib.cg_batch_add(h, 1, C.byref(f), srcs, tt, 3, reg, reg, None, -0.5, cg.OP_LAGRANGE,
                                cg.VERTEX_CENTERED, 5, 2, 0) == 1
        nt = (C.c_double * 3)(1.0, 0.0, -1.0)
        assert lib.cg_batch_set_time(h, 1, nt, 3, 0.5) == -14  # transfer 0 was built on one level
        assert "time levels" in lib.cg_last_error_message().decode()
        assert lib.cg_batch_set_time(h, 1, nt, 3, 3.0) == -9  # extrapolation
        assert lib.cg_batch_set_time(h, 1, nt, 3, 1.0) == 2  # exact matches: both stay in class
        assert lib.cg_batch_set_time(h, 2, nt, 3, 0.5) == 0  # no transfer in phase 2
        assert lib.cg_batch_set_time(h, 1, nt, 2, 0.5) == 0  # none added with 2 levels
    finally:
        lib.cg_batch_destroy(h)


def test_batch_phase_mode_arguments():
    """cg_batch_set_phase_mode takes 1 (side by side), 0 (one after the
    other) or -1 (automatic) for any phase, before or after finalize; other
    modes are CG_ERR_ARG.  Host logic only."""
    h = C.c_void_p()
    assert lib.cg_batch_create(C.byref(h)) == 0
    try:
        for mode in (1, 0, -1):
            assert lib.cg_batch_set_phase_mode(h, 3, mode) == 0
        assert lib.cg_batch_set_phase_mode(h, 3, 2) == -1
        assert "phase mode" in lib.cg_last_error_message().decode()
        assert lib.cg_batch_set_phase_mode(None, 3, 1) == -1
    finally:
        lib.cg_batch_destroy(h)


def test_vc_restriction_and_dgfe_dispatch():
    """A cell-centred restriction whose slab is_centered is not one of the
    reference's four cases is refused (data.cc:1106-1129 asserts), and with
    use_dgfe set the DGFE transfers are refused instead of running
    Lagrange (data.cc:617-627, 1055-1062)."""
    from carpet_amd._lib import GData, Islab
    cbox = [2, 2, 2, 2 + 4 * 9, 2 + 4 * 9, 2 + 4 * 9, 4, 4, 4]
    c = GData()
    c.storage = FAKE
    c.padext[:] = [10, 10, 10]
    c.ext[:] = [10, 10, 10]
    c.bbox[:] = cbox
    f = GData()
    f.storage = FAKE
    f.padext[:] = [24, 24, 24]
    f.ext[:] = [24, 24, 24]
    f.bbox[:] = [-1, -1, -1, -1 + 2 * 23, -1 + 2 * 23, -1 + 2 * 23, 2, 2, 2]
    tt = (C.c_double * 1)(0.0)
    creg = cg.Box((6, 6, 6), (18, 18, 18), (4, 4, 4)).as_c()
    fsrc = (C.POINTER(GData) * 1)(C.pointer(f))

    def rest(cent, op=cg.OP_LAGRANGE):
        slab = Islab(I3(0, 0, 0), I3(*cent))
        return lib.cg_transfer_from(C.byref(c), fsrc, tt, 1, creg, creg, C.byref(slab), 0.0, op,
                                    cg.CELL_CENTERED, 3, 0, None)

    for bad in [(0, 0, 1), (1, 0, 0), (0, 0, 0)]:
        assert rest(bad) == -11, bad
    assert rest((2, 1, 1)) == -1
    assert lib.cg_restrict_3d_vc_rf2_f64(0, 1, 2, FAKE, I3(24, 24, 24), I3(24, 24, 24), FAKE, I3(10, 10, 10),
                                         I3(10, 10, 10), f.bbox, c.bbox, creg, creg, None) == -1
    # refinement factor and footprint checks of the per-call entry
    assert lib.cg_restrict_3d_vc_rf2_f64(0, 1, 1, FAKE, I3(10, 10, 10), I3(10, 10, 10), FAKE, I3(10, 10, 10),
                                         I3(10, 10, 10), c.bbox, c.bbox, creg, creg, None) == -2
    far = cg.Box((6, 6, 6), (46, 18, 18), (4, 4, 4)).as_c()
    assert lib.cg_restrict_3d_vc_rf2_f64(0, 1, 1, FAKE, I3(24, 24, 24), I3(24, 24, 24), FAKE, I3(10, 10, 10),
                                         I3(10, 10, 10), f.bbox, c.bbox, far, far, None) == -6
    cg.set_dgfe(True)
    try:
        # full-cell restriction and cell-centred Lagrange prolongation
        assert lib.cg_transfer_from(C.byref(c), fsrc, tt, 1, creg, creg, None, 0.0, cg.OP_LAGRANGE,
                                    cg.CELL_CENTERED, 3, 0, None) == -11
        assert "dgfe" in lib.cg_last_error_message().decode().lower()
        csrc = (C.POINTER(GData) * 1)(C.pointer(c))
        freg = cg.Box((11, 11, 11), (21, 21, 21), (2, 2, 2)).as_c()
        assert lib.cg_transfer_from(C.byref(f), csrc, tt, 1, freg, freg, None, 0.0, cg.OP_LAGRANGE,
                                    cg.CELL_CENTERED, 3, 0, None) == -11
        # the vc restriction does not go through DGFE (data.cc:1106-1127)
        assert rest((0, 1, 1)) != -11
    finally:
        cg.set_dgfe(False)
