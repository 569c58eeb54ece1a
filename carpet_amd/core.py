"""Host-side mirror of CarpetLib's transfer interface over libcarpet_amd.so.

Names and argument meaning follow the reference operators
(CarpetLib/src/operator_prototypes_3d.hh:36-302) and the data<T> transfer
dispatch (CarpetLib/src/data.cc:464-1314); errors raise CarpetError with the
library's CG_ERR_* code where the reference would abort.

Host arrays use numpy's C order with shape (z, y, x): x is fastest, exactly
like the device layout (i + padext_x*(j + padext_y*k)).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from ._lib import GData, I3, I9, Islab, check, lib

# CarpetLib enums (operators.hh:8-25, defs.hh:134)
OP_ERROR, OP_NONE, OP_SYNC, OP_RESTRICT, OP_COPY, OP_LAGRANGE, OP_ENO, OP_WENO, OP_TVD, OP_LAGRANGE_MONOTONE = range(10)
OP_STAGGER011, OP_STAGGER101, OP_STAGGER110, OP_STAGGER111 = range(10, 14)
# staggered directions of the staggered operators: bit d set = staggered in
# direction d (CarpetLib's op_STAGGERxyz)
STAGGER011, STAGGER101, STAGGER110, STAGGER111 = 6, 5, 3, 7
VERTEX_CENTERED, CELL_CENTERED = 1, 2
PREC_F64, PREC_MIXED, PREC_MIXED_DOT2 = 0, 1, 2


@dataclass(frozen=True)
class Box:
    """bbox<int,3>: inclusive lower/upper bounds and stride (bbox.hh:34-70)."""

    lower: tuple
    upper: tuple
    stride: tuple

    def __post_init__(self):
        object.__setattr__(self, "lower", tuple(int(v) for v in self.lower))
        object.__setattr__(self, "upper", tuple(int(v) for v in self.upper))
        object.__setattr__(self, "stride", tuple(int(v) for v in self.stride))

    def empty(self) -> bool:
        return any(l > u for l, u in zip(self.lower, self.upper))

    def shape(self):
        if self.empty():
            return (0, 0, 0)
        return tuple(u - l + s for l, u, s in zip(self.lower, self.upper, self.stride))

    def sizes(self):
        return tuple(sh // s for sh, s in zip(self.shape(), self.stride))

    def size(self) -> int:
        a, b, c = self.sizes()
        return a * b * c

    def expand(self, lo, hi) -> "Box":
        return Box(tuple(l - a * s for l, a, s in zip(self.lower, lo, self.stride)),
                   tuple(u + b * s for u, b, s in zip(self.upper, hi, self.stride)), self.stride)

    def shift(self, v) -> "Box":
        return self.expand(tuple(-x for x in v), v)

    def contained_in(self, b: "Box") -> bool:
        if self.empty():
            return True
        return all(l >= bl and u <= bu for l, u, bl, bu in zip(self.lower, self.upper, b.lower, b.upper))

    def intersect(self, b: "Box") -> "Box":
        return Box(tuple(max(x, y) for x, y in zip(self.lower, b.lower)),
                   tuple(min(x, y) for x, y in zip(self.upper, b.upper)), self.stride)

    def as_c(self):
        # built once per box (the batch of a large hierarchy passes the same
        # region boxes for every grid function)
        c = self.__dict__.get("_c")
        if c is None:
            c = I9(*self.lower, *self.upper, *self.stride)
            object.__setattr__(self, "_c", c)
        return c


def pad_shape(shape, offset, vec_size: int, linesize: int, align_interior: bool = True):
    """pad_shape (CarpetLib/src/cacheinfo.cc:40-209) through the library."""
    ps, po = I3(), I3()
    check(lib.cg_pad_shape(I3(*shape), I3(*offset), vec_size, linesize, int(align_interior), ps, po),
          "cg_pad_shape")
    return tuple(ps), tuple(po)


class Stream:
    """A HIP stream owned by the library."""

    def __init__(self):
        h = C.c_void_p()
        check(lib.cg_stream_create(C.byref(h)), "cg_stream_create")
        self.handle = h.value

    def synchronize(self):
        check(lib.cg_stream_synchronize(self.handle), "cg_stream_synchronize")

    def close(self):
        if self.handle:
            lib.cg_stream_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _sh(stream) -> Optional[int]:
    return None if stream is None else (stream.handle if isinstance(stream, Stream) else stream)


class Event:
    def __init__(self):
        h = C.c_void_p()
        check(lib.cg_event_create(C.byref(h)), "cg_event_create")
        self.handle = h.value

    def record(self, stream=None):
        check(lib.cg_event_record(self.handle, _sh(stream)), "cg_event_record")

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        check(lib.cg_event_elapsed_ms(C.byref(ms), self.handle, end.handle), "cg_event_elapsed_ms")
        return ms.value

    def __del__(self):
        try:
            lib.cg_event_destroy(self.handle)
        except Exception:
            pass


def synchronize():
    check(lib.cg_device_synchronize(), "cg_device_synchronize")


class DeviceArray:
    """A padded device array for one data<T> block (data.cc:355-394).

    box is the array's bbox; owned (optional) the owned interior, which the
    padding aligns to a 128-byte row.  dtype float64 (drop-in) or float32
    (mixed-precision prolongation).
    """

    def __init__(self, box: Box, owned: Optional[Box] = None, dtype=np.float64,
                 padext=None):
        self.box = box
        self.dtype = np.dtype(dtype)
        self.g = GData()
        if padext is None:
            check(lib.cg_gdata_alloc(C.byref(self.g), box.as_c(), owned.as_c() if owned else None,
                                     self.dtype.itemsize), "cg_gdata_alloc")
            self._gdata_alloc = True
        else:
            # explicit padding (tests of arbitrary padded extents)
            n = int(np.prod(padext)) * self.dtype.itemsize
            p = C.c_void_p()
            check(lib.cg_dev_malloc(C.byref(p), n), "cg_dev_malloc")
            self.g.storage = p.value
            self.g.padext[:] = list(padext)
            self.g.ext[:] = list(box.sizes())
            self.g.bbox[:] = list(box.lower + box.upper + box.stride)
            self._gdata_alloc = False

    @classmethod
    def view(cls, box: Box, ptr: int, padext, dtype=np.float64) -> "DeviceArray":
        """An array over memory the caller owns (e.g. a pool)."""
        a = cls.__new__(cls)
        a.box = box
        a.dtype = np.dtype(dtype)
        a.g = GData()
        a.g.storage = ptr
        a.g.padext[:] = list(padext)
        a.g.ext[:] = list(box.sizes())
        a.g.bbox[:] = list(box.lower + box.upper + box.stride)
        a._gdata_alloc = None
        return a

    @property
    def ptr(self) -> int:
        return self.g.storage

    @property
    def padext(self):
        return tuple(self.g.padext)

    @property
    def ext(self):
        return tuple(self.g.ext)

    @property
    def nbytes_padded(self) -> int:
        return int(np.prod(self.padext)) * self.dtype.itemsize

    def padded_shape_zyx(self):
        p = self.padext
        return (p[2], p[1], p[0])

    def upload(self, a: np.ndarray, fill=0.0):
        """a: (ext_z, ext_y, ext_x) array (or the padded shape)."""
        e = self.ext
        a = np.asarray(a, dtype=self.dtype)
        if a.shape == self.padded_shape_zyx():
            host = np.ascontiguousarray(a)
        else:
            assert a.shape == (e[2], e[1], e[0]), (a.shape, e)
            host = np.full(self.padded_shape_zyx(), fill, dtype=self.dtype)
            host[: e[2], : e[1], : e[0]] = a
        check(lib.cg_memcpy_htod(self.ptr, host.ctypes.data, host.nbytes), "cg_memcpy_htod")

    def download_padded(self) -> np.ndarray:
        host = np.empty(self.padded_shape_zyx(), dtype=self.dtype)
        check(lib.cg_memcpy_dtoh(host.ctypes.data, self.ptr, host.nbytes), "cg_memcpy_dtoh")
        return host

    def download(self) -> np.ndarray:
        e = self.ext
        return self.download_padded()[: e[2], : e[1], : e[0]].copy()

    def fill(self, v: float, stream=None):
        if self.dtype != np.float64:
            raise TypeError("fill is for float64 arrays")
        check(lib.cg_fill_f64(self.ptr, float(v), int(np.prod(self.padext)), _sh(stream)), "cg_fill_f64")

    def fill_hash(self, seed: int, lo: float = -1.0, hi: float = 1.0, stream=None):
        """Seeded uniform data in the extent (cg_fill_hash_f64); the host
        equivalent is carpet_amd.workloads.hash_field."""
        check(lib.cg_fill_hash_f64(self.ptr, I3(*self.padext), I3(*self.ext), seed & (2**64 - 1),
                                   lo, hi, _sh(stream)), "cg_fill_hash_f64")

    def fill_fourier(self, scale: float, amp: float, kvec, phase, stream=None):
        """amp * sum_m sin(2 pi k_m . (X*scale) + phase_m) at global index X."""
        n = len(phase)
        kv = (C.c_int * (3 * n))(*[int(v) for k in kvec for v in k])
        ph = (C.c_double * n)(*phase)
        check(lib.cg_fill_fourier_f64(C.byref(self.g), scale, amp, n, kv, ph, _sh(stream)),
              "cg_fill_fourier_f64")

    def free(self):
        if self.g.storage:
            if self._gdata_alloc:
                lib.cg_gdata_free(C.byref(self.g))
            elif self._gdata_alloc is False:
                lib.cg_dev_free(self.g.storage)
            self.g.storage = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def fill_hash_many(fills, lo: float = -1.0, hi: float = 1.0, stream=None) -> int:
    """Seeded uniform data in many arrays at once: fills = [(DeviceArray,
    seed), ...] (cg_fill_hash_many_f64, one launch); returns once done."""
    from ._lib import HashFill
    n = len(fills)
    arr = (HashFill * max(1, n))()
    for k, (a, seed) in enumerate(fills):
        arr[k].dst = a.ptr
        arr[k].padext[:] = list(a.padext)
        arr[k].ext[:] = list(a.ext)
        arr[k].seed = seed & (2**64 - 1)
        arr[k].lo = lo
        arr[k].hi = hi
    check(lib.cg_fill_hash_many_f64(n, arr, _sh(stream)), "cg_fill_hash_many_f64")
    return 1


def _arr(a: DeviceArray):
    return (a.ptr, I3(*a.padext), I3(*a.ext))


# ---- per-call operators (operator_prototypes_3d.hh) ----------------------
def copy_3d(src: DeviceArray, dst: DeviceArray, srcregbbox: Box, dstregbbox: Box,
            slab_offset=None, stream=None):
    slab = None
    if slab_offset is not None:
        slab = Islab(I3(*slab_offset), I3(1, 1, 1))
    check(lib.cg_copy_3d_f64(*_arr(src), *_arr(dst), src.box.as_c(), dst.box.as_c(),
                             srcregbbox.as_c(), dstregbbox.as_c(),
                             C.byref(slab) if slab else None, _sh(stream)), "copy_3d")


def prolongate_3d_rf2(order: int, src: DeviceArray, dst: DeviceArray, dstregbbox: Box,
                      srcregbbox: Optional[Box] = None, stream=None):
    check(lib.cg_prolongate_3d_rf2_f64(order, *_arr(src), *_arr(dst), src.box.as_c(), dst.box.as_c(),
                                       (srcregbbox or dstregbbox).as_c(), dstregbbox.as_c(), _sh(stream)),
          "prolongate_3d_rf2")


def prolongate_3d_cc_rf2(order: int, src: DeviceArray, dst: DeviceArray, dstregbbox: Box,
                         srcregbbox: Optional[Box] = None, stream=None):
    check(lib.cg_prolongate_3d_cc_rf2_f64(order, *_arr(src), *_arr(dst), src.box.as_c(), dst.box.as_c(),
                                          (srcregbbox or dstregbbox).as_c(), dstregbbox.as_c(), _sh(stream)),
          "prolongate_3d_cc_rf2")


def prolongate_3d_cc_eno_rf2(order: int, src: DeviceArray, dst: DeviceArray, dstregbbox: Box,
                             srcregbbox: Optional[Box] = None, stream=None):
    check(lib.cg_prolongate_3d_cc_eno_rf2_f64(order, *_arr(src), *_arr(dst), src.box.as_c(),
                                              dst.box.as_c(), (srcregbbox or dstregbbox).as_c(),
                                              dstregbbox.as_c(), _sh(stream)),
          "prolongate_3d_cc_eno_rf2")


def prolongate_3d_eno(src: DeviceArray, dst: DeviceArray, dstregbbox: Box,
                      srcregbbox: Optional[Box] = None, stream=None):
    """Vertex-centred ENO (prolongate_3d_real8_eno.F90)."""
    check(lib.cg_prolongate_3d_eno_f64(*_arr(src), *_arr(dst), src.box.as_c(), dst.box.as_c(),
                                       (srcregbbox or dstregbbox).as_c(), dstregbbox.as_c(), _sh(stream)),
          "prolongate_3d_eno")


def prolongate_3d_o5_monotone_rf2(src: DeviceArray, dst: DeviceArray, dstregbbox: Box,
                                  srcregbbox: Optional[Box] = None, stream=None):
    """Limited order-5 vertex Lagrange (prolongate_3d_o5_monotone_rf2.cc)."""
    check(lib.cg_prolongate_3d_o5_monotone_rf2_f64(*_arr(src), *_arr(dst), src.box.as_c(), dst.box.as_c(),
                                                   (srcregbbox or dstregbbox).as_c(), dstregbbox.as_c(),
                                                   _sh(stream)),
          "prolongate_3d_o5_monotone_rf2")


def prolongate_3d_cc_tvd(src: DeviceArray, dst: DeviceArray, dstregbbox: Box,
                         srcregbbox: Optional[Box] = None, stream=None):
    check(lib.cg_prolongate_3d_cc_tvd_f64(*_arr(src), *_arr(dst), src.box.as_c(), dst.box.as_c(),
                                          (srcregbbox or dstregbbox).as_c(), dstregbbox.as_c(), _sh(stream)),
          "prolongate_3d_cc_tvd")


def restrict_3d_rf2(src: DeviceArray, dst: DeviceArray, dstregbbox: Box,
                    srcregbbox: Optional[Box] = None, stream=None):
    check(lib.cg_restrict_3d_rf2_f64(*_arr(src), *_arr(dst), src.box.as_c(), dst.box.as_c(),
                                     (srcregbbox or dstregbbox).as_c(), dstregbbox.as_c(), _sh(stream)),
          "restrict_3d_rf2")


def restrict_3d_cc_rf2(src: DeviceArray, dst: DeviceArray, dstregbbox: Box,
                       srcregbbox: Optional[Box] = None, stream=None):
    check(lib.cg_restrict_3d_cc_rf2_f64(*_arr(src), *_arr(dst), src.box.as_c(), dst.box.as_c(),
                                        (srcregbbox or dstregbbox).as_c(), dstregbbox.as_c(), _sh(stream)),
          "restrict_3d_cc_rf2")


def restrict_3d_cc_o3_rf2(src: DeviceArray, dst: DeviceArray, dstregbbox: Box,
                          srcregbbox: Optional[Box] = None, stream=None):
    """Cubic cell-centred restriction (CarpetLib/src/restrict_3d_cc_o3_rf2.cc);
    the source must hold one fine cell beyond the region's footprint."""
    check(lib.cg_restrict_3d_cc_o3_rf2_f64(*_arr(src), *_arr(dst), src.box.as_c(), dst.box.as_c(),
                                           (srcregbbox or dstregbbox).as_c(), dstregbbox.as_c(),
                                           _sh(stream)),
          "restrict_3d_cc_o3_rf2")


def restrict_3d_cc_o5_rf2(src: DeviceArray, dst: DeviceArray, dstregbbox: Box,
                          srcregbbox: Optional[Box] = None, stream=None):
    """Quintic cell-centred restriction (CarpetLib/src/restrict_3d_cc_o5_rf2.cc);
    the source must hold two fine cells beyond the region's footprint."""
    check(lib.cg_restrict_3d_cc_o5_rf2_f64(*_arr(src), *_arr(dst), src.box.as_c(), dst.box.as_c(),
                                           (srcregbbox or dstregbbox).as_c(), dstregbbox.as_c(),
                                           _sh(stream)),
          "restrict_3d_cc_o5_rf2")


def prolongate_3d_stagger(stagger: int, order: int, src: DeviceArray, dst: DeviceArray, dstregbbox: Box,
                          srcregbbox: Optional[Box] = None, stream=None):
    """Staggered prolongation (CarpetLib/src/prolongate_3d_stagger{011,101,110,111}.cc),
    stagger one of STAGGER011..STAGGER111, order 2 or 3; needs
    set_staggered_operators(True)."""
    check(lib.cg_prolongate_3d_stagger_f64(int(stagger), int(order), *_arr(src), *_arr(dst), src.box.as_c(),
                                           dst.box.as_c(), (srcregbbox or dstregbbox).as_c(), dstregbbox.as_c(),
                                           _sh(stream)),
          "prolongate_3d_stagger")


def restrict_3d_stagger(stagger: int, src: DeviceArray, dst: DeviceArray, dstregbbox: Box,
                        srcregbbox: Optional[Box] = None, stream=None):
    """Staggered restriction (CarpetLib/src/restrict_3d_stagger{011,101,110,111}.cc);
    needs set_staggered_operators(True)."""
    check(lib.cg_restrict_3d_stagger_f64(int(stagger), *_arr(src), *_arr(dst), src.box.as_c(), dst.box.as_c(),
                                         (srcregbbox or dstregbbox).as_c(), dstregbbox.as_c(), _sh(stream)),
          "restrict_3d_stagger")


def set_staggered_operators(support_staggered_operators: bool):
    """CarpetLib's support_staggered_operators parameter (param.ccl:65): the
    staggered operators refuse to run unless it is set."""
    check(lib.cg_set_staggered_operators(int(bool(support_staggered_operators))), "set_staggered_operators")


def restrict_3d_vc_rf2(cent, src: DeviceArray, dst: DeviceArray, dstregbbox: Box, stream=None):
    """restrict_3d_vc_rf2<T, centi, centj, centk>; cent = (centi, centj, centk)."""
    check(lib.cg_restrict_3d_vc_rf2_f64(int(cent[0]), int(cent[1]), int(cent[2]), *_arr(src), *_arr(dst),
                                        src.box.as_c(), dst.box.as_c(), dstregbbox.as_c(), dstregbbox.as_c(),
                                        _sh(stream)),
          "restrict_3d_vc_rf2")


def set_dgfe(use_dgfe: bool):
    """CarpetLib::use_dgfe: the DGFE operators are refused while it is set."""
    check(lib.cg_set_dgfe(1 if use_dgfe else 0), "cg_set_dgfe")


def set_restriction(use_higher_order_restriction: bool, restriction_order_space: int = 3):
    """CarpetLib's use_higher_order_restriction / restriction_order_space
    parameters (CarpetLib/param.ccl) as seen by transfer_from: cell-centred
    restrictions of non-ENO/WENO groups then use the o3 or o5 stencil."""
    check(lib.cg_set_restriction(int(bool(use_higher_order_restriction)), int(restriction_order_space)),
          "set_restriction")


def interpolate_3d_2tl(src1: DeviceArray, t1: float, src2: DeviceArray, t2: float,
                       dst: DeviceArray, t: float, dstregbbox: Box, stream=None):
    check(lib.cg_interpolate_3d_2tl_f64(src1.ptr, t1, src2.ptr, t2, I3(*src1.padext), I3(*src1.ext),
                                        dst.ptr, t, I3(*dst.padext), I3(*dst.ext), src1.box.as_c(),
                                        dst.box.as_c(), dstregbbox.as_c(), dstregbbox.as_c(), _sh(stream)),
          "interpolate_3d_2tl")


def interpolate_3d_3tl(src1: DeviceArray, t1: float, src2: DeviceArray, t2: float,
                       src3: DeviceArray, t3: float, dst: DeviceArray, t: float,
                       dstregbbox: Box, stream=None):
    check(lib.cg_interpolate_3d_3tl_f64(src1.ptr, t1, src2.ptr, t2, src3.ptr, t3, I3(*src1.padext),
                                        I3(*src1.ext), dst.ptr, t, I3(*dst.padext), I3(*dst.ext),
                                        src1.box.as_c(), dst.box.as_c(), dstregbbox.as_c(),
                                        dstregbbox.as_c(), _sh(stream)),
          "interpolate_3d_3tl")


def interpolate_3d_ntl(srcs: Sequence[DeviceArray], times: Sequence[float], dst: DeviceArray, t: float,
                       dstregbbox: Box, stream=None):
    """interpolate_3d_4tl / interpolate_3d_5tl (4 or 5 levels)."""
    n = len(srcs)
    if n not in (4, 5) or len(times) != n:
        raise ValueError("interpolate_3d_ntl takes 4 or 5 levels and their times")
    args = []
    for a, tt in zip(srcs, times):
        args += [a.ptr, tt]
    s0 = srcs[0]
    f = lib.cg_interpolate_3d_4tl_f64 if n == 4 else lib.cg_interpolate_3d_5tl_f64
    check(f(*args, I3(*s0.padext), I3(*s0.ext), dst.ptr, t, I3(*dst.padext), I3(*dst.ext), s0.box.as_c(),
            dst.box.as_c(), dstregbbox.as_c(), dstregbbox.as_c(), _sh(stream)),
          f"interpolate_3d_{n}tl")


def interpolate_eno_3d_3tl(src1: DeviceArray, t1: float, src2: DeviceArray, t2: float,
                           src3: DeviceArray, t3: float, dst: DeviceArray, t: float,
                           dstregbbox: Box, stream=None):
    check(lib.cg_interpolate_eno_3d_3tl_f64(src1.ptr, t1, src2.ptr, t2, src3.ptr, t3, I3(*src1.padext),
                                            I3(*src1.ext), dst.ptr, t, I3(*dst.padext), I3(*dst.ext),
                                            src1.box.as_c(), dst.box.as_c(), dstregbbox.as_c(),
                                            dstregbbox.as_c(), _sh(stream)),
          "interpolate_eno_3d_3tl")


def _gptrs(srcs: Sequence[DeviceArray]):
    arr = (C.POINTER(GData) * len(srcs))(*[C.pointer(s.g) for s in srcs])
    return arr


def _slab(slab_offset, is_centered):
    if slab_offset is None and is_centered is None:
        return None
    return Islab(I3(*(slab_offset or (0, 0, 0))), I3(*(is_centered or (1, 1, 1))))


def transfer_from(dst: DeviceArray, srcs: Sequence[DeviceArray], times: Sequence[float],
                  dstregbbox: Box, srcregbbox: Box, time: float, transport_operator: int,
                  centering: int, order_space: int, order_time: int, slab_offset=None,
                  stream=None, is_centered=None):
    """data<T>::transfer_from_innerloop with gdata::find_source_timelevel;
    a slab (offset, is_centered) is passed when either is given."""
    slab = _slab(slab_offset, is_centered)
    tt = (C.c_double * len(times))(*times)
    check(lib.cg_transfer_from(C.byref(dst.g), _gptrs(srcs), tt, len(times), dstregbbox.as_c(),
                               srcregbbox.as_c(), C.byref(slab) if slab else None, time,
                               transport_operator, centering, order_space, order_time, _sh(stream)),
          "transfer_from")


def transfer_from_noalloc(dst: DeviceArray, srcs: Sequence[DeviceArray], times: Sequence[float],
                          dstregbbox: Box, srcregbbox: Box, time: float, transport_operator: int,
                          centering: int, order_space: int, order_time: int, slab_offset=None,
                          stream=None):
    """transfer_from that allocates no device memory: time-level temporaries
    live in the library's per-GPU scratch array (cg_transfer_from_noalloc)."""
    slab = _slab(slab_offset, None)
    tt = (C.c_double * len(times))(*times)
    check(lib.cg_transfer_from_noalloc(C.byref(dst.g), _gptrs(srcs), tt, len(times), dstregbbox.as_c(),
                                       srcregbbox.as_c(), C.byref(slab) if slab else None, time,
                                       transport_operator, centering, order_space, order_time, _sh(stream)),
          "transfer_from_noalloc")


class Batch:
    """A batched level step (replaces ggf::transfer_from_all's region walk)."""

    def __init__(self):
        h = C.c_void_p()
        check(lib.cg_batch_create(C.byref(h)), "cg_batch_create")
        self.handle = h.value
        self._keep = []

    def add(self, phase: int, dst: DeviceArray, srcs: Sequence[DeviceArray], times: Sequence[float],
            dstregbbox: Box, srcregbbox: Box, time: float, transport_operator: int, centering: int,
            order_space: int, order_time: int, slab_offset=None, prec: int = PREC_F64,
            is_centered=None) -> int:
        slab = _slab(slab_offset, is_centered)
        tt = (C.c_double * len(times))(*times)
        return check(lib.cg_batch_add(self.handle, phase, C.byref(dst.g), _gptrs(srcs), tt, len(times),
                                      dstregbbox.as_c(), srcregbbox.as_c(),
                                      C.byref(slab) if slab else None, time, transport_operator,
                                      centering, order_space, order_time, prec), "cg_batch_add")

    def finalize(self, tile_cfg: int = -1):
        check(lib.cg_batch_finalize(self.handle, tile_cfg), "cg_batch_finalize")

    def run(self, stream=None):
        check(lib.cg_batch_run(self.handle, _sh(stream)), "cg_batch_run")

    def run_plain(self, stream=None):
        check(lib.cg_batch_run_plain(self.handle, _sh(stream)), "cg_batch_run_plain")

    def capture_graph(self, stream: Stream):
        check(lib.cg_batch_capture_graph(self.handle, _sh(stream)), "cg_batch_capture_graph")

    def launch_graph(self, stream: Stream):
        check(lib.cg_batch_launch_graph(self.handle, _sh(stream)), "cg_batch_launch_graph")

    def set_time(self, phase: int, times: Sequence[float], time: float) -> int:
        """New level times and target time for the transfers of `phase`
        added with len(times) source levels (cg_batch_set_time)."""
        tt = (C.c_double * len(times))(*times)
        return check(lib.cg_batch_set_time(self.handle, phase, tt, len(times), time), "cg_batch_set_time")

    def set_storage(self, pairs) -> int:
        """Rebinds storage: [(from DeviceArray, to DeviceArray), ...], all at
        once (cg_batch_set_storage; ggf::cycle_all's rotation)."""
        n = len(pairs)
        fr = (GData * max(1, n))(*[a.g for a, _ in pairs])
        to = (GData * max(1, n))(*[b.g for _, b in pairs])
        return check(lib.cg_batch_set_storage(self.handle, n, fr, to), "cg_batch_set_storage")

    def set_phase_mode(self, phase: int, mode: int):
        """How the launches of `phase` run: 1 side by side, 0 one after the
        other, -1 automatic (cg_batch_set_phase_mode)."""
        check(lib.cg_batch_set_phase_mode(self.handle, phase, mode), "cg_batch_set_phase_mode")

    def autotune(self, stream=None, reps: int = 3, cache_path: Optional[str] = None):
        """Searches the work-item sizes per launch class (cg_batch_autotune)."""
        check(lib.cg_batch_autotune(self.handle, _sh(stream), reps,
                                    cache_path.encode() if cache_path else None), "cg_batch_autotune")

    def describe(self) -> str:
        buf = C.create_string_buffer(1 << 16)
        check(lib.cg_batch_describe(self.handle, buf, len(buf)), "cg_batch_describe")
        return buf.value.decode()

    def launch_count(self) -> int:
        return check(lib.cg_batch_launch_count(self.handle), "cg_batch_launch_count")

    def run_launch(self, idx: int, stream=None):
        """Launch idx of the finalized batch on its own (per-kernel timing)."""
        check(lib.cg_batch_run_launch(self.handle, idx, _sh(stream)), "cg_batch_run_launch")

    def launch_stats(self, idx: int):
        """(phase, destination points, algorithmic bytes) of launch idx."""
        ph, p, b = C.c_int(), C.c_longlong(), C.c_longlong()
        check(lib.cg_batch_launch_stats(self.handle, idx, C.byref(ph), C.byref(p), C.byref(b)),
              "cg_batch_launch_stats")
        return ph.value, p.value, b.value

    def stats(self):
        n, p, b = C.c_longlong(), C.c_longlong(), C.c_longlong()
        check(lib.cg_batch_stats(self.handle, C.byref(n), C.byref(p), C.byref(b)), "cg_batch_stats")
        return {"transfers": n.value, "dst_points": p.value, "alg_bytes": b.value}

    def close(self):
        if self.handle:
            lib.cg_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- mixed precision (FP32 storage, compensated FP32 sums, FP64 time) -----
def prolongate_3d_rf2_mixed(order: int, src: DeviceArray, dst: DeviceArray, dstregbbox: Box,
                            stream=None):
    assert src.dtype == np.float32 and dst.dtype == np.float32
    check(lib.cg_prolongate_3d_rf2_mixed(order, *_arr(src), *_arr(dst), src.box.as_c(), dst.box.as_c(),
                                         dstregbbox.as_c(), dstregbbox.as_c(), _sh(stream)),
          "prolongate_3d_rf2_mixed")


def prolongate_3d_cc_rf2_mixed(order: int, src: DeviceArray, dst: DeviceArray, dstregbbox: Box,
                               stream=None):
    assert src.dtype == np.float32 and dst.dtype == np.float32
    check(lib.cg_prolongate_3d_cc_rf2_mixed(order, *_arr(src), *_arr(dst), src.box.as_c(),
                                            dst.box.as_c(), dstregbbox.as_c(), dstregbbox.as_c(),
                                            _sh(stream)),
          "prolongate_3d_cc_rf2_mixed")


def transfer_from_mixed(dst: DeviceArray, srcs: Sequence[DeviceArray], times: Sequence[float],
                        dstregbbox: Box, srcregbbox: Box, time: float, transport_operator: int,
                        centering: int, order_space: int, order_time: int, stream=None):
    tt = (C.c_double * len(times))(*times)
    check(lib.cg_transfer_from_mixed(C.byref(dst.g), _gptrs(srcs), tt, len(times), dstregbbox.as_c(),
                                     srcregbbox.as_c(), time, transport_operator, centering,
                                     order_space, order_time, _sh(stream)),
          "transfer_from_mixed")


def relative_errors(approx: np.ndarray, exact: np.ndarray):
    """Normwise max and RMS relative error of approx against exact."""
    a = np.asarray(approx, dtype=np.float64)
    e = np.asarray(exact, dtype=np.float64)
    scale = float(np.max(np.abs(e))) or 1.0
    d = a - e
    return float(np.max(np.abs(d)) / scale), float(np.sqrt(np.mean(d * d)) / (np.sqrt(np.mean(e * e)) or 1.0))
