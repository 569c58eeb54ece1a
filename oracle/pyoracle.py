"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of the CPU oracle (liboracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the checker or the timed CPU baseline.  Host arrays are
padded numpy arrays of shape (padext_z, padext_y, padext_x), x fastest.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


if not LIB.exists():
    build()
_lib = C.CDLL(str(LIB))

I3 = C.c_int * 3
I9 = C.c_int * 9
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_OP = (_dp, _ip, _ip, _dp, _ip, _ip, _ip, _ip, _ip, _ip)


def _decl(name, restype, *args):
    f = getattr(_lib, name)
    f.restype = restype
    f.argtypes = list(args)
    return f


_decl("orc_set_openmp", None, C.c_int)
_decl("orc_num_threads", C.c_int)
_decl("orc_vc_coeff_selftest", C.c_double, C.c_int)
_decl("orc_cc_coeff_selftest", C.c_double, C.c_int, C.c_int)
_decl("orc_eno_coeff_selftest", C.c_double, C.c_int, C.c_int, C.c_int)
_decl("orc_eno1", C.c_double, C.c_int, C.c_int, _dp)
_decl("orc_eno_window", C.c_int, C.c_int, C.c_int, _ip, _ip)
_decl("orc_copy_3d", C.c_int, *_OP, _ip)
_decl("orc_prolongate_3d_rf2", C.c_int, C.c_int, *_OP)
_decl("orc_prolongate_3d_cc_rf2", C.c_int, C.c_int, *_OP)
_decl("orc_prolongate_3d_cc_eno_rf2", C.c_int, C.c_int, *_OP)
_decl("orc_prolongate_3d_cc_tvd", C.c_int, *_OP)
_decl("orc_prolongate_3d_eno", C.c_int, *_OP)
_decl("orc_eno1d", C.c_double, _dp)
_decl("orc_prolongate_3d_o5_monotone_rf2", C.c_int, *_OP)
_decl("orc_prolongate_3d_stagger", C.c_int, C.c_int, C.c_int, *_OP)
_decl("orc_restrict_3d_stagger", C.c_int, C.c_int, *_OP)
_decl("orc_stagger_coeff", C.c_double, C.c_int, C.c_int, C.c_int)
_decl("orc_set_stagger", None, C.c_int)
_decl("orc_restrict_3d_rf2", C.c_int, *_OP)
_decl("orc_restrict_3d_cc_rf2", C.c_int, *_OP)
_decl("orc_restrict_3d_cc_ho_rf2", C.c_int, C.c_int, *_OP)
_decl("orc_restrict_3d_vc_rf2", C.c_int, C.c_int, *_OP)
_decl("orc_time_factors", C.c_int, C.c_int, _dp, C.c_double, _dp)
_decl("orc_interpolate_3d", C.c_int, C.c_int, C.POINTER(_dp), _dp, _ip, _ip, _dp, C.c_double,
      _ip, _ip, _ip, _ip, _ip, _ip)
_decl("orc_eno_time_factors", C.c_int, _dp, C.c_double, _dp, _dp, _ip)
_decl("orc_interpolate_eno_3d_3tl", C.c_int, C.POINTER(_dp), _dp, _ip, _ip, _dp, C.c_double,
      _ip, _ip, _ip, _ip, _ip, _ip)
_decl("orc_find_source_timelevel", C.c_int, _dp, C.c_int, C.c_double, C.c_int, C.c_int, _ip, _ip)
_decl("orc_prolongate_time", C.c_int, C.c_int, C.c_int, C.POINTER(_dp), _dp, C.c_int, _ip, _ip, _dp,
      C.c_double, _ip, _ip, _ip, _ip, _ip, C.c_int)

KIND_VLAG, KIND_CLAG, KIND_ENO, KIND_TVD, KIND_VENO, KIND_MONO5, KIND_STAGGER = 0, 1, 2, 3, 4, 5, 6
# staggered directions, bit d = direction d (CarpetLib's op_STAGGERxyz)
STAGGER011, STAGGER101, STAGGER110, STAGGER111 = 6, 5, 3, 7


@dataclass
class HostArr:
    """A padded host array with its bbox (lower, upper, stride)."""

    box: tuple  # (lower, upper, stride) each a 3-tuple
    data: np.ndarray  # (padz, pady, padx) float64

    @property
    def padext(self):
        return (self.data.shape[2], self.data.shape[1], self.data.shape[0])

    @property
    def ext(self):
        lo, up, st = self.box
        return tuple((u - l + s) // s if u >= l else 0 for l, u, s in zip(lo, up, st))

    def c_box(self):
        lo, up, st = self.box
        return I9(*lo, *up, *st)

    def view(self):
        e = self.ext
        return self.data[: e[2], : e[1], : e[0]]


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _b(box):
    lo, up, st = box
    return I9(*lo, *up, *st)


def _args(src: HostArr, dst: HostArr, srcreg, dstreg):
    return (_p(src.data), I3(*src.padext), I3(*src.ext), _p(dst.data), I3(*dst.padext), I3(*dst.ext),
            src.c_box(), dst.c_box(), _b(srcreg), _b(dstreg))


def set_openmp(on: bool):
    _lib.orc_set_openmp(int(on))


def num_threads() -> int:
    return _lib.orc_num_threads()


def copy_3d(src, dst, srcreg, dstreg, slab_offset=None) -> int:
    so = I3(*slab_offset) if slab_offset is not None else None
    return _lib.orc_copy_3d(*_args(src, dst, srcreg, dstreg), so)


def prolongate_3d_rf2(order, src, dst, dstreg) -> int:
    return _lib.orc_prolongate_3d_rf2(order, *_args(src, dst, dstreg, dstreg))


def prolongate_3d_cc_rf2(order, src, dst, dstreg) -> int:
    return _lib.orc_prolongate_3d_cc_rf2(order, *_args(src, dst, dstreg, dstreg))


def prolongate_3d_cc_eno_rf2(order, src, dst, dstreg) -> int:
    return _lib.orc_prolongate_3d_cc_eno_rf2(order, *_args(src, dst, dstreg, dstreg))


def prolongate_3d_cc_tvd(src, dst, dstreg) -> int:
    return _lib.orc_prolongate_3d_cc_tvd(*_args(src, dst, dstreg, dstreg))


def prolongate_3d_eno(src, dst, dstreg) -> int:
    """prolongate_3d_real8_eno (vertex-centred ENO)"""
    return _lib.orc_prolongate_3d_eno(*_args(src, dst, dstreg, dstreg))


def prolongate_3d_o5_monotone_rf2(src, dst, dstreg) -> int:
    """prolongate_3d_o5_monotone_rf2 (vertex-centred, order 5, limited)"""
    return _lib.orc_prolongate_3d_o5_monotone_rf2(*_args(src, dst, dstreg, dstreg))


def prolongate_3d_stagger(stag, order, src, dst, dstreg) -> int:
    """prolongate_3d_stagger{011,101,110,111}<ORDER> (stag: bit d = staggered in d)"""
    return _lib.orc_prolongate_3d_stagger(stag, order, *_args(src, dst, dstreg, dstreg))


def restrict_3d_stagger(stag, src, dst, dstreg) -> int:
    """restrict_3d_stagger{011,101,110,111}"""
    return _lib.orc_restrict_3d_stagger(stag, *_args(src, dst, dstreg, dstreg))


def stagger_coeff(order, which, tap) -> float:
    """coeff[which][tap + 1] of coeffs_prolongate_3d_stagger.hh"""
    return _lib.orc_stagger_coeff(order, which, tap)


def set_stagger(stag):
    """directions used by prolongate_time's KIND_STAGGER"""
    _lib.orc_set_stagger(stag)


def eno1d(q) -> float:
    """eno1d of prolongate_3d_real8_eno.F90 on 4 values"""
    q = np.ascontiguousarray(q, dtype=np.float64)
    assert q.size == 4
    return _lib.orc_eno1d(_p(q))


def restrict_3d_rf2(src, dst, dstreg) -> int:
    return _lib.orc_restrict_3d_rf2(*_args(src, dst, dstreg, dstreg))


def restrict_3d_cc_rf2(src, dst, dstreg) -> int:
    return _lib.orc_restrict_3d_cc_rf2(*_args(src, dst, dstreg, dstreg))


def restrict_3d_vc_rf2(cent, src, dst, dstreg) -> int:
    """restrict_3d_vc_rf2<T, centi, centj, centk>; cent = (centi, centj, centk)"""
    cmask = int(cent[0]) | (int(cent[1]) << 1) | (int(cent[2]) << 2)
    return _lib.orc_restrict_3d_vc_rf2(cmask, *_args(src, dst, dstreg, dstreg))


def restrict_3d_cc_ho_rf2(order, src, dst, dstreg) -> int:
    """restrict_3d_cc_o3_rf2 (order 3) / restrict_3d_cc_o5_rf2 (order 5)"""
    return _lib.orc_restrict_3d_cc_ho_rf2(order, *_args(src, dst, dstreg, dstreg))


def time_factors(times, t):
    n = len(times)
    s = (C.c_double * 5)()
    tt = (C.c_double * n)(*times)
    rc = _lib.orc_time_factors(n, tt, t, s)
    return rc, list(s)[:n]


def interpolate_3d(srcs, times, dst, t, dstreg) -> int:
    n = len(srcs)
    ptrs = (_dp * n)(*[_p(s.data) for s in srcs])
    tt = (C.c_double * n)(*times)
    s0 = srcs[0]
    return _lib.orc_interpolate_3d(n, ptrs, tt, I3(*s0.padext), I3(*s0.ext), _p(dst.data), t,
                                   I3(*dst.padext), I3(*dst.ext), s0.c_box(), dst.c_box(),
                                   _b(dstreg), _b(dstreg))


def eno_time_factors(times, t):
    """(rc, 3-point weights, 2-point weights, first level of the 2-point pair)"""
    tt = (C.c_double * 3)(*times)
    s3, s2, l0 = (C.c_double * 3)(), (C.c_double * 2)(), C.c_int()
    rc = _lib.orc_eno_time_factors(tt, t, s3, s2, C.byref(l0))
    return rc, list(s3), list(s2), l0.value


def interpolate_eno_3d_3tl(srcs, times, dst, t, dstreg) -> int:
    assert len(srcs) == 3
    ptrs = (_dp * 3)(*[_p(s.data) for s in srcs])
    tt = (C.c_double * 3)(*times)
    s0 = srcs[0]
    return _lib.orc_interpolate_eno_3d_3tl(ptrs, tt, I3(*s0.padext), I3(*s0.ext), _p(dst.data), t,
                                           I3(*dst.padext), I3(*dst.ext), s0.c_box(), dst.c_box(),
                                           _b(dstreg), _b(dstreg))


def find_source_timelevel(times, time, order_time, op):
    tl0, ntl = C.c_int(), C.c_int()
    tt = (C.c_double * len(times))(*times)
    rc = _lib.orc_find_source_timelevel(tt, len(times), time, order_time, op, C.byref(tl0), C.byref(ntl))
    return rc, tl0.value, ntl.value


def prolongate_time(kind, order, srcs, times, dst, time, dstreg, order_time) -> int:
    """data<T>::transfer_time for a prolongation (temporaries + time interpolation)."""
    n = len(srcs)
    ptrs = (_dp * n)(*[_p(s.data) for s in srcs])
    tt = (C.c_double * n)(*times)
    s0 = srcs[0]
    return _lib.orc_prolongate_time(kind, order, ptrs, tt, n, I3(*s0.padext), I3(*s0.ext),
                                    _p(dst.data), time, I3(*dst.padext), I3(*dst.ext), s0.c_box(),
                                    dst.c_box(), _b(dstreg), order_time)


def eno1(order, di, f) -> float:
    f = np.ascontiguousarray(f, dtype=np.float64)
    return _lib.orc_eno1(order, di, _p(f))


def eno_window(order, di):
    a, b = C.c_int(), C.c_int()
    rc = _lib.orc_eno_window(order, di, C.byref(a), C.byref(b))
    return rc, a.value, b.value


def vc_coeff_selftest(order) -> float:
    return _lib.orc_vc_coeff_selftest(order)


def cc_coeff_selftest(order, di) -> float:
    return _lib.orc_cc_coeff_selftest(order, di)


def eno_coeff_selftest(order, di, offset) -> float:
    return _lib.orc_eno_coeff_selftest(order, di, offset)


class Task(C.Structure):
    """orc_task: one transfer of a vertex-centred level step (op 0 copy_3d,
    1 restrict_3d_rf2, 2 prolongate_3d_rf2 + time interpolation)."""
    _fields_ = [("op", C.c_int), ("order", C.c_int), ("ntl", C.c_int), ("src", C.c_void_p * 3),
                ("spad", C.c_int * 3), ("sext", C.c_int * 3), ("sb", C.c_int * 9), ("dst", C.c_void_p),
                ("dpad", C.c_int * 3), ("dext", C.c_int * 3), ("db", C.c_int * 9), ("rb", C.c_int * 9),
                ("times", C.c_double * 3), ("time", C.c_double)]


_decl("orc_replay", C.c_int, C.c_int, C.c_int, C.POINTER(Task))


def level_step_tasks(steps, host):
    """orc_task list for workload transfers (copies, rf2 restrictions and
    vertex Lagrange prolongations with 1-3 time levels) over HostArr arrays;
    keeps the arrays alive through the returned object."""
    arr = (Task * len(steps))()
    for k, t in enumerate(steps):
        d = host[t.dst]
        ss = [host[x] for x in t.srcs]
        sst, dst_ = ss[0].box[2], d.box[2]
        x = arr[k]
        if tuple(sst) == tuple(dst_):
            x.op = 0
        elif all(a > b for a, b in zip(sst, dst_)):
            x.op = 2
        else:
            x.op = 1
        x.order = t.order_space
        x.ntl = len(ss) if x.op == 2 else 1
        for l, h in enumerate(ss[:3]):
            x.src[l] = h.data.ctypes.data
        s0 = ss[0]
        x.spad[:] = list(s0.padext)
        x.sext[:] = list(s0.ext)
        x.sb[:] = list(s0.box[0]) + list(s0.box[1]) + list(s0.box[2])
        x.dst = d.data.ctypes.data
        x.dpad[:] = list(d.padext)
        x.dext[:] = list(d.ext)
        x.db[:] = list(d.box[0]) + list(d.box[1]) + list(d.box[2])
        r = t.dstreg
        x.rb[:] = list(r.lower) + list(r.upper) + list(r.stride)
        for l in range(min(3, len(t.times))):
            x.times[l] = t.times[l]
        x.time = t.time
    return arr


def replay(mode, tasks) -> int:
    """orc_replay: mode 0 one operator call per transfer (the reference's
    shape), mode 1 all transfers in one parallel loop, time interpolation
    fused."""
    return _lib.orc_replay(mode, len(tasks), tasks)
