"""ctypes binding of libcarpet_amd.so (declarations in include/carpet_amd.h).

The shared library is the product; this module only declares argument
types.  Importing fails loudly if the library has not been built: there is
no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libcarpet_amd.so"

if not LIB_PATH.exists():
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
        "or `python carpet_amd/build.py` (needs hipcc)"
    )

lib = C.CDLL(str(LIB_PATH))

I3 = C.c_int * 3
I9 = C.c_int * 9
vp = C.c_void_p
ip = C.POINTER(C.c_int)
dp = C.c_double
i = C.c_int


class Islab(C.Structure):
    _fields_ = [("offset", C.c_int * 3), ("is_centered", C.c_int * 3)]


class GData(C.Structure):
    _fields_ = [("storage", C.c_void_p), ("padext", C.c_int * 3), ("ext", C.c_int * 3), ("bbox", C.c_int * 9)]


def _decl(name, restype, *argtypes):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = list(argtypes)
    return f


_decl("cg_version", i)
_decl("cg_error_string", C.c_char_p, i)
_decl("cg_last_error_message", C.c_char_p)
_decl("cg_set_device", i, i)
_decl("cg_device_count", i, ip)
_decl("cg_pad_shape", i, ip, ip, i, i, i, ip, ip)
_decl("cg_gdata_alloc", i, C.POINTER(GData), ip, ip, C.c_size_t)
_decl("cg_gdata_free", i, C.POINTER(GData))
_decl("cg_dev_malloc", i, C.POINTER(C.c_void_p), C.c_size_t)
_decl("cg_dev_free", i, vp)
_decl("cg_memcpy_htod", i, vp, vp, C.c_size_t)
_decl("cg_memcpy_dtoh", i, vp, vp, C.c_size_t)
_decl("cg_memcpy_dtod", i, vp, vp, C.c_size_t, vp)
_decl("cg_memset_d8", i, vp, i, C.c_size_t, vp)
_decl("cg_fill_f64", i, vp, dp, C.c_size_t, vp)
_decl("cg_stream_copy_f64", i, vp, vp, C.c_size_t, vp)
_decl("cg_fill_hash_f64", i, vp, ip, ip, C.c_ulonglong, dp, dp, vp)
_decl("cg_fill_fourier_f64", i, C.POINTER(GData), dp, dp, i, ip, C.POINTER(C.c_double), vp)


class HashFill(C.Structure):
    _fields_ = [("dst", C.c_void_p), ("padext", C.c_int * 3), ("ext", C.c_int * 3), ("seed", C.c_ulonglong),
                ("lo", C.c_double), ("hi", C.c_double)]


_decl("cg_fill_hash_many_f64", i, i, C.POINTER(HashFill), vp)
_decl("cg_stream_create", i, C.POINTER(C.c_void_p))
_decl("cg_stream_destroy", i, vp)
_decl("cg_stream_synchronize", i, vp)
_decl("cg_device_synchronize", i)
_decl("cg_event_create", i, C.POINTER(C.c_void_p))
_decl("cg_event_destroy", i, vp)
_decl("cg_event_record", i, vp, vp)
_decl("cg_event_elapsed_ms", i, C.POINTER(C.c_float), vp, vp)

_OPARGS = (vp, ip, ip, vp, ip, ip, ip, ip, ip, ip)
_decl("cg_copy_3d_f64", i, *_OPARGS, C.POINTER(Islab), vp)
_decl("cg_prolongate_3d_rf2_f64", i, i, *_OPARGS, vp)
_decl("cg_prolongate_3d_cc_rf2_f64", i, i, *_OPARGS, vp)
_decl("cg_prolongate_3d_cc_eno_rf2_f64", i, i, *_OPARGS, vp)
_decl("cg_prolongate_3d_cc_tvd_f64", i, *_OPARGS, vp)
_decl("cg_restrict_3d_rf2_f64", i, *_OPARGS, vp)
_decl("cg_restrict_3d_cc_rf2_f64", i, *_OPARGS, vp)
_decl("cg_prolongate_3d_eno_f64", i, *_OPARGS, vp)
_decl("cg_prolongate_3d_o5_monotone_rf2_f64", i, *_OPARGS, vp)
_decl("cg_restrict_3d_cc_o3_rf2_f64", i, *_OPARGS, vp)
_decl("cg_restrict_3d_cc_o5_rf2_f64", i, *_OPARGS, vp)
_decl("cg_set_restriction", i, i, i)
_decl("cg_restrict_3d_vc_rf2_f64", i, i, i, i, *_OPARGS, vp)
_decl("cg_set_dgfe", i, i)
_decl("cg_prolongate_3d_stagger_f64", i, i, i, *_OPARGS, vp)
_decl("cg_restrict_3d_stagger_f64", i, i, *_OPARGS, vp)
_decl("cg_set_staggered_operators", i, i)
_decl("cg_interpolate_3d_2tl_f64", i, vp, dp, vp, dp, ip, ip, vp, dp, ip, ip, ip, ip, ip, ip, vp)
_decl("cg_interpolate_3d_3tl_f64", i, vp, dp, vp, dp, vp, dp, ip, ip, vp, dp, ip, ip, ip, ip, ip, ip, vp)
_decl("cg_interpolate_3d_4tl_f64", i, vp, dp, vp, dp, vp, dp, vp, dp, ip, ip, vp, dp, ip, ip, ip, ip, ip, ip, vp)
_decl("cg_interpolate_3d_5tl_f64", i, vp, dp, vp, dp, vp, dp, vp, dp, vp, dp, ip, ip, vp, dp, ip, ip, ip, ip, ip,
      ip, vp)
_decl("cg_interpolate_eno_3d_3tl_f64", i, vp, dp, vp, dp, vp, dp, ip, ip, vp, dp, ip, ip, ip, ip, ip, ip, vp)
_GP = C.POINTER(GData)
_decl("cg_transfer_from", i, _GP, C.POINTER(_GP), C.POINTER(C.c_double), i, ip, ip,
      C.POINTER(Islab), dp, i, i, i, i, vp)
_decl("cg_transfer_from_noalloc", i, _GP, C.POINTER(_GP), C.POINTER(C.c_double), i, ip, ip,
      C.POINTER(Islab), dp, i, i, i, i, vp)
_decl("cg_batch_create", i, C.POINTER(C.c_void_p))
_decl("cg_batch_destroy", i, vp)
_decl("cg_batch_add", i, vp, i, _GP, C.POINTER(_GP), C.POINTER(C.c_double), i, ip, ip,
      C.POINTER(Islab), dp, i, i, i, i, i)
_decl("cg_batch_finalize", i, vp, i)
_decl("cg_batch_run", i, vp, vp)
_decl("cg_batch_run_plain", i, vp, vp)
_decl("cg_batch_capture_graph", i, vp, vp)
_decl("cg_batch_launch_graph", i, vp, vp)
_decl("cg_batch_set_time", i, vp, i, C.POINTER(C.c_double), i, dp)
_decl("cg_batch_set_storage", i, vp, i, _GP, _GP)
_decl("cg_batch_set_phase_mode", i, vp, i, i)
_decl("cg_batch_launch_count", i, vp)
_decl("cg_batch_run_launch", i, vp, i, vp)
_decl("cg_batch_launch_stats", i, vp, i, ip, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong))
_decl("cg_batch_stats", i, vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), C.POINTER(C.c_longlong))


class CarpetError(RuntimeError):
    """A libcarpet_amd call returned a negative CG_ERR_* code."""

    def __init__(self, code: int, where: str):
        self.code = code
        msg = lib.cg_error_string(code).decode()
        detail = lib.cg_last_error_message().decode()
        super().__init__(f"{where}: {msg} ({code}): {detail}")


def check(code: int, where: str) -> int:
    if code < 0:
        raise CarpetError(code, where)
    return code

_decl("cg_prolongate_3d_rf2_mixed", i, i, *_OPARGS, vp)
_decl("cg_prolongate_3d_cc_rf2_mixed", i, i, *_OPARGS, vp)
_decl("cg_transfer_from_mixed", i, _GP, C.POINTER(_GP), C.POINTER(C.c_double), i, ip, ip, dp, i, i, i, i, vp)
_decl("cg_batch_autotune", i, vp, vp, i, C.c_char_p)
_decl("cg_batch_describe", i, vp, C.c_char_p, C.c_size_t)


class LevelSpec(C.Structure):
    _fields_ = [("baseextent", C.c_int * 9), ("ncomp", C.c_int), ("interior", C.POINTER(C.c_int)),
                ("reffact", C.c_int * 3), ("ghost_width", C.c_int * 6), ("buffer_width", C.c_int * 6),
                ("overlap_width", C.c_int * 6), ("prolongation_order", C.c_int)]


_decl("cg_schedule_build", i, i, ip, i, C.POINTER(LevelSpec), i, i, C.POINTER(vp))
_decl("cg_schedule_free", None, vp)
_decl("cg_schedule_count", i, vp, i, i)
_decl("cg_schedule_region", i, vp, i, i, i, ip, ip, ip, ip)
_decl("cg_schedule_component", i, vp, i, i, i, ip)
_decl("cg_schedule_set", i, vp, i, i, i, i, ip)
_decl("cg_schedule_regrid", i, vp, vp)
