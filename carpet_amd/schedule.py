"""Communication schedule of a grid hierarchy: dh::regrid's region lists
(CarpetLib/src/dh.cc:211-1080) computed by libcarpet_amd
(cg_schedule_build), and a level step built from them.

A hierarchy is given the way CarpetLib's gh holds it: per refinement level
the domain exterior (h.baseextent), the components' boxes from the
regridder (h.extent), the refinement factor, ghost / buffer / overlap widths
and the prolongation order; plus the centering and the outer boundary
width.  The lists name transfers (source component, send box, destination
component, receive box) that cg_transfer_from / Batch.add execute.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Dict, List, Sequence, Tuple

from ._lib import I9, LevelSpec, check, lib
from .core import Box

REF_PROL, SYNC, REF_BND_PROL, REF_REST, OLD2NEW_SYNC, OLD2NEW_PROL = range(6)
LIST_NAMES = {REF_PROL: "ref_prol", SYNC: "sync", REF_BND_PROL: "ref_bnd_prol", REF_REST: "ref_rest",
              OLD2NEW_SYNC: "old2new_sync", OLD2NEW_PROL: "old2new_prol"}
EXTERIOR, INTERIOR, OWNED, COMMUNICATED = range(4)
GHOSTS, BUFFERS, OVERLAPS, ACTIVE, SYNCED, BNDREF, BOUNDARIES = range(10, 17)


def _box_tuple(b) -> Tuple[tuple, tuple, tuple]:
    if isinstance(b, Box):
        return (tuple(b.lower), tuple(b.upper), tuple(b.stride))
    lo, up, st = b
    return (tuple(lo), tuple(up), tuple(st))


def _i9(b) -> I9:
    lo, up, st = _box_tuple(b)
    return I9(*lo, *up, *st)


def _from9(a) -> Tuple[tuple, tuple, tuple]:
    return (tuple(a[0:3]), tuple(a[3:6]), tuple(a[6:9]))


@dataclass
class LevelDesc:
    """One refinement level: domain exterior, component boxes, refinement
    factor relative to level 0, widths ((lower x,y,z), (upper x,y,z)) and
    the prolongation order."""
    domain: tuple
    components: List[tuple]
    reffact: Tuple[int, int, int]
    ghost_width: tuple = ((0, 0, 0), (0, 0, 0))
    buffer_width: tuple = ((0, 0, 0), (0, 0, 0))
    overlap_width: tuple = ((0, 0, 0), (0, 0, 0))
    prolongation_order: int = 1


@dataclass
class Region:
    src: int
    send: tuple
    dst: int
    recv: tuple


class Schedule:
    """cg_schedule_build over a list of LevelDesc.  lists[rl][kind] are the
    Region entries of level rl (restriction entries live on the coarse,
    receiving level, as in the reference)."""

    def __init__(self, centering: int, boundary_width, levels: Sequence[LevelDesc],
                 use_higher_order_restriction: bool = False, restriction_order_space: int = 3):
        self.centering = centering
        self.levels = list(levels)
        specs = (LevelSpec * len(levels))()
        self._keep = []
        for sp, L in zip(specs, levels):
            sp.baseextent = _i9(L.domain)
            sp.ncomp = len(L.components)
            arr = (C.c_int * (9 * max(1, len(L.components))))()
            for c, b in enumerate(L.components):
                lo, up, st = _box_tuple(b)
                arr[9 * c:9 * c + 9] = [*lo, *up, *st]
            self._keep.append(arr)
            sp.interior = C.cast(arr, C.POINTER(C.c_int))
            sp.reffact = (C.c_int * 3)(*L.reffact)
            sp.ghost_width = (C.c_int * 6)(*L.ghost_width[0], *L.ghost_width[1])
            sp.buffer_width = (C.c_int * 6)(*L.buffer_width[0], *L.buffer_width[1])
            sp.overlap_width = (C.c_int * 6)(*L.overlap_width[0], *L.overlap_width[1])
            sp.prolongation_order = L.prolongation_order
        bw = (C.c_int * 6)(*boundary_width[0], *boundary_width[1])
        h = C.c_void_p()
        check(lib.cg_schedule_build(centering, bw, len(levels), specs, int(bool(use_higher_order_restriction)),
                                    restriction_order_space, C.byref(h)), "cg_schedule_build")
        self._h = h
        self._read_lists()

    def _read_lists(self):
        h = self._h
        self.lists: List[Dict[int, List[Region]]] = []
        for rl in range(len(self.levels)):
            per = {}
            for kind in LIST_NAMES:
                n = check(lib.cg_schedule_count(h, rl, kind), "cg_schedule_count")
                out = []
                for k in range(n):
                    s, d = C.c_int(), C.c_int()
                    sb, rb = I9(), I9()
                    check(lib.cg_schedule_region(h, rl, kind, k, C.byref(s), sb, C.byref(d), rb),
                          "cg_schedule_region")
                    out.append(Region(s.value, _from9(sb), d.value, _from9(rb)))
                per[kind] = out
            self.lists.append(per)

    def regrid(self, old: "Schedule | None"):
        """The regridding lists of this hierarchy after a regrid from old
        (cg_schedule_regrid): OLD2NEW_SYNC from old's components,
        OLD2NEW_PROL from this hierarchy's next coarser level."""
        check(lib.cg_schedule_regrid(self._h, old._h if old is not None else None), "cg_schedule_regrid")
        self._read_lists()

    def component(self, rl: int, c: int, which: int = EXTERIOR) -> tuple:
        b = I9()
        check(lib.cg_schedule_component(self._h, rl, c, which, b), "cg_schedule_component")
        return _from9(b)

    def region_set(self, rl: int, c: int, which: int) -> List[tuple]:
        n = check(lib.cg_schedule_set(self._h, rl, c, which, 0, None), "cg_schedule_set")
        arr = (C.c_int * (9 * max(1, n)))()
        check(lib.cg_schedule_set(self._h, rl, c, which, n, C.cast(arr, C.POINTER(C.c_int))), "cg_schedule_set")
        return [_from9(arr[9 * k:9 * k + 9]) for k in range(n)]

    def close(self):
        if self._h:
            lib.cg_schedule_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
