"""Shared test helpers: seeded geometry and data generators, comparisons.

Geometry follows CarpetLib's conventions: vertex-centred fine stride sf and
coarse stride 2*sf with bounds on multiples of the stride; cell-centred
coordinates put cell centres at odd multiples of half the stride (coarse
cell c at sc/2 + c*sc, fine cell f at sf/2 + f*sf), as in the diagrams of
CarpetLib/src/prolongate_3d_cc_rf2.cc:15-38.
"""
from __future__ import annotations

import numpy as np

from pyoracle import HostArr

# stencil windows per parity (coarse offsets relative to the base b=g>>1),
# from the reference's coeffs1d definitions
def windows(kind: str, order: int):
    if kind == "vlag":
        n = order + 1
        return [(0, 0), (-n // 2 + 1, -n // 2 + n)]
    if kind == "clag":
        n = order + 1
        t0 = -(n // 2)
        t1 = t0 + (1 if order % 2 else 0)
        return [(t0, t0 + n - 1), (t1, t1 + n - 1)]
    if kind == "eno":
        w = {2: [(-2, 1), (-1, 2)], 3: [(-3, 2), (-2, 3)]}
        return w[order]
    if kind == "tvd":
        return [(-1, 1), (-1, 1)]
    if kind in ("veno", "weno"):  # vertex ENO: b for coincident, b-1..b+2 for midpoints
        return [(0, 0), (-1, 2)]
    if kind == "mono":  # monotone order 5: b, or b-2..b+3
        return [(0, 0), (-2, 3)]
    if kind == "stag":  # staggered operators: the union over both direction kinds
        return [(-1, 1), (-1, 1)] if order == 2 else [(-2, 1), (-1, 2)]
    raise ValueError(kind)


def footprint(win, off, n):
    lo, hi = 10**9, -10**9
    for f in {0, 1, n - 2, n - 1}:
        if 0 <= f < n:
            g = off + f
            b, p = g >> 1, g & 1
            lo = min(lo, b + win[p][0])
            hi = max(hi, b + win[p][1])
    return lo, hi


def pad_of(ext, rng, maxpad=3):
    return tuple(int(e + rng.integers(0, maxpad + 1)) for e in ext)


def host_array(box, pad, fill=np.nan):
    lo, up, st = box
    data = np.full((pad[2], pad[1], pad[0]), fill, dtype=np.float64)
    return HostArr(box, data)


def prolong_case(rng, kind, order, n=None, scale=1, maxn=40):
    """Random prolongation geometry.  Returns (srcbox, dstbox, regbox)."""
    cell = kind in ("clag", "eno", "tvd")
    win = windows(kind, order)
    sf = (2 if cell else 1) * scale
    sc = 2 * sf
    if n is None:
        n = tuple(int(rng.integers(1, maxn + 1)) for _ in range(3))
    slo, sup, dlo, dup, rlo, rup = [], [], [], [], [], []
    for d in range(3):
        C0 = int(rng.integers(-5, 6))  # coarse lower index
        tmin = min(w[0] for w in win)
        off = 2 * (-tmin) + int(rng.integers(0, 4))
        lo, hi = footprint(win, off, n[d])
        assert lo >= 0
        CE = hi + 1 + int(rng.integers(0, 3))
        F0 = 2 * C0 + off  # global fine index of the region start
        if cell:
            slo.append(sc // 2 + C0 * sc)
            sup.append(sc // 2 + (C0 + CE - 1) * sc)
            rlo.append(sf // 2 + F0 * sf)
            rup.append(sf // 2 + (F0 + n[d] - 1) * sf)
        else:
            slo.append(C0 * sc)
            sup.append((C0 + CE - 1) * sc)
            rlo.append(F0 * sf)
            rup.append((F0 + n[d] - 1) * sf)
        m1, m2 = int(rng.integers(0, 3)), int(rng.integers(0, 3))
        dlo.append(rlo[-1] - m1 * sf)
        dup.append(rup[-1] + m2 * sf)
    st_c, st_f = (sc,) * 3, (sf,) * 3
    return (tuple(slo), tuple(sup), st_c), (tuple(dlo), tuple(dup), st_f), (tuple(rlo), tuple(rup), st_f)


def sizes(box):
    lo, up, st = box
    return tuple((u - l + s) // s for l, u, s in zip(lo, up, st))


def field(rng, ext, kind="mixed"):
    """Test data on an (ext_x, ext_y, ext_z) array, returned as (z,y,x)."""
    nz, ny, nx = ext[2], ext[1], ext[0]
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    if kind == "random":
        return rng.uniform(-1, 1, size=(nz, ny, nx))
    smooth = np.sin(0.37 * x + 0.1) * np.cos(0.23 * y - 0.4) * np.sin(0.31 * z + 0.7) + 0.01 * x
    if kind == "smooth":
        return smooth
    if kind == "mixed":
        jump = np.where(0.3 * x + 0.5 * y - 0.4 * z > 0.5 * nx * 0.3, 10.0, 1.0)
        return smooth * jump + 1e-3 * rng.uniform(-1, 1, size=(nz, ny, nx))
    raise ValueError(kind)


def same(a: np.ndarray, b: np.ndarray) -> bool:
    """Bit-level agreement up to the sign of zero and NaN payloads."""
    return a.shape == b.shape and bool(np.array_equal(a, b, equal_nan=True))


def first_diff(a, b):
    m = ~((a == b) | (np.isnan(a) & np.isnan(b)))
    idx = np.argwhere(m)
    if len(idx) == 0:
        return None
    k = tuple(idx[0])
    return k, a[k], b[k], int(m.sum())
