"""Seeded synthetic hierarchies for the transfer operators.

Each workload describes its arrays (boxes, padding, host input data) and the
ordered list of transfers of one level step, and runs that step on the GPU
through the library.  The same description is what the tests replay on the
CPU oracle.

Transfers are tuples (name, dst, srcs, times, dstreg, srcreg, time, op,
centering, order_space, order_time, slab_offset) where dst/srcs name arrays.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List

import numpy as np

from .core import (CELL_CENTERED, OP_ENO, OP_LAGRANGE, OP_RESTRICT, OP_SYNC, OP_TVD, PREC_F64,
                   PREC_MIXED, VERTEX_CENTERED, Batch, Box, DeviceArray, fill_hash_many, pad_shape,
                   transfer_from)
from ._lib import check, lib

GOLDEN = 0x9E3779B97F4A7C15


def splitmix64(x: np.ndarray) -> np.ndarray:
    """Counter-based splitmix64 (uint64 in, uint64 out)."""
    with np.errstate(over="ignore"):
        z = (x + np.uint64(GOLDEN)).astype(np.uint64)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, n: int, lo: float, hi: float) -> np.ndarray:
    z = splitmix64(np.arange(n, dtype=np.uint64) + np.uint64(seed))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


def hash_field(seed: int, ext, lo: float = -1.0, hi: float = 1.0) -> np.ndarray:
    """Host twin of cg_fill_hash_f64: (ext_z, ext_y, ext_x) array."""
    n = int(np.prod(ext))
    z = splitmix64(np.arange(n, dtype=np.uint64) + np.uint64(seed & (2**64 - 1)))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return (lo + (hi - lo) * u).reshape(ext[2], ext[1], ext[0])


@dataclass
class ArraySpec:
    box: Box
    owned: Box | None = None
    init: np.ndarray | None = None  # (ext_z, ext_y, ext_x); None -> NaN poison
    seed: int | None = None  # device-side seeded fill (large workloads)


@dataclass
class Transfer:
    name: str
    dst: str
    srcs: List[str]
    times: List[float]
    dstreg: Box
    srcreg: Box
    time: float
    op: int
    cent: int
    order_space: int
    order_time: int
    slab: tuple | None = None
    phase: int = 0


@dataclass
class Workload:
    name: str
    arrays: Dict[str, ArraySpec] = field(default_factory=dict)
    steps: List[Transfer] = field(default_factory=list)
    outputs: List[str] = field(default_factory=list)

    def allocate(self, pool: bool = False):
        """Device arrays for every ArraySpec.  pool=True carves them out of one
        allocation (thousands of small arrays); seeded arrays are filled on
        the device from a shared seeded buffer."""
        dev = {}
        if pool:
            specs = list(self.arrays.items())
            layout = []
            total = 0
            for k, spec in specs:
                shape = spec.box.sizes()
                off = (0, 0, 0)
                if spec.owned is not None:
                    off = tuple((o - l) // s for o, l, s in zip(spec.owned.lower, spec.box.lower, spec.box.stride))
                pad, poff = pad_shape(shape, off, 16, 16, spec.owned is not None)
                start = (total + 255) // 256 * 256 + poff[0] * 8
                n = int(np.prod(pad)) * 8
                layout.append((k, spec, pad, start))
                total = start + n
            import ctypes as C
            base = C.c_void_p()
            check(lib.cg_dev_malloc(C.byref(base), total), "cg_dev_malloc")
            self._pool = base.value
            self.pool_bytes = total
            for k, spec, pad, start in layout:
                dev[k] = DeviceArray.view(spec.box, self._pool + start, pad)
        else:
            for k, spec in self.arrays.items():
                dev[k] = DeviceArray(spec.box, owned=spec.owned)
        if pool:
            # the whole pool poisoned with NaN (padding and unseeded arrays
            # included), then every seeded array in one launch
            check(lib.cg_fill_f64(self._pool, float("nan"), self.pool_bytes // 8, None), "cg_fill_f64")
            seeded = [(dev[k], spec.seed) for k, spec in self.arrays.items()
                      if spec.init is None and spec.seed is not None]
            self.setup_launches = 1 + fill_hash_many(seeded)
            for k, spec in self.arrays.items():
                if spec.init is not None:
                    dev[k].upload(spec.init, fill=np.nan)
                    self.setup_launches += 1
            return dev
        for k, spec in self.arrays.items():
            a = dev[k]
            if spec.init is not None:
                a.upload(spec.init, fill=np.nan)
            elif spec.seed is not None:
                a.fill(np.nan)
                a.fill_hash(spec.seed)
            else:
                a.fill(np.nan)
        return dev

    def host_inputs(self) -> Dict[str, np.ndarray]:
        """Host copies of every array's initial extent (NaN where poisoned)."""
        out = {}
        for k, spec in self.arrays.items():
            e = spec.box.sizes()
            if spec.init is not None:
                out[k] = np.asarray(spec.init, dtype=np.float64).reshape(e[2], e[1], e[0])
            elif spec.seed is not None:
                out[k] = hash_field(spec.seed, e)
            else:
                out[k] = np.full((e[2], e[1], e[0]), np.nan)
        return out

    def free(self, dev):
        for a in dev.values():
            a.free()
        if getattr(self, "_pool", None):
            lib.cg_dev_free(self._pool)
            self._pool = None

    def batch(self, dev, prec: int = PREC_F64) -> Batch:
        b = Batch()
        for t in self.steps:
            b.add(t.phase, dev[t.dst], [dev[s] for s in t.srcs], t.times, t.dstreg, t.srcreg, t.time,
                  t.op, t.cent, t.order_space, t.order_time, slab_offset=t.slab, prec=prec)
        return b

    def run_gpu(self, cg=None, stream=None) -> Dict[str, np.ndarray]:
        dev = self.allocate()
        self.run_steps(dev, stream)
        out = {k: dev[k].download() for k in self.outputs}
        self.free(dev)
        return out

    def run_steps(self, dev, stream=None):
        """One transfer_from call per step, in phase order."""
        for t in sorted(self.steps, key=lambda t: t.phase):
            transfer_from(dev[t.dst], [dev[s] for s in t.srcs], t.times, t.dstreg, t.srcreg, t.time,
                          t.op, t.cent, t.order_space, t.order_time, slab_offset=t.slab, stream=stream)


def box3(lo, up, st) -> Box:
    return Box((lo,) * 3 if np.isscalar(lo) else lo, (up,) * 3 if np.isscalar(up) else up,
               (st,) * 3 if np.isscalar(st) else st)


def config1(seed: int = 20261202) -> Workload:
    """FP64 vertex-centred level step on a 33^3 coarse box (+3 ghosts): order-3
    prolongation onto the 64^3 fine region [2,65]^3 held by two 32x64x64
    components, rf2 restriction of the covered coarse points [1,32]^3 back,
    and the copy sync of the two components' ghost slabs (width 3).
    Values: sin(2 pi x) sin(2 pi y) sin(2 pi z), x = coarse index / 32, plus
    uniform noise of amplitude 1e-3 (splitmix64 counter stream)."""
    w = Workload("config1")
    g = 3
    # coarse: interior 0..32 (coarse index), stride 2 in finest units
    cbox = box3(-2 * g, 2 * (32 + g), 2)
    n = cbox.sizes()[0]
    idx = np.arange(-g, 32 + g + 1, dtype=np.float64) / 32.0
    s = np.sin(2 * np.pi * idx)
    data = s[:, None, None] * s[None, :, None] * s[None, None, :]
    data = data + uniform(seed, n**3, -1e-3, 1e-3).reshape(n, n, n)
    w.arrays["coarse"] = ArraySpec(cbox, owned=box3(0, 64, 2), init=data)
    # fine: two components split in x, interior [2,33] and [34,65], ghosts 3
    fa = Box((2 - g, 2 - g, 2 - g), (33 + g, 65 + g, 65 + g), (1, 1, 1))
    fb = Box((34 - g, 2 - g, 2 - g), (65 + g, 65 + g, 65 + g), (1, 1, 1))
    w.arrays["fineA"] = ArraySpec(fa, owned=Box((2, 2, 2), (33, 65, 65), (1, 1, 1)))
    w.arrays["fineB"] = ArraySpec(fb, owned=Box((34, 2, 2), (65, 65, 65), (1, 1, 1)))
    ra = Box((2, 2, 2), (33, 65, 65), (1, 1, 1))
    rb = Box((34, 2, 2), (65, 65, 65), (1, 1, 1))
    for nm, r in (("fineA", ra), ("fineB", rb)):
        w.steps.append(Transfer(f"prolong->{nm}", nm, ["coarse"], [0.0], r, r, 0.0, OP_LAGRANGE,
                                VERTEX_CENTERED, 3, 0, phase=0))
    # restriction of coarse points [1,32]^3 (finest units 2..64), split by owner
    w.steps.append(Transfer("restrict<-fineA", "coarse", ["fineA"], [0.0],
                            Box((2, 2, 2), (32, 64, 64), (2, 2, 2)), Box((2, 2, 2), (32, 64, 64), (2, 2, 2)),
                            0.0, OP_LAGRANGE, VERTEX_CENTERED, 3, 0, phase=1))
    w.steps.append(Transfer("restrict<-fineB", "coarse", ["fineB"], [0.0],
                            Box((34, 2, 2), (64, 64, 64), (2, 2, 2)), Box((34, 2, 2), (64, 64, 64), (2, 2, 2)),
                            0.0, OP_LAGRANGE, VERTEX_CENTERED, 3, 0, phase=1))
    # sync: A's upper x ghosts from B, B's lower x ghosts from A
    ga = Box((34, 2, 2), (36, 65, 65), (1, 1, 1))
    gb = Box((31, 2, 2), (33, 65, 65), (1, 1, 1))
    w.steps.append(Transfer("sync A<-B", "fineA", ["fineB"], [0.0], ga, ga, 0.0, OP_SYNC,
                            VERTEX_CENTERED, 3, 0, phase=2))
    w.steps.append(Transfer("sync B<-A", "fineB", ["fineA"], [0.0], gb, gb, 0.0, OP_SYNC,
                            VERTEX_CENTERED, 3, 0, phase=2))
    w.outputs = ["coarse", "fineA", "fineB"]
    return w


def config4(n_regions: int = 2000, n_gf: int = 24, edge=(8, 96), seed: int = 4,
            levels: int = 5, order: int = 5) -> Workload:
    """Batched uneven level step: n_regions destination regions per grid
    function with edges uniform in [edge[0], edge[1]] fine points per axis,
    split evenly between copies (copy_3d of a whole region box from another
    array whose box is offset from it, the kernel of a sync), order-5
    Lagrange prolongation with 3-level time interpolation (sources at t = 0,
    -dt, -2dt interpolated to t = -dt/2) and rf2 restriction.  Regions are
    whole boxes, not the thin ghost and buffer shells of a real hierarchy
    (those come from its communication schedule: hierarchy_step); each
    destination array holds its region plus 3 ghost points per side.  Regions belong to one of
    `levels` refinement levels (stride 2^(levels-1-rl) in finest units).
    Every region has its own destination and source arrays, so the
    destinations of a phase are disjoint, as dh's lists guarantee.
    Phases: 0 restriction, 1 prolongation, 2 sync (prolongation precedes
    sync, CarpetLib's SyncProlongateGroups order)."""
    rng = np.random.default_rng(seed)
    w = Workload(f"config4[{n_regions}x{n_gf}]")
    lo_e, hi_e = edge
    half = order // 2 + 1
    times = [0.0, -1.0, -2.0]
    for r in range(n_regions):
        n = tuple(int(v) for v in rng.integers(lo_e, hi_e + 1, size=3))
        rl = int(rng.integers(0, levels))
        sf = 2 ** (levels - 1 - rl)
        kind = r % 3
        base = tuple(int(v) for v in rng.integers(-40, 40, size=3))
        for g in range(n_gf):
            tag = f"r{r}g{g}"
            sd = seed * 1000003 + r * 1009 + g
            if kind == 0:  # copy: the region from another array offset from it
                lo = tuple(b * sf for b in base)
                reg = Box(lo, tuple(l + (k - 1) * sf for l, k in zip(lo, n)), (sf,) * 3)
                dbox = reg.expand((3, 3, 3), (3, 3, 3))
                sbox = reg.expand((1, 0, 2), (0, 2, 1))
                w.arrays[tag + "d"] = ArraySpec(dbox, owned=reg, seed=sd)
                w.arrays[tag + "s"] = ArraySpec(sbox, seed=sd + 7)
                w.steps.append(Transfer(tag, tag + "d", [tag + "s"], [0.0], reg, reg, 0.0, OP_LAGRANGE,
                                        VERTEX_CENTERED, order, 2, phase=2))
            elif kind == 1:  # prolongation + 3tl
                sc = 2 * sf
                lo = tuple((2 * b + int(rng.integers(0, 2))) * sf for b in base)
                reg = Box(lo, tuple(l + (k - 1) * sf for l, k in zip(lo, n)), (sf,) * 3)
                clo = tuple((l // sc) - half for l in reg.lower)
                chi = tuple((u // sc) + half + 1 for u in reg.upper)
                sbox = Box(tuple(c * sc for c in clo), tuple(c * sc for c in chi), (sc,) * 3)
                dbox = reg.expand((3, 3, 3), (3, 3, 3))
                w.arrays[tag + "d"] = ArraySpec(dbox, owned=reg, seed=sd)
                for tl in range(3):
                    w.arrays[tag + f"s{tl}"] = ArraySpec(sbox, seed=sd + 11 * (tl + 1))
                w.steps.append(Transfer(tag, tag + "d", [tag + f"s{tl}" for tl in range(3)], times, reg, reg,
                                        -0.5, OP_LAGRANGE, VERTEX_CENTERED, order, 2, phase=1))
            else:  # restriction: coarse region from a fine component
                sc = 2 * sf
                m = tuple(max(1, k // 2) for k in n)
                lo = tuple(b * sc for b in base)
                reg = Box(lo, tuple(l + (k - 1) * sc for l, k in zip(lo, m)), (sc,) * 3)
                sbox = Box(tuple(l - 2 * sf for l in reg.lower), tuple(u + 2 * sf for u in reg.upper), (sf,) * 3)
                dbox = reg.expand((3, 3, 3), (3, 3, 3))
                w.arrays[tag + "d"] = ArraySpec(dbox, owned=reg, seed=sd)
                w.arrays[tag + "s"] = ArraySpec(sbox, seed=sd + 5)
                w.steps.append(Transfer(tag, tag + "d", [tag + "s"], [0.0], reg, reg, 0.0, OP_LAGRANGE,
                                        VERTEX_CENTERED, order, 2, phase=0))
            w.outputs.append(tag + "d")
    return w


@dataclass
class Config2:
    """Single large patch, FP64 vertex-centred: n_gf grid functions on a
    coarse 257^3 box (+3 ghosts, 263^3 allocated, stride 2), each with 3 time
    levels at t = 1, 0, -1; the fine region [2,513]^3 (512^3 points, stride 1)
    is prolongated with prolongate_3d_rf2<5> fused with interpolate_3d_3tl at
    the target times t = 1/3 and 2/3.  Each grid function is a sum of 8
    Fourier modes, amplitude 1/8, wavenumbers in {1..8}^3 and phases uniform
    in [0, 2 pi) drawn from seed 1000+gf; time level t shifts each phase by
    0.1 t |k|."""

    n_gf: int = 24
    n_coarse: int = 257
    ghosts: int = 3
    order: int = 5
    targets: tuple = (1.0 / 3.0, 2.0 / 3.0)
    times: tuple = (1.0, 0.0, -1.0)

    def boxes(self):
        g = self.ghosts
        c = Box((-2 * g,) * 3, (2 * (self.n_coarse - 1 + g),) * 3, (2, 2, 2))
        n_f = 2 * (self.n_coarse - 1)
        reg = Box((2,) * 3, (n_f + 1,) * 3, (1, 1, 1))
        fine = reg.expand((g, g, g), (g, g, g))
        return c, fine, reg

    def dst_points(self) -> int:
        return self.boxes()[2].size() * self.n_gf * len(self.targets)

    def alg_bytes(self) -> int:
        """8 B per fine point written + 8 B per unique coarse source point per
        time level (the order-5 footprint of the region)."""
        c, _, reg = self.boxes()
        n = reg.sizes()[0]
        foot = (n // 2 + 5) ** 3  # bases 1..256 plus taps -2..+3
        per = 8 * reg.size() + 8 * len(self.times) * foot
        return per * self.n_gf * len(self.targets)

    def allocate(self, stream=None):
        cbox, fbox, reg = self.boxes()
        owned_c = Box((0,) * 3, (2 * (self.n_coarse - 1),) * 3, (2, 2, 2))
        src, dst = {}, {}
        for gf in range(self.n_gf):
            rng = np.random.default_rng(1000 + gf)
            kvec = rng.integers(1, 9, size=(8, 3))
            phase = rng.uniform(0, 2 * np.pi, size=8)
            for tl, t in enumerate(self.times):
                a = DeviceArray(cbox, owned=owned_c)
                ph = phase + 0.1 * t * np.linalg.norm(kvec, axis=1)
                # x = coarse index / (n_coarse-1): global index / (2*(n_coarse-1))
                a.fill_fourier(1.0 / (2 * (self.n_coarse - 1)), 1.0 / 8, kvec.tolist(), ph.tolist(), stream)
                src[(gf, tl)] = a
            for ti in range(len(self.targets)):
                d = DeviceArray(fbox, owned=reg)
                d.fill(np.nan, stream)
                dst[(gf, ti)] = d
        return src, dst

    def transfers(self, src, dst):
        _, _, reg = self.boxes()
        out = []
        for gf in range(self.n_gf):
            for ti, t in enumerate(self.targets):
                out.append((dst[(gf, ti)], [src[(gf, tl)] for tl in range(len(self.times))],
                            list(self.times), reg, t))
        return out

    def batch(self, src, dst, prec=PREC_F64) -> Batch:
        b = Batch()
        for d, ss, times, reg, t in self.transfers(src, dst):
            b.add(0, d, ss, times, reg, reg, t, OP_LAGRANGE, VERTEX_CENTERED, self.order, 2, prec=prec)
        return b


@dataclass
class Config3:
    """Cell-centred conservative set: n_var FP64 variables on a coarse 192^3
    box (+3 ghosts, 198^3 allocated; cell c centred at 2+4c in finest
    units) refined x2 onto the fine 384^3 region (all children of the
    interior cells) by prolongate_3d_cc_tvd and prolongate_3d_cc_eno_rf2
    (order_space 5 -> ENO<3>, 3 -> ENO<2>), then restricted back with
    restrict_3d_cc_rf2.  Variable 0 is rho = 1 + 0.1 prod_d sin(2 pi x_d)
    times 10 on one side of the plane n.x = d (n uniform on the sphere,
    d ~ U(0.3, 0.7), seed 3); the others are smooth products of sines."""

    n_var: int = 8
    n: int = 192
    ghosts: int = 3

    def boxes(self):
        g, n = self.ghosts, self.n
        c = Box((2 - 4 * g,) * 3, (2 + 4 * (n - 1 + g),) * 3, (4, 4, 4))
        owned = Box((2,) * 3, (2 + 4 * (n - 1),) * 3, (4, 4, 4))
        reg = Box((1,) * 3, (1 + 2 * (2 * n - 1),) * 3, (2, 2, 2))
        fine = reg.expand((g, g, g), (g, g, g))
        return c, owned, fine, reg

    def coarse_data(self, var: int) -> np.ndarray:
        c, _, _, _ = self.boxes()
        m = c.sizes()[0]
        x = (np.arange(m) - self.ghosts + 0.5) / self.n
        if var == 0:
            s = np.sin(2 * np.pi * x)
            rho = 1 + 0.1 * s[:, None, None] * s[None, :, None] * s[None, None, :]
            rng = np.random.default_rng(3)
            nv = rng.normal(size=3)
            nv /= np.linalg.norm(nv)
            d = rng.uniform(0.3, 0.7)
            # n.x with x = (x, y, z) on the (z, y, x) array
            plane = nv[0] * x[None, None, :] + nv[1] * x[None, :, None] + nv[2] * x[:, None, None]
            return np.where(plane > d, 10.0 * rho, rho)
        k = 1 + var % 4
        a = np.sin(2 * np.pi * k * x + 0.3 * var)
        b = np.cos(2 * np.pi * (k + 1) * x - 0.1 * var)
        return 1.0 + 0.2 * a[:, None, None] * b[None, :, None] * a[None, None, :]


    def allocate(self, stream=None):
        """Device arrays: coarse sources (uploaded), fine destinations and
        restriction targets (NaN), per variable."""
        cbox, owned, fbox, reg = self.boxes()
        coarse, fine, back = [], [], []
        for var in range(self.n_var):
            c = DeviceArray(cbox, owned=owned)
            c.upload(self.coarse_data(var))
            coarse.append(c)
            f = DeviceArray(fbox, owned=reg)
            f.fill(np.nan, stream)
            fine.append(f)
            r = DeviceArray(cbox, owned=owned)
            r.fill(np.nan, stream)
            back.append(r)
        return coarse, fine, back

    def prolong_batch(self, coarse, fine, op: int, order_space: int) -> Batch:
        _, _, _, reg = self.boxes()
        b = Batch()
        for c, f in zip(coarse, fine):
            b.add(0, f, [c], [0.0], reg, reg, 0.0, op, CELL_CENTERED, order_space, 0)
        return b

    def restrict_batch(self, fine, back) -> Batch:
        _, owned, _, _ = self.boxes()
        b = Batch()
        for f, r in zip(fine, back):
            b.add(0, r, [f], [0.0], owned, owned, 0.0, OP_TVD, CELL_CENTERED, 1, 0)
        return b


@dataclass
class Config5:
    """Mixed precision on the config-1 and config-3 shapes: n_var vertex-
    centred grid functions on a coarse n^3 box (+3 ghosts, stride 2), 3 time
    levels at t = 1, 0, -1, prolongated by prolongate_3d_rf2<5> fused with
    interpolate_3d_3tl at t = 1/3 onto the fine region [2, 2n-1]^3 (bases
    1..n-1, order-5 taps -2..+3 inside the ghosts).  Shape "cfg1": n = 33
    (fine 64^3), 1 variable; shape "cfg3": n = 193 (fine 384^3), 8 variables.
    Data: sums of 8 Fourier modes as in the single-patch configuration.  The
    same values feed the FP64 path (float64 arrays) and the mixed path
    (float32 storage of sources and destination; compensated FP32 sums; FP64
    time interpolation)."""

    shape: str = "cfg3"
    order: int = 5
    target: float = 1.0 / 3.0
    times: tuple = (1.0, 0.0, -1.0)
    ghosts: int = 3

    @property
    def n(self) -> int:
        return 33 if self.shape == "cfg1" else 193

    @property
    def n_var(self) -> int:
        return 1 if self.shape == "cfg1" else 8

    def boxes(self):
        g, n = self.ghosts, self.n
        c = Box((-2 * g,) * 3, (2 * (n - 1 + g),) * 3, (2, 2, 2))
        owned = Box((0,) * 3, (2 * (n - 1),) * 3, (2, 2, 2))
        reg = Box((2,) * 3, (2 * n - 1,) * 3, (1, 1, 1))
        fine = reg.expand((g, g, g), (g, g, g))
        return c, owned, fine, reg

    def allocate(self, stream=None):
        """Per precision: sources [var][tl] and destinations [var].  The FP64
        sources hold the FP32-rounded values when rounded=True (isolates the
        arithmetic), else the original FP64 values."""
        cbox, owned, fbox, reg = self.boxes()
        src64, src32, src64r, dst64, dst32 = [], [], [], [], []
        for var in range(self.n_var):
            rng = np.random.default_rng(5000 + var)
            kvec = rng.integers(1, 9, size=(8, 3))
            phase = rng.uniform(0, 2 * np.pi, size=8)
            s64, s32, s64r = [], [], []
            for t in self.times:
                a = DeviceArray(cbox, owned=owned)
                ph = phase + 0.1 * t * np.linalg.norm(kvec, axis=1)
                a.fill_fourier(1.0 / (2 * (self.n - 1)), 1.0 / 8, kvec.tolist(), ph.tolist(), stream)
                if stream is not None:
                    stream.synchronize()
                h = a.download()
                b32 = DeviceArray(cbox, owned=owned, dtype=np.float32)
                b32.upload(h.astype(np.float32))
                b64 = DeviceArray(cbox, owned=owned)
                b64.upload(h.astype(np.float32).astype(np.float64))
                s64.append(a)
                s32.append(b32)
                s64r.append(b64)
            src64.append(s64)
            src32.append(s32)
            src64r.append(s64r)
            d64 = DeviceArray(fbox, owned=reg)
            d64.fill(np.nan, stream)
            dst64.append(d64)
            d32 = DeviceArray(fbox, owned=reg, dtype=np.float32)
            d32.upload(np.full(d32.padded_shape_zyx(), np.nan, np.float32))
            dst32.append(d32)
        return src64, src32, src64r, dst64, dst32

    def batch(self, srcs, dsts, prec) -> Batch:
        _, _, _, reg = self.boxes()
        b = Batch()
        for ss, d in zip(srcs, dsts):
            b.add(0, d, ss, list(self.times), reg, reg, self.target, OP_LAGRANGE, VERTEX_CENTERED,
                  self.order, 2, prec=prec)
        return b


def nested_levels(ncoarse: int = 32, nlevels: int = 3, split=(2, 1, 1), centering: int = VERTEX_CENTERED,
                  ghost: int = 3, buffer: int = 6, order: int = 5, fraction: float = 0.5, maxlevels: int = 0):
    """A nested hierarchy as CarpetLib's gh would hold it (schedule.LevelDesc
    per level): level 0 covers the domain of ncoarse coarse points per axis,
    every finer level (refinement factor 2) the central `fraction` of the
    previous level's region; each level's region is tiled into
    split[0] x split[1] x split[2] components.  Vertex centring puts level rl
    on stride 2^(nlevels-1-rl); cell centring on stride 2^(nlevels-rl) with
    cell centres at odd multiples of half the stride.  maxlevels (default
    nlevels) fixes the stride of the finest possible level, as Carpet's
    maximum number of refinement levels does, so hierarchies with different
    numbers of levels share one index space."""
    from .schedule import LevelDesc
    top = max(maxlevels, nlevels)
    ns = 2 ** (top - 1) if centering == VERTEX_CENTERED else 2 ** top
    levels = []
    plo, phi = None, None
    for rl in range(nlevels):
        s = ns // 2 ** rl
        off = 0 if centering == VERTEX_CENTERED else s // 2
        npts = ncoarse * 2 ** rl + (1 if centering == VERTEX_CENTERED else 0)
        dom = ((off,) * 3, (off + s * (npts - 1),) * 3, (s,) * 3)
        if rl == 0:
            lo, hi = [0] * 3, [npts - 1] * 3
        else:
            lo, hi = [], []
            for d in range(3):
                a, b = 2 * plo[d], 2 * phi[d] + (0 if centering == VERTEX_CENTERED else 1)
                n = b - a + 1
                keep = max(8, int(round(n * fraction)))
                c0 = a + (n - keep) // 2
                lo.append(c0)
                hi.append(c0 + keep - 1)
        comps = []
        for ix in range(split[0]):
            for iy in range(split[1]):
                for iz in range(split[2]):
                    blo, bhi = [], []
                    for d, (k, nk) in enumerate(zip((ix, iy, iz), split)):
                        n = hi[d] - lo[d] + 1
                        a = lo[d] + n * k // nk
                        b = lo[d] + n * (k + 1) // nk - 1
                        blo.append(off + s * a)
                        bhi.append(off + s * b)
                    comps.append((tuple(blo), tuple(bhi), (s,) * 3))
        g = ((ghost,) * 3, (ghost,) * 3)
        bw = ((buffer,) * 3, (buffer,) * 3) if rl > 0 else ((0,) * 3, (0,) * 3)
        levels.append(LevelDesc(dom, comps, (2 ** rl,) * 3, g, bw, ((0,) * 3, (0,) * 3), order))
        plo, phi = lo, hi
    return levels


def tiled_levels(counts=((4, 4, 4), (5, 5, 5), (7, 7, 6), (8, 8, 8), (10, 10, 10)), edge=(8, 96), seed: int = 7,
                 ghost: int = 3, buffer: int = 6, order: int = 5, margin: int = 24):
    """A nested vertex-centred hierarchy whose levels are tiled by boxes
    with edges uniform in [edge[0], edge[1]] points per axis (counts[rl] =
    boxes per axis on level rl; the default makes 1995 boxes on 5 levels).
    Level rl's region is the union of its boxes, centred inside level rl-1's
    region (refinement factor 2) with `margin` fine points to spare on every
    side for ghost, buffer and prolongation stencil; level 0's region is the
    domain.  Returns schedule.LevelDesc per level (vertex centring: level rl
    on stride 2^(L-1-rl))."""
    from .schedule import LevelDesc
    rng = np.random.default_rng(seed)
    nl = len(counts)
    widths = []
    for rl in range(nl):
        for _ in range(1000):
            w = [[int(v) for v in rng.integers(edge[0], edge[1] + 1, size=counts[rl][d])] for d in range(3)]
            if rl == 0 or all(sum(w[d]) + 2 * margin <= 2 * (sum(widths[-1][d]) - 1) + 1 for d in range(3)):
                break
        else:
            raise ValueError(f"level {rl}: no tiling fits inside level {rl - 1}")
        widths.append(w)
    levels = []
    lo_prev = [0, 0, 0]
    for rl in range(nl):
        s = 2 ** (nl - 1 - rl)
        w = widths[rl]
        L = [sum(w[d]) for d in range(3)]
        if rl == 0:
            lo = [0, 0, 0]
        else:
            # level rl-1's region in level-rl points: 2*lo_prev .. 2*(lo_prev+Lp-1)
            Lp = [sum(widths[rl - 1][d]) for d in range(3)]
            lo = [2 * lo_prev[d] + (2 * (Lp[d] - 1) + 1 - L[d]) // 2 for d in range(3)]
        comps = []
        starts = []
        for d in range(3):
            acc, st = lo[d], []
            for k in w[d]:
                st.append(acc)
                acc += k
            starts.append(st)
        for iz in range(counts[rl][2]):
            for iy in range(counts[rl][1]):
                for ix in range(counts[rl][0]):
                    a = (starts[0][ix], starts[1][iy], starts[2][iz])
                    b = (a[0] + w[0][ix] - 1, a[1] + w[1][iy] - 1, a[2] + w[2][iz] - 1)
                    comps.append((tuple(s * v for v in a), tuple(s * v for v in b), (s,) * 3))
        if rl == 0:
            dom = ((0,) * 3, tuple(s * (L[d] - 1) for d in range(3)), (s,) * 3)
        else:
            dom = levels[0].domain
            # each level's domain in its own stride: the level-0 domain refined
            dom = ((0,) * 3, tuple(dom[1][d] for d in range(3)), (s,) * 3)
        g = ((ghost,) * 3, (ghost,) * 3)
        bw = ((buffer,) * 3, (buffer,) * 3) if rl > 0 else ((0,) * 3, (0,) * 3)
        levels.append(LevelDesc(dom, comps, (2 ** rl,) * 3, g, bw, ((0,) * 3, (0,) * 3), order))
        lo_prev = lo
    return levels


def hierarchy_step(levels, centering: int = VERTEX_CENTERED, n_gf: int = 2, seed: int = 6,
                   boundary=((0, 0, 0), (0, 0, 0)), op: int = OP_LAGRANGE, only_used: bool = False) -> Workload:
    """One level step of a real hierarchy, from its communication schedule
    (schedule.Schedule, dh's lists): for each level rl in turn, restriction
    from rl+1 into rl (phase 3rl), boundary prolongation from rl-1 with
    3-level time interpolation to the intermediate fine time (phase 3rl+1),
    then synchronisation (phase 3rl+2).  Arrays: every (level, component,
    grid function, time level) on the component's exterior, seeded; the
    outputs are the current time level of every component."""
    from .schedule import REF_BND_PROL, REF_REST, SYNC, Schedule
    sch = Schedule(centering, boundary, levels)
    w = Workload(f"hierarchy[{len(levels)} levels, {[len(L.components) for L in levels]} components x {n_gf}]")
    times = [0.0, -1.0, -2.0]
    for rl, L in enumerate(levels):
        for c in range(len(L.components)):
            ext = Box(*sch.component(rl, c, 0))
            own = Box(*sch.component(rl, c, 2))
            for g in range(n_gf):
                for tl in range(3):
                    nm = f"l{rl}c{c}g{g}t{tl}"
                    w.arrays[nm] = ArraySpec(ext, owned=own, seed=seed * 7919 + hash(nm) % 1000003)
                w.outputs.append(f"l{rl}c{c}g{g}t0")
    order = {rl: L.prolongation_order for rl, L in enumerate(levels)}
    for rl in range(len(levels)):
        # the region boxes are the same for every grid function
        rest = [(r, Box(*r.recv), Box(*r.send)) for r in sch.lists[rl][REF_REST]]
        bnd = [(r, Box(*r.recv), Box(*r.send)) for r in sch.lists[rl][REF_BND_PROL]]
        syn = [(r, Box(*r.recv), Box(*r.send)) for r in sch.lists[rl][SYNC]]
        for g in range(n_gf):
            for r, recv, send in rest:  # fine rl+1 -> coarse rl
                w.steps.append(Transfer(f"rest l{rl + 1}c{r.src}->l{rl}c{r.dst} g{g}", f"l{rl}c{r.dst}g{g}t0",
                                        [f"l{rl + 1}c{r.src}g{g}t0"], [0.0], recv, send, 0.0, op,
                                        centering, order[rl + 1], 0, phase=3 * rl))
            for r, recv, send in bnd:  # coarse rl-1 -> fine rl
                w.steps.append(Transfer(f"bnd l{rl - 1}c{r.src}->l{rl}c{r.dst} g{g}", f"l{rl}c{r.dst}g{g}t0",
                                        [f"l{rl - 1}c{r.src}g{g}t{tl}" for tl in range(3)], times,
                                        recv, send, -0.5, op, centering, order[rl], 2,
                                        phase=3 * rl + 1))
            for r, recv, send in syn:
                w.steps.append(Transfer(f"sync l{rl}c{r.src}->c{r.dst} g{g}", f"l{rl}c{r.dst}g{g}t0",
                                        [f"l{rl}c{r.src}g{g}t0"], [0.0], recv, send, 0.0, op,
                                        centering, order[rl], 0, phase=3 * rl + 2))
    if only_used:
        # arrays no transfer reads or writes (the finest level's past time
        # levels) are not allocated
        used = {t.dst for t in w.steps} | {x for t in w.steps for x in t.srcs} | set(w.outputs)
        w.arrays = {k: v for k, v in w.arrays.items() if k in used}
    w.schedule = sch
    return w


def recompose_step(old_levels, new_levels, centering: int = VERTEX_CENTERED, n_gf: int = 2, seed: int = 8,
                   boundary=((0, 0, 0), (0, 0, 0)), op: int = OP_LAGRANGE) -> Workload:
    """The fill of a new hierarchy's storage after a regrid
    (ggf::recompose_fill, CarpetLib/src/ggf.cc:175-215) from the regridding
    lists (schedule.Schedule.regrid): for every new level in turn and each of
    its 3 time levels, the copies from the old components of that level
    (phase 2rl), then prolongation from the new coarser level, whose 3 time
    levels at t = 0, -1, -2 are interpolated to the fine time level's time
    0, -1/2, -1 (phase 2rl+1).  Old and new arrays are seeded (new ones stand
    for freshly allocated storage); the outputs are the new arrays."""
    from .schedule import OLD2NEW_PROL, OLD2NEW_SYNC, Schedule
    s_old = Schedule(centering, boundary, old_levels)
    s_new = Schedule(centering, boundary, new_levels)
    s_new.regrid(s_old)
    w = Workload(f"recompose[{len(old_levels)} -> {len(new_levels)} levels x {n_gf}]")
    for tag, sch, levels in (("o", s_old, old_levels), ("n", s_new, new_levels)):
        for rl, L in enumerate(levels):
            for c in range(len(L.components)):
                ext = Box(*sch.component(rl, c, 0))
                own = Box(*sch.component(rl, c, 2))
                for g in range(n_gf):
                    for tl in range(3):
                        nm = f"{tag}l{rl}c{c}g{g}t{tl}"
                        w.arrays[nm] = ArraySpec(ext, owned=own, seed=seed * 7919 + hash(nm) % 1000003)
                        if tag == "n":
                            w.outputs.append(nm)
    ctimes = [0.0, -1.0, -2.0]
    for rl, L in enumerate(new_levels):
        for g in range(n_gf):
            for tl in range(3):
                for r in s_new.lists[rl][OLD2NEW_SYNC]:
                    w.steps.append(Transfer(f"old2new l{rl} c{r.src}->c{r.dst} g{g} t{tl}",
                                            f"nl{rl}c{r.dst}g{g}t{tl}", [f"ol{rl}c{r.src}g{g}t{tl}"], [0.0],
                                            Box(*r.recv), Box(*r.send), 0.0, op, centering, L.prolongation_order,
                                            0, phase=2 * rl))
                for r in s_new.lists[rl][OLD2NEW_PROL]:
                    w.steps.append(Transfer(f"regrid prol l{rl - 1}c{r.src}->l{rl}c{r.dst} g{g} t{tl}",
                                            f"nl{rl}c{r.dst}g{g}t{tl}",
                                            [f"nl{rl - 1}c{r.src}g{g}t{k}" for k in range(3)], ctimes,
                                            Box(*r.recv), Box(*r.send), -0.5 * tl, op, centering,
                                            L.prolongation_order, 2, phase=2 * rl + 1))
    w.schedule = (s_old, s_new)
    return w
