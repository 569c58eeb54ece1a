"""Benchmark of the transfer operators on one MI355X.

Default workload (the batched uneven level step): 2000 regions per grid
function with edges uniform in [8,96] fine points per axis, spread over 5
refinement levels, 24 FP64 vertex-centred grid functions x 3 time levels;
a third of the regions are copies (a component's box from a neighbour's
array, copy_3d), a third order-5 Lagrange prolongation fused with 3-level
time interpolation, a third rf2 restriction.  One step = all 48000
(region, grid function) transfers of the level step.

The batched path (one launch per phase and kernel class over the
device-resident descriptor list, replayed as a captured hipGraph) gives
`value`; the plain path (one launch per transfer, the reference's
per-region call shape) is timed beside it.  Kernel times are measured with
HIP events on the stream the work runs on; the dominant launch of the step
gives `roofline`.  `cpu_baseline` times the CPU oracle (the reference's
algorithm restated, OpenMP over the host cores) on a bounded sample of the
same transfers, in the reference's shape (one operator call per region,
time interpolation through temporaries) and fused (one parallel loop over
all regions, time interpolation in place).  `copy_ceiling` is the measured
flat-copy rate of the same GPU.  Beside the headline line: the single
512^3 patch, the level steps of a nested 3-level hierarchy and of a
5-level hierarchy tiled by the same 2000 boxes (regions from the
communication schedule), and the mixed-precision comparison.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "prolongate+restrict+sync Gpoints/s & HBM GB/s per level step, batched vs plain"


def _profile_traffic(name: str, key: str = None):
    """HBM bytes per launch measured with rocprofv3 FETCH_SIZE/WRITE_SIZE
    (profiles/<name>.json; with key, the launch of that class key in its
    "launches" list), or None."""
    p = ROOT / "profiles" / f"{name}.json"
    if p.exists():
        try:
            d = json.loads(p.read_text())
            if key is None:
                return d.get("traffic_bytes_per_launch")
            for l in d.get("launches", []):
                if l.get("key") == key:
                    return l.get("traffic_bytes_per_launch")
        except Exception:
            return None
    return None


def cpu_baseline(target_seconds: float = 12.0):
    """The oracle on a bounded sample: 1 grid function, 1 target time, a
    256 x 256 x 128 fine sub-region of the same geometry (3 time levels,
    temporaries + time interpolation, as data<T>::transfer_time does)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import numpy as np

    import pyoracle as orc

    rng = np.random.default_rng(1000)
    n = (256, 256, 128)
    # coarse box covering the sub-region's order-5 footprint (+ghosts)
    nc = tuple(v // 2 + 8 for v in n)
    sb = ((-6, -6, -6), tuple(-6 + 2 * (k - 1) for k in nc), (2, 2, 2))
    rb = ((2, 2, 2), tuple(2 + k - 1 for k in n), (1, 1, 1))
    srcs = []
    for t in (1.0, 0.0, -1.0):
        h = orc.HostArr(sb, np.ascontiguousarray(rng.uniform(-1, 1, size=(nc[2], nc[1], nc[0]))))
        srcs.append(h)
    dst = orc.HostArr(rb, np.zeros((n[2], n[1], n[0])))
    orc.set_openmp(True)
    assert orc.prolongate_time(orc.KIND_VLAG, 5, srcs, [1.0, 0.0, -1.0], dst, 1 / 3, rb, 2) == 0
    reps, t0 = 0, time.perf_counter()
    while True:
        assert orc.prolongate_time(orc.KIND_VLAG, 5, srcs, [1.0, 0.0, -1.0], dst, 1 / 3, rb, 2) == 0
        reps += 1
        el = time.perf_counter() - t0
        if el >= target_seconds or reps >= 200:
            break
    pts = n[0] * n[1] * n[2] * reps
    return {"value": pts / el / 1e9, "unit": "Gpoints/s", "cores": orc.num_threads(), "kind": "port",
            "sample": f"oracle prolongate_3d_rf2<5> + interpolate_3d_3tl via temporaries, 1 grid function, "
                      f"{n[0]}x{n[1]}x{n[2]} fine points, {reps} repetitions in {el:.1f} s"}


def copy_ceiling(n: int = 640, reps: int = 40) -> dict:
    """GB/s (bytes read + written) of flat copies on this GPU, the STREAM
    copy ceiling beside the 8 TB/s peak: the library's flat stream copy
    (cg_stream_copy_f64), the HIP runtime's device-to-device hipMemcpyAsync
    and the library's copy_3d of a whole n^3 FP64 array into another; HIP
    events on the stream they run on, over reps calls after a warmup."""
    import ctypes as C

    import carpet_amd as cg
    from carpet_amd._lib import check, lib

    box = cg.Box((0, 0, 0), (n - 1,) * 3, (1, 1, 1))
    a, b = cg.DeviceArray(box), cg.DeviceArray(box)
    nbytes = 8 * n**3
    st = cg.Stream()
    a.fill(1.0)
    b.fill(0.0)
    out = {}
    for name in ("stream_copy", "hip_memcpy", "copy_3d"):
        def once():
            if name == "stream_copy":
                check(lib.cg_stream_copy_f64(C.c_void_p(b.ptr), C.c_void_p(a.ptr), n**3, st.handle),
                      "cg_stream_copy_f64")
            elif name == "hip_memcpy":
                check(lib.cg_memcpy_dtod(C.c_void_p(b.ptr), C.c_void_p(a.ptr), nbytes, st.handle),
                      "cg_memcpy_dtod")
            else:
                cg.copy_3d(a, b, box, box, stream=st)
        for _ in range(5):
            once()
        e0, e1 = cg.Event(), cg.Event()
        e0.record(st)
        for _ in range(reps):
            once()
        e1.record(st)
        st.synchronize()
        out[name] = round(2 * nbytes / (e0.elapsed_ms(e1) / reps) / 1e6, 1)
    a.free()
    b.free()
    return out


def calib_copy(n: int = 512):
    """One copy_3d of an n^3 FP64 array (known traffic: 8 n^3 bytes read and
    written) so that a profiled run can scale its HBM counters."""
    import numpy as np

    import carpet_amd as cg

    box = cg.Box((0, 0, 0), (n - 1,) * 3, (1, 1, 1))
    a, b = cg.DeviceArray(box), cg.DeviceArray(box)
    a.fill(1.0)
    b.fill(0.0)
    cg.copy_3d(a, b, box, box)
    cg.synchronize()
    a.free()
    b.free()
    return 8 * n**3


def run_config2(steps: int, warmup: int, n_gf: int, plain_steps: int, tune_cache=None):
    import carpet_amd as cg
    from carpet_amd.workloads import Config2

    cfg = Config2(n_gf=n_gf)
    st = cg.Stream()
    src, dst = cfg.allocate(stream=st)
    st.synchronize()
    b = cfg.batch(src, dst)
    b.finalize()
    if tune_cache:
        b.autotune(st, reps=3, cache_path=tune_cache)
    stats = b.stats()
    print(b.describe(), file=sys.stderr, flush=True)
    b.capture_graph(st)
    for _ in range(warmup):
        b.launch_graph(st)
    st.synchronize()
    ev = [cg.Event() for _ in range(steps + 1)]
    cg.synchronize()
    t0 = time.perf_counter()
    ev[0].record(st)
    for k in range(steps):
        b.launch_graph(st)
        ev[k + 1].record(st)
    st.synchronize()
    wall = time.perf_counter() - t0
    per_launch = [ev[k].elapsed_ms(ev[k + 1]) for k in range(steps)]
    ms_step = sum(per_launch) / steps
    # plain path: one launch per transfer
    ms_plain = float("nan")
    if plain_steps > 0:
        for _ in range(max(1, warmup // 2)):
            b.run_plain(st)
        st.synchronize()
        e0, e1 = cg.Event(), cg.Event()
        e0.record(st)
        for _ in range(plain_steps):
            b.run_plain(st)
        e1.record(st)
        st.synchronize()
        ms_plain = e0.elapsed_ms(e1) / plain_steps
    b.close()
    for a in list(src.values()) + list(dst.values()):
        a.free()
    return cfg, stats, ms_step, wall, ms_plain


def _launch_times(b, st, steps: int):
    """Per-launch ms (HIP events on st around each launch run on its own,
    averaged over steps) with each launch's phase, points and algorithmic
    bytes."""
    import carpet_amd as cg

    n = b.launch_count()
    # class keys in launch order (cg_batch_describe lists the launches in order)
    keys = re.findall(r"^phase -?\d+ \[([^\]]+)\]", b.describe(), re.M)
    ev = [cg.Event() for _ in range(n + 1)]
    tot = [0.0] * n
    for _ in range(steps):
        ev[0].record(st)
        for i in range(n):
            b.run_launch(i, st)
            ev[i + 1].record(st)
        st.synchronize()
        for i in range(n):
            tot[i] += ev[i].elapsed_ms(ev[i + 1])
    out = []
    for i in range(n):
        ph, pts, alg = b.launch_stats(i)
        out.append({"launch": i, "phase": ph, "key": keys[i] if i < len(keys) else None,
                    "ms": round(tot[i] / steps, 4), "dst_points": pts,
                    "algorithmic_bytes": alg, "gbytes_per_s": round(alg / (tot[i] / steps) / 1e6, 1)})
    return out


def run_config4(n_regions: int, n_gf: int, steps: int, warmup: int, plain_steps: int,
                tune_cache=None, cpu_seconds: float = 0.0):
    """Batched uneven level step: n_regions regions (copies / O=5 prolongation
    with 3 time levels / rf2 restriction, edges uniform in [8,96]) per grid
    function; batched graph replay vs one launch per transfer, per-launch
    kernel times, the step advancing in time through the per-step
    parameter path, and the two paths' outputs compared bit for bit on a
    sample of destinations."""
    import numpy as np

    import carpet_amd as cg
    from carpet_amd.workloads import config4

    w = config4(n_regions=n_regions, n_gf=n_gf)
    t0 = time.perf_counter()
    dev = w.allocate(pool=True)
    pool = getattr(w, "pool_bytes", None)
    print(f"# config4: pool [{w._pool:#x}, {w._pool + pool:#x}) {pool} bytes, {len(w.arrays)} arrays, "
          f"{w.setup_launches} setup launches", file=sys.stderr, flush=True)
    st = cg.Stream()
    b = w.batch(dev)
    b.finalize()
    if tune_cache:
        b.autotune(st, reps=3, cache_path=tune_cache)
    desc = b.describe()
    print(f"# config4: {len(w.steps)} transfers, set up in {time.perf_counter() - t0:.1f} s\n" + desc,
          file=sys.stderr, flush=True)
    if tune_cache and os.path.exists(tune_cache):
        print("# tuner cache:\n" + open(tune_cache).read(), file=sys.stderr, flush=True)
    stats = b.stats()
    b.capture_graph(st)
    for _ in range(warmup):
        b.launch_graph(st)
    st.synchronize()
    e0, e1 = cg.Event(), cg.Event()
    e0.record(st)
    for _ in range(steps):
        b.launch_graph(st)
    e1.record(st)
    st.synchronize()
    ms_b = e0.elapsed_ms(e1) / steps
    launches = _launch_times(b, st, max(2, steps // 4))
    print("# config4 launches: " + json.dumps(launches), file=sys.stderr, flush=True)
    # the step advancing in time: the prolongation phase alternates between
    # the intermediate fine time (-1/2, 3 levels) and the coarse-aligned time
    # (0, exact match), new weights uploaded before each replay
    prol_phase = 1
    e0.record(st)
    for k in range(steps):
        b.set_time(prol_phase, [0.0, -1.0, -2.0], 0.0 if k % 2 else -0.5)
        b.launch_graph(st)
    e1.record(st)
    st.synchronize()
    ms_retimed = e0.elapsed_ms(e1) / steps
    b.set_time(prol_phase, [0.0, -1.0, -2.0], -0.5)
    sample = w.outputs[:: max(1, len(w.outputs) // 64)]
    same, ms_p = None, float("nan")
    if plain_steps > 0:
        b.launch_graph(st)
        st.synchronize()
        got_b = {k: dev[k].download() for k in sample}
        for k in sample:  # back to the initial state (NaN, then the seeded values)
            dev[k].fill(np.nan)
            if w.arrays[k].seed is not None:
                dev[k].fill_hash(w.arrays[k].seed)
        st.synchronize()
        b.run_plain(st)
        st.synchronize()
        same = all(np.array_equal(dev[k].download(), got_b[k], equal_nan=True) for k in sample)
        e0.record(st)
        for _ in range(plain_steps):
            b.run_plain(st)
        e1.record(st)
        st.synchronize()
        ms_p = e0.elapsed_ms(e1) / plain_steps
    b.close()
    w.free(dev)
    desc_bytes = int(desc.split("device descriptor bytes ")[1].split()[0]) if "descriptor bytes" in desc else None
    pts, alg = stats["dst_points"], stats["alg_bytes"]
    out = {"regions": n_regions, "grid_functions": n_gf, "transfers": stats["transfers"],
           "dst_points_per_step": pts, "algorithmic_bytes_per_step": alg,
           "resident_array_bytes": pool, "device_descriptor_bytes": desc_bytes,
           "batched_ms": round(ms_b, 4), "batched_gpoints_per_s": round(pts / ms_b / 1e6, 3),
           "batched_gbytes_per_s": round(alg / ms_b / 1e6, 1),
           "batched_roofline_frac": round(alg / ms_b / 1e6 / HBM_PEAK_GBS, 4),
           "launches": launches,
           "retimed_step_ms": round(ms_retimed, 4),
           "plain_ms": round(ms_p, 4), "plain_gpoints_per_s": round(pts / ms_p / 1e6, 3),
           "plain_launches_per_step": stats["transfers"], "batched_vs_plain": round(ms_p / ms_b, 3),
           "batched_equals_plain_on_sample": same, "sample_destinations": len(sample)}
    if cpu_seconds > 0:
        out["cpu"] = cpu_baseline_config4(w, cpu_seconds)
    return out


def cpu_baseline_config4(w, target_seconds: float):
    """The oracle on the first regions of the level step (one grid function,
    the three kinds in turn) until about target_seconds of CPU work per
    shape: the reference's shape (one operator call per region, each
    parallelised inside, time interpolation through temporaries) and fused
    (all regions in one parallel loop, time interpolation in place); both
    must give the same bits."""
    sys.path.insert(0, str(ROOT / "oracle"))
    sys.path.insert(0, str(ROOT / "tests"))
    import numpy as np

    import pyoracle as orc
    from workload_check import _host

    steps = [t for t in w.steps if t.name.endswith("g0")]
    # the first transfers, up to about 4e7 destination points
    pick, pts = [], 0
    for t in steps:
        if pts > 4e7:
            break
        pick.append(t)
        pts += t.dstreg.size()
    names = sorted({t.dst for t in pick} | {s for t in pick for s in t.srcs})
    from carpet_amd.workloads import hash_field
    host = {}
    for k in names:
        spec = w.arrays[k]
        e = spec.box.sizes()
        host[k] = _host(spec, hash_field(spec.seed, e) if spec.seed is not None else np.full(e[::-1], np.nan))
    tasks = orc.level_step_tasks(pick, host)
    res = {}
    outs = {}
    orc.set_openmp(True)
    for mode, key in ((0, "reference_shaped"), (1, "fused")):
        assert orc.replay(mode, tasks) == 0
        reps, t0 = 0, time.perf_counter()
        while True:
            assert orc.replay(mode, tasks) == 0
            reps += 1
            el = time.perf_counter() - t0
            if el >= target_seconds / 2 or reps >= 500:
                break
        outs[key] = {t.dst: host[t.dst].data.copy() for t in pick}
        res[key] = {"value": pts * reps / el / 1e9, "seconds": round(el, 2), "repetitions": reps}
    same = all(np.array_equal(outs["reference_shaped"][k], outs["fused"][k], equal_nan=True)
               for k in outs["fused"])
    kinds = {"copy": 0, "prolongation": 0, "restriction": 0}
    for t in pick:
        kinds["copy" if t.phase == 2 else "prolongation" if t.phase == 1 else "restriction"] += 1
    return {"value": res["reference_shaped"]["value"], "unit": "Gpoints/s", "cores": orc.num_threads(),
            "kind": "port",
            "sample": f"oracle replay of the first {len(pick)} transfers of grid function 0 ({kinds}), "
                      f"{pts} destination points, one operator call per region as the reference does "
                      f"(OpenMP inside each call, time interpolation through temporaries); "
                      f"{res['reference_shaped']['repetitions']} repetitions in "
                      f"{res['reference_shaped']['seconds']} s",
            "best_cpu_fused": {"value": res["fused"]["value"], "unit": "Gpoints/s",
                               "repetitions": res["fused"]["repetitions"], "seconds": res["fused"]["seconds"],
                               "same_bits_as_reference_shaped": bool(same)}}


def run_config6(n_gf: int, steps: int, warmup: int, plain_steps: int, ncoarse: int = 128, tune_cache=None):
    """Level step of a nested 3-level hierarchy (ncoarse^3 coarse points,
    each finer level the central half of the previous, 8 components per
    level, ghost 3, buffer 6, order-5 vertex Lagrange with 3-level time
    interpolation) built from its communication schedule: restriction,
    boundary prolongation and sync lists of dh, batched (one launch per
    phase and class, hipGraph) vs one cg_transfer_from per region."""
    import numpy as np

    import carpet_amd as cg
    from carpet_amd.workloads import hierarchy_step, nested_levels

    t0 = time.perf_counter()
    lv = nested_levels(ncoarse=ncoarse, nlevels=3, split=(2, 2, 2), centering=cg.VERTEX_CENTERED, buffer=6,
                       order=5)
    w = hierarchy_step(lv, cg.VERTEX_CENTERED, n_gf=n_gf)
    dev = w.allocate(pool=True)
    st = cg.Stream()
    b = w.batch(dev)
    b.finalize()
    if tune_cache:
        b.autotune(st, reps=3, cache_path=tune_cache)
    desc = b.describe()
    print(f"# config6: {w.name}: {len(w.steps)} transfers, {len(w.arrays)} arrays, set up in "
          f"{time.perf_counter() - t0:.1f} s\n" + desc, file=sys.stderr, flush=True)
    stats = b.stats()
    ms_b = _time_graph(b, st, steps, warmup)
    launches = _launch_times(b, st, 2)
    print("# config6 launches: " + json.dumps(launches), file=sys.stderr, flush=True)

    def reseed():
        # later phases read what earlier ones write (restricted coarse points
        # feed the next level's prolongation), so each comparison starts
        # from the seeded state and runs one step
        for k in w.outputs:
            dev[k].fill(np.nan, stream=st)
            dev[k].fill_hash(w.arrays[k].seed, stream=st)

    sample = w.outputs[:: max(1, len(w.outputs) // 48)]
    same, ms_p = None, float("nan")
    if plain_steps > 0:
        reseed()
        b.run(st)
        st.synchronize()
        got_b = {k: dev[k].download() for k in sample}
        reseed()
        w.run_steps(dev, st)
        st.synchronize()
        same = all(np.array_equal(dev[k].download(), got_b[k], equal_nan=True) for k in sample)
        e0, e1 = cg.Event(), cg.Event()
        e0.record(st)
        for _ in range(plain_steps):
            w.run_steps(dev, st)
        e1.record(st)
        st.synchronize()
        ms_p = e0.elapsed_ms(e1) / plain_steps
    b.close()
    w.free(dev)
    pts, alg = stats["dst_points"], stats["alg_bytes"]
    return {"transfers": stats["transfers"], "dst_points_per_step": pts, "algorithmic_bytes_per_step": alg,
            "batched_ms": round(ms_b, 4), "batched_gpoints_per_s": round(pts / ms_b / 1e6, 3),
            "batched_gbytes_per_s": round(alg / ms_b / 1e6, 1),
            "batched_roofline_frac": round(alg / ms_b / 1e6 / HBM_PEAK_GBS, 4), "launches": launches,
            "plain_ms": round(ms_p, 4), "plain_gpoints_per_s": round(pts / ms_p / 1e6, 3),
            "batched_vs_plain": round(ms_p / ms_b, 3), "batched_equals_plain_on_sample": same,
            "sample_destinations": len(sample)}


def run_config7(n_gf: int, steps: int, warmup: int, plain_steps: int, tune_cache=None):
    """Level step of a 5-level vertex-centred hierarchy of 1995 boxes with
    edges uniform in [8,96] points per axis (each level tiled by its boxes,
    nested inside the coarser level), ghost width 3, buffer width 6, order
    5 with 3 time levels: dh's lists from the library's communication
    schedule -- restriction, boundary prolongation (3-level time
    interpolation to the intermediate fine time) and sync of every level --
    batched (one launch per phase and kernel class, hipGraph) vs one
    cg_transfer_from per region, with the step's outputs compared on a
    sample."""
    import numpy as np

    import carpet_amd as cg
    from carpet_amd.workloads import hierarchy_step, tiled_levels

    t0 = time.perf_counter()
    lv = tiled_levels()
    w = hierarchy_step(lv, cg.VERTEX_CENTERED, n_gf=n_gf, only_used=True)
    t1 = time.perf_counter()
    print(f"# config7: schedule and {len(w.steps)} transfers in {t1 - t0:.1f} s", file=sys.stderr, flush=True)
    dev = w.allocate(pool=True)
    print(f"# config7: {len(w.arrays)} arrays allocated and seeded", file=sys.stderr, flush=True)
    st = cg.Stream()
    b = w.batch(dev)
    t2 = time.perf_counter()
    print(f"# config7: batch built in {t2 - t1:.1f} s", file=sys.stderr, flush=True)
    b.finalize()
    print(f"# config7: finalized in {time.perf_counter() - t2:.1f} s", file=sys.stderr, flush=True)
    if tune_cache:
        b.autotune(st, reps=3, cache_path=tune_cache)
    desc = b.describe()
    print(f"# config7: {w.name}: {len(w.steps)} transfers, {len(w.arrays)} arrays, pool {w.pool_bytes} bytes; "
          f"schedule and steps {t1 - t0:.1f} s, allocation and batch {t2 - t1:.1f} s, finalize and tuning "
          f"{time.perf_counter() - t2:.1f} s\n" + desc, file=sys.stderr, flush=True)
    stats = b.stats()
    ms_b = _time_graph(b, st, steps, warmup)
    launches = _launch_times(b, st, 2)
    print("# config7 launches: " + json.dumps(launches), file=sys.stderr, flush=True)

    def reseed():
        # every output back to its seeded values (one launch; transfers write
        # only inside the extents, which this overwrites)
        st.synchronize()
        cg.fill_hash_many([(dev[k], w.arrays[k].seed) for k in w.outputs], stream=st)

    sample = w.outputs[:: max(1, len(w.outputs) // 64)]
    same, ms_p = None, float("nan")
    if plain_steps > 0:
        reseed()
        b.run(st)
        st.synchronize()
        got_b = {k: dev[k].download() for k in sample}
        reseed()
        w.run_steps(dev, st)
        st.synchronize()
        print("# config7: plain step done", file=sys.stderr, flush=True)
        same = all(np.array_equal(dev[k].download(), got_b[k], equal_nan=True) for k in sample)
        e0, e1 = cg.Event(), cg.Event()
        e0.record(st)
        for _ in range(plain_steps):
            w.run_steps(dev, st)
        e1.record(st)
        st.synchronize()
        ms_p = e0.elapsed_ms(e1) / plain_steps
    b.close()
    w.free(dev)
    pts, alg = stats["dst_points"], stats["alg_bytes"]
    return {"boxes": sum(len(L.components) for L in lv), "levels": len(lv), "grid_functions": n_gf,
            "transfers": stats["transfers"], "dst_points_per_step": pts, "algorithmic_bytes_per_step": alg,
            "resident_array_bytes": w.pool_bytes,
            "batched_ms": round(ms_b, 4), "batched_gpoints_per_s": round(pts / ms_b / 1e6, 3),
            "batched_gbytes_per_s": round(alg / ms_b / 1e6, 1),
            "batched_roofline_frac": round(alg / ms_b / 1e6 / HBM_PEAK_GBS, 4), "launches": launches,
            "plain_ms": round(ms_p, 4), "batched_vs_plain": round(ms_p / ms_b, 3),
            "batched_equals_plain_on_sample": same, "sample_destinations": len(sample)}


def _time_graph(b, st, steps: int, warmup: int) -> float:
    """ms per replay of b's captured graph (HIP events on st)."""
    import carpet_amd as cg

    b.capture_graph(st)
    for _ in range(warmup):
        b.launch_graph(st)
    st.synchronize()
    e0, e1 = cg.Event(), cg.Event()
    e0.record(st)
    for _ in range(steps):
        b.launch_graph(st)
    e1.record(st)
    st.synchronize()
    return e0.elapsed_ms(e1) / steps


def run_config3(steps: int, warmup: int, n_var: int):
    """Cell-centred conservative set: per operator (cc TVD, cc ENO<3>,
    cc ENO<2>) the batched prolongation of n_var variables 192^3 -> 384^3,
    then restrict_3d_cc_rf2 back; conservation of every coarse cell."""
    import numpy as np

    import carpet_amd as cg
    from carpet_amd.workloads import Config3

    cfg = Config3(n_var=n_var)
    cbox, owned, fbox, reg = cfg.boxes()
    st = cg.Stream()
    coarse, fine, back = cfg.allocate(stream=st)
    st.synchronize()
    out = {}
    for name, op, order_space in (("cc_tvd", cg.OP_TVD, 1), ("cc_eno3", cg.OP_ENO, 5), ("cc_eno2", cg.OP_ENO, 3)):
        bp = cfg.prolong_batch(coarse, fine, op, order_space)
        bp.finalize()
        br = cfg.restrict_batch(fine, back)
        br.finalize()
        ms_p = _time_graph(bp, st, steps, warmup)
        ms_r = _time_graph(br, st, steps, warmup)
        sp, sr = bp.stats(), br.stats()
        worst = 0.0
        for v in range(n_var):
            o = [(l - dl) // s for l, dl, s in zip(owned.lower, cbox.lower, cbox.stride)]
            e = owned.sizes()
            sl = np.s_[o[2]:o[2] + e[2], o[1]:o[1] + e[1], o[0]:o[0] + e[0]]
            r = back[v].download()[sl]
            orig = cfg.coarse_data(v)[sl]
            worst = max(worst, float(np.max(np.abs(r - orig)) / np.max(np.abs(orig))))
        out[name] = {"prolong_ms": round(ms_p, 4), "prolong_gpoints_per_s": round(sp["dst_points"] / ms_p / 1e6, 3),
                     "prolong_gbytes_per_s": round(sp["alg_bytes"] / ms_p / 1e6, 1),
                     "restrict_ms": round(ms_r, 4), "restrict_gbytes_per_s": round(sr["alg_bytes"] / ms_r / 1e6, 1),
                     "step_ms": round(ms_p + ms_r, 4),
                     "step_gpoints_per_s": round((sp["dst_points"] + sr["dst_points"]) / (ms_p + ms_r) / 1e6, 3),
                     "step_gbytes_per_s": round((sp["alg_bytes"] + sr["alg_bytes"]) / (ms_p + ms_r) / 1e6, 1),
                     "conservation_max_rel": worst}
        print(f"# config3 {name}: {out[name]}", file=sys.stderr, flush=True)
        bp.close()
        br.close()
    for a in coarse + fine + back:
        a.free()
    return out


def run_config5(steps: int, warmup: int, tune_cache=None):
    """Mixed precision vs batched FP64 on the config-1 and config-3 shapes
    (vertex order 5 + 3 time levels), for both mixed modes (FP64 arithmetic
    on FP32 storage; compensated FP32 sums): times, speedups, and the
    normwise max / RMS relative error of the mixed result against FP64 (a) on
    the original FP64 sources (storage + arithmetic) and (b) on the
    FP32-rounded sources (arithmetic and the final rounding only)."""
    import numpy as np

    import carpet_amd as cg
    from carpet_amd.workloads import Config5

    out = {}
    for shape in ("cfg1", "cfg3"):
        cfg = Config5(shape=shape)
        _, _, fbox, reg = cfg.boxes()
        st = cg.Stream()
        src64, src32, src64r, dst64, dst32 = cfg.allocate(stream=st)
        st.synchronize()
        o = [(l - dl) // s for l, dl, s in zip(reg.lower, fbox.lower, fbox.stride)]
        e = reg.sizes()
        sl = np.s_[o[2]:o[2] + e[2], o[1]:o[1] + e[1], o[0]:o[0] + e[0]]
        b64 = cfg.batch(src64, dst64, cg.PREC_F64)
        b64.finalize()
        if tune_cache:
            b64.autotune(st, reps=3, cache_path=tune_cache)
            print(b64.describe(), file=sys.stderr, flush=True)
        ms64 = _time_graph(b64, st, steps, warmup)
        s64 = b64.stats()
        ref = [dst64[v].download()[sl] for v in range(cfg.n_var)]
        br = cfg.batch(src64r, dst64, cg.PREC_F64)
        br.finalize()
        br.run(st)
        st.synchronize()
        ref_r = [dst64[v].download()[sl] for v in range(cfg.n_var)]
        res = {"fine_points": s64["dst_points"], "fp64_ms": round(ms64, 4),
               "fp64_gbytes_per_s": round(s64["alg_bytes"] / ms64 / 1e6, 1)}
        for name, prec in (("mixed_fp64_arith", cg.PREC_MIXED), ("mixed_dot2", cg.PREC_MIXED_DOT2)):
            bm = cfg.batch(src32, dst32, prec)
            bm.finalize()
            if tune_cache:
                bm.autotune(st, reps=3, cache_path=tune_cache)
            print(bm.describe(), file=sys.stderr, flush=True)
            ms = _time_graph(bm, st, steps, warmup)
            sm = bm.stats()
            got = [dst32[v].download()[sl] for v in range(cfg.n_var)]
            full = [cg.relative_errors(g, r) for g, r in zip(got, ref)]
            arith = [cg.relative_errors(g, r) for g, r in zip(got, ref_r)]
            res[name] = {"ms": round(ms, 4), "speedup_vs_fp64": round(ms64 / ms, 3),
                         "gbytes_per_s": round(sm["alg_bytes"] / ms / 1e6, 1),
                         "vs_fp64_inputs": {"max_rel": max(x[0] for x in full), "rms_rel": max(x[1] for x in full)},
                         "vs_fp64_on_fp32_inputs": {"max_rel": max(x[0] for x in arith),
                                                    "rms_rel": max(x[1] for x in arith)}}
            bm.close()
        out[shape] = res
        print(f"# config5 {shape}: {res}", file=sys.stderr, flush=True)
        for b in (b64, br):
            b.close()
        for a in [x for v in src64 + src32 + src64r for x in v] + dst64 + dst32:
            a.free()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n-gf", type=int, default=24)
    ap.add_argument("--plain-steps", type=int, default=5,
                    help="steps of the plain path (one launch per transfer); 0 skips it")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-config2", action="store_true",
                    help="skip the single-patch prolongation reported beside the headline line")
    ap.add_argument("--no-config6", action="store_true",
                    help="skip the hierarchy level step reported beside the headline line")
    ap.add_argument("--no-config7", action="store_true",
                    help="skip the 2000-box hierarchy level step reported beside the headline line")
    ap.add_argument("--no-config5", action="store_true",
                    help="skip the mixed-precision comparison reported beside the headline line")
    ap.add_argument("--cpu-seconds", type=float, default=16.0)
    ap.add_argument("--config", type=int, default=4, choices=(2, 3, 4, 5, 6, 7),
                    help="4: batched uneven level step (default, the headline line); 2: single large patch; "
                         "3: cell-centred conservative set; 5: mixed precision vs FP64; 6: level step of a "
                         "nested hierarchy from its communication schedule; 7: level step of a 5-level "
                         "hierarchy of 1995 boxes with edges U[8,96] from its communication schedule")
    ap.add_argument("--n-var", type=int, default=8, help="config 3 variables")
    ap.add_argument("--n-regions", type=int, default=2000)
    ap.add_argument("--tune-cache", default=None,
                    help="batch tuner cache file: classes it names take its parameters, the others are "
                         "searched (cg_batch_autotune) before the timed steps; default: a fresh file, so "
                         "every class is searched")
    ap.add_argument("--no-tune", action="store_true",
                    help="time the default launch settings of cg_batch_finalize (no tuner search)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run one 1 GiB copy_3d first (HBM counter calibration under rocprofv3)")
    args = ap.parse_args(argv)
    explicit_cache = args.tune_cache is not None
    if args.no_tune or (args.config == 7 and not explicit_cache):
        # the 2000-box hierarchy (1.27 M transfers) is timed with the default
        # launch settings unless a tuner cache is named: every tuner trial
        # rebuilds its whole descriptor list
        args.tune_cache = None
    elif args.tune_cache is None:
        import tempfile
        fd, args.tune_cache = tempfile.mkstemp(prefix="carpet_amd_tune_", suffix=".txt")
        os.close(fd)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the transfer operators of a level step run on a single "
              "MI355X here (sharding across GPUs is a design note, not a measured path); "
              "run with --gpus 1", file=sys.stderr)
        return 2
    if args.calib_copy:
        nbytes = calib_copy()
        print(f"# calibration copy: {nbytes} bytes read and written", file=sys.stderr, flush=True)
    if args.config == 3:
        res = run_config3(args.steps, args.warmup, args.n_var)
        tvd = res["cc_tvd"]
        line = {"metric": METRIC, "value": tvd["step_gpoints_per_s"], "unit": "Gpoints/s", "n_gpus": 1,
                "steps": args.steps, "warmup": args.warmup, "ms_per_step": tvd["step_ms"],
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
                "data": "synthetic",
                "config": {"workload": f"cell-centred conservative set: {args.n_var} FP64 variables, coarse 192^3 "
                                       "(+3 ghosts) -> fine 384^3 by prolongate_3d_cc_tvd / cc_eno, then "
                                       "restrict_3d_cc_rf2 back; value = TVD prolong + restrict step",
                           "operators": res},
                "gbytes_per_s": tvd["step_gbytes_per_s"],
                # dominant kernel: the batched TVD prolongation
                "roofline": {"bound": "hbm", "achieved": tvd["prolong_gbytes_per_s"], "peak": HBM_PEAK_GBS,
                             "unit": "GB/s", "frac": round(tvd["prolong_gbytes_per_s"] / HBM_PEAK_GBS, 4),
                             "traffic": _profile_traffic("k_grid_point_tvd_config3")}}
        print(json.dumps(line))
        return 0
    if args.config == 5:
        res = run_config5(args.steps, args.warmup, tune_cache=args.tune_cache)
        c3 = res["cfg3"]
        mx = c3["mixed_fp64_arith"]
        pts = c3["fine_points"]
        line = {"metric": METRIC, "value": round(pts / mx["ms"] / 1e6, 3), "unit": "Gpoints/s",
                "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": mx["ms"],
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
                "data": "synthetic",
                "config": {"workload": "mixed precision: vertex prolongate_3d_rf2<5> + interpolate_3d_3tl at "
                                       "t=1/3, batched; sources and destination stored in FP32, time "
                                       "interpolation in FP64; stencil sums in FP64 with the time levels "
                                       "combined before the spatial passes (mixed_fp64_arith, the value) or "
                                       "compensated FP32 per level (mixed_dot2); config-1 shape (33^3 -> "
                                       "64^3, 1 variable) and config-3 shape (193^3 -> 384^3, 8 variables); "
                                       "value = config-3 shape",
                           "destination_precision": "fp32", "shapes": res},
                "speedup_vs_batched_fp64": mx["speedup_vs_fp64"],
                "roofline": {"bound": "hbm", "achieved": mx["gbytes_per_s"], "peak": HBM_PEAK_GBS,
                             "unit": "GB/s", "frac": round(mx["gbytes_per_s"] / HBM_PEAK_GBS, 4),
                             "traffic": _profile_traffic("k_persist_tensor_vlag5_mixed_config3_shape")}}
        print(json.dumps(line))
        return 0
    if args.config == 6:
        r = run_config6(args.n_gf, args.steps, args.warmup, args.plain_steps, tune_cache=args.tune_cache)
        line = {"metric": METRIC, "value": r["batched_gpoints_per_s"], "unit": "Gpoints/s",
                "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": r["batched_ms"],
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
                "data": "synthetic",
                "config": {"workload": f"nested 3-level vertex hierarchy, 128^3 coarse points, 8 components per "
                                       f"level, {args.n_gf} grid functions x 3 time levels; level step from the "
                                       "communication schedule (restriction, boundary prolongation O5 + "
                                       "3-level time interpolation, sync)", **r},
                "roofline": {"bound": "hbm", "achieved": r["batched_gbytes_per_s"], "peak": HBM_PEAK_GBS,
                             "unit": "GB/s", "frac": r["batched_roofline_frac"], "traffic": None}}
        print(json.dumps(line))
        return 0
    if args.config == 7:
        r = run_config7(args.n_gf, args.steps, args.warmup, args.plain_steps, tune_cache=args.tune_cache)
        dom = max(r["launches"], key=lambda x: x["ms"])
        line = {"metric": METRIC, "value": r["batched_gpoints_per_s"], "unit": "Gpoints/s",
                "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": r["batched_ms"],
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
                "data": "synthetic",
                "config": {"workload": f"level step of a 5-level vertex hierarchy of {r['boxes']} boxes (edges "
                                       f"U[8,96] per axis, ghost 3, buffer 6), {args.n_gf} grid functions x 3 "
                                       "time levels; dh's restriction, boundary prolongation (O5 + 3-level time "
                                       "interpolation) and sync lists from the communication schedule", **r},
                "roofline": {"bound": "hbm", "achieved": dom["gbytes_per_s"], "peak": HBM_PEAK_GBS,
                             "unit": "GB/s", "frac": round(dom["gbytes_per_s"] / HBM_PEAK_GBS, 4),
                             "traffic": None, "kernel": f"launch {dom['launch']} (phase {dom['phase']})",
                             "ms": dom["ms"]}}
        print(json.dumps(line))
        return 0
    if args.config == 2:
        print(f"# config2: {args.n_gf} grid functions, {args.steps} steps", file=sys.stderr, flush=True)
        line = config2_line(args)
        print(json.dumps(line))
        return 0
    # default: the batched uneven level step (2000 regions x 24 grid functions)
    r = run_config4(args.n_regions, args.n_gf, args.steps, args.warmup, args.plain_steps, args.tune_cache,
                    cpu_seconds=0.0 if args.no_cpu else args.cpu_seconds)
    dom = max(r["launches"], key=lambda x: x["ms"])
    cpu = r.pop("cpu", None)
    line = {"metric": METRIC, "value": r["batched_gpoints_per_s"], "unit": "Gpoints/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": r["batched_ms"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": f"batched uneven level step: {args.n_regions} regions x {args.n_gf} grid "
                                   "functions (edges U[8,96] over 5 refinement levels; whole-box copies, O=5 "
                                   "prolongation + 3-level time interpolation, rf2 restriction), one launch "
                                   "per phase and kernel class + hipGraph", **r},
            "gbytes_per_s": r["batched_gbytes_per_s"],
            "plain": {"ms_per_step": r["plain_ms"], "gpoints_per_s": r["plain_gpoints_per_s"],
                      "launches_per_step": r["plain_launches_per_step"]},
            "batched_vs_plain": r["batched_vs_plain"],
            "step_roofline_frac": r["batched_roofline_frac"],
            # the dominant launch of the step (HIP events around each launch)
            "roofline": {"bound": "hbm", "achieved": dom["gbytes_per_s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(dom["gbytes_per_s"] / HBM_PEAK_GBS, 4),
                         "traffic": _profile_traffic("level_step_2000_regions_x24", dom["key"]),
                         "kernel": f"launch {dom['launch']} (phase {dom['phase']})", "ms": dom["ms"],
                         "algorithmic_bytes": dom["algorithmic_bytes"]}}
    # the measured flat-copy ceiling of this GPU beside the 8 TB/s peak
    cc = copy_ceiling()
    ceil = max(cc.values())
    line["copy_ceiling"] = {"gbytes_per_s": ceil, **cc,
                            "step_frac": round(r["batched_gbytes_per_s"] / ceil, 4),
                            "dominant_launch_frac": round(dom["gbytes_per_s"] / ceil, 4),
                            "how": "the fastest of the library's flat stream copy (cg_stream_copy_f64), the HIP "
                                   "runtime's device-to-device hipMemcpyAsync and copy_3d, each of a whole "
                                   "640^3 FP64 array into another (16 B per point), HIP events over 40 calls"}
    print(f"# copy ceiling: {cc} GB/s", file=sys.stderr, flush=True)
    if not args.no_config2:
        print("# config2 beside the headline line", file=sys.stderr, flush=True)
        line["single_patch_512"] = config2_line(args, side=True)
    if not args.no_config6:
        print("# config6 beside the headline line", file=sys.stderr, flush=True)
        line["hierarchy_step_3_levels"] = run_config6(args.n_gf, 10, 2, 1, tune_cache=args.tune_cache)
    if not args.no_config7:
        # the same 2000 boxes as the regions of a real level step (dh's lists)
        print("# config7 beside the headline line", file=sys.stderr, flush=True)
        line["hierarchy_step_2000_boxes"] = run_config7(args.n_gf, 5, 1, 1, tune_cache=args.tune_cache)
    if not args.no_config5:
        # FP32 storage (FP64 arithmetic, and compensated FP32 sums) against
        # batched FP64 on the 64^3 and 8 x 384^3 prolongation shapes
        print("# config5 beside the headline line", file=sys.stderr, flush=True)
        line["mixed_precision"] = run_config5(10, 2, tune_cache=args.tune_cache)
    line["tuned"] = args.tune_cache is not None
    if cpu is not None:
        line["cpu_baseline"] = cpu
    print(json.dumps(line))
    return 0


def config2_line(args, side: bool = False):
    cfg, stats, ms_step, wall, ms_plain = run_config2(args.steps, args.warmup, args.n_gf, args.plain_steps,
                                                      args.tune_cache)
    print(f"# config2 timed: {ms_step:.3f} ms/step batched, {ms_plain:.3f} ms/step plain", file=sys.stderr,
          flush=True)
    pts = cfg.dst_points()
    alg = cfg.alg_bytes()
    gpts = pts / (ms_step * 1e-3) / 1e9
    gbs = alg / (ms_step * 1e-3) / 1e9
    out = {"value": round(gpts, 3), "unit": "Gpoints/s", "ms_per_step": round(ms_step, 4),
           "config": {"workload": "single large patch: 24 FP64 vertex-centred grid functions, coarse 257^3 "
                                  "(+3 ghosts, 3 time levels t=1,0,-1) -> fine 512^3, prolongate_3d_rf2<5> fused "
                                  "with interpolate_3d_3tl at t=1/3 and 2/3",
                      "grid_functions": cfg.n_gf, "fine_points_per_step": pts, "algorithmic_bytes_per_step": alg},
           "gbytes_per_s": round(gbs, 1),
           "plain": {"ms_per_step": round(ms_plain, 4), "gpoints_per_s": round(pts / (ms_plain * 1e-3) / 1e9, 3),
                     "launches_per_step": stats["transfers"]},
           "batched_vs_plain": round(ms_plain / ms_step, 3),
           "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(gbs / HBM_PEAK_GBS, 4),
                        "traffic": _profile_traffic("k_persist_tensor_vlag5_3tl_config2")}}
    if side:
        return out
    out.update({"metric": METRIC, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
                "data": "synthetic"})
    out["config"]["path"] = ("one batched launch (resident workgroups taking 32x8-base x 128-plane tile "
                             "columns from per-XCD counters), hipGraph replay")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    return out


if __name__ == "__main__":
    sys.exit(main())
