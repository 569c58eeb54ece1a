"""Benchmark of the transfer operators on one MI355X.

Default workload (the single-large-patch configuration): 24 FP64
vertex-centred grid functions, coarse 257^3 (+3 ghosts) with 3 time levels
at t = 1, 0, -1, prolongated by prolongate_3d_rf2<5> fused with
interpolate_3d_3tl onto the fine region [2,513]^3 at t = 1/3 and 2/3.
One step = all 48 (grid function, target time) transfers of the level step.

The batched path (one launch over the device descriptor list, a workgroup
per tile column, replayed as a captured hipGraph) gives `value`; the plain
path (one launch per transfer, the reference's per-region call shape) is
timed beside it.
Kernel time is measured with HIP events on the stream the work runs on.
`cpu_baseline` times the CPU oracle (the reference's algorithm restated,
prolongating every time level into a temporary and interpolating in time,
OpenMP over the host cores) on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "prolongate+restrict+sync Gpoints/s & HBM GB/s per level step, batched vs plain"


def _profile_traffic(name: str):
    """HBM bytes per launch measured with rocprofv3 FETCH_SIZE/WRITE_SIZE
    (profiles/<name>.json), or None."""
    p = ROOT / "profiles" / f"{name}.json"
    if p.exists():
        try:
            return json.loads(p.read_text()).get("traffic_bytes_per_launch")
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


def run_config4(n_regions: int, n_gf: int, steps: int, warmup: int, plain_steps: int,
                tune_cache=None):
    """Batched uneven hierarchy: n_regions regions (copy sync / O=5 prolongation
    with 3 time levels / rf2 restriction, edges uniform in [8,96]) per grid
    function; batched graph replay vs one launch per transfer."""
    import carpet_amd as cg
    from carpet_amd.workloads import config4

    w = config4(n_regions=n_regions, n_gf=n_gf)
    t0 = time.perf_counter()
    dev = w.allocate(pool=True)
    st = cg.Stream()
    b = w.batch(dev)
    b.finalize()
    if tune_cache:
        b.autotune(st, reps=3, cache_path=tune_cache)
    print(f"# config4: {len(w.steps)} transfers, {len(w.arrays)} arrays, set up in "
          f"{time.perf_counter() - t0:.1f} s\n" + b.describe(), file=sys.stderr, flush=True)
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
    b.run_plain(st)
    st.synchronize()
    e0.record(st)
    for _ in range(plain_steps):
        b.run_plain(st)
    e1.record(st)
    st.synchronize()
    ms_p = e0.elapsed_ms(e1) / plain_steps
    b.close()
    w.free(dev)
    return stats, ms_b, ms_p


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n-gf", type=int, default=24)
    ap.add_argument("--plain-steps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--config", type=int, default=2, choices=(2, 4),
                    help="2: single large patch (default, the headline line); 4: batched uneven hierarchy")
    ap.add_argument("--n-regions", type=int, default=2000)
    ap.add_argument("--tune-cache", default=None,
                    help="batch tuner cache file: classes it names take its parameters, the others are "
                         "searched (cg_batch_autotune) before the timed steps")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run one 1 GiB copy_3d first (HBM counter calibration under rocprofv3)")
    args = ap.parse_args(argv)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the transfer operators of a level step run on a single "
              "MI355X here (sharding across GPUs is a design note, not a measured path); "
              "run with --gpus 1", file=sys.stderr)
        return 2
    if args.config == 4:
        n_gf = args.n_gf
        stats, ms_b, ms_p = run_config4(args.n_regions, n_gf, args.steps, args.warmup, args.plain_steps,
                                        args.tune_cache)
        pts, alg = stats["dst_points"], stats["alg_bytes"]
        line = {"metric": METRIC, "value": round(pts / (ms_b * 1e-3) / 1e9, 3), "unit": "Gpoints/s",
                "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_b, 4),
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
                "data": "synthetic",
                "config": {"workload": f"batched uneven level step: {args.n_regions} regions x {n_gf} grid "
                                       "functions (edges U[8,96]; copy sync, O=5 prolongation + 3 time levels, "
                                       "rf2 restriction), one batched launch per phase + hipGraph",
                           "transfers": stats["transfers"], "dst_points_per_step": pts,
                           "algorithmic_bytes_per_step": alg},
                "gbytes_per_s": round(alg / (ms_b * 1e-3) / 1e9, 1),
                "plain": {"ms_per_step": round(ms_p, 4), "gpoints_per_s": round(pts / (ms_p * 1e-3) / 1e9, 3),
                          "launches_per_step": stats["transfers"]},
                "batched_vs_plain": round(ms_p / ms_b, 3),
                "roofline": {"bound": "hbm", "achieved": round(alg / (ms_b * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                             "unit": "GB/s", "frac": round(alg / (ms_b * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                             "traffic": None}}
        print(json.dumps(line))
        return 0
    if args.calib_copy:
        nbytes = calib_copy()
        print(f"# calibration copy: {nbytes} bytes read and written", file=sys.stderr, flush=True)
    print(f"# config2: {args.n_gf} grid functions, {args.steps} steps", file=sys.stderr, flush=True)
    cfg, stats, ms_step, wall, ms_plain = run_config2(args.steps, args.warmup, args.n_gf, args.plain_steps,
                                                      args.tune_cache)
    print(f"# timed: {ms_step:.3f} ms/step batched, {ms_plain:.3f} ms/step plain", file=sys.stderr, flush=True)
    pts = cfg.dst_points()
    alg = cfg.alg_bytes()
    gpts = pts / (ms_step * 1e-3) / 1e9
    gbs = alg / (ms_step * 1e-3) / 1e9
    line = {
        "metric": METRIC,
        "value": round(gpts, 3),
        "unit": "Gpoints/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "single large patch: 24 FP64 vertex-centred grid functions, coarse 257^3 "
                               "(+3 ghosts, 3 time levels t=1,0,-1) -> fine 512^3, prolongate_3d_rf2<5> fused "
                               "with interpolate_3d_3tl at t=1/3 and 2/3",
                   "grid_functions": cfg.n_gf, "fine_points_per_step": pts,
                   "algorithmic_bytes_per_step": alg,
                   "path": "one batched launch (a workgroup per 32x8-base tile column), hipGraph replay"},
        "gbytes_per_s": round(gbs, 1),
        "plain": {"ms_per_step": round(ms_plain, 4), "gpoints_per_s": round(pts / (ms_plain * 1e-3) / 1e9, 3),
                  "launches_per_step": stats["transfers"]},
        "batched_vs_plain": round(ms_plain / ms_step, 3),
        "wall_s_timed": round(wall, 3),
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4),
                     "traffic": _profile_traffic("k_grid_tensor_vlag5_3tl_config2")},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
