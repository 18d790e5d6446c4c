"""Benchmark of the MI355X path-tracing hot path.

Default workload: the heterogeneous 100k-primitive field (60 % spheres, 30 %
boxes, 10 % affine), 1920x1080, 64 spp, 8 bounces, Russian roulette after
bounce 3, rendered by the per-pixel megakernel on one GPU.  A step is one full
frame (all 64 samples) with the scene resident in HBM and the frame left in
HBM.  Prints one JSON line.

Roofline: the dominant kernel is the render kernel; its algorithmic bytes per
frame come from the exact per-frame traversal counters of an instrumented run
of the same frames (same seeds, bit-identical paths), priced per path segment
as  24 N_aabb + 16 (N_internal + N_leaf) + 4 N_refs + 20 N_sphere + 28 N_box
+ 88 (N_affine) + 104 N_plane + 52 N_segments + 12 bytes per pixel, divided by
the kernel time measured with HIP events on the tracer's stream.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "master-rad_amd"))
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

import mrad  # noqa: E402
from mrad import configs  # noqa: E402

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(c: dict, pixels: int) -> float:
    n_refs = c["primitive_tests"] - c["plane_tests"]
    return (24.0 * c["aabb_tests"] + 16.0 * (c["internal_node_visits"] + c["leaf_node_visits"]) + 4.0 * n_refs
            + 20.0 * c["sphere_tests"] + 28.0 * c["box_tests"]
            + 88.0 * (c["affine_sphere_tests"] + c["affine_box_tests"]) + 104.0 * c["plane_tests"]
            + 52.0 * c["rays"] + 12.0 * pixels)


def load_traffic(workload: str):
    p = ROOT / "profiles" / f"traffic_{workload}.json"
    if not p.exists():
        return None
    try:
        d = json.loads(p.read_text())
        return float(d["hbm_bytes_per_launch"])
    except Exception:
        return None


def cpu_baseline(w: configs.Workload, seconds_hint: float = 15.0) -> dict:
    """FP64 CPU restatement of the reference CPU renderer, all host threads (max 16)."""
    from oracle import oracle as O
    threads = max(1, min(16, os.cpu_count() or 1))
    osc = O.Scene(w.scene.wire())
    cam = w.camera.packed()
    # calibrate on a small band, then size the sample to ~seconds_hint
    rows = 16
    t0 = time.perf_counter()
    O.render64(osc, cam, w.width, rows, w.max_depth, 1, 1, rr_start=w.rr_start, threads=threads)
    dt = max(1e-3, time.perf_counter() - t0)
    rate = w.width * rows / dt
    spp = 1
    height = int(min(w.height, max(rows, rate * seconds_hint / w.width)))
    t0 = time.perf_counter()
    O.render64(osc, cam, w.width, height, w.max_depth, spp, 2, rr_start=w.rr_start, threads=threads)
    dt = time.perf_counter() - t0
    paths = w.width * height * spp
    return {"value": paths / dt / 1e6, "unit": "Mpaths/s", "cores": threads, "kind": "port",
            "sample": f"{w.width}x{height} band of the {w.name} frame, {spp} spp, FP64 CPU renderer, {dt:.1f} s"}


def bench_megakernel(args) -> dict:
    w = configs.het(100_000)
    stats = w.scene.bvh_stats()
    t = mrad.GpuRayTracer(w.width, w.height, w.spp, True)
    if not t.is_available():
        raise SystemExit("no HIP device available: " + mrad._native.lib.mrad_last_error().decode())
    t.set_russian_roulette(w.rr_start)
    cam = w.camera.packed()
    seed = 0x4D524144
    for i in range(args.warmup):
        t.render_sample(None, w.scene, w.camera, w.max_depth, i, mrad.frame_seed(seed, i))
    kernel_ns = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        fi = args.warmup + i
        t.render_sample(None, w.scene, cam, w.max_depth, fi, mrad.frame_seed(seed, fi))
        kernel_ns.append(t.last_frame_stats()["kernel_nanos"])
    wall = time.perf_counter() - t0
    t.close()
    # instrumented replay of the same frames for the exact per-frame counters
    tm = mrad.GpuRayTracer(w.width, w.height, w.spp, True, collect_metrics=True)
    tm.set_russian_roulette(w.rr_start)
    totals = None
    for i in range(args.steps):
        fi = args.warmup + i
        tm.render_sample(None, w.scene, cam, w.max_depth, fi, mrad.frame_seed(seed, fi))
        st = tm.last_frame_stats()
        totals = st if totals is None else {k: totals[k] + st[k] if k != "max_stack_size" else max(totals[k], st[k])
                                            for k in st}
    tm.close()
    paths = w.paths * args.steps
    segs = totals["rays"]
    kernel_s = sum(kernel_ns) / 1e9
    achieved = algorithmic_bytes(totals, w.width * w.height * args.steps) / kernel_s / 1e9
    traffic = load_traffic("het-100k")
    res = {
        "metric": "Mpaths/s at fixed relMSE, wavefront batch vs megakernel; adjoint/fwd ratio",
        "value": paths / wall / 1e6,
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene generator)",
        "config": {"workload": "het-100k megakernel", "primitives": 100_000, "width": w.width, "height": w.height,
                   "spp": w.spp, "max_depth": w.max_depth, "russian_roulette_after": w.rr_start,
                   "bvh": f"weighted SAH leaf 8, {stats.node_count} nodes", "form": "megakernel"},
        "mrays_per_s": segs / wall / 1e6,
        "segments_per_path": segs / paths,
        "kernel_ms_per_step": kernel_s / args.steps * 1e3,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "counters_per_step": {k: totals[k] / args.steps for k in mrad.STAT_NAMES},
    }
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X: the path renders each frame on one GPU and this build runs no "
              "multi-GPU form; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    res = bench_megakernel(args)
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline(configs.het(100_000))
    else:
        res["cpu_baseline"] = None
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
