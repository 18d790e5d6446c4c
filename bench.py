"""Benchmark of the MI355X path-tracing hot path (one GPU, one JSON line).

Default workload (the configuration the headline metric is quoted on): the
64-scene batch -- 64 seeded heterogeneous fields of 1k..1M primitives at
256^2..1920x1080, 4/16/64 spp, 6 bounces, roulette after bounce 3 -- rendered
by the batched wavefront form.  A step is one render of the whole batch with
every scene resident in HBM and every image left in HBM.  The wavefront's
launch parameters (pool size, waves per CU, LDS split and depth of the
traversal stack) are autotuned on the batch first, within a time budget.
Reported beside it:

  * the same batch rendered as 64 sequential plain-megakernel frames, and the
    bit-identity of the two forms' images (so relMSE is equal by construction);
  * the batch without roulette (every path runs its 6 bounces);
  * the roofline of the dominant kernel (the wavefront trace kernel).  It is
    bound by VALU issue: wave-level VALU instructions per launch (rocprofv3
    counters of the same render, profiles/counters_batch-64.json) over the
    launch time measured here with HIP events, against MI355X's issue peak of
    one VALU instruction per CU per cycle (256 CUs at 2.4 GHz).  Beside it,
    the HBM bandwidth it draws (FETCH_SIZE + WRITE_SIZE per launch,
    profiles/traffic_batch-64.json) and the logical bytes of its node and
    primitive fetches (exact traversal counters priced per path segment as
    24 N_aabb + 16 (N_internal + N_leaf) + 4 N_refs + 20 N_sphere + 28 N_box
    + 88 N_affine + 104 N_plane + 52 N_segments, most of them cache hits);
  * the adjoint/forward time ratio on the differentiable configuration (10k
    primitives, 512^2, 32 spp, 4 bounces);
  * the reference's GI Test frame at its platform-comparison protocol
    (1280x720, 1 spp, depth 12, uniform SAH leaf 8, catalog camera) beside
    the published RTX 3080 kernel time and segment count;
  * the FP64 CPU renderer on this host: k fresh processes, one warm-up and
    three timed frames each, median of the frames, median over processes;
    and a relMSE check of the GPU image against a converged FP64 reference
    with the fraction of pixels inside the 3.29-sigma Monte-Carlo band.

--workload het (100k primitives, 1080p, 64 spp, 8 bounces) and --workload
clustered (1M clustered primitives, half affine, 1080p, 16 spp, 6 bounces) time
one full-size frame in the megakernel, the wavefront and the regrouped
wavefront forms.  --workload hero times the reference's own Layered Hero
frames (96 to 1M primitives, 1080p, 1 spp, 12 bounces) and its overlap-10k
gallery frame under the uniform and weighted SAH.
"""
from __future__ import annotations

import argparse
import json
import math
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
# VALU issue peak of MI355X: one wave64 VALU instruction per CU per cycle, 256
# CUs at the 2.4 GHz peak engine clock (the published 157.3 TFLOP/s FP32
# vector figure counts packed two-wide FMAs: 256 x 2.4e9 x 64 lanes x 2 x 2).
VALU_PEAK_GINST = 256 * 2.4
METRIC = "Mpaths/s at fixed relMSE, wavefront batch vs megakernel; adjoint/fwd ratio"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(c: dict) -> float:
    n_refs = c["primitive_tests"] - c["plane_tests"]
    return (24.0 * c["aabb_tests"] + 16.0 * (c["internal_node_visits"] + c["leaf_node_visits"]) + 4.0 * n_refs
            + 20.0 * c["sphere_tests"] + 28.0 * c["box_tests"]
            + 88.0 * (c["affine_sphere_tests"] + c["affine_box_tests"]) + 104.0 * c["plane_tests"]
            + 52.0 * c["rays"])


def load_traffic(name: str):
    p = ROOT / "profiles" / f"traffic_{name}.json"
    if not p.exists():
        return None
    try:
        return float(json.loads(p.read_text())["hbm_bytes_per_launch"])
    except Exception:
        return None


def load_counters(name: str):
    """Per-dispatch means of the trace kernel's SQ counters (profiles/counters_<name>.json, written by
    python -m mrad.profile_run summarize from rocprofv3 counter runs of the same render)."""
    p = ROOT / "profiles" / f"counters_{name}.json"
    if not p.exists():
        return None
    try:
        return json.loads(p.read_text())["per_dispatch"]
    except Exception:
        return None


def lscpu() -> dict:
    """Sockets, cores and SMT of this host (lscpu), and the CPUs this process may run on."""
    import subprocess
    info = {"model": "unknown CPU", "sockets": None, "cores_per_socket": None, "threads_per_core": None,
            "logical_cpus": os.cpu_count(), "cpus_available": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
    except Exception:
        out = ""
    keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
            "Thread(s) per core": "threads_per_core", "CPU(s)": "logical_cpus"}
    for line in out.splitlines():
        k, _, v = line.partition(":")
        if k.strip() in keys:
            v = v.strip()
            info[keys[k.strip()]] = int(v) if v.isdigit() else v
    return info


def _cpu_run(index: int, width: int, height: int, spp: int, threads: int, frames: int, timeout: float) -> dict:
    import subprocess
    cmd = [sys.executable, "-m", "oracle.cpu_baseline", "--scene-index", str(index), "--width", str(width),
           "--height", str(height), "--spp", str(spp), "--threads", str(threads), "--frames", str(frames)]
    env = dict(os.environ, PYTHONPATH=os.pathsep.join([str(ROOT / "master-rad_amd"), str(ROOT)]))
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, env=env, cwd=str(ROOT), check=True)
    return json.loads(out.stdout.strip().splitlines()[-1])


def cpu_baseline(w: configs.Workload, index: int, processes: int = 3) -> dict:
    """FP64 CPU restatement of the reference CPU renderer (parallel rows, oracle/oracle_fp64.c) timed as the
    reference's protocol times it: `processes` fresh processes, each one warm-up and three timed frames, the
    median of its frames, and the median over processes.  Each frame is one sample per pixel of the whole view
    (RendererFactory.java:46-51: the CPU renderers accumulate 1-spp frames), on this process's CPUs (at most 16),
    and one process on one thread at a quarter of the resolution in each axis for the single-thread figure."""
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    sw, sh, spp = w.width, w.height, 1
    per_proc = []
    for _ in range(processes):
        r = _cpu_run(index, sw, sh, spp, threads, 3, 600)
        per_proc.append(float(np.median(r["frame_seconds"])))
        log(f"cpu baseline process: frames {['%.2f' % t for t in r['frame_seconds']]} s")
    med = float(np.median(per_proc))
    paths = sw * sh * spp
    qw, qh = max(8, w.width // 4), max(8, w.height // 4)
    r1 = _cpu_run(index, qw, qh, spp, 1, 3, 600)
    t1 = float(np.median(r1["frame_seconds"]))
    return {"value": paths / med / 1e6, "unit": "Mpaths/s", "cores": threads, "kind": "port",
            "sample": f"{w.name} ({w.scene.primitive_count()} primitives) camera at {sw}x{sh}, {spp} spp, "
                      f"depth {w.max_depth}, roulette after {w.rr_start}: {paths} paths per frame, FP64 CPU renderer",
            "protocol": f"{processes} fresh processes x (1 warm-up + 3 timed frames); median of frames, median over "
                        "processes",
            "process_median_seconds": per_proc,
            "single_thread_value": qw * qh * spp / t1 / 1e6,
            "single_thread_sample": f"{qw}x{qh}, {spp} spp, one process, median of 3 frames: {t1:.2f} s",
            "host": lscpu()}


def relmse_check(ws, ref_spp: int = 1024) -> dict:
    """relMSE = mean (I - R)^2 / (R^2 + 1e-2) of the GPU image of the batch's
    smallest 256^2 scene against a converged FP64 CPU reference R of the same
    scene and camera (independent random numbers).  The wavefront and
    megakernel images are bit-identical, so this is the relMSE of both forms.
    `expected` is the Monte-Carlo expectation mean(var / (R^2 + 1e-2)) with
    var = s^2 (1/spp + 1/ref_spp) from the per-pixel sample variance s^2 of the
    reference run; `band_fraction` is the fraction of pixel channels with
    |I - R| <= 3.29 sqrt(var) plus the FP32 rounding of a spp-sample float sum
    (4 spp 2^-24 |R|).  At 4 spp the ratio of the two fluctuates widely between
    frames (heavy-tailed emitter hits), so the band is the convergence test."""
    from oracle import oracle as O
    w = min((w for w in ws if w.width == 256), key=lambda w: w.scene.primitive_count())
    t = mrad.GpuRayTracer(w.width, w.height, w.spp, True)
    t.set_russian_roulette(w.rr_start)
    img = np.zeros(w.width * w.height * 3, np.float32)
    t.render_sample(img, w.scene, w.camera, w.max_depth, 0, w.frame_seed)
    t.close()
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    t0 = time.perf_counter()
    ref, var = O.render64(O.Scene(w.scene.wire()), w.camera.packed(), w.width, w.height, w.max_depth, ref_spp, 0x5EED,
                          rr_start=w.rr_start, threads=threads, variance=True)
    dt = time.perf_counter() - t0
    ref, var = ref.reshape(-1), np.maximum(var.reshape(-1), 0.0) * (1.0 / w.spp + 1.0 / ref_spp)
    den = ref ** 2 + 1e-2
    diff = img.astype(np.float64) - ref
    rel = float(np.mean(diff ** 2 / den))
    expected = float(np.mean(var / den))
    band = float(np.mean(np.abs(diff) <= 3.29 * np.sqrt(var) + 4.0 * w.spp * 2.0 ** -24 * np.abs(ref)))
    return {"scene": f"{w.name} ({w.scene.primitive_count()} primitives) {w.width}x{w.height}", "gpu_spp": w.spp,
            "reference_spp": ref_spp, "relmse": rel, "expected_relmse": expected, "band_fraction": band,
            "reference_seconds": dt}


def gi_test_context(frames: int = 20) -> dict:
    """The reference's GI Test frame at its matched-platform protocol (1280x720, 1 spp, depth 12, uniform SAH
    leaf 8, SceneCatalog's camera; benchmarks/results/platform-comparison.csv:3) in the megakernel: median
    kernel time and the segments of a frame, beside the published RTX 3080 figures.  The published run's
    frame seeds are not recorded and it was made at an earlier revision of the reference, so the segment
    count is a statistical cross-check only."""
    from mrad import workloads as W
    sc = W.gi_test_scene(mrad.BvhBuildConfig.uniform(W.LEAF_SIZE))
    cam = W.gi_test_gpu_camera()
    wd, ht = W.GI_TEST_WIDTH, W.GI_TEST_HEIGHT
    with mrad.GpuRayTracer(wd, ht, 1, True) as t:
        t.render_sample(None, sc, cam, W.MAXIMUM_PATH_DEPTH, 0, mrad.frame_seed(0x617E, 0))
        kern = []
        for i in range(frames):
            t.render_sample(None, sc, cam, W.MAXIMUM_PATH_DEPTH, i, mrad.frame_seed(0x617E, i))
            kern.append(t.last_frame_stats()["kernel_nanos"] / 1e6)
    segs = []
    with mrad.GpuRayTracer(wd, ht, 1, True, collect_metrics=True) as t:
        for i in range(4):
            t.render_sample(None, sc, cam, W.MAXIMUM_PATH_DEPTH, i, mrad.frame_seed(0x617E, i))
            segs.append(t.last_frame_stats()["rays"])
    k = float(np.median(kern))
    return {"workload": "GI_TEST 1280x720, 1 spp, depth 12, uniform SAH leaf 8, catalog camera",
            "kernel_ms": k, "segments_per_frame": float(np.mean(segs)),
            "segments_per_frame_spread": float(np.std(segs)), "gsegments_per_s": float(np.mean(segs)) / (k * 1e6),
            "published_rtx3080_kernel_ms": 4.6029, "published_gpu_segments": 4215974,
            "published_cpu_segments": 4215129}


# The reference's published platform comparison (benchmarks/results/platform-comparison.csv:2-7, RTX 3080):
# scene, kernel mode, median kernel ms, GPU path segments of a frame.
PUBLISHED_PLATFORM = [("GI_TEST", "uniform_bvh", 4.6029, 4215974), ("GI_TEST", "linear", 3.5724, 4215974),
                      ("CITY_OF_NIGHT_V1", "uniform_bvh", 27.0259, 2552461),
                      ("SIGNAL_CHAMBER", "uniform_bvh", 37.79205, 4461552)]


def platform_comparison(frames: int = 10) -> list:
    """The catalog scenes of the reference's CPU-vs-GPU platform comparison at its protocol (1280x720, 1 spp,
    depth 12, uniform SAH leaf 8, SceneCatalog's cameras) in the megakernel, BVH or linear scan as the row says:
    median kernel time and mean segments of a frame beside the published RTX 3080 figures.  The published
    frame seeds are not recorded, so segments are a statistical cross-check."""
    from mrad import workloads as W
    make = {"GI_TEST": (W.gi_test_scene, W.gi_test_camera()),
            "CITY_OF_NIGHT_V1": (W.city_of_night_scene, W.city_of_night_camera()),
            "SIGNAL_CHAMBER": (W.signal_chamber_scene, W.signal_chamber_camera())}
    wd, ht = W.GI_TEST_WIDTH, W.GI_TEST_HEIGHT
    rows = []
    for name, mode, pub_ms, pub_segs in PUBLISHED_PLATFORM:
        scene_fn, camera = make[name]
        sc = scene_fn(mrad.BvhBuildConfig.uniform(W.LEAF_SIZE))
        cam = W.gpu_camera(camera, wd, ht)
        bvh = mode != "linear"
        kern, segs = [], []
        with mrad.GpuRayTracer(wd, ht, 1, bvh) as t:
            t.render_sample(None, sc, cam, W.MAXIMUM_PATH_DEPTH, 0, mrad.frame_seed(0x617E, 0))
            for i in range(frames):
                t.render_sample(None, sc, cam, W.MAXIMUM_PATH_DEPTH, i, mrad.frame_seed(0x617E, i))
                kern.append(t.last_frame_stats()["kernel_nanos"] / 1e6)
        with mrad.GpuRayTracer(wd, ht, 1, bvh, collect_metrics=True) as t:
            for i in range(2):
                t.render_sample(None, sc, cam, W.MAXIMUM_PATH_DEPTH, i, mrad.frame_seed(0x617E, i))
                segs.append(t.last_frame_stats()["rays"])
        k = float(np.median(kern))
        rows.append({"scene": name, "mode": mode, "kernel_ms": k, "segments_per_frame": float(np.mean(segs)),
                     "published_rtx3080_kernel_ms": pub_ms, "published_gpu_segments": pub_segs,
                     "kernel_time_ratio": pub_ms / k})
        log(f"platform comparison {name} {mode}: {k:.3f} ms (published RTX 3080 {pub_ms} ms)")
    return rows


def make_batch(ws, **cfg) -> mrad.BatchRenderer:
    b = mrad.BatchRenderer(mrad.WavefrontConfig(**cfg))
    for w in ws:
        b.add_scene(w.scene, w.camera, w.width, w.height, w.spp, w.max_depth, 0, w.frame_seed, w.rr_start)
    return b


def bench_batch(args) -> dict:
    t0 = time.perf_counter()
    ws = configs.batch(64, threads=16)
    total_paths = sum(w.paths for w in ws)
    prims = sum(w.scene.primitive_count() for w in ws)
    log(f"built 64 scenes ({prims} primitives, {total_paths} paths) in {time.perf_counter() - t0:.1f} s")
    cfg = dict(queue_capacity=args.queue_capacity, waves_per_cu=args.waves_per_cu, trace_passes=args.trace_passes,
               pass_pops=args.pass_pops, pass_dense_lanes=args.pass_dense_lanes)
    b = make_batch(ws, **cfg)
    tuned = None
    if args.autotune_budget > 0:
        from mrad.autotune import tune
        t1 = time.perf_counter()
        grid = {"queue_capacity": [1 << 25, 1 << 26, 1 << 27], "waves_per_cu": [24, 28, 32],
                "lds_stack_entries": [4, 10], "stack_storage": [32, 64], "trace_passes": [1, 2],
                "pass_pops": [64, 80, 96], "pass_dense_lanes": [-1, 16, 32]}
        best, table = tune(b, grid=grid, log=log, budget_s=args.autotune_budget)
        cfg = dict(queue_capacity=best.queue_capacity, waves_per_cu=best.waves_per_cu,
                   lds_stack_entries=best.lds_stack_entries, stack_storage=best.stack_storage,
                   trace_passes=best.trace_passes, pass_pops=best.pass_pops, pass_dense_lanes=best.pass_dense_lanes)
        tuned = {"best": cfg, "grid": {k: v for k, v in grid.items()}, "trials": table,
                 "budget_seconds": args.autotune_budget, "seconds": time.perf_counter() - t1}
        log(f"autotuned in {tuned['seconds']:.0f} s: {cfg}")
    for _ in range(args.warmup):
        b.render()
    wall, kern = [], []
    for i in range(args.steps):
        t1 = time.perf_counter()
        st = b.render()
        wall.append(time.perf_counter() - t1)
        kern.append(st["kernel_nanos"] / 1e9)
        log(f"wavefront step {i}: {wall[-1] * 1e3:.0f} ms, {st['steps']} bounce steps")
    t1 = time.perf_counter()
    wf_images = [b.image(k) for k in range(len(ws))]
    copy_s = time.perf_counter() - t1
    image_bytes = sum(img.nbytes for img in wf_images)
    b.close()
    value = total_paths * args.steps / sum(wall) / 1e6

    # 64 sequential plain-megakernel renders of the same frames
    tracers = []
    for w in ws:
        t = mrad.GpuRayTracer(w.width, w.height, w.spp, True)
        t.set_russian_roulette(w.rr_start)
        t.ensure_scene_resident(w.scene)
        tracers.append(t)
    mk_wall, mk_kern, identical = 0.0, 0.0, True
    for k, (w, t) in enumerate(zip(ws, tracers)):
        t.render_sample(None, w.scene, w.camera, w.max_depth, 0, w.frame_seed)   # warm
    for k, (w, t) in enumerate(zip(ws, tracers)):
        t1 = time.perf_counter()
        t.render_sample(None, w.scene, w.camera, w.max_depth, 0, w.frame_seed)
        mk_wall += time.perf_counter() - t1
        mk_kern += t.last_frame_stats()["kernel_nanos"] / 1e9
        img = np.zeros(w.width * w.height * 3, np.float32)
        t.copy_frame(img)
        identical &= bool(np.array_equal(img.view(np.uint32), wf_images[k].view(np.uint32)))
    for t in tracers:
        t.close()
    mk_value = total_paths / mk_wall / 1e6
    log(f"megakernel: {mk_wall * 1e3:.0f} ms for 64 renders, images bit-identical: {identical}")

    # dominant kernel: per-kernel event timing and exact counters of the same batch, in the configuration the
    # committed counter profile was taken with (its per-dispatch counters and these times describe the same
    # launches; the trace runs trace_passes launches per bounce step)
    from mrad.profile_run import BATCH_PROFILE_CONFIG
    prof_cfg = dict(BATCH_PROFILE_CONFIG)
    passes = prof_cfg.get("trace_passes", 1)
    bt = make_batch(ws, time_kernels=True, **prof_cfg)
    st_t = bt.render()
    bt.close()
    bm = make_batch(ws, collect_metrics=True, **cfg)
    st_m = bm.render()
    bm.close()
    trace_s = st_t["trace_nanos"] / 1e9
    launches = max(1, st_t["steps"] * passes)
    trace_kernel = "wfTracePass" if passes > 1 else "wfTrace"
    log(f"trace kernel {trace_s * 1e3:.0f} ms of {st_t['kernel_nanos'] / 1e6:.0f} ms over {launches} launches")

    # the same batch without roulette: every path runs its 6 bounces
    nr = [configs.Workload(w.name, w.scene, w.camera, w.width, w.height, w.spp, w.max_depth, -1, w.frame_seed)
          for w in ws]
    bn = make_batch(nr, **cfg)
    bn.render()
    t1 = time.perf_counter()
    bn.render()
    nr_s = time.perf_counter() - t1
    bn.close()
    bnm = make_batch(nr, collect_metrics=True, **cfg)
    nr_rays = bnm.render()["rays"]
    bnm.close()
    log(f"without roulette: {nr_s * 1e3:.0f} ms, {nr_rays / total_paths:.3f} segments per path")

    adj = bench_adjoint()
    gi = gi_test_context()
    log(f"GI_TEST context: {gi['kernel_ms']:.3f} ms kernel, {gi['segments_per_frame']:.0f} segments")
    platform = platform_comparison()
    return {
        "metric": METRIC,
        "value": value,
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sum(wall) / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene generator)",
        "config": {"workload": "batch-64 wavefront", "scenes": 64, "primitives": prims, "paths_per_step": total_paths,
                   "resolutions": "256^2..1920x1080", "spp": "4/16/64", "max_depth": 6, "russian_roulette_after": 3,
                   **cfg, "hipgraph": True},
        "autotune": tuned,
        "megakernel_mpaths_per_s": mk_value,
        "wavefront_over_megakernel": value / mk_value,
        "kernel_only": {"wavefront_ms": float(np.mean(kern)) * 1e3, "megakernel_ms": mk_kern * 1e3,
                        "ratio": mk_kern / float(np.mean(kern))},
        "images_bit_identical": identical,
        "host_copy": {"image_bytes": image_bytes, "seconds": copy_s,
                      "pcie_inclusive_mpaths_per_s": total_paths / (sum(wall) / args.steps + copy_s) / 1e6},
        "relmse_equal": identical,
        "mrays_per_s": st_m["rays"] / (sum(wall) / args.steps) / 1e6,
        "segments_per_path": st_m["rays"] / total_paths,
        "no_roulette": {"mpaths_per_s": total_paths / nr_s / 1e6, "ms": nr_s * 1e3,
                        "segments_per_path": nr_rays / total_paths, "mrays_per_s": nr_rays / nr_s / 1e6},
        "adjoint": adj,
        "adjoint_over_forward": adj["ratio"],
        "gi_test_context": gi,
        "platform_comparison": platform,
        "roofline": {**roofline_entry(trace_kernel, trace_s / launches, launches, trace_s * 1e3,
                                      algorithmic_bytes(st_m) / launches, load_traffic("batch-64"),
                                      load_counters("batch-64"), "batch-64"), "render_config": prof_cfg},
        "counters_per_step": {k: st_m[k] for k in mrad.STAT_NAMES},
    }, ws


def roofline_entry(kernel: str, launch_s: float, launches: int, kernel_ms_per_step: float,
                   logical_bytes_per_launch: float, traffic, counters, profile: str) -> dict:
    """The roofline object of the dominant kernel.  The roof that binds is VALU issue: `achieved` is the
    kernel's wave-level VALU instructions per launch (SQ_INSTS_VALU, rocprofv3 counters of the same render)
    over its average launch time measured here, in G instructions/s, against VALU_PEAK_GINST.
    `lane_efficiency` = SQ_THREAD_CYCLES_VALU / (64 SQ_ACTIVE_INST_VALU), the share of lanes doing work per
    VALU instruction.  `hbm` is the bandwidth the kernel draws (FETCH_SIZE + WRITE_SIZE bytes per launch,
    `traffic`) against 8 TB/s; `logical_bytes` prices every node and primitive fetch of the exact counters as
    if it came from HBM (most are L2 / Infinity Cache hits), so it is not a bandwidth roof."""
    r = {"bound": "valu_issue", "unit": "Ginst/s", "peak": VALU_PEAK_GINST, "achieved": None, "frac": None,
         "traffic": traffic, "kernel": kernel, "kernel_ms_per_step": kernel_ms_per_step,
         "launches_per_step": launches, "launch_ms": launch_s * 1e3}
    if counters and counters.get("SQ_INSTS_VALU"):
        ach = counters["SQ_INSTS_VALU"] / launch_s / 1e9
        r.update(achieved=ach, frac=ach / VALU_PEAK_GINST, valu_insts_per_launch=counters["SQ_INSTS_VALU"],
                 counters=f"profiles/counters_{profile}.json")
        if counters.get("SQ_ACTIVE_INST_VALU") and counters.get("SQ_THREAD_CYCLES_VALU"):
            r["lane_efficiency"] = counters["SQ_THREAD_CYCLES_VALU"] / (64.0 * counters["SQ_ACTIVE_INST_VALU"])
    if traffic is not None and launch_s > 0:
        r["hbm"] = {"gbs": traffic / launch_s / 1e9, "peak_gbs": HBM_PEAK_GBS,
                    "frac": traffic / launch_s / 1e9 / HBM_PEAK_GBS, "bytes_per_launch": traffic}
    r["logical_bytes"] = {"bytes_per_launch": logical_bytes_per_launch, "gbs": logical_bytes_per_launch / launch_s / 1e9,
                          "note": "node and primitive records priced per segment as if fetched from HBM"}
    return r


def bench_adjoint() -> dict:
    """Differentiable configuration: forward and path-replay backward kernel times."""
    from mrad.adjoint import PathReplayRenderer, l2_loss
    w = configs.diff(10_000)
    r = PathReplayRenderer(w.scene, w.camera, w.width, w.height, w.spp, w.max_depth, frame_seed=0xD1FF)
    rng = np.random.default_rng(9)
    m = r.base_materials.copy()
    m[:, 0:3] *= rng.uniform(0.8, 1.2, (len(m), 3)).astype(np.float32)
    m[:, 9:12] *= rng.uniform(0.8, 1.2, (len(m), 1)).astype(np.float32)
    target = r.forward(m)
    img = r.forward(r.base_materials)
    _, g = l2_loss(img, target)
    r.backward(g)
    fwd, bwd = [], []
    for _ in range(3):
        r.forward()
        fwd.append(r.last_kernel_nanos())
        r.backward(g)
        bwd.append(r.last_kernel_nanos())
    r.close()
    f, bk = float(np.median(fwd)) / 1e6, float(np.median(bwd)) / 1e6
    return {"workload": "diff-10k 512x512 32 spp 4 bounces", "forward_ms": f, "backward_ms": bk, "ratio": bk / f}


def bench_single(args, w: configs.Workload) -> dict:
    """One full-size frame: the plain megakernel (the line's value) beside the
    wavefront form and the wavefront form with leaf regrouping by primitive
    class, all three on the same frame (bit-identical images)."""
    stats = w.scene.bvh_stats()
    seed = 0x4D524144
    t = mrad.GpuRayTracer(w.width, w.height, w.spp, True)
    t.set_russian_roulette(w.rr_start)
    cam = w.camera.packed()
    for i in range(args.warmup):
        t.render_sample(None, w.scene, w.camera, w.max_depth, i, mrad.frame_seed(seed, i))
    kernel_ns = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        fi = args.warmup + i
        t.render_sample(None, w.scene, cam, w.max_depth, fi, mrad.frame_seed(seed, fi))
        kernel_ns.append(t.last_frame_stats()["kernel_nanos"])
    wall = time.perf_counter() - t0
    mk_img = np.zeros(w.width * w.height * 3, np.float32)
    t.copy_frame(mk_img)
    t.close()
    tm = mrad.GpuRayTracer(w.width, w.height, w.spp, True, collect_metrics=True)
    tm.set_russian_roulette(w.rr_start)
    totals = None
    for i in range(args.steps):
        fi = args.warmup + i
        tm.render_sample(None, w.scene, cam, w.max_depth, fi, mrad.frame_seed(seed, fi))
        st = tm.last_frame_stats()
        totals = st if totals is None else {k: (totals[k] + st[k]) if k != "max_stack_size"
                                            else max(totals[k], st[k]) for k in st}
    tm.close()
    log(f"megakernel {wall / args.steps * 1e3:.0f} ms/frame")
    # the last frame again in the wavefront form, plain and regrouped
    last = args.warmup + args.steps - 1
    forms = {}
    for name, regroup in (("wavefront", 0), ("wavefront_regroup", 1)):
        b = mrad.BatchRenderer(mrad.WavefrontConfig(queue_capacity=args.queue_capacity,
                                                    waves_per_cu=args.waves_per_cu, leaf_regroup=regroup))
        b.add_scene(w.scene, w.camera, w.width, w.height, w.spp, w.max_depth, last, mrad.frame_seed(seed, last),
                    w.rr_start)
        b.render()
        t1 = time.perf_counter()
        st = b.render()
        dt = time.perf_counter() - t1
        same = bool(np.array_equal(b.image(0).view(np.uint32), mk_img.view(np.uint32)))
        b.close()
        bt = mrad.BatchRenderer(mrad.WavefrontConfig(queue_capacity=args.queue_capacity,
                                                     waves_per_cu=args.waves_per_cu, leaf_regroup=regroup,
                                                     time_kernels=True))
        bt.add_scene(w.scene, w.camera, w.width, w.height, w.spp, w.max_depth, last, mrad.frame_seed(seed, last),
                     w.rr_start)
        stt = bt.render()
        bt.close()
        forms[name] = {"ms": dt * 1e3, "mpaths_per_s": w.paths / dt / 1e6, "trace_ms": stt["trace_nanos"] / 1e6,
                       "image_bit_identical_to_megakernel": same}
        log(f"{name} {dt * 1e3:.0f} ms/frame (trace {stt['trace_nanos'] / 1e6:.0f} ms), bit-identical: {same}")
    paths = w.paths * args.steps
    kernel_s = sum(kernel_ns) / 1e9
    achieved = (algorithmic_bytes(totals) + 12.0 * w.width * w.height * args.steps) / kernel_s / 1e9
    return {
        "metric": METRIC, "value": paths / wall / 1e6, "unit": "Mpaths/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded scene generator)",
        "config": {"workload": f"{w.name} megakernel", "primitives": w.scene.primitive_count(), "width": w.width,
                   "height": w.height, "spp": w.spp, "max_depth": w.max_depth, "russian_roulette_after": w.rr_start,
                   "bvh": f"{w.scene.config.mode.split('_')[0].lower()} SAH leaf {w.scene.config.leaf_size}, {stats.node_count} nodes"},
        "mrays_per_s": totals["rays"] / wall / 1e6, "segments_per_path": totals["rays"] / paths,
        "forms": forms,
        "roofline": roofline_entry("renderKernel", kernel_s / args.steps, 1, kernel_s / args.steps * 1e3,
                                   achieved * 1e9 * kernel_s / args.steps, load_traffic(w.name),
                                   load_counters(w.name), w.name),
        "counters_per_step": {k: totals[k] / args.steps for k in mrad.STAT_NAMES},
    }, [w]


def bench_random(args) -> list:
    """The random first-hit rows F1-F4 at 100 / 1k / 10k objects
    (RandomSingleRaySceneFactory, the study's look-at camera; 1920x1080, 1 spp,
    depth 12, megakernel without leaf batching), over the first
    ``--random-layouts`` layouts of each row.  Per layout, as the reference's
    analysis does (benchmarks/analyze_results.py:669-706, 1132-1140): the work
    reduction of weighted against uniform SAH from one counter frame each, and
    the time reduction as the median of four ABBA blocks, each
    100 (1 - sqrt(B1 B2 / (A1 A2))) over kernel times.  A row reports the
    medians over its layouts beside the published ones (RTX 3080, 100 layouts)."""
    from mrad import random_scenes as RS
    from mrad import workloads as W
    from mrad.recorder import traversal_work_of
    pub = {r["rowId"]: r for r in json.loads((ROOT / "tests" / "golden" / "random_reductions.json").read_text())}
    cam = RS.random_gpu_camera(W.WIDTH, W.HEIGHT)
    kw = dict(bvh_stack_size=W.STACK_CAPACITY)
    rows = []
    with mrad.GpuRayTracer(W.WIDTH, W.HEIGHT, 1, True, collect_metrics=True, **kw) as tm, \
            mrad.GpuRayTracer(W.WIDTH, W.HEIGHT, 1, True, **kw) as ta, \
            mrad.GpuRayTracer(W.WIDTH, W.HEIGHT, 1, True, **kw) as tb:
        for count in RS.OBJECT_COUNTS:
            cache: dict = {}
            for fam in RS.FAMILIES:
                rid = RS.row_id(fam, count)
                t_red, w_red, rays, kms = [], [], [], {"uniform": [], "weighted": []}
                for lid in range(args.random_layouts):
                    scenes, work = {}, {}
                    for method, cfg in (("uniform", mrad.BvhBuildConfig.uniform(W.LEAF_SIZE)),
                                        ("weighted", mrad.BvhBuildConfig.weighted(W.LEAF_SIZE, 1.0))):
                        scenes[method] = RS.random_scene(fam, count, lid, cfg, cache)
                        tm.render_sample(None, scenes[method], cam, W.MAXIMUM_PATH_DEPTH, 0, RS.metrics_seed(rid, lid))
                        st = tm.last_frame_stats()
                        work[method] = traversal_work_of(st)
                        if method == "uniform":
                            rays.append(st["rays"])
                    w_red.append(100.0 * (1.0 - work["weighted"] / work["uniform"]))
                    base = W.timing_seed(RS.workload_key(rid, lid), 0)

                    def frame(t, method, i):
                        t.render_sample(None, scenes[method], cam, W.MAXIMUM_PATH_DEPTH, i, mrad.frame_seed(base, i))
                        ns = t.last_frame_stats()["kernel_nanos"]
                        kms[method].append(ns / 1e6)
                        return ns

                    for t, method in ((ta, "uniform"), (tb, "weighted")):          # uploads and warm-up
                        t.render_sample(None, scenes[method], cam, W.MAXIMUM_PATH_DEPTH, 0, mrad.frame_seed(base, 0))
                    blocks = []
                    for blk in range(4):
                        s1, s2 = 1 + 2 * blk, 2 + 2 * blk
                        a1, b1, b2, a2 = frame(ta, "uniform", s1), frame(tb, "weighted", s1), \
                            frame(tb, "weighted", s2), frame(ta, "uniform", s2)
                        blocks.append(100.0 * (1.0 - math.sqrt(b1 * b2 / (a1 * a2))))
                    t_red.append(float(np.median(blocks)))
                row = {"row": rid, "primitives": count, "layouts": args.random_layouts,
                       "uniform_kernel_ms": float(np.median(kms["uniform"])),
                       "weighted_kernel_ms": float(np.median(kms["weighted"])),
                       "weighted_time_reduction_pct": float(np.median(t_red)),
                       "weighted_work_reduction_pct": float(np.median(w_red)),
                       "median_segments": float(np.median(rays)),
                       "published_rtx3080_time_reduction_pct": pub[rid]["timeReductionPercent"],
                       "published_work_reduction_pct": pub[rid]["workReductionPercent"],
                       "published_median_segments": pub[rid]["medianRays"]}
                log(f"{rid}: {row['uniform_kernel_ms']:.3f} / {row['weighted_kernel_ms']:.3f} ms kernel, reduction: "
                    f"time {row['weighted_time_reduction_pct']:+.2f} % (published {pub[rid]['timeReductionPercent']:+.2f} %), "
                    f"work {row['weighted_work_reduction_pct']:+.2f} % (published {pub[rid]['workReductionPercent']:+.2f} %), "
                    f"median segments {row['median_segments']:.0f} (published {pub[rid]['medianRays']:.0f})")
                rows.append(row)
    return rows


def bench_hero(args) -> dict:
    """The reference's own Layered Hero timing frames (LayeredHeroFamily, 96 to
    1M primitives; 1920x1080, 1 spp, depth 12, stack 32) in the megakernel,
    uniform and weighted (lambda 1) SAH at leaf 8.  Reported per scale: kernel
    ms per frame (median), segments per second, and the weighted-vs-uniform
    kernel-time reduction beside the reference's published one (RTX 3080,
    core8-detail.csv).  Each frame is timed with the megakernel's leaf
    batching off (the reduction columns, the reference's traversal order of
    work) and on (16 lanes; same images and counters); the line's value is the
    1M-primitive weighted frame in the faster of the two."""
    from mrad import workloads as W
    pub = {(r["rowId"], r["method"]): r
           for r in json.loads((ROOT / "tests" / "golden" / "hero_counters.json").read_text())}
    frames = max(5, 10 * args.steps)
    cam = W.hero_gpu_camera()
    rows, line = [], None
    for scale in W.HERO_SCALES:
        rid = W.scale_id(scale)
        key = W.workload_key(rid)
        res = {}
        for method, lam in (("uniform", 0.0), ("weighted", 1.0)):
            cfg = mrad.BvhBuildConfig.uniform(W.LEAF_SIZE) if lam == 0.0 else mrad.BvhBuildConfig.weighted(W.LEAF_SIZE, lam)
            sc = W.hero_scene(scale, cfg)
            base = W.timing_seed(key, 0)
            timed = {}
            for lanes in (0, 16):
                with mrad.GpuRayTracer(W.WIDTH, W.HEIGHT, 1, True, bvh_stack_size=W.STACK_CAPACITY) as t:
                    t.set_leaf_batch(lanes)
                    for i in range(max(1, args.warmup)):
                        t.render_sample(None, sc, cam, W.MAXIMUM_PATH_DEPTH, i, mrad.frame_seed(base, i))
                    kern = []
                    t0 = time.perf_counter()
                    for i in range(frames):
                        t.render_sample(None, sc, cam, W.MAXIMUM_PATH_DEPTH, i, mrad.frame_seed(base, i))
                        kern.append(t.last_frame_stats()["kernel_nanos"])
                    timed[lanes] = (float(np.median(kern)) / 1e6, (time.perf_counter() - t0) / frames)
            with mrad.GpuRayTracer(W.WIDTH, W.HEIGHT, 1, True, collect_metrics=True,
                                   bvh_stack_size=W.STACK_CAPACITY) as tm:
                tm.render_sample(None, sc, cam, W.MAXIMUM_PATH_DEPTH, 0, W.metrics_seed(key))
                st = tm.last_frame_stats()
            best = min(timed, key=lambda k: timed[k][0])
            k_ms, wall = timed[best]
            res[method] = {"kernel_ms": timed[0][0], "kernel_ms_leaf_batch": timed[16][0], "leaf_batch": best,
                           "best_kernel_ms": k_ms, "wall_ms": wall * 1e3, "segments": st["rays"],
                           "gsegments_per_s": st["rays"] / (k_ms * 1e6),
                           "roofline_gbs": (algorithmic_bytes(st) + 12.0 * W.WIDTH * W.HEIGHT) / (k_ms * 1e6),
                           "counters": st}
            log(f"{rid} {method}: {timed[0][0]:.3f} ms kernel ({timed[16][0]:.3f} ms with leaf batching), "
                f"{wall * 1e3:.3f} ms wall, {st['rays']} segments")
        red = 100.0 * (1.0 - res["weighted"]["kernel_ms"] / res["uniform"]["kernel_ms"])
        rows.append({"row": rid, "primitives": scale + 9,
                     "uniform_kernel_ms": res["uniform"]["kernel_ms"], "weighted_kernel_ms": res["weighted"]["kernel_ms"],
                     "uniform_kernel_ms_leaf_batch": res["uniform"]["kernel_ms_leaf_batch"],
                     "weighted_kernel_ms_leaf_batch": res["weighted"]["kernel_ms_leaf_batch"],
                     "weighted_gsegments_per_s": res["weighted"]["gsegments_per_s"],
                     "weighted_time_reduction_pct": red,
                     "published_rtx3080_time_reduction_pct": pub[(rid, "weighted")]["timeReductionPercent"]})
        if scale == W.HERO_SCALES[-1]:
            line = dict(res["weighted"], scene=sc)
    # the overlap-10k gallery frame (GalleryOverlapScene), uniform and weighted
    ovl = {(r["rowId"], r["method"]): r for r in json.loads((ROOT / "tests" / "golden" / "overlap_10k.json").read_text())}
    ocam = W.gpu_camera(W.overlap_camera(), W.WIDTH, W.HEIGHT)
    okey = W.workload_key("overlap-10k")
    oms, olb = {}, {}
    for method, lam in (("uniform", 0.0), ("weighted", 1.0)):
        cfg = mrad.BvhBuildConfig.uniform(W.LEAF_SIZE) if lam == 0.0 else mrad.BvhBuildConfig.weighted(W.LEAF_SIZE, lam)
        sc = W.overlap_scene(cfg)
        for lanes in (0, 16):
            with mrad.GpuRayTracer(W.WIDTH, W.HEIGHT, 1, True, bvh_stack_size=W.STACK_CAPACITY) as t:
                t.set_leaf_batch(lanes)
                t.render_sample(None, sc, ocam, W.MAXIMUM_PATH_DEPTH, 0, W.timing_seed(okey, 0))
                kern = []
                for i in range(max(3, frames // 4)):
                    t.render_sample(None, sc, ocam, W.MAXIMUM_PATH_DEPTH, i,
                                    mrad.frame_seed(W.timing_seed(okey, 0), i))
                    kern.append(t.last_frame_stats()["kernel_nanos"])
            (olb if lanes else oms)[method] = float(np.median(kern)) / 1e6
        log(f"overlap-10k {method}: {oms[method]:.2f} ms kernel ({olb[method]:.2f} ms with leaf batching)")
    rows.append({"row": "overlap-10k", "primitives": W.OVERLAP_OBJECTS, "uniform_kernel_ms": oms["uniform"],
                 "weighted_kernel_ms": oms["weighted"], "uniform_kernel_ms_leaf_batch": olb["uniform"],
                 "weighted_kernel_ms_leaf_batch": olb["weighted"],
                 "weighted_time_reduction_pct": 100.0 * (1.0 - oms["weighted"] / oms["uniform"]),
                 "published_rtx3080_time_reduction_pct": ovl[("overlap-10k", "weighted")]["timeReductionPercent"]})
    rows.extend(bench_random(args))
    paths = W.WIDTH * W.HEIGHT
    return {
        "metric": METRIC, "value": paths / (line["wall_ms"] / 1e3) / 1e6, "unit": "Mpaths/s", "n_gpus": 1,
        "steps": frames, "warmup": args.warmup, "ms_per_step": line["wall_ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (the reference's deterministic Layered Hero generator)",
        "config": {"workload": "hero-1m weighted megakernel (LayeredHeroFamily 1M + 9 environment primitives)",
                   "width": W.WIDTH, "height": W.HEIGHT, "spp": 1, "max_depth": W.MAXIMUM_PATH_DEPTH,
                   "bvh": "weighted SAH lambda 1 leaf 8", "stack": W.STACK_CAPACITY,
                   "leaf_batch": line["leaf_batch"]},
        "mrays_per_s": line["segments"] / (line["wall_ms"] / 1e3) / 1e6,
        "scales": rows,
        "roofline": roofline_entry("renderKernel", line["best_kernel_ms"] / 1e3, 1, line["best_kernel_ms"],
                                   line["roofline_gbs"] * line["best_kernel_ms"] * 1e6, load_traffic("hero-1m"),
                                   load_counters("hero-1m"), "hero-1m"),
        "counters_per_step": {k: line["counters"][k] for k in mrad.STAT_NAMES},
    }, [configs.Workload("hero-1m", line["scene"], cam, W.WIDTH, W.HEIGHT, 1, W.MAXIMUM_PATH_DEPTH)]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["batch", "het", "clustered", "hero"], default="batch")
    ap.add_argument("--queue-capacity", type=int, default=1 << 26)
    ap.add_argument("--waves-per-cu", type=int, default=28)
    ap.add_argument("--trace-passes", type=int, default=2,
                    help="closest-hit passes per bounce step of the batch (capped passes regroup long traversals)")
    ap.add_argument("--pass-pops", type=int, default=80, help="pop budget of a capped trace pass")
    ap.add_argument("--pass-dense-lanes", type=int, default=16,
                    help="lanes still traversing at the pop budget that keep a wave going (-1: always suspend)")
    ap.add_argument("--random-layouts", type=int, default=10,
                    help="layouts per random F1-F4 row timed by --workload hero (the study used 100)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--autotune-budget", type=float, default=40.0,
                    help="seconds for tuning pool size, waves per CU, LDS split and stack depth on the batch "
                         "before timing (0: no tuning)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X: every frame and batch renders on one GPU and the package has no "
              "multi-GPU form; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    index = None
    if args.workload == "batch":
        res, ws = bench_batch(args)
        # the largest 1080p scene of the batch
        index = max((i for i, w in enumerate(ws) if w.width == 1920), key=lambda i: ws[i].scene.primitive_count())
        sample = ws[index]
    elif args.workload == "hero":
        res, ws = bench_hero(args)
        sample = ws[0]
    else:
        res, ws = bench_single(args, configs.het(100_000) if args.workload == "het" else configs.clustered())
        sample = ws[0]
    if args.no_cpu_baseline or index is None:
        res["cpu_baseline"] = None
    else:
        res["cpu_baseline"] = cpu_baseline(sample, index)
        res["cpu_baseline"]["relmse"] = relmse_check(ws)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
