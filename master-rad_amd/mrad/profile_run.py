"""Short, fixed workloads to run under rocprofv3 (kernel trace or counters).

    python -m mrad.profile_run het        # 100k field, 1080p, 16 spp: megakernel frame + wavefront render
    python -m mrad.profile_run batch      # the 64-scene batch, one wavefront render
    python -m mrad.profile_run copy       # one device copy of 1 GiB (calibrates FETCH_SIZE / WRITE_SIZE)
    python -m mrad.profile_run adjoint    # diff-10k forward + backward
    python -m mrad.profile_run hero       # Layered Hero 1M weighted, 1080p, 1 spp, depth 12: 5 megakernel
                                          # frames with leaf batching (16 lanes)
    python -m mrad.profile_run random     # F4-10k layout 0 weighted, 1080p, 1 spp, depth 12: 5 megakernel frames
    python -m mrad.profile_run clustered  # clustered-1M, 1080p, 16 spp, 6 bounces: megakernel frame + wavefront

Run from master-rad_amd/ (or with it on sys.path).  Each prints one line per
step so a profiled run shows progress.

    python -m mrad.profile_run summarize OUT.json KERNEL_REGEX CSV [CSV ...]

averages the counters of every dispatch whose kernel name matches
KERNEL_REGEX over the rocprofv3 counter_collection.csv files given (one file
per counter pass) and writes them, per dispatch, to OUT.json.

    python -m mrad.profile_run traffic OUT.json KERNEL_REGEX WORKLOAD FETCH.csv WRITE.csv COPY_FETCH.csv COPY_WRITE.csv

turns FETCH_SIZE and WRITE_SIZE runs (KiB per dispatch) into HBM bytes per
launch of the matching kernel, each counter scaled by what it reads for the
1 GiB device copy of `copy` (bytes moved / counter reading).
"""
from __future__ import annotations

import sys
import time

import numpy as np


def _het(spp: int):
    import mrad
    from mrad import configs
    w = configs.het(100_000, spp=spp)
    t = mrad.GpuRayTracer(w.width, w.height, w.spp, True)
    t.set_russian_roulette(w.rr_start)
    t.render_sample(None, w.scene, w.camera, w.max_depth, 0, 7)
    print(f"megakernel frame {t.last_frame_stats()['kernel_nanos'] / 1e6:.1f} ms", flush=True)
    t.close()
    b = mrad.BatchRenderer(mrad.WavefrontConfig(queue_capacity=1 << 23))
    b.add_scene(w.scene, w.camera, w.width, w.height, w.spp, w.max_depth, 0, 7, w.rr_start)
    st = b.render()
    print(f"wavefront render {st['kernel_nanos'] / 1e6:.1f} ms, {st['steps']} steps", flush=True)
    b.close()


# The wavefront configuration of the batch-64 profiles (kernel trace, SQ
# counters, HBM traffic): bench.py times its roofline render with the same
# one, so per-dispatch counters and times describe the same kernel launches.
BATCH_PROFILE_CONFIG = dict(queue_capacity=1 << 26, waves_per_cu=28, trace_passes=2, pass_pops=80,
                            pass_dense_lanes=16)


def _batch():
    import mrad
    from mrad import configs
    ws = configs.batch(64, threads=16)
    print("batch built", flush=True)
    b = mrad.BatchRenderer(mrad.WavefrontConfig(**BATCH_PROFILE_CONFIG))
    for w in ws:
        b.add_scene(w.scene, w.camera, w.width, w.height, w.spp, w.max_depth, 0, w.frame_seed, w.rr_start)
    st = b.render()
    print(f"wavefront render {st['kernel_nanos'] / 1e6:.1f} ms, {st['steps']} steps", flush=True)
    b.close()


def _copy(nbytes: int = 1 << 30):
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    a.fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.copy_(a)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    print(f"copy {nbytes} bytes in {dt * 1e3:.2f} ms ({nbytes * 2 / dt / 1e9:.0f} GB/s read+write)", flush=True)


def _clustered():
    """The clustered configuration (1M primitives in 64 Gaussian clusters, half affine, a ground box, 0.5 %
    emitters, BVH leaf 4; 1920x1080, 16 spp, 6 bounces): one megakernel frame, then the same frame in the
    wavefront form."""
    import mrad
    from mrad import configs
    w = configs.clustered()
    w.scene.bvh_stats()
    print("clustered-1M built", flush=True)
    with mrad.GpuRayTracer(w.width, w.height, w.spp, True) as t:
        t.render_sample(None, w.scene, w.camera, w.max_depth, 0, 7)
        print(f"megakernel frame {t.last_frame_stats()['kernel_nanos'] / 1e6:.1f} ms", flush=True)
    b = mrad.BatchRenderer(mrad.WavefrontConfig(queue_capacity=1 << 25))
    b.add_scene(w.scene, w.camera, w.width, w.height, w.spp, w.max_depth, 0, 7)
    st = b.render()
    print(f"wavefront render {st['kernel_nanos'] / 1e6:.1f} ms, {st['steps']} steps", flush=True)
    b.close()


def _adjoint():
    from mrad import configs
    from mrad.adjoint import PathReplayRenderer, l2_loss
    w = configs.diff(10_000)
    r = PathReplayRenderer(w.scene, w.camera, w.width, w.height, w.spp, w.max_depth, frame_seed=0xD1FF)
    img = r.forward()
    _, g = l2_loss(img, img * np.float32(0.9))
    r.backward(g)
    print(f"backward {r.last_kernel_nanos() / 1e6:.1f} ms", flush=True)
    r.close()


def _hero(frames: int = 5):
    import mrad
    from mrad import workloads as W
    sc = W.hero_scene(1_000_000, mrad.BvhBuildConfig.weighted(W.LEAF_SIZE, 1.0))
    cam = W.hero_gpu_camera()
    print("hero-1m built", flush=True)
    with mrad.GpuRayTracer(W.WIDTH, W.HEIGHT, 1, True, bvh_stack_size=W.STACK_CAPACITY) as t:
        t.set_leaf_batch(16)
        for i in range(frames):
            t.render_sample(None, sc, cam, W.MAXIMUM_PATH_DEPTH, i, mrad.frame_seed(7, i))
            print(f"hero-1m frame {i}: {t.last_frame_stats()['kernel_nanos'] / 1e6:.3f} ms", flush=True)


def _random(frames: int = 5):
    """F4-10k layout 0 (10,000 affine boxes), weighted SAH, 1920x1080, 1 spp,
    depth 12, stack 32: megakernel frames without leaf batching."""
    import mrad
    from mrad import random_scenes as RS
    from mrad import workloads as W
    sc = RS.random_scene("F4", 10_000, 0, mrad.BvhBuildConfig.weighted(W.LEAF_SIZE, 1.0))
    cam = RS.random_gpu_camera(W.WIDTH, W.HEIGHT)
    with mrad.GpuRayTracer(W.WIDTH, W.HEIGHT, 1, True, bvh_stack_size=W.STACK_CAPACITY) as t:
        for i in range(frames):
            t.render_sample(None, sc, cam, W.MAXIMUM_PATH_DEPTH, i, mrad.frame_seed(7, i))
            print(f"F4-10k frame {i}: {t.last_frame_stats()['kernel_nanos'] / 1e6:.3f} ms", flush=True)


def summarize(out: str, kernel_regex: str, csvs: list[str]) -> dict:
    """Per-dispatch means of rocprofv3 counters (counter_collection.csv rows:
    one per dispatch, counter and, for per-XCD counters, instance) of the
    dispatches whose kernel name matches kernel_regex."""
    import csv
    import json
    import re
    pat = re.compile(kernel_regex)
    totals: dict[str, dict[str, float]] = {}
    kernels = set()
    for path in csvs:
        with open(path, newline="") as f:
            for row in csv.DictReader(f):
                name = row.get("Kernel_Name", "")
                if not pat.search(name):
                    continue
                kernels.add(name)
                key = f"{path}:{row.get('Dispatch_Id', row.get('Correlation_Id', ''))}"
                d = totals.setdefault(row["Counter_Name"], {})
                d[key] = d.get(key, 0.0) + float(row["Counter_Value"])
    res = {"kernel_regex": kernel_regex, "kernels": sorted(kernels), "files": csvs,
           "dispatches": {c: len(v) for c, v in totals.items()},
           "per_dispatch": {c: sum(v.values()) / len(v) for c, v in totals.items()}}
    with open(out, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res["per_dispatch"]), flush=True)
    return res


def traffic(out: str, kernel_regex: str, workload: str, fetch_csv: str, write_csv: str, copy_fetch_csv: str,
            copy_write_csv: str, copy_bytes: int = 1 << 30) -> dict:
    """HBM bytes per launch from one FETCH_SIZE and one WRITE_SIZE counter run of the workload and the same two
    counters on a plain device copy of copy_bytes (which reads and writes copy_bytes once each)."""
    import json
    import tempfile
    from pathlib import Path

    def mean(csv_path: str, regex: str, counter: str) -> tuple[float, int, list]:
        with tempfile.TemporaryDirectory() as d:
            r = summarize(str(Path(d) / "s.json"), regex, [csv_path])
        return r["per_dispatch"][counter], r["dispatches"][counter], r["kernels"]

    fetch, n, kernels = mean(fetch_csv, kernel_regex, "FETCH_SIZE")
    write, _, _ = mean(write_csv, kernel_regex, "WRITE_SIZE")
    copy_fetch, _, _ = mean(copy_fetch_csv, "copyBuffer", "FETCH_SIZE")
    copy_write, _, _ = mean(copy_write_csv, "copyBuffer", "WRITE_SIZE")
    fetch_scale = copy_bytes / 1024.0 / copy_fetch
    write_scale = copy_bytes / 1024.0 / copy_write
    res = {"kernel": kernels[0] if len(kernels) == 1 else kernels, "workload": workload,
           "fetch_kib_per_dispatch": fetch, "write_kib_per_dispatch": write,
           "fetch_scale": fetch_scale, "write_scale": write_scale,
           "hbm_bytes_per_launch": (fetch * fetch_scale + write * write_scale) * 1024.0, "dispatches": n}
    with open(out, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res), flush=True)
    return res


def main(argv):
    what = argv[1] if len(argv) > 1 else "het"
    if what == "summarize":
        summarize(argv[2], argv[3], argv[4:])
        return
    if what == "traffic":
        traffic(*argv[2:9])
        return
    if what == "het":
        _het(int(argv[2]) if len(argv) > 2 else 16)
    elif what == "batch":
        _batch()
    elif what == "copy":
        _copy()
    elif what == "adjoint":
        _adjoint()
    elif what == "hero":
        _hero()
    elif what == "random":
        _random()
    elif what == "clustered":
        _clustered()
    else:
        raise SystemExit(f"unknown workload {what}")


if __name__ == "__main__":
    main(sys.argv)
