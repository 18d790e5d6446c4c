"""One process of the CPU baseline: the FP64 CPU restatement of the reference
CPU renderer (oracle_fp64.c: PathTracer / BSDF / solids over parallel rows, as
playground/HeadlessCpuRenderer.java:39-81 renders) on one seeded batch scene.

TEST INFRASTRUCTURE ONLY: run as a child process by bench.py's cpu_baseline
leg, which starts several fresh processes and takes medians; it is the
reported CPU baseline, never the thing the GPU path runs.

    python -m oracle.cpu_baseline --scene-index I --width W --height H --spp S --threads T --frames N

builds batch scene I, renders one warm-up frame and N timed frames (frame
seeds 0..N), and prints one JSON line with the frame times.
"""
from __future__ import annotations

import argparse
import json
import time

from mrad import configs
from oracle import oracle as O


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--scene-index", type=int, required=True)
    ap.add_argument("--width", type=int, required=True)
    ap.add_argument("--height", type=int, required=True)
    ap.add_argument("--spp", type=int, default=1)
    ap.add_argument("--threads", type=int, default=1)
    ap.add_argument("--frames", type=int, default=3)
    a = ap.parse_args()
    t0 = time.perf_counter()
    w = configs.batch_scene(configs.batch_specs(64)[a.scene_index])
    osc = O.Scene(w.scene.wire())
    build = time.perf_counter() - t0
    cam = w.camera.packed()
    times = []
    for f in range(a.frames + 1):
        t1 = time.perf_counter()
        O.render64(osc, cam, a.width, a.height, w.max_depth, a.spp, f, rr_start=w.rr_start, threads=a.threads)
        if f > 0:
            times.append(time.perf_counter() - t1)
    print(json.dumps({"scene": w.name, "build_seconds": build, "frame_seconds": times,
                      "paths_per_frame": a.width * a.height * a.spp, "threads": a.threads}), flush=True)


if __name__ == "__main__":
    main()
