"""GPU: the reference's own Layered Hero metrics frames (1920x1080, 1 spp,
depth 12, stack 32, frame seed BenchmarkProtocol.metricsSeed), rendered by
the megakernel (with and without leaf batching) and by the wavefront form, against the oracle bit for bit:
every pixel's float bits and all 14 frame counters.  These scenes exercise
the mirror, glass and emissive lobes and thousands of affine boxes."""
from __future__ import annotations

import numpy as np
import pytest

import mrad
from mrad import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CASES = [("hero-96", 96, 0.0), ("hero-10k", 10_000, 1.0), ("hero-100k", 100_000, 1.0)]


def _cfg(lam):
    return mrad.BvhBuildConfig.uniform(8) if lam == 0.0 else mrad.BvhBuildConfig.weighted(8, lam)


@pytest.mark.parametrize("row,scale,lam", CASES, ids=[c[0] for c in CASES])
def test_hero_metrics_frame_bit_exact(require_gpu, row, scale, lam):
    sc = W.hero_scene(scale, _cfg(lam))
    cam = W.hero_gpu_camera()
    seed = W.metrics_seed(W.workload_key(row))
    print(f"hero frame {row} lambda={lam} seed={seed:#x}", flush=True)
    ref, ref_st = O.render_pixels(O.Scene(sc.wire(), stack_size=W.STACK_CAPACITY), cam.packed(), W.WIDTH, W.HEIGHT,
                                  W.MAXIMUM_PATH_DEPTH, 0, seed, 1)
    ref_stats = dict(zip(mrad.STAT_NAMES, map(int, ref_st)))
    # the megakernel without and with leaf batching (16 lanes), with and without counters
    for lanes, metrics in ((0, True), (16, True), (16, False)):
        img = np.zeros(W.WIDTH * W.HEIGHT * 3, np.float32)
        with mrad.GpuRayTracer(W.WIDTH, W.HEIGHT, W.SAMPLES_PER_PIXEL, True, collect_metrics=metrics,
                               bvh_stack_size=W.STACK_CAPACITY) as t:
            t.set_leaf_batch(lanes)
            t.render_sample(img, sc, cam, W.MAXIMUM_PATH_DEPTH, 0, seed)
            st = t.last_frame_stats()
        np.testing.assert_array_equal(img.view(np.uint32), ref.reshape(-1).view(np.uint32))
        if metrics:
            for k in mrad.STAT_NAMES:
                assert st[k] == ref_stats[k], (lanes, k)

    # the same frame in the wavefront form, with and without counters
    for metrics in (True, False):
        cfg = mrad.WavefrontConfig(queue_capacity=1 << 20, collect_metrics=metrics, bvh_stack_size=W.STACK_CAPACITY)
        with mrad.BatchRenderer(cfg) as b:
            b.add_scene(sc, cam, W.WIDTH, W.HEIGHT, 1, W.MAXIMUM_PATH_DEPTH, 0, seed)
            bst = b.render()
            np.testing.assert_array_equal(b.image(0).view(np.uint32), ref.reshape(-1).view(np.uint32))
        if metrics:
            for k in mrad.STAT_NAMES:
                assert bst[k] == ref_stats[k], k


def test_overlap_frame_against_oracle_and_published(require_gpu):
    """The overlap-10k metrics frame (3,000 spheres, 7,000 ribbons; ~2,560 AABB and
    ~1,830 affine-box tests per segment): the megakernel's image equals the
    oracle's on 4,096 sampled pixels bit for bit, the wavefront form equals the
    megakernel on the whole frame, and the GPU frame counters are within 2e-3
    of the published core8-detail.csv row (statistical, as for the Hero rows)."""
    import json
    from pathlib import Path
    row = json.loads((Path(__file__).resolve().parent / "golden" / "overlap_10k.json").read_text())[0]
    sc = W.overlap_scene(_cfg(0.0))
    cam = W.gpu_camera(W.overlap_camera(), W.WIDTH, W.HEIGHT)
    seed = W.metrics_seed(W.workload_key("overlap-10k"))
    print(f"overlap frame seed={seed:#x}", flush=True)
    img = np.zeros(W.WIDTH * W.HEIGHT * 3, np.float32)
    with mrad.GpuRayTracer(W.WIDTH, W.HEIGHT, 1, True, collect_metrics=True, bvh_stack_size=W.STACK_CAPACITY) as t:
        t.render_sample(img, sc, cam, W.MAXIMUM_PATH_DEPTH, 0, seed)
        st = t.last_frame_stats()
    for name, pub in row["counters"].items():
        assert abs(st[name] - pub) <= 2e-3 * max(pub, 1), (name, st[name], pub)
    pixels = np.random.default_rng(5).choice(W.WIDTH * W.HEIGHT, 4096, replace=False).astype(np.int32)
    ref, _ = O.render_pixels(O.Scene(sc.wire(), stack_size=W.STACK_CAPACITY), cam.packed(), W.WIDTH, W.HEIGHT,
                             W.MAXIMUM_PATH_DEPTH, 0, seed, 1, pixels=pixels)
    np.testing.assert_array_equal(img.reshape(-1, 3)[pixels].view(np.uint32), ref.view(np.uint32))
    with mrad.BatchRenderer(mrad.WavefrontConfig(queue_capacity=1 << 21, bvh_stack_size=W.STACK_CAPACITY)) as b:
        b.add_scene(sc, cam, W.WIDTH, W.HEIGHT, 1, W.MAXIMUM_PATH_DEPTH, 0, seed)
        b.render()
        np.testing.assert_array_equal(b.image(0).view(np.uint32), img.view(np.uint32))
