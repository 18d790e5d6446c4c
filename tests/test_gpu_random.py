"""GPU: random-family frames (benchmark/RandomSingleRaySceneFactory.java) with
the study's camera (-3, 0, 0) towards the origin and the layout's metrics
seed, depth 12, 1 spp, against the oracle bit for bit (pixels and the 14
counters).  F2 holds all four bounded primitive kinds in one BVH, spheres,
axis-aligned boxes, affine spheres and affine boxes; under uniform SAH F1 at
1,000 objects puts spheres and affine boxes in mixed leaves."""
from __future__ import annotations

import numpy as np
import pytest

import mrad
from mrad import random_scenes as RS
from mrad import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CASES = [("F2", 100, 0, 0.0), ("F2", 100, 41, 1.0), ("F1", 1_000, 7, 0.0)]


@pytest.mark.parametrize("family,count,layout_id,lam", CASES, ids=[f"{c[0]}-{c[1]}-L{c[2]}" for c in CASES])
def test_random_frame_bit_exact(require_gpu, family, count, layout_id, lam):
    width, height = 480, 270
    cfg = mrad.BvhBuildConfig.uniform(8) if lam == 0.0 else mrad.BvhBuildConfig.weighted(8, lam)
    sc = RS.random_scene(family, count, layout_id, cfg)
    cam = RS.random_gpu_camera(width, height)
    seed = RS.metrics_seed(RS.row_id(family, count), layout_id)
    print(f"random frame {family}-{count} layout {layout_id} lambda={lam} seed={seed:#x}", flush=True)
    ref, ref_st = O.render_pixels(O.Scene(sc.wire(), stack_size=W.STACK_CAPACITY), cam.packed(), width, height,
                                  W.MAXIMUM_PATH_DEPTH, 0, seed, 1)
    ref_stats = dict(zip(mrad.STAT_NAMES, map(int, ref_st)))
    assert ref_stats["aabb_tests"] > 0 and ref_stats["primitive_tests"] > 0
    for lanes in (0, 16):
        img = np.zeros(width * height * 3, np.float32)
        with mrad.GpuRayTracer(width, height, 1, True, collect_metrics=True, bvh_stack_size=W.STACK_CAPACITY) as t:
            t.set_leaf_batch(lanes)
            t.render_sample(img, sc, cam, W.MAXIMUM_PATH_DEPTH, 0, seed)
            st = t.last_frame_stats()
        np.testing.assert_array_equal(img.view(np.uint32), ref.reshape(-1).view(np.uint32))
        for k in mrad.STAT_NAMES:
            assert st[k] == ref_stats[k], (lanes, k)
    wcfg = mrad.WavefrontConfig(queue_capacity=1 << 18, collect_metrics=True, bvh_stack_size=W.STACK_CAPACITY)
    with mrad.BatchRenderer(wcfg) as b:
        b.add_scene(sc, cam, width, height, 1, W.MAXIMUM_PATH_DEPTH, 0, seed)
        bst = b.render()
        np.testing.assert_array_equal(b.image(0).view(np.uint32), ref.reshape(-1).view(np.uint32))
    for k in mrad.STAT_NAMES:
        assert bst[k] == ref_stats[k], k
