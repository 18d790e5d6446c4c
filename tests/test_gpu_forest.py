"""GPU: a per-type forest (one tree per primitive class, four roots) given through the boundary's wire arrays:
closest-hit replay, the megakernel (with and without leaf batching) and the wavefront form equal the oracle bit
for bit, with every per-ray and frame counter.  The forest is built on the CPU side as a caller would
(tests/helpers.per_type_forest); the root insertion order of GpuKernelSources.java:863-885 decides the traversal."""
from __future__ import annotations

import numpy as np
import pytest

import mrad
from mrad import configs
from oracle import oracle as O

from helpers import assert_replay_equal, camera_grid_rays, per_type_forest, random_rays
from scenes import per_type_parts

pytestmark = pytest.mark.gpu


def test_forest_replay_and_frames_equal_oracle(require_gpu):
    full, parts = per_type_parts()
    forest = per_type_forest(full, parts)
    osc = O.Scene(forest.wire())
    cam = configs.field_camera()
    rays = np.concatenate([camera_grid_rays(cam.packed(), 64, 36), random_rays(np.random.default_rng(3), 2048)])
    print("forest replay", flush=True)
    ref = O.trace_replay(osc, rays)
    with mrad.GpuRayTracer(8, 8, 1, True, collect_metrics=True) as t:
        res = t.trace_replay(forest, rays, collect_metrics=True, per_ray_counters=True)
    assert_replay_equal(res, ref)
    w, h, spp, depth = 96, 54, 2, 8
    seed = mrad.frame_seed(0xF0, 0)
    print("forest frame", flush=True)
    img_ref, stats_ref = O.render_pixels(osc, cam.packed(), w, h, depth, 0, seed, spp)
    img_ref = img_ref.reshape(-1)
    for lanes in (0, 16):
        with mrad.GpuRayTracer(w, h, spp, True, collect_metrics=True) as t:
            t.set_leaf_batch(lanes)
            img = np.zeros(w * h * 3, np.float32)
            t.render_sample(img, forest, cam, depth, 0, seed)
            stats = t.last_frame_stats()
        np.testing.assert_array_equal(img.view(np.uint32), img_ref.view(np.uint32))
        np.testing.assert_array_equal(np.array([stats[k] for k in mrad.STAT_NAMES], np.uint64), stats_ref)
    with mrad.BatchRenderer(mrad.WavefrontConfig(queue_capacity=1 << 14, collect_metrics=True)) as b:
        b.add_scene(forest, cam, w, h, spp, depth, 0, seed)
        st = b.render()
        np.testing.assert_array_equal(b.image(0).view(np.uint32), img_ref.view(np.uint32))
    np.testing.assert_array_equal(np.array([st[k] for k in mrad.STAT_NAMES], np.uint64), stats_ref)
