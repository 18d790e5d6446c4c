"""GPU: the reference's BVH correctness gate (benchmark/GpuBvhCorrectnessCheck.java:49-132) on its scenario list,
against the oracle bit for bit, and whole frames of the catalog's City of Night and Signal Chamber scenes.

For every scenario (the catalog's introductory presets, Layered Hero 96, the gallery overlap and the 100-object
random rows F1..F4) and both object-SAH methods, the 32x18 primary grid is replayed on the GPU twice:
* the linear path;
* the BVH path, with metrics and per-ray counters.

Both must equal the oracle's results bit for bit: hit flags, distance bits, primitive orders, materials, normal bits
and every per-ray counter. That is stronger than the gate's 1e-4 distance tolerance between the BVH and linear GPU
paths. The BVH path must also show no stack overflow and a maximum stack below 32. The FP32-vs-FP64 half of the gate
runs on the CPU (tests/test_correctness_gate.py)."""
from __future__ import annotations

import numpy as np
import pytest

import mrad
from mrad import workloads as W
from oracle import oracle as O

from helpers import assert_replay_equal
from scenes import gate_methods, gate_rays, gate_scenarios

pytestmark = pytest.mark.gpu

SCENARIOS = gate_scenarios()


@pytest.mark.parametrize("name,make,camera", SCENARIOS, ids=[s[0] for s in SCENARIOS])
def test_gate_replay_equals_oracle(require_gpu, name, make, camera):
    rays = gate_rays(camera)
    for method, cfg in gate_methods():
        sc = make(cfg)
        print(f"gate {name} {method}: {len(rays)} rays, {sc.primitive_count()} primitives", flush=True)
        for use_bvh in (False, True):
            ref = O.trace_replay(O.Scene(sc.wire(), use_bvh=use_bvh, stack_size=32), rays)
            with mrad.GpuRayTracer(1, 1, 1, use_bvh, collect_metrics=True, bvh_stack_size=32) as t:
                res = t.trace_replay(sc, rays, collect_metrics=True, per_ray_counters=True)
            assert_replay_equal(res, ref)
            if use_bvh:
                assert res.stats["stack_overflows"] == 0, (name, method)
                assert res.stats["max_stack_size"] < 32, (name, method)


@pytest.mark.parametrize("name", ["CITY_OF_NIGHT_V1", "SIGNAL_CHAMBER"])
def test_catalog_frames_equal_oracle(require_gpu, name):
    """A 128x72 frame at 2 spp and depth 12 (the benchmark protocol's depth and stack) with the catalog camera:
    the megakernel, with and without leaf batching, and the wavefront form equal the oracle's frame and its 14
    counters."""
    make, cam = ((W.city_of_night_scene, W.city_of_night_camera()) if name == "CITY_OF_NIGHT_V1"
                 else (W.signal_chamber_scene, W.signal_chamber_camera()))
    sc = make(mrad.BvhBuildConfig.weighted(W.LEAF_SIZE, 1.0))
    w, h, spp, depth = 128, 72, 2, W.MAXIMUM_PATH_DEPTH
    gcam = W.gpu_camera(cam, w, h)
    seed = mrad.frame_seed(0xCA7A, 0)
    print(f"{name} {w}x{h} {spp} spp: oracle frame", flush=True)
    ref, ref_stats = O.render_pixels(O.Scene(sc.wire()), gcam.packed(), w, h, depth, 0, seed, spp)
    ref = ref.reshape(-1)
    assert ref.max() > 0
    for lanes in (0, 16):
        with mrad.GpuRayTracer(w, h, spp, True, collect_metrics=True, bvh_stack_size=W.STACK_CAPACITY) as t:
            t.set_leaf_batch(lanes)
            img = np.zeros(w * h * 3, np.float32)
            t.render_sample(img, sc, gcam, depth, 0, seed)
            stats = t.last_frame_stats()
        np.testing.assert_array_equal(img.view(np.uint32), ref.view(np.uint32))
        np.testing.assert_array_equal(np.array([stats[k] for k in mrad.STAT_NAMES], np.uint64), ref_stats)
    with mrad.BatchRenderer(mrad.WavefrontConfig(queue_capacity=1 << 16, collect_metrics=True)) as b:
        b.add_scene(sc, gcam, w, h, spp, depth, 0, seed)
        st = b.render()
        np.testing.assert_array_equal(b.image(0).view(np.uint32), ref.view(np.uint32))
    np.testing.assert_array_equal(np.array([st[k] for k in mrad.STAT_NAMES], np.uint64), ref_stats)
