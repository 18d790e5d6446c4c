"""GPU: the two full-size benchmark configurations against the oracle.

* batch-64 (the headline workload): the 64 seeded scenes of uneven size,
  resolution and sample count rendered in one wavefront batch; sampled pixels
  of every scene, and more of the largest, equal the oracle's bit for bit.
* clustered-1M (1M primitives in 64 Gaussian clusters, half of them affine,
  BVH leaf 4, 1920x1080, 6 bounces): a whole 1-sample frame in the megakernel
  and in the wavefront form equals the oracle's frame bit for bit, with all
  14 frame counters; at the configuration's 16 samples, sampled pixels of the
  megakernel frame equal the oracle's.

The reference's own correctness gate compares closest hits on ray grids
(benchmark/GpuBvhCorrectnessCheck.java:49-132); these compare whole paths.
"""
from __future__ import annotations

import numpy as np
import pytest

import mrad
from mrad import configs
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch64():
    print("building the 64-scene batch", flush=True)
    return configs.batch(64)


def test_batch64_sampled_pixels_equal_oracle(require_gpu, batch64):
    ws = batch64
    classes = {(w.width, w.height, w.spp) for w in ws}
    assert len(classes) >= 8                       # the draw covers most (resolution, spp) classes
    largest = max(range(len(ws)), key=lambda i: ws[i].scene.primitive_count())
    cfg = mrad.WavefrontConfig(queue_capacity=1 << 24)
    with mrad.BatchRenderer(cfg) as b:
        for w in ws:
            b.add_scene(w.scene, w.camera, w.width, w.height, w.spp, w.max_depth, 0, w.frame_seed, w.rr_start)
        print(f"rendering batch-64: {sum(w.paths for w in ws)} paths", flush=True)
        st = b.render()
        assert st["paths"] == sum(w.paths for w in ws)
        rng = np.random.default_rng(64)
        checked = 0
        for i, w in enumerate(ws):
            n = 640 if i == largest else 64
            pix = rng.choice(w.width * w.height, n, replace=False).astype(np.int32)
            ref, _ = O.render_pixels(O.Scene(w.scene.wire()), w.camera.packed(), w.width, w.height, w.max_depth, 0,
                                     w.frame_seed, w.spp, pixels=pix, rr_start=w.rr_start)
            got = b.image(i).reshape(-1, 3)[pix]
            np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32), err_msg=w.name)
            checked += n
        assert checked >= 4096
    print(f"batch-64: {checked} pixels of {len(ws)} scenes ({len(classes)} classes) bit-exact", flush=True)


@pytest.fixture(scope="module")
def clustered1m():
    print("building clustered-1M", flush=True)
    w = configs.clustered()
    w.scene.bvh_stats()
    return w, O.Scene(w.scene.wire())


def test_clustered1m_frame_and_counters_equal_oracle(require_gpu, clustered1m):
    w, osc = clustered1m
    seed = mrad.frame_seed(0xC3, 0)
    print("clustered-1M 1920x1080 1 spp: oracle frame", flush=True)
    ref, ref_stats = O.render_pixels(osc, w.camera.packed(), w.width, w.height, w.max_depth, 0, seed, 1)
    ref = ref.reshape(-1)
    assert int(ref_stats[mrad.STAT_NAMES.index("affine_box_tests")]) > 0
    with mrad.GpuRayTracer(w.width, w.height, 1, True, collect_metrics=True) as t:
        img = np.zeros(w.width * w.height * 3, np.float32)
        t.render_sample(img, w.scene, w.camera, w.max_depth, 0, seed)
        stats = t.last_frame_stats()
    np.testing.assert_array_equal(img.view(np.uint32), ref.view(np.uint32))
    np.testing.assert_array_equal(np.array([stats[k] for k in mrad.STAT_NAMES], np.uint64), ref_stats)
    print("clustered-1M: wavefront form, same frame", flush=True)
    with mrad.BatchRenderer(mrad.WavefrontConfig(queue_capacity=1 << 22, collect_metrics=True)) as b:
        b.add_scene(w.scene, w.camera, w.width, w.height, 1, w.max_depth, 0, seed)
        st = b.render()
        np.testing.assert_array_equal(b.image(0).view(np.uint32), ref.view(np.uint32))
    np.testing.assert_array_equal(np.array([st[k] for k in mrad.STAT_NAMES], np.uint64), ref_stats)


def test_clustered1m_full_spp_sampled_pixels(require_gpu, clustered1m):
    w, osc = clustered1m
    seed = mrad.frame_seed(0xC3, 1)
    print(f"clustered-1M 1920x1080 {w.spp} spp megakernel", flush=True)
    with mrad.GpuRayTracer(w.width, w.height, w.spp, True) as t:
        img = np.zeros(w.width * w.height * 3, np.float32)
        t.render_sample(img, w.scene, w.camera, w.max_depth, 1, seed)
    pix = np.random.default_rng(33).choice(w.width * w.height, 2048, replace=False).astype(np.int32)
    ref, _ = O.render_pixels(osc, w.camera.packed(), w.width, w.height, w.max_depth, 1, seed, w.spp, pixels=pix)
    np.testing.assert_array_equal(img.reshape(-1, 3)[pix].view(np.uint32), ref.view(np.uint32))
