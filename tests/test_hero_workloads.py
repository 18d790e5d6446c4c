"""The reference's own Layered Hero workloads (CPU).

* Scene generator + SAH builder: the packed BVH of LayeredHeroFamily.create(n)
  reproduces the published topology SHA-256 of every (scale, leaf size,
  lambda) cell of benchmarks/results/main-bvh-study/hero-sensitivity.csv and
  of the uniform / weighted rows of core8-detail.csv.  The hash covers every
  node bound's float bits, so it pins the generator's double arithmetic, the
  FP32 packing with conservative bounds, the affine inverse quantization and
  the builder's splits, bit for bit.
* Metrics frame: the oracle's per-frame counters of the reference's metrics
  frame (1920x1080, 1 spp, depth 12, stack 32, frame seed
  BenchmarkProtocol.metricsSeed(row ordinal)) against the published
  core8-detail.csv counters.  This is a statistical cross-check only: frames
  with other seeds scatter by 4e-5..5e-4 relative (one standard deviation)
  and fall inside the same band, so it pins the integrator's and the
  traversal's expected work per frame, not the seed chain.  Measured: every
  counter within 1.2e-3 relative; the test allows 2e-3.
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

import mrad
from mrad import workloads as W
from oracle import oracle as O

from helpers import topology_sha256

GOLDEN = Path(__file__).resolve().parent / "golden"
CELLS = json.loads((GOLDEN / "hero_topology.json").read_text())
COUNTER_ROWS = json.loads((GOLDEN / "hero_counters.json").read_text())


def _config(leaf: int, lam: float) -> mrad.BvhBuildConfig:
    return mrad.BvhBuildConfig.uniform(leaf) if lam == 0.0 else mrad.BvhBuildConfig.weighted(leaf, lam)


def _scene(prims, cfg) -> mrad.GpuScene:
    sc = mrad.GpuScene(cfg)
    sc.set_background(prims["background"])
    sc.add_spheres(*prims["spheres"])
    sc.add_boxes(*prims["boxes"])
    sc.add_affine_transforms("box", *prims["affine_boxes"])
    return sc


def test_hero_population_matches_reference_counts():
    """BenchmarkWorkloads.fixed (:55-61): 3n/4 + 4 spheres, 5 boxes, n/4 affine boxes."""
    for n in (96, 1_000, 10_000):
        c = W.hero_scene(n).counts()
        assert (c["spheres"], c["boxes"], c["affine_spheres"], c["affine_boxes"], c["planes"]) == \
            (n * 3 // 4 + 4, 5, 0, n // 4, 0)
        assert c["materials"] == 8
    with pytest.raises(ValueError):
        W.hero_primitives(100)


@pytest.mark.parametrize("scale", [96, 1_000, 10_000])
def test_hero_topology_hashes_every_sensitivity_cell(scale):
    prims = W.hero_primitives(scale)
    cells = [c for c in CELLS if c["scale"] == scale]
    assert len(cells) == 79
    bad = [c for c in cells if topology_sha256(_scene(prims, _config(c["leafSize"], c["lambda"])))
           != c["topologySha256"]]
    assert not bad, bad[:3]


@pytest.mark.parametrize("scale", [100_000, 1_000_000])
def test_hero_topology_hashes_large_scales(scale):
    """Leaf 8 uniform and weighted (lambda 1) plus the leaf 1 / lambda 100 extremes."""
    prims = W.hero_primitives(scale)
    cells = [c for c in CELLS if c["scale"] == scale and
             ((c["leafSize"] == 8 and c["lambda"] in (0.0, 1.0)) or (scale == 100_000 and c["leafSize"] in (1, 64)
                                                                      and c["lambda"] in (0.0, 100.0)))]
    assert len(cells) >= 2
    for c in cells:
        assert topology_sha256(_scene(prims, _config(c["leafSize"], c["lambda"]))) == c["topologySha256"], c


def test_metrics_seed_properties():
    """BenchmarkProtocol.seedDomain/seed (:385-408).  No seed value is published,
    so only the properties the protocol relies on are checked: 64-bit values,
    distinct per domain, row and comparison, never zero."""
    seeds = {W.metrics_seed(k) for k in range(6)} | {W.timing_seed(k, c) for k in range(6) for c in range(7)}
    assert len(seeds) == 6 + 42 and 0 not in seeds
    assert all(0 < s < (1 << 64) for s in seeds)


def test_hero_gpu_camera_is_benchmark_linearization():
    """BenchmarkRays.gpuCamera of SceneSupport.camera(0, 0, 0): Perspective(1/3) looks
    along +z with focal length 3, unit right and up vectors."""
    cam = W.hero_gpu_camera()
    assert tuple(map(float, cam.position)) == (0.0, 0.0, 0.0)
    assert tuple(map(float, cam.forward)) == (0.0, 0.0, 3.0)
    assert tuple(map(float, cam.right)) == (1.0, 0.0, 0.0)
    assert tuple(map(float, cam.up)) == (0.0, 1.0, 0.0)


@pytest.mark.parametrize("row", [r for r in COUNTER_ROWS if r["scale"] <= 10_000],
                         ids=lambda r: f"{r['rowId']}-{r['method']}")
def test_hero_metrics_frame_counters_near_published(row):
    sc = W.hero_scene(row["scale"], _config(8, 0.0 if row["method"] == "uniform" else 1.0))
    assert topology_sha256(sc) == row["topologySha256"]
    seed = W.metrics_seed(W.workload_key(row["rowId"]))
    _, st = O.render_pixels(O.Scene(sc.wire(), stack_size=W.STACK_CAPACITY), W.hero_gpu_camera().packed(),
                            W.WIDTH, W.HEIGHT, W.MAXIMUM_PATH_DEPTH, 0, seed, W.SAMPLES_PER_PIXEL)
    ours = dict(zip(mrad.STAT_NAMES, map(int, st)))
    for name, pub in row["counters"].items():
        if pub == 0:
            assert ours[name] == 0, name
        else:
            assert abs(ours[name] - pub) <= 2e-3 * pub, (name, ours[name], pub)
    # every camera pixel starts exactly one segment
    assert ours["rays"] > W.WIDTH * W.HEIGHT
    assert ours["stack_overflows"] == 0 and ours["max_stack_size"] <= W.STACK_CAPACITY


OVERLAP_ROWS = json.loads((GOLDEN / "overlap_10k.json").read_text())


def test_overlap_population_and_topology_hashes():
    """GalleryOverlapScene.create (:47-126) over GalleryOverlapSource (:62-151):
    3,000 spheres and 7,000 affine ribbons; the published uniform and weighted
    topology SHA-256 of core8-detail.csv."""
    assert len(OVERLAP_ROWS) == 2
    prims = W.overlap_primitives()
    for row in OVERLAP_ROWS:
        sc = mrad.GpuScene(_config(8, 0.0 if row["method"] == "uniform" else 1.0))
        sc.set_background(prims["background"])
        sc.add_spheres(*prims["spheres"])
        sc.add_affine_transforms("box", *prims["affine_boxes"])
        c = sc.counts()
        assert (c["spheres"], c["boxes"], c["affine_boxes"], c["planes"]) == (3000, 0, 7000, 0)
        assert topology_sha256(sc) == row["topologySha256"], row["method"]


def test_overlap_camera_rolls_an_eighth_turn():
    """SceneSupport.camera(0, 0, 0, 0.125): the benchmark camera's right and up
    vectors are the unit axes turned by 45 degrees about the view direction."""
    cam = W.gpu_camera(W.overlap_camera(), W.WIDTH, W.HEIGHT)
    s = np.sqrt(0.5)
    np.testing.assert_allclose(cam.right, (s, s, 0.0), atol=1e-12)
    np.testing.assert_allclose(cam.up, (-s, s, 0.0), atol=1e-12)
    np.testing.assert_allclose(cam.forward, (0.0, 0.0, 3.0), atol=1e-12)
