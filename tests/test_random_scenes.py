"""Random first-hit families F1-F4 (benchmark/RandomSingleRaySceneFactory.java)
on the CPU.

* SplittableRandom: SplitMix64's published output for seed 1234567 (the
  reference C implementation's sequence), and the range and bucket balance of
  the derived draws.
* Layouts: every pair of centres at least (2 + gap) radii apart and every
  envelope inside [-1, 1]^3, checked by brute force rather than the grid the
  generator uses.
* Populations: the F1-F4 shares and shared centres; affine transforms are a
  proper rotation of the stated axes.
* Construction: medians over the 100 layouts of node count, depth and bytes
  against the published construction matrices (bvh-*-matrix.csv): equal at 100
  objects; at 1,000 objects the depths are equal and nodes and bytes are within
  0.3 %.  The published data holds only medians, so the layouts themselves are
  not pinned beyond this.
"""
from __future__ import annotations

import json
import math
import statistics
from pathlib import Path

import numpy as np
import pytest

import mrad
from mrad import random_scenes as RS

GOLDEN = Path(__file__).resolve().parent / "golden"
PUBLISHED = {(r["rowId"], r["method"]): r for r in json.loads((GOLDEN / "bvh_matrix.json").read_text())}


def test_splitmix64_known_answers():
    r = RS.SplittableRandom(1234567)
    assert [r.next_long() for _ in range(5)] == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                                                 4593380528125082431, 16408922859458223821]


def test_derived_draws():
    r = RS.SplittableRandom(99)
    draws = [r.next_int_bounded(3) for _ in range(30_000)]
    counts = np.bincount(draws, minlength=3)
    assert counts.sum() == 30_000 and np.all(np.abs(counts - 10_000) < 400)
    a, b = RS.SplittableRandom(5), RS.SplittableRandom(5)
    assert all(a.next_int_bounded(64) == b.next_int() & 63 for _ in range(100))     # power of two: low bits
    r = RS.SplittableRandom(7)
    xs = [r.next_double_in(-0.25, 0.25) for _ in range(10_000)]
    assert min(xs) >= -0.25 and max(xs) < 0.25 and abs(statistics.fmean(xs)) < 0.01
    r = RS.SplittableRandom(11)
    assert 4_700 < sum(r.next_boolean() for _ in range(10_000)) < 5_300


@pytest.mark.parametrize("count,layout_id", [(100, 0), (100, 57), (1_000, 3)])
def test_layout_separation_and_containment(count, layout_id):
    seed = RS.layout_seed(RS.LAYOUT_EXPERIMENT_SEED, count, layout_id)
    centres, radius = RS.layout(count, seed)
    c = np.array(centres)
    assert c.shape == (count, 3) and radius == pytest.approx(0.58 / count ** (1 / 3), rel=1e-15)
    assert np.max(np.abs(c)) + radius <= 1.0
    d2 = ((c[:, None, :] - c[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(d2, np.inf)
    assert np.sqrt(d2.min()) >= radius * (2.0 + RS.GAP_FRACTION)
    assert (centres, radius) == RS.layout(count, seed)                        # deterministic


def test_populations_share_the_layout():
    seed = RS.layout_seed(RS.LAYOUT_EXPERIMENT_SEED, 100, 4)
    centres, radius = RS.layout(100, seed)
    kinds = {f: RS.population_kinds(100, seed, RS.POPULATION[f]) for f in RS.FAMILIES}
    assert sorted(np.bincount(kinds["F1"], minlength=4).tolist()) == [0, 0, 50, 50]
    assert kinds["F1"].count(RS.SPHERE) == 50 and kinds["F1"].count(RS.AFFINE_BOX) == 50
    assert np.bincount(kinds["F2"], minlength=4).tolist() == [25, 25, 25, 25]
    assert set(kinds["F3"]) == {RS.SPHERE} and set(kinds["F4"]) == {RS.AFFINE_BOX}
    _, prims = RS.primitives(centres, radius, seed, RS.FOUR_TYPES_EQUAL)
    for (x, y, z, r) in prims[RS.SPHERE]:
        assert (x, y, z) in centres and r == radius
    for kind in (RS.AFFINE_SPHERE, RS.AFFINE_BOX):
        for t in prims[kind]:
            m = np.array(t).reshape(3, 4)
            assert tuple(m[:, 3]) in centres
            axes = np.linalg.norm(m[:, :3], axis=0)                  # columns: rotation * diag(axes)
            rot = m[:, :3] / axes
            np.testing.assert_allclose(rot.T @ rot, np.eye(3), atol=1e-12)
            assert np.linalg.det(rot) == pytest.approx(1.0, abs=1e-12)
            assert np.all(axes <= radius * (1 + 1e-12))
            if kind == RS.AFFINE_SPHERE:                             # (r, r a, r b)
                assert axes[0] == pytest.approx(radius, rel=1e-12)
            else:                                                    # proportions scaled to length r
                assert math.sqrt((axes ** 2).sum()) == pytest.approx(radius, rel=1e-12)


def test_scene_counts_and_seeds():
    sc = RS.random_scene("F2", 100, 0)
    c = sc.counts()
    assert (c["spheres"], c["boxes"], c["affine_spheres"], c["affine_boxes"]) == (25, 25, 25, 25)
    assert RS.workload_key("F1-100", 0) == 6 and RS.workload_key("F4-10k", 99) == 6 + 11 * 100 + 99
    with pytest.raises(ValueError):
        RS.workload_key("F1-100", 100)


def _medians(family: str, count: int):
    cache: dict = {}
    out = {}
    for method, cfg in (("uniform", mrad.BvhBuildConfig.uniform(8)), ("weighted", mrad.BvhBuildConfig.weighted(8))):
        stats = [RS.random_scene(family, count, lid, cfg, cache).bvh_stats() for lid in RS.LAYOUT_IDS]
        out[method] = {"nodeCount": statistics.median(s.node_count for s in stats),
                       "maxDepth": statistics.median(s.max_depth for s in stats),
                       "bytes": statistics.median(s.bytes for s in stats)}
    return out


@pytest.mark.parametrize("family", RS.FAMILIES)
def test_construction_medians_at_100_objects_equal_published(family):
    ours = _medians(family, 100)
    for method in ("uniform", "weighted"):
        pub = PUBLISHED[(RS.row_id(family, 100), method)]
        assert ours[method] == {k: pub[k] for k in ("nodeCount", "maxDepth", "bytes")}, method


@pytest.mark.parametrize("family", RS.FAMILIES)
def test_construction_medians_at_1k_objects_near_published(family):
    ours = _medians(family, 1_000)
    for method in ("uniform", "weighted"):
        pub = PUBLISHED[(RS.row_id(family, 1_000), method)]
        assert ours[method]["maxDepth"] == pub["maxDepth"]
        for k in ("nodeCount", "bytes"):
            assert abs(ours[method][k] - pub[k]) <= 3e-3 * pub[k], (method, k, ours[method][k], pub[k])
