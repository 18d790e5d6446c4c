"""Scene packer and BVH builder (CPU): published topology hashes, an
independent literal restatement of the reference builder, structural
invariants and input validation."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

import mrad
from mrad import configs
from oracle import bvh_ref

from helpers import topology_sha256
from scenes import gi_test_scene

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_gi_test_topology_hashes_match_published():
    """benchmarks/results/main-bvh-study/leaf-size-time.csv: GI_TEST rows (uniform SAH)."""
    rows = json.loads((GOLDEN / "gi_test_topology.json").read_text())
    assert len(rows) == 7
    for row in rows:
        sc = gi_test_scene(leaf_size=row["leafSize"])
        st = sc.bvh_stats()
        assert topology_sha256(sc) == row["topologySha256"], row
        assert st.node_count == row["nodeCount"] and st.max_depth == row["maxDepth"]


@pytest.mark.parametrize("scene_id", ["CITY_OF_NIGHT_V1", "SIGNAL_CHAMBER"])
def test_catalog_topology_hashes_match_published(scene_id):
    """leaf-size-time.csv's City of Night and Signal Chamber rows (LeafSizeStudy.java:127-140: object SAH with the
    calibrated base at lambda 0, which is the uniform cost): every leaf size's topology SHA-256, node count and
    depth, under both the uniform and the lambda-0 weighted build."""
    from mrad import workloads as W
    make = W.city_of_night_scene if scene_id == "CITY_OF_NIGHT_V1" else W.signal_chamber_scene
    rows = [r for r in json.loads((GOLDEN / "catalog_topology.json").read_text()) if r["sceneId"] == scene_id]
    assert len(rows) == 7
    for row in rows:
        for cfg in (mrad.BvhBuildConfig.uniform(row["leafSize"]), mrad.BvhBuildConfig.weighted(row["leafSize"], 0.0)):
            sc = make(cfg)
            st = sc.bvh_stats()
            assert topology_sha256(sc) == row["topologySha256"], (row, cfg.mode)
            assert st.node_count == row["nodeCount"] and st.max_depth == row["maxDepth"], row


def _random_small(rng, n_s, n_b, n_as, n_ab):
    spheres = np.concatenate([rng.uniform(-1, 1, (n_s, 3)), rng.uniform(0.02, 0.3, (n_s, 1))], 1)
    c = rng.uniform(-1, 1, (n_b, 3))
    e = rng.uniform(0.02, 0.3, (n_b, 3))
    boxes = np.concatenate([c - e, c + e], 1)
    # a few exact duplicates and shared centroids exercise the (centroid, order) tie rule
    if n_s > 4:
        spheres[1] = spheres[0]
        spheres[3, :3] = spheres[2, :3]
    aff = []
    for n in (n_as, n_ab):
        cc = rng.uniform(-1, 1, (n, 3))
        ee = rng.uniform(0.05, 0.3, (n, 3))
        aff.append(np.concatenate([cc - ee, cc + ee], 1))
    return spheres, boxes, aff[0], aff[1]


def _native_scene(cfg, spheres, boxes, abnd_s, abnd_b):
    sc = mrad.GpuScene(cfg)
    m = lambda n: mrad.material_array((0.5, 0.5, 0.5), n=n)  # noqa: E731
    sc.add_spheres(spheres, m(len(spheres)))
    sc.add_boxes(boxes, m(len(boxes)))
    for kind, b in (("sphere", abnd_s), ("box", abnd_b)):
        if len(b) == 0:
            continue
        n = len(b)
        ident = np.tile(np.array([1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0, 0]), (n, 1))
        centre = (b[:, :3] + b[:, 3:]) / 2
        sc.add_affine(kind, ident, ident, centre, b, m(n))
    return sc


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("weighted,leaf", [(False, 1), (False, 4), (True, 8), (True, 2)])
def test_builder_matches_literal_restatement(seed, weighted, leaf):
    rng = np.random.default_rng(seed)
    sizes = [(40, 30, 10, 10), (7, 3, 0, 2), (120, 60, 20, 20), (1, 0, 0, 0), (0, 5, 0, 0), (33, 0, 17, 0)][seed]
    spheres, boxes, a_s, a_b = _random_small(rng, *sizes)
    cfg = mrad.BvhBuildConfig.weighted(leaf) if weighted else mrad.BvhBuildConfig.uniform(leaf)
    sc = _native_scene(cfg, spheres, boxes, a_s, a_b)
    q = [cfg.sphere_weight, cfg.box_weight, cfg.affine_sphere_weight, cfg.affine_box_weight]
    q = [1.0 + cfg.lam * (w - 1.0) if weighted else 1.0 for w in q]
    prims = bvh_ref.primitives(spheres, boxes, a_s, a_b, q)
    nb, nd, refs, roots = bvh_ref.build(prims, leaf, cfg.traversal_weight)
    np.testing.assert_array_equal(sc.bvh_node_bounds().view(np.uint32), nb.view(np.uint32))
    np.testing.assert_array_equal(sc.bvh_node_data(), nd)
    np.testing.assert_array_equal(sc.bvh_primitive_refs(), refs)
    np.testing.assert_array_equal(sc.bvh_root_indices(), roots)


def test_structural_invariants_field():
    sc = configs.field_scene(99, 20_000)
    nb, nd, refs = sc.bvh_node_bounds(), sc.bvh_node_data(), sc.bvh_primitive_refs().view(np.uint32)
    st = sc.bvh_stats()
    leaves = nd[:, 0] < 0
    assert st.node_count == 2 * st.leaf_count - 1 == len(nb)
    assert leaves.sum() == st.leaf_count
    # every primitive referenced exactly once, leaf ranges tile the ref array in DFS order
    counts = sc.counts()
    n = counts["spheres"] + counts["boxes"] + counts["affine_spheres"] + counts["affine_boxes"]
    assert len(refs) == n
    assert sorted(nd[leaves, 2].tolist()) == np.concatenate([[0], np.cumsum(nd[leaves, 3])[:-1]]).tolist()
    # children lie inside parents; left child is the next node (pre-order)
    inner = np.where(~leaves)[0]
    assert np.all(nd[inner, 0] == inner + 1)
    for k in (0, 1):
        ch = nd[inner, k]
        assert np.all(nb[ch, :3] >= nb[inner, :3]) and np.all(nb[ch, 3:] <= nb[inner, 3:])
    # leaf kind matches its refs
    for i in np.where(leaves)[0][:2000]:
        types = set(int(r) >> 30 for r in refs[nd[i, 2]:nd[i, 2] + nd[i, 3]])
        assert nd[i, 0] == (-1 if len(types) > 1 else -2 - types.pop())
    assert st.max_leaf_occupancy <= 8 or st.max_leaf_occupancy >= 1


def test_weighted_lambda_zero_is_uniform():
    spheres, boxes, a_s, a_b = _random_small(np.random.default_rng(3), 50, 50, 10, 10)
    u = _native_scene(mrad.BvhBuildConfig.uniform(4), spheres, boxes, a_s, a_b)
    w0 = _native_scene(mrad.BvhBuildConfig.weighted(4, lam=0.0), spheres, boxes, a_s, a_b)
    assert topology_sha256(u) == topology_sha256(w0)


def test_material_dedup_follows_double_compare():
    sc = mrad.GpuScene()
    m = np.zeros((4, 13))
    m[:, 0:3] = 0.5
    m[2, 1] = -0.0   # Double.compare(-0.0, 0.0) != 0: a distinct material
    m[3, 1] = -0.0
    m[1, 1] = 0.5
    m[2, 0:3] = [0.5, -0.0, 0.5]
    m[3, 0:3] = [0.5, -0.0, 0.5]
    m[0, 1] = 0.0
    m[0, 0:3] = [0.5, 0.0, 0.5]
    m[1, 0:3] = [0.5, 0.0, 0.5]
    sc.add_spheres([[0, 0, 0, 0.1], [1, 0, 0, 0.1], [2, 0, 0, 0.1], [3, 0, 0, 0.1]], m)
    mats = sc.sphere_data()[:, 4]
    assert mats.tolist() == [0.0, 0.0, 1.0, 1.0]
    assert sc.counts()["materials"] == 2


@pytest.mark.parametrize("bad", [
    lambda sc: sc.add_spheres([[0, 0, 0, -1.0]], mrad.material_array((1, 1, 1))),
    lambda sc: sc.add_spheres([[np.inf, 0, 0, 1.0]], mrad.material_array((1, 1, 1))),
    lambda sc: sc.add_boxes([[0, 0, 0, 1, 1, 0]], mrad.material_array((1, 1, 1))),
    lambda sc: sc.add_planes([[0, 0, 0, 0, 0, 0]], mrad.material_array((1, 1, 1))),
    lambda sc: sc.add_affine_transforms("box", [[1e-6, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]], mrad.material_array((1, 1, 1))),
])
def test_invalid_primitives_raise_invalid_argument(bad):
    with pytest.raises(mrad.InvalidArgument):
        bad(mrad.GpuScene())


def test_unweighted_mode_requires_lambda_zero():
    with pytest.raises(mrad.InvalidArgument):
        mrad.GpuScene(mrad.BvhBuildConfig("UNIFORM_SAH", 8, lam=0.5))
    with pytest.raises(ValueError):
        mrad.GpuScene(mrad.BvhBuildConfig("SAH_ROTATIONS", 8))


def test_affine_payload_matches_builder_rules():
    """GpuSceneBuilder.gpuAffine: float-quantized inverse, transpose as normal matrix, centre in slots 3/7/11."""
    sc = mrad.GpuScene()
    t = np.array([[0.0, -2.0, 0.0, 0.3, 1.0, 0.0, 0.0, -0.2, 0.0, 0.0, 0.5, 0.1]])
    sc.add_affine_transforms("sphere", t, mrad.material_array((0.5, 0.5, 0.5)))
    d = sc.affine_sphere_data()[0]
    inv = np.linalg.inv(t[0].reshape(3, 4)[:, :3]).astype(np.float32)
    np.testing.assert_allclose(d[[0, 1, 2, 4, 5, 6, 8, 9, 10]], inv.reshape(-1), rtol=0, atol=1e-7)
    np.testing.assert_array_equal(d[[3, 7, 11]], np.float32([0.3, -0.2, 0.1]))
    np.testing.assert_array_equal(d[12:21], d[[0, 4, 8, 1, 5, 9, 2, 6, 10]])
    nb = sc.bvh_node_bounds()[0]
    # sphere bounds: centre +- row lengths of the forward linear map (2, 1, 0.5), widened by 2 ulp
    np.testing.assert_allclose(nb, [0.3 - 2, -0.2 - 1, 0.1 - 0.5, 0.3 + 2, -0.2 + 1, 0.1 + 0.5], atol=1e-6)
    assert nb[0] < np.float32(0.3 - 2) and nb[3] > np.float32(0.3 + 2)


def test_material_registry_and_packing_limits():
    """materialCount / materialIndexOf (GpuScene.java:493-500): materials are registered once, in first-use order;
    MIN_PACKED_PRIMITIVE_EXTENT (:20-21) is the smallest radius a sphere may keep after float packing."""
    from mrad import scene as S
    red, grey = mrad.material_array((0.8, 0.1, 0.1)), mrad.material_array((0.5, 0.5, 0.5))
    sc = mrad.GpuScene()
    sc.add_spheres([[0, 0, 0, 1], [2, 0, 0, 1], [4, 0, 0, 1]], np.concatenate([red, grey, red]))
    assert sc.material_count() == 2
    assert sc.material_index_of(red[0]) == 0 and sc.material_index_of(grey[0]) == 1
    assert sc.material_index_of(mrad.material_array((0.1, 0.2, 0.3))[0]) == -1
    assert S.MIN_PACKED_PRIMITIVE_EXTENT == float(np.float32(np.sqrt(np.float32(2.0 ** -126))))
    sc.add_spheres([[0, 0, 0, S.MIN_PACKED_PRIMITIVE_EXTENT]], grey)       # at the origin its bounds stay apart
    with pytest.raises(mrad.InvalidArgument, match="Sphere radius"):
        sc.add_spheres([[0, 0, 0, S.MIN_PACKED_PRIMITIVE_EXTENT / 2]], grey)
    with pytest.raises(mrad.InvalidArgument, match="Sphere bounds"):     # at y = 5 they collapse in float
        sc.add_spheres([[0, 5, 0, S.MIN_PACKED_PRIMITIVE_EXTENT]], grey)
    t = mrad.GpuRayTracer(4, 4, 1, True, collect_metrics=True)
    assert t.collects_metrics()
    t.close()
