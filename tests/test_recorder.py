"""BenchmarkRecorder mirror (CPU): schema, Java number formatting, the
recorder's checks, the traversal-work cost against the published study, and
BVH statistics against the published node-count, depth and byte matrices.

* traversal_work: core8-detail.csv's medianModeledWork of every single-layout
  row (48 rows, eight methods) is TraversalWorkMetric.cost of that row's
  published counters; ours must agree to double rounding.  A wrong weight or a
  counter paired with the wrong weight moves it by orders of magnitude more.
* BVH statistics: bvh-node-count-matrix.csv, bvh-depth-matrix.csv and
  bvh-bytes-matrix.csv (uniform and weighted SAH, leaf 8) for the Layered Hero
  scales and overlap-10k.
* Number formatting: Java's Formatter rounds the shortest decimal digits of a
  double half-up (String.format("%.1f", 0.15) is "0.2", "%.2f" of 1.005 is
  "1.01", of 0.125 is "0.13"), where Python's % rounds the binary value half-even.
"""
from __future__ import annotations

import json
import math
from pathlib import Path

import pytest

import mrad
from mrad import recorder as R
from mrad import workloads as W

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_java_fixed_rounds_shortest_digits_half_up():
    assert R.java_fixed(0.15, 1) == "0.2"
    assert R.java_fixed(1.005, 2) == "1.01"
    assert R.java_fixed(0.125, 2) == "0.13"
    assert R.java_fixed(1.03125, 4) == "1.0313"
    assert R.java_fixed(2.5, 4) == "2.5000"
    assert R.java_fixed(-0.0, 4) == "-0.0000"
    assert R.java_fixed(-1e-9, 4) == "-0.0000"
    assert R.java_fixed(1e30, 4) == "1000000000000000000000000000000.0000"
    assert R.java_fixed(float("nan"), 4) == "NaN" and R.java_fixed(float("inf"), 4) == "Infinity"


def test_java_double_string():
    cases = [(1e7, "1.0E7"), (1.2345e-5, "1.2345E-5"), (0.001, "0.001"), (123.0, "123.0"),
             (9999999.0, "9999999.0"), (1.1428571428571428, "1.1428571428571428"), (-2.5e-4, "-2.5E-4"),
             (123456789.0, "1.23456789E8"), (0.0, "0.0"), (-0.0, "-0.0")]
    for v, text in cases:
        assert R.java_double_string(v) == text, v


def test_schema_columns():
    assert len(R.COLUMNS) == 34 and len(set(R.COLUMNS)) == 34
    assert R.COLUMNS[:2] == ("count", "totalMs") and R.COLUMNS[-2:] == ("stackOverflows", "maxStackSize")


@pytest.mark.parametrize("row", json.loads((GOLDEN / "modeled_work.json").read_text()),
                         ids=lambda r: f"{r['rowId']}-{r['method']}")
def test_traversal_work_matches_published_modeled_work(row):
    ours = R.traversal_work(row["medianAabbTests"], row["medianSphereTests"], row["medianBoxTests"],
                            row["medianPlaneTests"], row["medianAffineSphereTests"], row["medianAffineBoxTests"])
    assert math.isclose(ours, float(row["medianModeledWork"]), rel_tol=1e-15, abs_tol=0.0)


def _scene(row_id: str, method: str):
    cfg = mrad.BvhBuildConfig.uniform(8) if method == "uniform" else mrad.BvhBuildConfig.weighted(8, 1.0)
    if row_id == "overlap-10k":
        return W.overlap_scene(cfg)
    return W.hero_scene({"hero-96": 96, "hero-1k": 1_000, "hero-10k": 10_000, "hero-100k": 100_000}[row_id], cfg)


@pytest.mark.parametrize("row", [r for r in json.loads((GOLDEN / "bvh_matrix.json").read_text())
                                 if r["rowId"] in ("hero-96", "hero-1k", "hero-10k", "hero-100k", "overlap-10k")],
                         ids=lambda r: f"{r['rowId']}-{r['method']}")
def test_bvh_stats_match_published_matrices(row):
    st = _scene(row["rowId"], row["method"]).bvh_stats()
    assert (st.node_count, st.max_depth, st.bytes) == (row["nodeCount"], row["maxDepth"], row["bytes"])


def test_frame_metrics_checks():
    with pytest.raises(mrad.InvalidArgument):
        R.FrameMetrics(1, 10, kernel_nanos=5, maximum_physical_kernel_nanos=6)
    with pytest.raises(mrad.InvalidArgument):
        R.FrameMetrics(1, 10, kernel_nanos=-1, maximum_physical_kernel_nanos=0)
    with pytest.raises(mrad.InvalidArgument):
        R.FrameMetrics(1, 10, leaf_node_visits=5, homogeneous_leaf_node_visits=3, mixed_leaf_node_visits=1)
    R.FrameMetrics(1, 10, kernel_nanos=5, maximum_physical_kernel_nanos=5, leaf_node_visits=4,
                   homogeneous_leaf_node_visits=3, mixed_leaf_node_visits=1)
    with pytest.raises(mrad.IllegalState):
        R.root_aabb_tests({"aabb_tests": 9, "internal_node_visits": 5})
    assert R.root_aabb_tests({"aabb_tests": 11, "internal_node_visits": 5}) == 1


def _frame(**kw):
    base = dict(count=4, total_nanos=12_345_678, kernel_nanos=10_000_050, maximum_physical_kernel_nanos=2_500_000,
                copy_nanos=125_000, upload_nanos=0, width=64, height=32, camera_samples=8192, path_segments=20480,
                primitive_tests=30000, aabb_tests=400000, sphere_tests=20000, box_tests=6000, plane_tests=0,
                affine_sphere_tests=3000, affine_box_tests=1000, root_aabb_tests=20480, traversal_work=None,
                stack_overflows=0, max_stack_size=12, internal_node_visits=189760, leaf_node_visits=25000,
                homogeneous_leaf_node_visits=25000, mixed_leaf_node_visits=0, backend_nanos=11_000_000)
    base.update(kw)
    if base["traversal_work"] is None:
        base["traversal_work"] = R.traversal_work(base["aabb_tests"], base["sphere_tests"], base["box_tests"],
                                                  base["plane_tests"], base["affine_sphere_tests"],
                                                  base["affine_box_tests"])
    return R.FrameMetrics(**base)


def test_recorder_file_rows_and_state(tmp_path):
    out = tmp_path / "run" / "bench.csv"
    rec = R.BenchmarkRecorder(out, flush_every=2, cpu_info="AMD EPYC", gpu_info="MI355X", version="gpu-bvh",
                              scene_name="hero-96", configuration="1920x1080", echo=False)
    with pytest.raises(mrad.InvalidArgument):
        rec.set_compiled_code_object_sha256("abc")
    rec.set_compiled_code_object_sha256("AB" * 32)
    rec.set_run_metadata(R.RunMetadata(scene_bytes=100, bvh_bytes=7504, mean_leaf_occupancy=1.1797752808988764,
                                       uniform_sah_cost=12345678.9, acceleration_config="mode=LINEAR"))
    rec.record(_frame())
    assert not out.exists()                       # buffered until flush_every rows
    rec.record(_frame(camera_samples=None, path_segments=None, traversal_work=None))
    lines = out.read_text().splitlines()
    assert lines[0] == ("# cpu=AMD EPYC; gpu=MI355X; version=gpu-bvh; scene=hero-96; config=1920x1080; "
                        "metricSchemaVersion=5; sceneBytes=100; bvhBytes=7504; meanLeafOccupancy=1.1797752808988764; "
                        "uniformSahCost=1.23456789E7; accelerationConfig=mode=LINEAR; compiledPtxSha256=" + "ab" * 32)
    assert lines[1] == ",".join(R.COLUMNS)
    row = dict(zip(R.COLUMNS, lines[2].split(",")))
    assert len(lines[2].split(",")) == 34
    assert row["totalMs"] == "12.3457" and row["kernelMs"] == "10.0001" and row["copyMs"] == "0.1250"
    assert row["backendMs"] == "11.0000" and row["uploadMs"] == "0.0000"
    assert row["secondaryPathSegments"] == "12288"
    assert row["primitiveTestsPerPathSegment"] == R.java_fixed(30000 / 20480, 4) == "1.4648"
    assert row["aabbTestsPerCameraSample"] == R.java_fixed(400000 / 8192, 4)
    assert row["rootAabbTestsPerPathSegment"] == "1.0000"
    tw = R.traversal_work(400000, 20000, 6000, 0, 3000, 1000)
    assert row["traversalWork"] == R.java_fixed(tw, 4)
    assert row["traversalWorkPerCameraSample"] == R.java_fixed(tw / 8192, 4)
    row2 = dict(zip(R.COLUMNS, lines[3].split(",")))
    assert row2["cameraSamples"] == row2["pathSegments"] == row2["secondaryPathSegments"] == ""
    assert row2["primitiveTestsPerPathSegment"] == "" and row2["traversalWork"] != ""
    with pytest.raises(mrad.IllegalState):
        rec.set_run_metadata(R.RunMetadata())
    with pytest.raises(mrad.IllegalState):
        rec.set_compiled_code_object_sha256("cd" * 32)
    rec.record(_frame(count=1))
    rec.close()
    assert len(out.read_text().splitlines()) == 5


def test_bench_line_and_timing_record(capsys):
    rec = R.BenchmarkRecorder(None)
    rec.record_timing(1, 8, 4, 2_000_000, 1_500_000, None)
    rec.record(_frame())
    out = capsys.readouterr().out.splitlines()
    assert out[0] == "[BENCH]: count=1 total=2.000 ms, kernel=1.500 ms (8x4)"
    assert out[1] == ("[BENCH]: count=4 total=12.346 ms, backend=11.000 ms, kernel=10.000 ms, max-kernel=2.500 ms, "
                      "copy=0.125 ms, upload=0.000 ms, prim/segment=1.46, aabb/segment=19.53 (64x32)")


def test_bvh_config_summary():
    assert mrad.BvhBuildConfig.uniform(8).summary() == (
        "mode=UNIFORM_SAH,leafSize=8,interiorTraversal=1.275577,sphere=1.000000,box=1.288952,"
        "affineSphere=2.552744,affineBox=2.728115,lambda=0.000000")
    assert mrad.BvhBuildConfig.weighted(4, 1.0).summary().endswith(",lambda=1.000000")


def test_scene_metrics_of_hero_96():
    sc = W.hero_scene(96)
    m = R.scene_metrics(sc)
    c = sc.counts()
    assert m.primitive_count == 105 and m.sphere_count == c["spheres"] and m.material_count == c["materials"]
    assert m.scene_bytes == 4 * (13 * c["materials"] + 5 * c["spheres"] + 7 * c["boxes"] + 22 * c["affine_boxes"])
