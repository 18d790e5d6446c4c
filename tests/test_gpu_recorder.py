"""GPU: HeadlessGpuRenderer with a BenchmarkRecorder, as
HeadlessGpuRenderer.render drives it (playground/HeadlessGpuRenderer.java:58-148):
run metadata from the scene and its BVH, the code-object hash, and one row per
measured frame whose counters are the oracle's for that frame's seed."""
from __future__ import annotations

import numpy as np
import pytest

import mrad
from mrad import recorder as R
from mrad import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_headless_renderer_writes_reference_rows(require_gpu, tmp_path):
    sc = W.hero_scene(96, mrad.BvhBuildConfig.weighted(8, 1.0))
    cam = W.hero_gpu_camera()
    width, height, seed = 320, 180, 7
    out = tmp_path / "hero-96.csv"
    print("headless hero-96 320x180, 1 warm-up + 2 recorded frames", flush=True)
    rec = R.BenchmarkRecorder(out, flush_every=1, gpu_info="AMD Instinct MI355X", scene_name="hero-96", echo=False)
    r = mrad.HeadlessGpuRenderer(sc, cam, width, height, W.MAXIMUM_PATH_DEPTH, seed, collect_metrics=True,
                                 bvh_stack_size=W.STACK_CAPACITY)
    img = r.render(2, warmup_frames=1, recorder=rec)
    last = r.tracer.last_frame_stats()
    rec.close()
    assert np.isfinite(img).all()

    lines = out.read_text().splitlines()
    st = sc.bvh_stats()
    head = lines[0]
    for item in ("metricSchemaVersion=5", f"bvhBytes={st.bytes}", f"accelerationNodeCount={st.node_count}",
                 f"leafCount={st.leaf_count}", f"maxTreeDepth={st.max_depth}", "primitiveCount=105",
                 f"accelerationConfig={sc.config.summary()}",
                 f"compiledPtxSha256={mrad.GpuRayTracer.compiled_code_object_sha256()}"):
        assert item in head, item
    assert lines[1] == ",".join(R.COLUMNS)
    rows = [dict(zip(R.COLUMNS, line.split(","))) for line in lines[2:]]
    assert len(rows) == 2 and all(len(line.split(",")) == 34 for line in lines[2:])

    # the last row is frame 2: its counters are the oracle's for that frame
    _, ref = O.render_pixels(O.Scene(sc.wire(), stack_size=W.STACK_CAPACITY), cam.packed(), width, height,
                             W.MAXIMUM_PATH_DEPTH, 2, mrad.frame_seed(seed, 2), 1)
    ref = dict(zip(mrad.STAT_NAMES, map(int, ref)))
    assert {k: last[k] for k in mrad.STAT_NAMES} == ref
    row = rows[-1]
    for col, key in (("pathSegments", "rays"), ("primitiveTests", "primitive_tests"), ("aabbTests", "aabb_tests"),
                     ("internalNodeVisits", "internal_node_visits"), ("leafNodeVisits", "leaf_node_visits"),
                     ("sphereTests", "sphere_tests"), ("boxTests", "box_tests"),
                     ("affineBoxTests", "affine_box_tests"), ("maxStackSize", "max_stack_size")):
        assert int(row[col]) == ref[key], col
    assert int(row["cameraSamples"]) == width * height
    assert int(row["secondaryPathSegments"]) == ref["rays"] - width * height
    assert int(row["rootAabbTests"]) == ref["aabb_tests"] - 2 * ref["internal_node_visits"]
    assert row["traversalWork"] == R.java_fixed(R.traversal_work_of(ref), 4)
    assert row["uploadMs"] == "0.0000" and float(row["kernelMs"]) > 0.0


def test_benchmark_runner_writes_one_file_per_pair(require_gpu, tmp_path, capsys):
    """The command-line driver over two versions and one scene, counters on:
    one CSV per pair with a row per measured frame, exit status 0."""
    from mrad import benchmark_runner as B
    props = {"gfxlab.frames": "2", "gfxlab.warmupFrames": "1", "gfxlab.width": "64", "gfxlab.height": "32",
             "gfxlab.maxDepth": "12", "gfxlab.progressEvery": "1", "gfxlab.seed": "7", "gfxlab.bvhLeafSize": "8",
             "gfxlab.benchmarkProcessId": "1", "gfxlab.benchmarkPhase": "counters",
             "gfxlab.gpu.samplesPerFrame": "2", "gfxlab.gpu.collectMetrics": "true", "gfxlab.gpu.leafBatch": "16",
             "gfxlab.versions": "gpu_linear,gpu_weighted_bvh", "gfxlab.scenes": "hero-96",
             "gfxlab.benchmarkDir": str(tmp_path)}
    print("benchmark runner: hero-96 64x32, GPU_LINEAR and GPU_WEIGHTED_BVH", flush=True)
    assert B.main([f"-D{k}={v}" for k, v in props.items()]) == 0
    files = sorted(tmp_path.glob("*.csv"))
    assert [f.name.rsplit("-", 3)[0] for f in files] == ["hero-96-gpu_linear", "hero-96-gpu_weighted_bvh"]
    for f, mode in zip(files, ("mode=LINEAR", "mode=WEIGHTED_SAH")):
        lines = f.read_text().splitlines()
        assert f"accelerationConfig={mode}" in lines[0] and "config=phase:counters," in lines[0]
        assert ",samplesPerPhysicalKernel:1,physicalKernelLaunchesPerLogicalFrame:2," in lines[0]
        assert lines[1] == ",".join(R.COLUMNS) and len(lines) == 4
        rows = [dict(zip(R.COLUMNS, line.split(","))) for line in lines[2:]]
        assert all(r["count"] == "2" and int(r["cameraSamples"]) == 64 * 32 * 2 for r in rows)
    linear, bvh = ([dict(zip(R.COLUMNS, f.read_text().splitlines()[3].split(",")))] for f in files)
    assert int(linear[0]["aabbTests"]) == 0 and int(bvh[0]["aabbTests"]) > 0
    assert "=== Running GPU_WEIGHTED_BVH on scene hero-96" in capsys.readouterr().out


def test_benchmark_runner_random_row(require_gpu, tmp_path):
    """A random row through the driver: the layout named by gfxlab.layoutId,
    its identifier in the configuration, and the frame's counters equal the
    oracle's for that layout and frame seed."""
    from mrad import benchmark_runner as B
    from mrad import random_scenes as RS
    props = {"gfxlab.frames": "1", "gfxlab.warmupFrames": "0", "gfxlab.width": "96", "gfxlab.height": "54",
             "gfxlab.maxDepth": "12", "gfxlab.progressEvery": "1", "gfxlab.seed": "11", "gfxlab.bvhLeafSize": "8",
             "gfxlab.benchmarkProcessId": "1", "gfxlab.benchmarkPhase": "counters",
             "gfxlab.gpu.samplesPerFrame": "1", "gfxlab.gpu.collectMetrics": "true", "gfxlab.layoutId": "3",
             "gfxlab.versions": "gpu_bvh", "gfxlab.scenes": "f2-100", "gfxlab.benchmarkDir": str(tmp_path)}
    print("benchmark runner: f2-100 layout 3, 96x54", flush=True)
    assert B.main([f"-D{k}={v}" for k, v in props.items()]) == 0
    (f,) = tmp_path.glob("f2-100-gpu_bvh-*.csv")
    lines = f.read_text().splitlines()
    assert ",layoutId:3" in lines[0] and len(lines) == 3
    row = dict(zip(R.COLUMNS, lines[2].split(",")))
    sc = RS.random_scene("F2", 100, 3, mrad.BvhBuildConfig.uniform(8))
    _, ref = O.render_pixels(O.Scene(sc.wire(), stack_size=32), RS.random_gpu_camera(96, 54).packed(), 96, 54,
                             12, 0, mrad.frame_seed(11, 0), 1)
    ref = dict(zip(mrad.STAT_NAMES, map(int, ref)))
    assert int(row["pathSegments"]) == ref["rays"] and int(row["aabbTests"]) == ref["aabb_tests"]
    assert int(row["affineSphereTests"]) == ref["affine_sphere_tests"] > 0
