"""Command-line driver (CPU): the -Dgfxlab.* property checks of
BenchmarkRunner (playground/BenchmarkRunner.java:18-41, 86-188), Long.decode
of the seed, the configuration string, the file name and the BVH build
configuration read from properties (gpu/BvhBuildConfig.java:52-64)."""
from __future__ import annotations

import datetime as dt

import pytest

import mrad
from mrad import benchmark_runner as B

BASE = {"gfxlab.frames": "3", "gfxlab.warmupFrames": "1", "gfxlab.width": "64", "gfxlab.height": "32",
        "gfxlab.maxDepth": "12", "gfxlab.progressEvery": "1", "gfxlab.seed": "0x5eed", "gfxlab.bvhLeafSize": "8",
        "gfxlab.benchmarkProcessId": "2", "gfxlab.benchmarkPhase": "timing", "gfxlab.gpu.samplesPerFrame": "4",
        "gfxlab.versions": "gpu_bvh,GPU_WEIGHTED_BVH", "gfxlab.scenes": "hero-96,OVERLAP_10K",
        "gfxlab.benchmarkDir": "runs"}


def _argv(props):
    return [f"-D{k}={v}" for k, v in props.items()]


def test_plan_from_properties():
    plan = B.Plan(B.parse_properties(_argv(BASE)))
    assert (plan.frames, plan.warmup_frames, plan.width, plan.height, plan.max_depth) == (3, 1, 64, 32, 12)
    assert plan.seed == 0x5EED and plan.versions == ["GPU_BVH", "GPU_WEIGHTED_BVH"]
    assert plan.scenes == ["hero-96", "overlap-10k"] and plan.stack_size == 32 and plan.leaf_batch == 0
    assert not plan.collect_metrics
    text = plan.configuration()
    assert text.startswith("phase:timing,processId:2,width:64,height:32,maxDepth:12,bvhLeafSize:8,"
                           "warmupFrames:1,measuredFrames:3,seed:0x5eed,python:")
    assert text.endswith(",logicalSamplesPerFrame:4,renderPixelsPerLaunch:262144,replayRaysPerLaunch:262144,"
                         "samplesPerPhysicalKernel:1,physicalKernelLaunchesPerLogicalFrame:4,"
                         "kernelNanosSemantics:" + B.KERNEL_NANOS_SEMANTICS)
    assert plan.configs["GPU_BVH"] == mrad.BvhBuildConfig.uniform(8)
    assert plan.configs["GPU_WEIGHTED_BVH"] == mrad.BvhBuildConfig.weighted(8, 1.0)


@pytest.mark.parametrize("key,value,message", [
    ("gfxlab.frames", "0", "must be positive"),
    ("gfxlab.frames", "2.5", "must be an integer"),
    ("gfxlab.frames", "2147483648", "must be an integer"),
    ("gfxlab.warmupFrames", "-1", "must be nonnegative"),
    ("gfxlab.width", "", "Missing required system property gfxlab.width"),
    ("gfxlab.seed", "0x", "signed 64-bit"),
    ("gfxlab.seed", "08", "signed 64-bit"),
    ("gfxlab.seed", "0x8000000000000000", "signed 64-bit"),
    ("gfxlab.benchmarkPhase", "warm", "timing or counters"),
    ("gfxlab.versions", "gpu_bvh,,gpu_linear", "empty, malformed, or repeated"),
    ("gfxlab.versions", "gpu_bvh,gpu_bvh", "empty, malformed, or repeated"),
    ("gfxlab.versions", "cpu_bvh", "runs on the CPU"),
    ("gfxlab.versions", "gpu_fast", "Unknown renderer version"),
    ("gfxlab.scenes", "gi_test", "Unknown scene preset"),
    ("gfxlab.scenes", "hero-1k,HERO_1K", "empty, malformed, or repeated"),
    ("gfxlab.gpu.samplesPerFrame", "0", "must be positive"),
    ("gfxlab.gpu.bvhStackSize", "4", "between 8 and 256"),
    ("gfxlab.gpu.leafBatch", "8", "0 or 16"),
    ("gfxlab.gpu.replayRaysPerLaunch", "262145", "between 1 and 262144"),
    ("gfxlab.gpu.renderPixelsPerLaunch", "0", "between 1 and 262144"),
    ("gfxlab.bvh.lambda", "-1", "nonnegative"),
    ("gfxlab.bvh.boxWeight", "0", "positive"),
    ("gfxlab.benchmarkDir", " ", "Missing required system property"),
])
def test_property_errors(key, value, message):
    props = dict(BASE, **{key: value})
    with pytest.raises(mrad.InvalidArgument, match=message):
        B.Plan(B.parse_properties(_argv(props)))


def test_missing_property_and_positional_argument():
    props = dict(BASE)
    del props["gfxlab.maxDepth"]
    with pytest.raises(mrad.InvalidArgument, match="gfxlab.maxDepth"):
        B.Plan(B.parse_properties(_argv(props)))
    with pytest.raises(mrad.InvalidArgument, match="no positional arguments"):
        B.parse_properties(["hero-96"])


@pytest.mark.parametrize("raw,value", [("0", 0), ("42", 42), ("-42", -42), ("+7", 7), ("0x5EED", 0x5EED),
                                       ("#ff", 255), ("017", 15), ("-0x8000000000000000", -(1 << 63)),
                                       ("9223372036854775807", (1 << 63) - 1)])
def test_seed_long_decode(raw, value):
    assert B.required_long({"gfxlab.seed": raw}, "gfxlab.seed") == value


def test_negative_seed_prints_twos_complement():
    text = B.base_configuration("counters", 1, 8, 8, 4, 8, 0, 1, -1)
    assert ",seed:0xffffffffffffffff," in text


def test_weighted_config_from_properties_and_unweighted_lambda():
    props = dict(BASE, **{"gfxlab.bvh.lambda": "0.5", "gfxlab.bvh.boxWeight": "2.0", "gfxlab.bvhLeafSize": "4"})
    plan = B.Plan(B.parse_properties(_argv(props)))
    w, u = plan.configs["GPU_WEIGHTED_BVH"], plan.configs["GPU_BVH"]
    assert (w.mode, w.leaf_size, w.lam, w.box_weight) == ("WEIGHTED_SAH", 4, 0.5, 2.0)
    assert (u.mode, u.lam, u.box_weight) == ("UNIFORM_SAH", 0.0, 2.0)
    assert w.summary() == ("mode=WEIGHTED_SAH,leafSize=4,interiorTraversal=1.275577,sphere=1.000000,"
                           "box=2.000000,affineSphere=2.552744,affineBox=2.728115,lambda=0.500000")


def test_benchmark_file_name():
    now = dt.datetime(2026, 3, 4, 5, 6, 7, 89_123)
    assert B.benchmark_file("runs", "hero-1m", "GPU_WEIGHTED_BVH", now).as_posix() == \
        "runs/hero-1m-gpu_weighted_bvh-20260304-050607-089.csv"


def test_boolean_property():
    assert B.java_boolean({"k": "TRUE"}, "k") and not B.java_boolean({"k": "1"}, "k") and not B.java_boolean({}, "k")


def test_random_rows_and_layout_id():
    props = dict(BASE, **{"gfxlab.scenes": "F2-100,f4-1k", "gfxlab.layoutId": "7"})
    plan = B.Plan(B.parse_properties(_argv(props)))
    assert plan.scenes == ["f2-100", "f4-1k"] and plan.layout_id == 7
    assert plan.configuration("f2-100").endswith(",layoutId:7") and "layoutId" not in plan.configuration("hero-96")
    sc, cam = B.scene_setup("f2-100", mrad.BvhBuildConfig.uniform(8), 64, 32, 7)
    c = sc.counts()
    assert (c["spheres"], c["boxes"], c["affine_spheres"], c["affine_boxes"]) == (25, 25, 25, 25)
    with pytest.raises(mrad.InvalidArgument, match="layoutId"):
        B.Plan(B.parse_properties(_argv(dict(BASE, **{"gfxlab.layoutId": "100"}))))


def test_failed_pairs_exit_with_status_2(tmp_path, capsys):
    """Without a usable device every (version, scene) pair fails with a message
    on stderr and the run ends with status 2 (BenchmarkRunner.java:71-84)."""
    with mrad.GpuRayTracer(1, 1, 1, True) as t:
        if t.is_available():
            pytest.skip("a GPU is present; the failure path is exercised on the CPU")
    props = dict(BASE, **{"gfxlab.versions": "gpu_bvh", "gfxlab.scenes": "hero-96", "gfxlab.benchmarkDir": str(tmp_path)})
    assert B.main(_argv(props)) == 2
    captured = capsys.readouterr()
    assert "=== Running GPU_BVH on scene hero-96" in captured.out
    assert "Failed to run GPU_BVH on hero-96:" in captured.err


def test_launch_plan_is_the_reference_tile_layout():
    """GpuLaunchProvenance.renderTileLayout: a 1920x1080 frame at the default 262,144 pixels per launch is 8
    tiles of 1920x136 rows, so one sample is 8 physical launches; a cap of 1000 pixels gives 2 x 1080 tiles."""
    assert B.render_tile_count(1920, 1080, 262_144) == 8
    assert B.render_tile_count(1920, 1080, 1000) == 2 * 1080
    assert B.render_tile_count(64, 32, 262_144) == 1
    assert "physicalKernelLaunchesPerLogicalFrame:16," in B.gpu_configuration(1920, 1080, 2)
    assert ",replayRaysPerLaunch:4096," in B.gpu_configuration(1920, 1080, 2, replay_rays_per_launch=4096)
