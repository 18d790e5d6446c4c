"""Command-line benchmark driver, after BenchmarkRunner.main
(playground/BenchmarkRunner.java:18-84): the same ``-Dgfxlab.*`` properties,
one BenchmarkRecorder CSV per (renderer version, scene) pair, flushed every
20 rows, and exit status 2 when any pair failed.

    python -m mrad.benchmark_runner -Dgfxlab.frames=8 -Dgfxlab.warmupFrames=2 \\
        -Dgfxlab.width=1920 -Dgfxlab.height=1080 -Dgfxlab.maxDepth=12 \\
        -Dgfxlab.progressEvery=4 -Dgfxlab.seed=0x5eed -Dgfxlab.bvhLeafSize=8 \\
        -Dgfxlab.benchmarkProcessId=1 -Dgfxlab.benchmarkPhase=timing \\
        -Dgfxlab.gpu.samplesPerFrame=1 -Dgfxlab.versions=gpu_weighted_bvh \\
        -Dgfxlab.scenes=hero-1m -Dgfxlab.benchmarkDir=runs/

Renderer versions are the GPU ones of RendererFactory.Version
(playground/RendererFactory.java:8-14): GPU_LINEAR, GPU_BVH (uniform SAH) and
GPU_WEIGHTED_BVH (weighted SAH, BvhBuildConfig.fromProperties with the
``gfxlab.bvh.*`` weights, gpu/BvhBuildConfig.java:52-64).  CPU_LINEAR and
CPU_BVH are rejected: this package renders on the GPU only.  The scenes are
the benchmark study's workloads (hero-96 to hero-1m, overlap-10k and the
random rows f1-100 to f4-10k) with their published cameras, in place of the
interactive SceneCatalog presets.

``gfxlab.gpu.collectMetrics=true`` turns on the traversal counters (the
reference's counters phase), ``gfxlab.gpu.bvhStackSize`` sets the traversal
stack (8..256, GpuRayTracer.java:1100-1103) and ``gfxlab.gpu.leafBatch`` (0 or
16) the megakernel's leaf-batched traversal.  The random rows (f1-100 to f4-10k)
render the layout that ``gfxlab.layoutId`` names (0..99, default 0).  Every
frame is one launch over the whole image, so the launch fields of the
configuration string describe that.
"""
from __future__ import annotations

import datetime as _dt
import os
import platform
import re
import sys
from pathlib import Path

from . import _native as N
from .recorder import BenchmarkRecorder
from .scene import BvhBuildConfig
from .tracer import GpuRayTracer, HeadlessGpuRenderer

BENCHMARK_FLUSH_EVERY = 20                   # BenchmarkRunner.java:13
GPU_VERSIONS = ("GPU_LINEAR", "GPU_BVH", "GPU_WEIGHTED_BVH")
CPU_VERSIONS = ("CPU_LINEAR", "CPU_BVH")
SCENES = ("hero-96", "hero-1k", "hero-10k", "hero-100k", "hero-1m", "overlap-10k",
          *(f"{f.lower()}-{n}" for f in ("F1", "F2", "F3", "F4") for n in ("100", "1k", "10k")))
SAMPLES_PER_PHYSICAL_KERNEL = 1
KERNEL_NANOS_SEMANTICS = "one HIP-event interval around the complete physical-launch sequence of a logical frame"
_INT = re.compile(r"[+-]?[0-9]+")


def _invalid(message: str) -> N.InvalidArgument:
    return N.InvalidArgument(N.MRAD_ERR_INVALID_ARGUMENT, message)


def parse_properties(argv) -> dict:
    """``-Dkey=value`` arguments; anything else is an error, as BenchmarkRunner
    takes no positional arguments (BenchmarkRunner.java:19-22)."""
    props = {}
    for arg in argv:
        if not arg.startswith("-D") or "=" not in arg:
            raise _invalid("BenchmarkRunner accepts no positional arguments; use gfxlab system properties.")
        key, value = arg[2:].split("=", 1)
        props[key] = value
    return props


def required(props: dict, key: str) -> str:
    """BenchmarkRunner.requiredProperty (:142-149)."""
    value = props.get(key)
    if value is None or not value.strip():
        raise _invalid(f"Missing required system property {key}.")
    return value.strip()


def _int32(props: dict, key: str) -> int:
    raw = required(props, key)
    if not _INT.fullmatch(raw) or not -(1 << 31) <= int(raw) < (1 << 31):
        raise _invalid(f"{key} must be an integer.")
    return int(raw)


def required_positive_int(props: dict, key: str) -> int:
    value = _int32(props, key)
    if value <= 0:
        raise _invalid(f"{key} must be positive.")
    return value


def required_non_negative_int(props: dict, key: str) -> int:
    value = _int32(props, key)
    if value < 0:
        raise _invalid(f"{key} must be nonnegative.")
    return value


def optional_int(props: dict, key: str, default: int) -> int:
    """Integer.getInteger: the default when the property is absent."""
    return default if key not in props else _int32(props, key)


def required_long(props: dict, key: str) -> int:
    """Long.decode (:180-188): optional sign, then decimal, 0x / 0X / # hex or
    0-prefixed octal; the result must fit a signed 64-bit integer."""
    raw = required(props, key)
    m = re.fullmatch(r"([+-]?)(0[xX][0-9a-fA-F]+|#[0-9a-fA-F]+|0[0-7]*|[1-9][0-9]*)", raw)
    if m is None:
        raise _invalid(f"{key} must be a signed 64-bit integer.")
    sign, body = m.groups()
    if body[:2] in ("0x", "0X"):
        value = int(body[2:], 16)
    elif body[0] == "#":
        value = int(body[1:], 16)
    elif len(body) > 1 and body[0] == "0":
        value = int(body[1:], 8)
    else:
        value = int(body)
    value = -value if sign == "-" else value
    if not -(1 << 63) <= value < (1 << 63):
        raise _invalid(f"{key} must be a signed 64-bit integer.")
    return value


def split_unique(configured: str, label: str) -> list:
    """BenchmarkRunner.splitUnique (:127-140)."""
    values = [v.strip() for v in configured.split(",")]
    if not values or any(not v for v in values) or len(set(values)) != len(values):
        raise _invalid(f"Configured {label} list is empty, malformed, or repeated: {configured}")
    return values


def configured_versions(props: dict) -> list:
    out = []
    for name in split_unique(required(props, "gfxlab.versions"), "renderer version"):
        version = name.upper()
        if version in CPU_VERSIONS:
            raise _invalid(f"Renderer version {name} runs on the CPU; this package provides {', '.join(GPU_VERSIONS)}")
        if version not in GPU_VERSIONS:
            raise _invalid(f"Unknown renderer version: {name}")
        out.append(version)
    return out


def configured_scenes(props: dict) -> list:
    out = []
    for name in split_unique(required(props, "gfxlab.scenes"), "scene preset"):
        scene = name.lower().replace("_", "-")
        if scene not in SCENES:
            raise _invalid(f"Unknown scene preset: {name}")
        out.append(scene)
    if len(set(out)) != len(out):
        raise _invalid(f"Configured scene preset list is empty, malformed, or repeated: {props['gfxlab.scenes']}")
    return out


def _positive_float(props: dict, key: str, default: float, allow_zero: bool = False) -> float:
    raw = props.get(key)
    if raw is None:
        return default
    try:
        value = float(raw.strip())
    except ValueError:
        raise _invalid(f"{key} must be a number.") from None
    if not (value >= 0.0 if allow_zero else value > 0.0) or value != value or value == float("inf"):
        raise _invalid(f"{key} must be {'nonnegative' if allow_zero else 'positive'}.")
    return value


def bvh_config(props: dict, version: str, leaf_size: int) -> BvhBuildConfig:
    """BvhBuildConfig.fromProperties (BvhBuildConfig.java:52-64) for the
    version's build mode; lambda is zero for an unweighted build."""
    d = BvhBuildConfig()
    weighted = version == "GPU_WEIGHTED_BVH"
    return BvhBuildConfig(
        "WEIGHTED_SAH" if weighted else "UNIFORM_SAH", leaf_size,
        traversal_weight=_positive_float(props, "gfxlab.bvh.traversalWeight", d.traversal_weight),
        sphere_weight=_positive_float(props, "gfxlab.bvh.sphereWeight", d.sphere_weight),
        box_weight=_positive_float(props, "gfxlab.bvh.boxWeight", d.box_weight),
        affine_sphere_weight=_positive_float(props, "gfxlab.bvh.affineSphereWeight", d.affine_sphere_weight),
        affine_box_weight=_positive_float(props, "gfxlab.bvh.affineBoxWeight", d.affine_box_weight),
        lam=_positive_float(props, "gfxlab.bvh.lambda", 1.0, allow_zero=True) if weighted else 0.0)


def java_boolean(props: dict, key: str) -> bool:
    """Boolean.getBoolean: true only for the text "true", any case."""
    return props.get(key, "").strip().lower() == "true"


def base_configuration(phase: str, process_id: int, width: int, height: int, max_depth: int, leaf_size: int,
                       warmup_frames: int, frames: int, seed: int) -> str:
    """The configuration string of BenchmarkRunner.java:38-41; the runtime
    field names the Python interpreter that drives the run."""
    return (f"phase:{phase},processId:{process_id},width:{width},height:{height},maxDepth:{max_depth},"
            f"bvhLeafSize:{leaf_size},warmupFrames:{warmup_frames},measuredFrames:{frames},"
            f"seed:0x{seed & 0xFFFFFFFFFFFFFFFF:x},python:{platform.python_version()}")


MAX_RENDER_PIXELS_PER_LAUNCH = 262_144      # GpuLaunchProvenance.java:5
MAX_REPLAY_RAYS_PER_LAUNCH = 262_144        # GpuLaunchProvenance.java:6


def render_tile_count(width: int, height: int, pixels_per_launch: int) -> int:
    """GpuLaunchProvenance.renderTileLayout (:44-60): tiles of min(W, n) x max(1, n / tile width) pixels."""
    tw = min(width, pixels_per_launch)
    th = min(height, max(1, pixels_per_launch // tw))
    return (1 + (width - 1) // tw) * (1 + (height - 1) // th)


def gpu_configuration(width: int, height: int, samples_per_frame: int,
                      pixels_per_launch: int = MAX_RENDER_PIXELS_PER_LAUNCH,
                      replay_rays_per_launch: int = MAX_REPLAY_RAYS_PER_LAUNCH) -> str:
    """The launch provenance appended for GPU versions (BenchmarkRunner.java:48-59).  As in the reference
    (GpuLaunchProvenance: SAMPLES_PER_PHYSICAL_KERNEL = 1, physicalRenderLaunches = tiles x samples), each
    physical launch renders one sample of one pixel tile of at most renderPixelsPerLaunch pixels; run()
    builds its renderers with samples_per_launch = 1 and that tile size to match."""
    launches = render_tile_count(width, height, pixels_per_launch) * samples_per_frame
    return (f",logicalSamplesPerFrame:{samples_per_frame},renderPixelsPerLaunch:{pixels_per_launch},"
            f"replayRaysPerLaunch:{replay_rays_per_launch},samplesPerPhysicalKernel:{SAMPLES_PER_PHYSICAL_KERNEL},"
            f"physicalKernelLaunchesPerLogicalFrame:{launches},kernelNanosSemantics:{KERNEL_NANOS_SEMANTICS}")


def benchmark_file(directory: str, scene: str, version: str, now: _dt.datetime | None = None) -> Path:
    """BenchmarkRunner.buildBenchmarkFile (:86-91):
    <scene>-<version>-<yyyyMMdd-HHmmss-SSS>.csv under gfxlab.benchmarkDir."""
    now = now or _dt.datetime.now()
    stamp = now.strftime("%Y%m%d-%H%M%S-") + f"{now.microsecond // 1000:03d}"
    return Path(directory) / f"{scene}-{version.lower()}-{stamp}.csv"


def scene_setup(name: str, config: BvhBuildConfig, width: int, height: int, layout_id: int = 0):
    """The scene and camera of one benchmark workload; a random row (f1-100 ..
    f4-10k) takes its layout identifier."""
    from . import workloads as W
    if name.startswith("f"):
        from . import random_scenes as RS
        family, size = name.upper().split("-")
        count = {"100": 100, "1K": 1_000, "10K": 10_000}[size]
        return RS.random_scene(family, count, layout_id, config), RS.random_gpu_camera(width, height)
    if name == "overlap-10k":
        return W.overlap_scene(config), W.gpu_camera(W.overlap_camera(), width, height)
    scale = {"hero-96": 96, "hero-1k": 1_000, "hero-10k": 10_000, "hero-100k": 100_000, "hero-1m": 1_000_000}[name]
    return W.hero_scene(scale, config), W.hero_gpu_camera(width, height)


def host_cpu() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown CPU"


def host_gpu() -> str:
    """Name of HIP device 0 from the library's own mrad_device_info; "unknown GPU" without one."""
    try:
        with GpuRayTracer(1, 1, 1, True) as t:
            if t.is_available():
                name, arch = t.device_info()
                return name or arch or "unknown GPU"
    except N.MradError:
        pass
    return "unknown GPU"


class Plan:
    """Every property checked before the first render, in BenchmarkRunner's order."""

    def __init__(self, props: dict):
        self.frames = required_positive_int(props, "gfxlab.frames")
        self.warmup_frames = required_non_negative_int(props, "gfxlab.warmupFrames")
        self.width = required_positive_int(props, "gfxlab.width")
        self.height = required_positive_int(props, "gfxlab.height")
        self.max_depth = required_non_negative_int(props, "gfxlab.maxDepth")
        self.progress_every = required_positive_int(props, "gfxlab.progressEvery")
        self.seed = required_long(props, "gfxlab.seed")
        self.leaf_size = required_positive_int(props, "gfxlab.bvhLeafSize")
        self.process_id = required_positive_int(props, "gfxlab.benchmarkProcessId")
        self.phase = required(props, "gfxlab.benchmarkPhase")
        if self.phase not in ("timing", "counters"):
            raise _invalid("gfxlab.benchmarkPhase must be timing or counters.")
        self.base = base_configuration(self.phase, self.process_id, self.width, self.height, self.max_depth,
                                       self.leaf_size, self.warmup_frames, self.frames, self.seed)
        self.versions = configured_versions(props)
        self.scenes = configured_scenes(props)
        self.samples_per_frame = required_positive_int(props, "gfxlab.gpu.samplesPerFrame")
        self.collect_metrics = java_boolean(props, "gfxlab.gpu.collectMetrics")
        self.stack_size = optional_int(props, "gfxlab.gpu.bvhStackSize", 32)
        self.pixels_per_launch = optional_int(props, "gfxlab.gpu.renderPixelsPerLaunch", MAX_RENDER_PIXELS_PER_LAUNCH)
        if not 1 <= self.pixels_per_launch <= MAX_RENDER_PIXELS_PER_LAUNCH:
            raise _invalid(f"gfxlab.gpu.renderPixelsPerLaunch must be between 1 and {MAX_RENDER_PIXELS_PER_LAUNCH}")
        self.replay_rays_per_launch = optional_int(props, "gfxlab.gpu.replayRaysPerLaunch", MAX_REPLAY_RAYS_PER_LAUNCH)
        if not 1 <= self.replay_rays_per_launch <= MAX_REPLAY_RAYS_PER_LAUNCH:
            raise _invalid(f"gfxlab.gpu.replayRaysPerLaunch must be between 1 and {MAX_REPLAY_RAYS_PER_LAUNCH}")
        if not 8 <= self.stack_size <= 256:
            raise _invalid("gfxlab.gpu.bvhStackSize must be between 8 and 256")
        self.leaf_batch = optional_int(props, "gfxlab.gpu.leafBatch", 0)
        if self.leaf_batch not in (0, 16):
            raise _invalid("gfxlab.gpu.leafBatch must be 0 or 16")
        self.layout_id = optional_int(props, "gfxlab.layoutId", 0)
        if not 0 <= self.layout_id <= 99:
            raise _invalid("gfxlab.layoutId must be between 0 and 99")
        self.directory = required(props, "gfxlab.benchmarkDir")
        self.configs = {v: bvh_config(props, v, self.leaf_size) for v in self.versions}

    def configuration(self, scene_name: str = "") -> str:
        text = self.base + gpu_configuration(self.width, self.height, self.samples_per_frame, self.pixels_per_launch,
                                             self.replay_rays_per_launch)
        return text + f",layoutId:{self.layout_id}" if scene_name.startswith("f") else text


def run(plan: Plan, out=None, err=None) -> int:
    """The version x scene loop of BenchmarkRunner.java:44-84; returns the
    number of failed pairs.  ``out`` and ``err`` default to the streams
    current at the call."""
    out, err = out or sys.stdout, err or sys.stderr
    cpu, gpu = host_cpu(), host_gpu()
    failures = 0
    for version in plan.versions:
        for scene_name in plan.scenes:
            print(f"=== Running {version} on scene {scene_name} ({plan.warmup_frames} warm-up + {plan.frames} "
                  f"measured frames, {plan.width}x{plan.height}) ===", file=out, flush=True)
            try:
                scene, camera = scene_setup(scene_name, plan.configs[version], plan.width, plan.height,
                                            plan.layout_id)
                path = benchmark_file(plan.directory, scene_name, version)
                with BenchmarkRecorder(path, BENCHMARK_FLUSH_EVERY, cpu, gpu, version.lower(), scene_name,
                                       plan.configuration(scene_name)) as recorder:
                    renderer = HeadlessGpuRenderer(scene, camera, plan.width, plan.height, plan.max_depth,
                                                   plan.seed, plan.samples_per_frame, version != "GPU_LINEAR",
                                                   collect_metrics=plan.collect_metrics,
                                                   bvh_stack_size=plan.stack_size,
                                                   samples_per_launch=SAMPLES_PER_PHYSICAL_KERNEL,
                                                   pixels_per_launch=plan.pixels_per_launch)
                    try:
                        if plan.leaf_batch:
                            renderer.tracer.set_leaf_batch(plan.leaf_batch)
                        renderer.render(plan.frames, plan.warmup_frames, plan.progress_every, recorder)
                    finally:
                        renderer.tracer.close()
            except (N.MradError, OSError, ValueError) as ex:
                print(f"Failed to run {version} on {scene_name}: {ex}", file=err, flush=True)
                failures += 1
    return failures


def main(argv=None) -> int:
    plan = Plan(parse_properties(sys.argv[1:] if argv is None else argv))
    os.makedirs(plan.directory, exist_ok=True)
    return 2 if run(plan) > 0 else 0


if __name__ == "__main__":
    sys.exit(main())
