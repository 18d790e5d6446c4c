"""Per-frame benchmark rows in the reference's CSV schema.

BenchmarkRecorder mirrors benchmark/BenchmarkRecorder.java (metricSchemaVersion
5): a `# key=value; ...` provenance line, the 34-column header, one row per
recorded frame with the derived per-segment and per-sample ratios, and the
same argument and state checks.  traversal_work restates
gpu/TraversalWorkMetric.java, scene_metrics benchmark/SceneBenchmarkMetrics.java,
and root_aabb_tests HeadlessGpuRenderer.rootAabbTests
(playground/HeadlessGpuRenderer.java:150-156), so a run of this library's
HeadlessGpuRenderer writes the file the reference's would.

Numbers are written as Java writes them: `String.format(Locale.ROOT, "%.4f")`
rounds the shortest decimal digits of the double half-up (so 0.15 at one
decimal is "0.2"), and metadata values use Double.toString (scientific
notation outside [1e-3, 1e7)).
"""
from __future__ import annotations

import math
import sys
from dataclasses import dataclass, fields
from decimal import ROUND_HALF_UP, Decimal, localcontext
from pathlib import Path

from . import _native as N
from .scene import AFFINE_BOX, AFFINE_SPHERE, BOX, NODE_AABB, PLANE, SPHERE

METRIC_SCHEMA_VERSION = 5

# BenchmarkRecorder.CSV_HEADER (BenchmarkRecorder.java:17-24); the column order
# is the file format.
COLUMNS = (
    "count", "totalMs", "backendMs", "kernelMs", "maximumPhysicalKernelMs", "copyMs", "uploadMs", "width", "height",
    "cameraSamples", "pathSegments", "secondaryPathSegments", "primitiveTests", "aabbTests", "rootAabbTests",
    "internalNodeVisits", "leafNodeVisits", "homogeneousLeafNodeVisits", "mixedLeafNodeVisits", "sphereTests",
    "boxTests", "planeTests", "affineSphereTests", "affineBoxTests", "traversalWork", "primitiveTestsPerPathSegment",
    "primitiveTestsPerCameraSample", "aabbTestsPerPathSegment", "aabbTestsPerCameraSample",
    "rootAabbTestsPerPathSegment", "traversalWorkPerPathSegment", "traversalWorkPerCameraSample", "stackOverflows",
    "maxStackSize")


def _invalid(message: str) -> N.InvalidArgument:
    return N.InvalidArgument(N.MRAD_ERR_INVALID_ARGUMENT, message)


def _illegal(message: str) -> N.IllegalState:
    return N.IllegalState(N.MRAD_ERR_ILLEGAL_STATE, message)


def java_fixed(value: float, decimals: int) -> str:
    """String.format(Locale.ROOT, "%.<decimals>f", value): the shortest decimal
    digits of the double, rounded half-up."""
    if math.isnan(value):
        return "NaN"
    if math.isinf(value):
        return "Infinity" if value > 0 else "-Infinity"
    with localcontext() as ctx:
        ctx.prec = 800
        q = Decimal(repr(float(value))).quantize(Decimal(1).scaleb(-decimals), rounding=ROUND_HALF_UP)
    text = f"{q:f}"
    if q == 0 and math.copysign(1.0, value) < 0 and not text.startswith("-"):
        text = "-" + text
    return text


def java_double_string(value: float) -> str:
    """Double.toString: plain decimal with at least one fraction digit for
    1e-3 <= |v| < 1e7, else d.ddddE[-]n, from the shortest round-trip digits
    (those of Java 19 and later)."""
    v = float(value)
    if math.isnan(v):
        return "NaN"
    if math.isinf(v):
        return "Infinity" if v > 0 else "-Infinity"
    if v == 0.0:
        return "-0.0" if math.copysign(1.0, v) < 0 else "0.0"
    sign = "-" if v < 0 else ""
    d = Decimal(repr(abs(v)))
    if Decimal("1e-3") <= d < Decimal("1e7"):
        text = f"{d:f}"
        if "." not in text:
            text += ".0"
        whole, frac = text.split(".")
        frac = frac.rstrip("0") or "0"
        return f"{sign}{whole}.{frac}"
    digits = d.as_tuple().digits
    exponent = d.adjusted()
    mantissa = "".join(map(str, digits)).rstrip("0") or "0"
    return f"{sign}{mantissa[0]}.{mantissa[1:] or '0'}E{exponent}"


def _format(value) -> str:
    if value is None:
        return ""
    if isinstance(value, float):
        return java_fixed(value, 4)
    return str(int(value))


def traversal_work(aabb_tests: int, sphere_tests: int, box_tests: int, plane_tests: int, affine_sphere_tests: int,
                   affine_box_tests: int) -> float:
    """TraversalWorkMetric.cost (TraversalWorkMetric.java:14-28): the counters
    weighted by PrimitiveCostModel, summed in the reference's order."""
    return (NODE_AABB * aabb_tests + SPHERE * sphere_tests + BOX * box_tests + PLANE * plane_tests
            + AFFINE_SPHERE * affine_sphere_tests + AFFINE_BOX * affine_box_tests)


def traversal_work_of(stats: dict) -> float:
    return traversal_work(stats["aabb_tests"], stats["sphere_tests"], stats["box_tests"], stats["plane_tests"],
                          stats["affine_sphere_tests"], stats["affine_box_tests"])


def root_aabb_tests(stats: dict) -> int:
    """AABB tests of the tree roots: every interior visit tests two children
    (HeadlessGpuRenderer.java:150-156)."""
    n = stats["aabb_tests"] - 2 * stats["internal_node_visits"]
    if n < 0:
        raise _illegal("AABB and interior-node counters are inconsistent")
    return n


@dataclass(frozen=True)
class SceneMetrics:
    """SceneBenchmarkMetrics.fromGpuScene (SceneBenchmarkMetrics.java:21-33)."""
    sphere_count: int
    box_count: int
    plane_count: int
    affine_sphere_count: int
    affine_box_count: int
    primitive_count: int
    material_count: int
    scene_bytes: int


def scene_metrics(scene) -> SceneMetrics:
    c = scene.counts()
    floats = (scene.material_data().size + scene.sphere_data().size + scene.box_data().size
              + scene.plane_data().size + scene.affine_sphere_data().size + scene.affine_box_data().size)
    primitives = c["spheres"] + c["boxes"] + c["planes"] + c["affine_spheres"] + c["affine_boxes"]
    return SceneMetrics(c["spheres"], c["boxes"], c["planes"], c["affine_spheres"], c["affine_boxes"], primitives,
                        c["materials"], 4 * floats)


@dataclass(frozen=True)
class RunMetadata:
    """BenchmarkRecorder.RunMetadata (BenchmarkRecorder.java:26-50); None
    leaves a key out of the provenance line."""
    scene_bytes: int | None = None
    bvh_bytes: int | None = None
    host_bytes: int | None = None
    estimated_device_bytes: int | None = None
    primitive_count: int | None = None
    sphere_count: int | None = None
    box_count: int | None = None
    plane_count: int | None = None
    affine_sphere_count: int | None = None
    affine_box_count: int | None = None
    material_count: int | None = None
    acceleration_node_count: int | None = None
    acceleration_reference_count: int | None = None
    leaf_count: int | None = None
    max_tree_depth: int | None = None
    leaf_size: int | None = None
    min_leaf_occupancy: int | None = None
    max_leaf_occupancy: int | None = None
    mean_leaf_occupancy: float | None = None
    generalized_sah_cost: float | None = None
    uniform_sah_cost: float | None = None
    weighted_sah_cost: float | None = None
    acceleration_config: str | None = None


# provenance keys in the reference's order (BenchmarkRecorder.java:325-347)
_METADATA_KEYS = ("sceneBytes", "bvhBytes", "hostBytes", "estimatedDeviceBytes", "primitiveCount", "sphereCount",
                  "boxCount", "planeCount", "affineSphereCount", "affineBoxCount", "materialCount",
                  "accelerationNodeCount", "accelerationReferenceCount", "leafCount", "maxTreeDepth", "leafSize",
                  "minLeafOccupancy", "maxLeafOccupancy", "meanLeafOccupancy", "generalizedSahCost",
                  "uniformSahCost", "weightedSahCost", "accelerationConfig")


@dataclass(frozen=True)
class FrameMetrics:
    """BenchmarkRecorder.FrameMetrics (BenchmarkRecorder.java:52-104), with
    its two consistency checks."""
    count: int
    total_nanos: int
    kernel_nanos: int | None = None
    maximum_physical_kernel_nanos: int | None = None
    copy_nanos: int | None = None
    upload_nanos: int | None = None
    width: int = 0
    height: int = 0
    camera_samples: int | None = None
    path_segments: int | None = None
    primitive_tests: int | None = None
    aabb_tests: int | None = None
    sphere_tests: int | None = None
    box_tests: int | None = None
    plane_tests: int | None = None
    affine_sphere_tests: int | None = None
    affine_box_tests: int | None = None
    root_aabb_tests: int | None = None
    traversal_work: float | None = None
    stack_overflows: int | None = None
    max_stack_size: int | None = None
    internal_node_visits: int | None = None
    leaf_node_visits: int | None = None
    homogeneous_leaf_node_visits: int | None = None
    mixed_leaf_node_visits: int | None = None
    backend_nanos: int | None = None

    def __post_init__(self):
        k, m = self.kernel_nanos, self.maximum_physical_kernel_nanos
        if k is not None and m is not None and (k < 0 or m < 0 or m > k):
            raise _invalid("Maximum physical-kernel time must be bounded by total kernel time")
        lv, hv, mv = self.leaf_node_visits, self.homogeneous_leaf_node_visits, self.mixed_leaf_node_visits
        if lv is not None and hv is not None and mv is not None and lv != hv + mv:
            raise _invalid("Leaf-kind visits must partition total leaf visits")


def _ms(nanos):
    return None if nanos is None else nanos / 1_000_000.0


def _ratio(num, den):
    return None if num is None or den is None or den == 0 else num / float(den)


def _row(m: FrameMetrics) -> str:
    """FrameSample.toCsv (BenchmarkRecorder.java:133-177)."""
    secondary = (None if m.camera_samples is None or m.path_segments is None
                 else max(0, m.path_segments - m.camera_samples))
    values = (
        m.count, java_fixed(_ms(m.total_nanos), 4), _ms(m.backend_nanos), _ms(m.kernel_nanos),
        _ms(m.maximum_physical_kernel_nanos), _ms(m.copy_nanos), _ms(m.upload_nanos), m.width, m.height,
        m.camera_samples, m.path_segments, secondary, m.primitive_tests, m.aabb_tests, m.root_aabb_tests,
        m.internal_node_visits, m.leaf_node_visits, m.homogeneous_leaf_node_visits, m.mixed_leaf_node_visits,
        m.sphere_tests, m.box_tests, m.plane_tests, m.affine_sphere_tests, m.affine_box_tests,
        None if m.traversal_work is None else float(m.traversal_work),
        _ratio(m.primitive_tests, m.path_segments), _ratio(m.primitive_tests, m.camera_samples),
        _ratio(m.aabb_tests, m.path_segments), _ratio(m.aabb_tests, m.camera_samples),
        _ratio(m.root_aabb_tests, m.path_segments), _ratio(m.traversal_work, m.path_segments),
        _ratio(m.traversal_work, m.camera_samples), m.stack_overflows, m.max_stack_size)
    return ",".join(v if isinstance(v, str) else _format(v) for v in values) + "\n"


def _bench_line(m: FrameMetrics) -> str:
    """printSample (BenchmarkRecorder.java:257-285)."""
    def part(label, nanos):
        return "" if nanos is None else f", {label}={java_fixed(_ms(nanos), 3)} ms"
    prim = ("" if m.primitive_tests is None or not m.path_segments
            else f", prim/segment={java_fixed(m.primitive_tests / m.path_segments, 2)}")
    aabb = ("" if m.aabb_tests is None or not m.path_segments
            else f", aabb/segment={java_fixed(m.aabb_tests / m.path_segments, 2)}")
    return (f"[BENCH]: count={m.count} total={java_fixed(_ms(m.total_nanos), 3)} ms"
            f"{part('backend', m.backend_nanos)}{part('kernel', m.kernel_nanos)}"
            f"{part('max-kernel', m.maximum_physical_kernel_nanos)}{part('copy', m.copy_nanos)}"
            f"{part('upload', m.upload_nanos)}{prim}{aabb} ({m.width}x{m.height})")


class BenchmarkRecorder:
    """BenchmarkRecorder (BenchmarkRecorder.java:192-365): rows are buffered and
    written every `flush_every` frames and on close; the provenance line and
    header go out with the first flush, after which the run metadata and the
    code-object hash can no longer change."""

    def __init__(self, output_file: str | Path | None, flush_every: int = 1, cpu_info: str | None = None,
                 gpu_info: str | None = None, version: str | None = None, scene_name: str | None = None,
                 configuration: str | None = None, echo: bool = True):
        self.output_file = None if output_file is None else Path(output_file)
        self.flush_every = max(1, int(flush_every))
        self.cpu_info, self.gpu_info = cpu_info or "", gpu_info or ""
        self.version, self.scene_name, self.configuration = version or "", scene_name or "", configuration or ""
        self.echo = echo
        self._buffer: list[FrameMetrics] = []
        self._header_written = False
        self._metadata: RunMetadata | None = None
        self._code_sha: str | None = None

    def record(self, metrics: FrameMetrics) -> None:
        self._buffer.append(metrics)
        if self.echo:
            print(_bench_line(metrics), flush=True)
        if self.output_file is not None and len(self._buffer) >= self.flush_every:
            self.flush()

    def record_timing(self, count: int, width: int, height: int, total_nanos: int, kernel_nanos: int | None,
                      copy_nanos: int | None) -> None:
        """record(countPerFrame, width, height, totalNanos, kernelNanos, copyNanos) (:219-222)."""
        self.record(FrameMetrics(count, total_nanos, kernel_nanos, None, copy_nanos, None, width, height))

    def set_run_metadata(self, metadata: RunMetadata) -> None:
        if self._header_written:
            raise _illegal("Run metadata must be set before the first benchmark row is flushed.")
        self._metadata = metadata

    def set_compiled_code_object_sha256(self, sha: str) -> None:
        """setCompiledPtxSha256 (:245-255): the SHA-256 of the gfx950 code
        object stands in for the PTX hash."""
        if self._header_written:
            raise _illegal("Compiled PTX hash must be set before the first benchmark row is flushed.")
        if sha is None or len(sha) != 64 or any(ch not in "0123456789abcdefABCDEF" for ch in sha):
            raise _invalid("Compiled PTX hash must be a 64-digit SHA-256 value.")
        self._code_sha = sha.lower()

    def header(self) -> str:
        """buildHeader (:318-352)."""
        parts = [f"# cpu={self.cpu_info}; gpu={self.gpu_info}; version={self.version}; scene={self.scene_name}"
                 f"; config={self.configuration}; metricSchemaVersion={METRIC_SCHEMA_VERSION}"]
        if self._metadata is not None:
            for key, f in zip(_METADATA_KEYS, fields(RunMetadata)):
                v = getattr(self._metadata, f.name)
                if v is not None:
                    parts.append(f"{key}={java_double_string(v) if isinstance(v, float) else v}")
        if self._code_sha is not None:
            parts.append(f"compiledPtxSha256={self._code_sha}")
        return "; ".join(parts) + "\n" + ",".join(COLUMNS) + "\n"

    def flush(self) -> None:
        if self.output_file is None or not self._buffer:
            return
        try:
            self.output_file.parent.mkdir(parents=True, exist_ok=True)
            if not self._header_written:
                self.output_file.write_text(self.header(), encoding="utf-8")
                self._header_written = True
            with self.output_file.open("a", encoding="utf-8") as f:
                f.write("".join(_row(m) for m in self._buffer))
            self._buffer.clear()
        except OSError as ex:
            print(f"Failed to write benchmark data: {ex}", file=sys.stderr)

    def close(self) -> None:
        self.flush()

    def __enter__(self) -> "BenchmarkRecorder":
        return self

    def __exit__(self, *exc) -> None:
        self.close()
