"""GpuRayTracer and HeadlessGpuRenderer mirrors over libmrad.

Method names, argument meaning and error classes follow
src/xyz/marsavic/gfxlab/gpu/GpuRayTracer.java and
playground/HeadlessGpuRenderer.java; the work is done by the HIP kernels in
libmrad.so.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N
from .scene import GpuScene

MIN_FULL_PATH_DEPTH = 0
MAX_FULL_PATH_DEPTH = 32


@dataclass
class GpuCamera:
    """GpuCamera (gpu/GpuCamera.java:8-22): ray dir = forward + right*u + up*v."""
    position: tuple
    forward: tuple
    right: tuple
    up: tuple

    def packed(self) -> np.ndarray:
        return np.array([*self.position, *self.forward, *self.right, *self.up], dtype=np.float32)

    @staticmethod
    def look_at(position, target, vertical_fov_deg: float, world_up=(0.0, 1.0, 0.0)) -> "GpuCamera":
        """Pinhole camera: v in [-1, 1] spans the vertical field of view, u spans aspect*v."""
        p = np.asarray(position, np.float64)
        f = np.asarray(target, np.float64) - p
        f /= np.linalg.norm(f)
        r = np.cross(np.asarray(world_up, np.float64), f)
        r /= np.linalg.norm(r)
        u = np.cross(f, r)
        k = np.tan(np.radians(vertical_fov_deg) / 2.0)
        # right-handed image: +u to the right of the viewer looking along f
        return GpuCamera(tuple(p), tuple(f), tuple(-r * k), tuple(u * k))


@dataclass
class TraceReplayResult:
    """TraceReplayResult (GpuRayTracer.java:119-126) plus per-ray counters."""
    stats: dict
    hit_flags: np.ndarray
    hit_distances: np.ndarray
    hit_primitive_orders: np.ndarray
    hit_material_indices: np.ndarray
    hit_normals: np.ndarray
    per_ray_counters: np.ndarray | None


@dataclass(frozen=True)
class ConditioningStats:
    """ConditioningStats (GpuRayTracer.java:101-117): the diagnostics of a
    conditioning sample, validated like the reference record."""
    physical_launches: int
    upload_nanos: int
    aggregate_physical_kernel_nanos: int
    maximum_physical_kernel_nanos: int
    total_nanos: int

    def __post_init__(self):
        if (self.physical_launches < 1 or self.upload_nanos < 0 or self.aggregate_physical_kernel_nanos <= 0
                or self.maximum_physical_kernel_nanos <= 0
                or self.maximum_physical_kernel_nanos > self.aggregate_physical_kernel_nanos
                or self.total_nanos <= 0):
            raise N.InvalidArgument(N.MRAD_ERR_INVALID_ARGUMENT, "Invalid conditioning diagnostics")


def validate_full_path_depth(max_depth: int) -> None:
    """GpuRayTracer.validateFullPathDepth (GpuRayTracer.java:470-476)."""
    N.check(N.lib.mrad_validate_full_path_depth(int(max_depth)))


def mix64(z: int) -> int:
    """SplitMix64 finalizer of HeadlessGpuRenderer.mix64 (HeadlessGpuRenderer.java:200-205)."""
    m = (1 << 64) - 1
    z = (z + 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return z ^ (z >> 31)


def frame_seed(seed: int, frame_index: int) -> int:
    """frameSeed = mix64(seed ^ frameIndex * 0x9E3779B97F4A7C15) (HeadlessGpuRenderer.java:99)."""
    m = (1 << 64) - 1
    return mix64((seed ^ ((frame_index * 0x9E3779B97F4A7C15) & m)) & m)


def _stats_dict(fs: N.FrameStats) -> dict:
    d = {"total_nanos": fs.total_nanos, "kernel_nanos": fs.kernel_nanos,
         "maximum_physical_kernel_nanos": fs.maximum_physical_kernel_nanos, "copy_nanos": fs.copy_nanos,
         "upload_nanos": fs.upload_nanos, "physical_launches": fs.physical_launches}
    for i, name in enumerate(N.STAT_NAMES):
        d[name] = int(fs.counters[i])
    return d


class GpuRayTracer:
    """GpuRayTracer (GpuRayTracer.java:56-1317) on one MI355X."""

    def __init__(self, width: int, height: int, samples_per_frame: int, bvh_traversal: bool,
                 collect_metrics: bool = False, bvh_stack_size: int = 32, samples_per_launch: int = 0,
                 device: int = 0, pixels_per_launch: int = 0, max_physical_kernel_nanos: int = 0,
                 replay_rays_per_launch: int = 0):
        """pixels_per_launch is gfxlab.gpu.renderPixelsPerLaunch (0: the whole image per launch),
        replay_rays_per_launch gfxlab.gpu.replayRaysPerLaunch (0: 262144) and max_physical_kernel_nanos the
        per-launch duration guard (0: none; the reference uses 250 ms)."""
        cfg = N.TracerConfig(int(width), int(height), int(samples_per_frame), 1 if bvh_traversal else 0,
                             1 if collect_metrics else 0, int(bvh_stack_size), int(samples_per_launch), int(device),
                             int(pixels_per_launch), int(replay_rays_per_launch), int(max_physical_kernel_nanos))
        h = C.c_void_p()
        N.check(N.lib.mrad_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self.width = max(1, int(width))
        self.height = max(1, int(height))
        self.samples_per_frame = max(1, int(samples_per_frame))
        self.collect_metrics = bool(collect_metrics)
        self.render_form = N.MRAD_RENDER_FORM_MEGAKERNEL

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            N.lib.mrad_destroy(self._h)
        self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()

    def _handle(self):
        if self._h is None:
            raise N.IllegalState(N.MRAD_ERR_ILLEGAL_STATE, "tracer is closed")
        return self._h

    def is_available(self) -> bool:
        return bool(N.lib.mrad_is_available(self._handle()))

    def collects_metrics(self) -> bool:
        """collectsMetrics (GpuRayTracer.java:289-292)."""
        return self.collect_metrics

    def set_russian_roulette(self, start_depth: int) -> None:
        N.check(N.lib.mrad_set_russian_roulette(self._handle(), int(start_depth)))

    def set_leaf_batch(self, lanes: int) -> None:
        """0 (default) or 16: the megakernel tests popped leaves 16 lanes at a
        time; same images and counters, faster on leaf-heavy scenes."""
        N.check(N.lib.mrad_set_leaf_batch(self._handle(), int(lanes)))

    def set_render_form(self, form) -> None:
        """"megakernel" (0, the default: the reference's per-pixel kernel and launch contract) or "wavefront" (1):
        the same frame, bit for bit with the same counters, rendered by the wavefront form's bounded path pool over
        the resident scene; its physical launches are bounce steps.  No counterpart in the reference."""
        forms = {"megakernel": N.MRAD_RENDER_FORM_MEGAKERNEL, "wavefront": N.MRAD_RENDER_FORM_WAVEFRONT}
        value = forms.get(form, -1) if isinstance(form, str) else int(form)
        N.check(N.lib.mrad_set_render_form(self._handle(), value))
        self.render_form = value

    def ensure_scene_resident(self, scene: GpuScene) -> None:
        N.check(N.lib.mrad_upload_scene(self._handle(), C.byref(scene.arrays()), C.c_uint64(scene.key)))

    def render_sample(self, rgb: np.ndarray | None, scene: GpuScene, camera: GpuCamera, max_depth: int,
                      frame_index: int, frame_seed: int) -> None:
        """renderSample (GpuRayTracer.java:317-432); rgb=None keeps the frame on the device."""
        validate_full_path_depth(max_depth)
        if rgb is not None:
            if rgb.dtype != np.float32 or not rgb.flags.c_contiguous or rgb.size < self.width * self.height * 3:
                raise N.InvalidArgument(N.MRAD_ERR_INVALID_ARGUMENT,
                                        f"RGB output must contain at least {self.width * self.height * 3} floats")
        if scene is None or camera is None:
            raise N.InvalidArgument(N.MRAD_ERR_INVALID_ARGUMENT, "GPU scene and camera must not be null")
        if self.render_form == N.MRAD_RENDER_FORM_MEGAKERNEL:
            self.render_launch_count()
        self.ensure_scene_resident(scene)
        cam = camera.packed() if isinstance(camera, GpuCamera) else np.ascontiguousarray(camera, np.float32)
        N.check(N.lib.mrad_render_sample(self._handle(), N.ptr(rgb, N.fp), N.ptr(cam, N.fp), int(max_depth),
                                         int(frame_index), C.c_uint64(frame_seed & ((1 << 64) - 1))))

    def render_launch_count(self) -> int:
        """renderLaunchCount (GpuRayTracer.java:556-564): pixel tiles x sample ranges of one frame; more than
        8192 raises InvalidArgument.  Needs no device."""
        n = C.c_int()
        N.check(N.lib.mrad_render_launch_count(self._handle(), C.byref(n)))
        return n.value

    def condition_steady_state_sample(self, scene: GpuScene, camera: GpuCamera, max_depth: int, frame_index: int,
                                      frame_seed: int) -> ConditioningStats:
        """conditionSteadyStateSample (GpuRayTracer.java:434-453): one sample
        into a private buffer (uploading the scene if needed), its diagnostics."""
        if getattr(self, "_conditioning", None) is None:
            self._conditioning = np.zeros(self.width * self.height * 3, np.float32)
        self.render_sample(self._conditioning, scene, camera, max_depth, frame_index, frame_seed)
        st = self.last_frame_stats()
        return ConditioningStats(st["physical_launches"], st["upload_nanos"], st["kernel_nanos"],
                                 st["maximum_physical_kernel_nanos"], st["total_nanos"])

    def render_steady_state_sample(self, rgb: np.ndarray | None, scene: GpuScene, camera: GpuCamera, max_depth: int,
                                   frame_index: int, frame_seed: int) -> None:
        """renderSteadyStateSample (GpuRayTracer.java:455-467): a sample that
        must find the scene resident."""
        self.render_sample(rgb, scene, camera, max_depth, frame_index, frame_seed)
        if self.last_frame_stats()["upload_nanos"] != 0:
            raise N.IllegalState(N.MRAD_ERR_ILLEGAL_STATE, "A steady-state sample uploaded scene data")

    def last_physical_kernel_launch_count(self) -> int:
        """lastPhysicalKernelLaunchCount (GpuRayTracer.java:294-296)."""
        return int(self.last_frame_stats()["physical_launches"])

    def copy_frame(self, rgb: np.ndarray) -> None:
        N.check(N.lib.mrad_copy_frame(self._handle(), N.ptr(rgb, N.fp)))

    def trace_replay(self, scene: GpuScene, rays: np.ndarray, ray_count: int | None = None,
                     collect_metrics: bool = False, per_ray_counters: bool = False) -> TraceReplayResult:
        """traceReplay (GpuRayTracer.java:593-716); rays are origin.xyz, direction.xyz."""
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1)
        n = len(rays) // 6 if ray_count is None else int(ray_count)
        if n <= 0 or len(rays) < 6 * n:
            raise N.InvalidArgument(N.MRAD_ERR_INVALID_ARGUMENT, "Replay rays must contain origin.xyz and direction.xyz.")
        if scene is None:                                               # requireReplayScene (:725-729)
            raise N.InvalidArgument(N.MRAD_ERR_INVALID_ARGUMENT, "GPU scene must not be null")
        self.ensure_scene_resident(scene)
        flags = np.zeros(n, np.int32)
        dist = np.zeros(n, np.float32)
        order = np.zeros(n, np.int32)
        mat = np.zeros(n, np.int32)
        normals = np.zeros(3 * n, np.float32)
        per = np.zeros((n, N.STAT_COUNT), np.uint32) if per_ray_counters else None
        N.check(N.lib.mrad_trace_replay(self._handle(), N.ptr(rays, N.fp), n, 1 if collect_metrics else 0,
                                        N.ptr(flags, N.ip), N.ptr(dist, N.fp), N.ptr(order, N.ip),
                                        N.ptr(mat, N.ip), N.ptr(normals, N.fp), N.ptr(per, N.up)))
        return TraceReplayResult(self.last_frame_stats(), flags, dist, order, mat, normals, per)

    def last_frame_stats(self) -> dict:
        fs = N.FrameStats()
        N.check(N.lib.mrad_get_frame_stats(self._handle(), C.byref(fs)))
        return _stats_dict(fs)

    def device_info(self) -> tuple[str, str]:
        name = C.create_string_buffer(256)
        arch = C.create_string_buffer(256)
        N.check(N.lib.mrad_device_info(self._handle(), name, 256, arch, 256))
        return name.value.decode(), arch.value.decode()

    @staticmethod
    def compiled_code_object_sha256() -> str:
        buf = C.create_string_buffer(65)
        N.check(N.lib.mrad_code_object_sha256(buf))
        return buf.value.decode()


class HeadlessGpuRenderer:
    """HeadlessGpuRenderer.render (HeadlessGpuRenderer.java:58-148): per-frame
    seeds, one renderSample per frame, FP64 accumulation of frame * spp, and
    with a recorder the run metadata and one benchmark row per measured frame."""

    def __init__(self, scene: GpuScene, camera: GpuCamera, width: int, height: int, max_depth: int, seed: int,
                 samples_per_frame: int = 1, bvh_traversal: bool = True, rr_start: int = -1, **tracer_kw):
        validate_full_path_depth(max_depth)
        self.scene, self.camera = scene, camera
        self.width, self.height, self.max_depth, self.seed = width, height, max_depth, seed
        self.samples_per_frame = samples_per_frame
        self.bvh_traversal = bvh_traversal
        self.collect_metrics = bool(tracer_kw.get("collect_metrics", False))
        self.tracer = GpuRayTracer(width, height, samples_per_frame, bvh_traversal, **tracer_kw)
        if rr_start >= 0:
            self.tracer.set_russian_roulette(rr_start)

    def _run_metadata(self, frame_floats: int):
        """The RunMetadata of HeadlessGpuRenderer.java:65-88."""
        from .recorder import RunMetadata, scene_metrics
        m = scene_metrics(self.scene)
        st = self.scene.bvh_stats() if self.bvh_traversal else None
        bvh_bytes = 0 if st is None else int(st.bytes)
        return RunMetadata(
            m.scene_bytes, bvh_bytes, frame_floats * 8 + m.scene_bytes + bvh_bytes,
            frame_floats * 4 + m.scene_bytes + bvh_bytes + 14 * 8,
            m.primitive_count, m.sphere_count, m.box_count, m.plane_count, m.affine_sphere_count, m.affine_box_count,
            m.material_count, 0 if st is None else st.node_count, 0 if st is None else st.primitive_ref_count,
            0 if st is None else st.leaf_count, 0 if st is None else st.max_depth,
            None if st is None else st.leaf_size, None if st is None else st.min_leaf_occupancy,
            None if st is None else st.max_leaf_occupancy, None if st is None else float(st.mean_leaf_occupancy),
            None if st is None else float(st.generalized_sah_cost), None if st is None else float(st.uniform_sah_cost),
            None if st is None else float(st.weighted_sah_cost),
            self.scene.config.summary() if self.bvh_traversal else "mode=LINEAR")

    def render(self, frames: int, warmup_frames: int = 0, progress_every: int = 0, recorder=None) -> np.ndarray:
        import time
        from .recorder import FrameMetrics, root_aabb_tests, traversal_work_of
        frame = np.zeros(self.width * self.height * 3, np.float32)
        sums = np.zeros(self.width * self.height * 3, np.float64)
        if recorder is not None:
            recorder.set_run_metadata(self._run_metadata(frame.size))
            recorder.set_compiled_code_object_sha256(self.compiled_code_object_sha256())
        for fi in range(max(0, warmup_frames) + max(0, frames)):
            t0 = time.perf_counter_ns()
            self.tracer.render_sample(frame, self.scene, self.camera, self.max_depth, fi, frame_seed(self.seed, fi))
            measured = fi >= warmup_frames
            if measured:
                sums += frame.astype(np.float64) * self.samples_per_frame
            wall = time.perf_counter_ns() - t0
            st = self.tracer.last_frame_stats()
            if measured and warmup_frames > 0 and st["upload_nanos"] != 0:
                raise N.IllegalState(N.MRAD_ERR_ILLEGAL_STATE,
                                     "Measured steady-state GPU frame unexpectedly uploaded scene data")
            if measured and recorder is not None:
                m = self.collect_metrics
                pick = (lambda k: st[k] if m else None)
                recorder.record(FrameMetrics(
                    self.samples_per_frame, wall, st["kernel_nanos"], st["maximum_physical_kernel_nanos"],
                    st["copy_nanos"], st["upload_nanos"], self.width, self.height,
                    self.width * self.height * self.samples_per_frame if m else None, pick("rays"),
                    pick("primitive_tests"), pick("aabb_tests"), pick("sphere_tests"), pick("box_tests"),
                    pick("plane_tests"), pick("affine_sphere_tests"), pick("affine_box_tests"),
                    root_aabb_tests(st) if m else None, traversal_work_of(st) if m else None,
                    pick("stack_overflows"), pick("max_stack_size"), pick("internal_node_visits"),
                    pick("leaf_node_visits"), pick("homogeneous_leaf_node_visits"), pick("mixed_leaf_node_visits"),
                    st["total_nanos"]))
            done = fi - warmup_frames + 1
            if measured and progress_every > 0 and done % progress_every == 0:
                print(f"  progress {done}/{frames}", flush=True)
        return sums / max(1, frames * self.samples_per_frame)

    @staticmethod
    def compiled_code_object_sha256() -> str:
        return GpuRayTracer.compiled_code_object_sha256()
