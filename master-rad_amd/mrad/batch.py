"""Batched wavefront renderer: many scenes, one pool of in-flight paths.

The Python face of mrad_batch_* (include/mrad.h).  Each scene's image is
bit-identical to GpuRayTracer.render_sample for the same view; the batch form
exists for throughput on many scenes of uneven size.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N
from .scene import GpuScene
from .tracer import GpuCamera


@dataclass
class WavefrontConfig:
    queue_capacity: int = 1 << 22
    waves_per_cu: int = 28
    bvh_stack_size: int = 32
    linear_scan: bool = False
    collect_metrics: bool = False
    use_graph: bool = True
    time_kernels: bool = False
    lds_stack_entries: int = 10
    leaf_regroup: bool = False
    result_chunk_paths: int = 0
    stack_storage: int = 0
    trace_passes: int = 0        # closest-hit passes per bounce step (0 = the library default)
    pass_pops: int = 0           # pops of a capped pass (0 = the library default)
    suspend_capacity: int = 0    # suspended queries per pass (0 = half the queue)
    pass_dense_lanes: int = 0    # lanes still traversing that keep a wave going at the budget (0 = the library
                                 # default, -1 = every wave suspends)
    device: int = 0

    def native(self) -> N.WavefrontConfig:
        return N.WavefrontConfig(self.device, self.queue_capacity, self.waves_per_cu, self.bvh_stack_size,
                                 1 if self.linear_scan else 0, 1 if self.collect_metrics else 0,
                                 1 if self.use_graph else 0, 1 if self.time_kernels else 0,
                                 self.lds_stack_entries, 1 if self.leaf_regroup else 0, self.result_chunk_paths,
                                 self.stack_storage, self.trace_passes, self.pass_pops, self.suspend_capacity,
                                 self.pass_dense_lanes)


def make_view(camera, width: int, height: int, samples: int, max_depth: int, frame_index: int, frame_seed: int,
              rr_start: int = -1) -> N.View:
    v = N.View()
    v.width, v.height, v.samples, v.max_depth = int(width), int(height), int(samples), int(max_depth)
    v.rr_start, v.frame_index = int(rr_start), int(frame_index)
    v.frame_seed = frame_seed & ((1 << 64) - 1)
    cam = camera.packed() if isinstance(camera, GpuCamera) else np.asarray(camera, np.float32)
    for i in range(12):
        v.camera[i] = float(cam[i])
    return v


class BatchRenderer:
    def __init__(self, config: WavefrontConfig | None = None):
        self.config = config or WavefrontConfig()
        h = C.c_void_p()
        N.check(N.lib.mrad_batch_create(C.byref(self.config.native()), C.byref(h)))
        self._h = h
        self.views: list[N.View] = []

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            N.lib.mrad_batch_destroy(self._h)
        self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()

    def add_scene(self, scene: GpuScene, camera, width: int, height: int, samples: int, max_depth: int,
                  frame_index: int = 0, frame_seed: int = 0, rr_start: int = -1) -> int:
        v = make_view(camera, width, height, samples, max_depth, frame_index, frame_seed, rr_start)
        idx = C.c_int()
        N.check(N.lib.mrad_batch_add_scene(self._h, C.byref(scene.arrays()), C.byref(v), C.byref(idx)))
        self.views.append(v)
        return idx.value

    def render(self) -> dict:
        N.check(N.lib.mrad_batch_render(self._h))
        return self.stats()

    def configure(self, config: WavefrontConfig) -> None:
        """New launch parameters, scenes stay resident (mrad_batch_configure)."""
        N.check(N.lib.mrad_batch_configure(self._h, C.byref(config.native())))
        self.config = config

    def effective_config(self) -> dict:
        """The launch parameters in use, every default resolved (mrad_batch_get_config)."""
        c = N.WavefrontConfig()
        N.check(N.lib.mrad_batch_get_config(self._h, C.byref(c)))
        return {name: getattr(c, name) for name, _ in N.WavefrontConfig._fields_}

    def image(self, scene: int) -> np.ndarray:
        v = self.views[scene]
        out = np.zeros(v.width * v.height * 3, np.float32)
        N.check(N.lib.mrad_batch_copy_image(self._h, int(scene), N.ptr(out, N.fp)))
        return out

    def stats(self) -> dict:
        s = N.BatchStats()
        N.check(N.lib.mrad_batch_get_stats(self._h, C.byref(s)))
        d = {"total_nanos": s.total_nanos, "kernel_nanos": s.kernel_nanos, "upload_nanos": s.upload_nanos,
             "steps": s.steps, "paths": s.paths, "trace_nanos": s.trace_nanos, "shade_nanos": s.shade_nanos,
             "generate_nanos": s.generate_nanos}
        for i, name in enumerate(N.STAT_NAMES):
            d[name] = int(s.counters[i])
        return d
