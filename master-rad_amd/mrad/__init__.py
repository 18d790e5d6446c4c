"""mrad: MI355X-native GPU path-tracing hot path of the master-rad renderer.

Host-side mirror of the reference's GpuScene / GpuRayTracer API over the
C-ABI library libmrad.so (include/mrad.h).
"""
from ._native import MradError, InvalidArgument, IllegalState, STAT_NAMES  # noqa: F401
from .scene import BvhBuildConfig, GpuScene, material_array  # noqa: F401
from .tracer import (ConditioningStats, GpuCamera, GpuRayTracer, HeadlessGpuRenderer, TraceReplayResult,  # noqa: F401
                     frame_seed, mix64, validate_full_path_depth)
from .batch import BatchRenderer, WavefrontConfig  # noqa: F401
