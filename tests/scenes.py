"""Small parity scenes shared by CPU and GPU tests."""
from __future__ import annotations

import mrad
from mrad import configs
from mrad import workloads as W


def gi_test_scene(leaf_size: int = 8, mode: str = "UNIFORM_SAH") -> mrad.GpuScene:
    """GI Test (graphics3d/scene/v1/GITestPlain.java:15-49): 4 spheres, 5 half-spaces,
    glass and mirror-like materials, emissive ceiling plane."""
    cfg = mrad.BvhBuildConfig.uniform(leaf_size) if mode == "UNIFORM_SAH" else mrad.BvhBuildConfig.weighted(leaf_size)
    return W.gi_test_scene(cfg)


def gi_camera() -> mrad.GpuCamera:
    return mrad.GpuCamera.look_at((0.0, 0.0, -2.6), (0.0, 0.0, 0.0), 45.0)


def small_field(n: int = 3000, seed: int = 11, clustered: bool = False) -> mrad.GpuScene:
    """A seeded field; the clustered one keeps the reference's weights (its
    deeper tree overflows an 8-entry traversal stack), the other uses the
    fields' default gfx950 weights."""
    if clustered:
        return configs.field_scene(seed, n, clustered=True, leaf_size=4, emissive_fraction=0.005, ground=True,
                                   affine_fraction=0.5, box_fraction=0.0, bvh=mrad.BvhBuildConfig.weighted(4))
    return configs.field_scene(seed, n)
