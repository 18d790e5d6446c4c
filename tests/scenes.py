"""Small parity scenes shared by CPU and GPU tests."""
from __future__ import annotations

import colorsys

import numpy as np

import mrad
from mrad import configs


def gi_test_scene(leaf_size: int = 8, mode: str = "UNIFORM_SAH") -> mrad.GpuScene:
    """GI Test (graphics3d/scene/v1/GITestPlain.java:15-49): 4 spheres, 5 half-spaces,
    glass and mirror-like materials, emissive ceiling plane."""
    cfg = mrad.BvhBuildConfig.uniform(leaf_size) if mode == "UNIFORM_SAH" else mrad.BvhBuildConfig.weighted(leaf_size)
    sc = mrad.GpuScene(cfg)
    sc.set_background((0.0, 0.0, 0.0))
    left = colorsys.hsv_to_rgb(0.0, 0.5, 0.7)
    right = colorsys.hsv_to_rgb(1.0 / 3.0, 0.5, 0.7)
    grey = (0.7, 0.7, 0.7)

    def m(diffuse=(0, 0, 0), reflective=(0, 0, 0), refractive=(0, 0, 0), emit=(0, 0, 0), ior=1.5):
        return mrad.material_array(diffuse, reflective, refractive, emit, ior)[0]

    spheres = [[-0.2, -0.5, 0.0, 0.3], [0.5, -0.5, -0.3, 0.3], [0.0, 0.2, 0.0, 0.2], [-0.4, 0.5, 0.1, 0.2]]
    smats = [m(reflective=(0.05,) * 3, refractive=(0.95,) * 3, ior=1.4), m(reflective=(1, 1, 1), ior=1.5),
             m(diffuse=grey), m(reflective=(0.9,) * 3, ior=1.5)]
    sc.add_spheres(spheres, np.array(smats))
    planes = [[-1, 0, 0, 1, 0, 0], [1, 0, 0, -1, 0, 0], [0, -1, 0, 0, 1, 0], [0, 1, 0, 0, -1, 0], [0, 0, 1, 0, 0, -1]]
    pmats = [m(diffuse=left), m(diffuse=right), m(diffuse=grey), m(diffuse=(1, 1, 1), emit=(1, 1, 1)), m(diffuse=grey)]
    sc.add_planes(planes, np.array(pmats))
    return sc


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
