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


# The scenario list of the reference's BVH correctness gate (benchmark/GpuBvhCorrectnessCheck.java:151-171): the
# catalog's introductory presets, Layered Hero 96, the gallery overlap (front camera) and the first layout of the
# 100-object random rows F1..F4 (publication rows 6, 9, 12, 15); each traced on a 32x18 primary grid at pixel
# phase (0.5, 0.5) (:173-176).
GATE_GRID = (32, 18)


def gate_scenarios():
    """[(name, build(config) -> GpuScene, camera)] in the gate's order."""
    from mrad import random_scenes as RS
    out = [("GI_TEST", W.gi_test_scene, W.gi_test_camera()),
           ("CITY_OF_NIGHT_V1", W.city_of_night_scene, W.city_of_night_camera()),
           ("SIGNAL_CHAMBER", W.signal_chamber_scene, W.signal_chamber_camera()),
           ("LAYERED_HERO_96", lambda cfg: W.hero_scene(96, cfg), W.hero_camera()),
           ("GALLERY_OVERLAP", W.overlap_scene, W.overlap_camera())]
    for fam in ("F1", "F2", "F3", "F4"):
        out.append((f"{fam}-100", (lambda cfg, f=fam: RS.random_scene(f, 100, 0, cfg)), RS.random_camera()))
    return out


def gate_methods():
    """The object-SAH methods of MethodCatalog.FAMILIES this build has (uniform and weighted, leaf 8, lambda 1,
    the reference's calibrated weights: BenchmarkProtocol.LEAF_SIZE / WEIGHTED_LAMBDA)."""
    return [("uniform", mrad.BvhBuildConfig.uniform(W.LEAF_SIZE)),
            ("weighted", mrad.BvhBuildConfig.weighted(W.LEAF_SIZE, 1.0))]


def gate_rays(camera):
    return W.primary_grid(camera, *GATE_GRID, 0.5, 0.5)
