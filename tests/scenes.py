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


def per_type_parts(seed: int = 17, n_sph: int = 300, n_box: int = 200, n_as: int = 60, n_ab: int = 100):
    """A field of spheres, boxes, affine spheres and affine boxes as one scene (`full`, one BVH) and as one scene
    per class (`parts`, same primitives in the same order), from which tests build per-type forests."""
    import numpy as np
    rng = np.random.default_rng(seed)
    mats = lambda n: mrad.material_array(rng.uniform(0.1, 0.9, (n, 3)), emittance=np.where(
        rng.uniform(size=(n, 1)) < 0.03, rng.uniform(5, 20, (n, 1)), 0.0) * np.ones((1, 3)), n=n)
    sph = np.concatenate([rng.uniform(-1, 1, (n_sph, 3)), rng.uniform(0.02, 0.08, (n_sph, 1))], 1)
    c, e = rng.uniform(-1, 1, (n_box, 3)), rng.uniform(0.01, 0.06, (n_box, 3))
    box = np.concatenate([c - e, c + e], 1)
    tas = configs._transforms(configs._rotations(rng, n_as), rng.uniform(0.02, 0.08, (n_as, 3)),
                              rng.uniform(-1, 1, (n_as, 3)))
    tab = configs._transforms(configs._rotations(rng, n_ab), rng.uniform(0.02, 0.08, (n_ab, 3)),
                              rng.uniform(-1, 1, (n_ab, 3)))
    ms, mb, mas, mab = mats(n_sph), mats(n_box), mats(n_as), mats(n_ab)
    cfg = mrad.BvhBuildConfig.weighted(4)
    full = mrad.GpuScene(cfg)
    full.add_spheres(sph, ms).add_boxes(box, mb).add_affine_transforms("sphere", tas, mas)
    full.add_affine_transforms("box", tab, mab)
    parts = [mrad.GpuScene(cfg).add_spheres(sph, ms), mrad.GpuScene(cfg).add_boxes(box, mb),
             mrad.GpuScene(cfg).add_affine_transforms("sphere", tas, mas),
             mrad.GpuScene(cfg).add_affine_transforms("box", tab, mab)]
    for s in (full, *parts):
        s.set_background((0.18, 0.22, 0.30))
    return full, parts
