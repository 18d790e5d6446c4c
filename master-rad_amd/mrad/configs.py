"""Synthetic scenes of the five benchmark configurations.

None of these scenes exists in the reference; they are defined here and
generated from fixed seeds.  Scene k draws from numpy's PCG64 seeded with
splitmix64(0x4D49333535580000 + k); batch scene i uses splitmix64(s_4 + i).

  room         closed unit room: 5 wall boxes 0.05 thick, emissive box light
               0.3 x 0.01 x 0.3 (radiance 10), 2 spheres r = 0.2, 1 affine box
               (half-extents 0.15, 0.3, 0.15 rotated 30 deg about y); fixed
               albedos in [0.2, 0.8]; 256^2, 16 spp, 4 bounces.
  het          heterogeneous field, centres U[-1,1]^3: 60 % spheres (radius
               lognormal, median 0.004, sigma 0.5), 30 % boxes (edges
               U[0.002, 0.01]), 10 % affine (half spheres, half boxes) whose
               base size is drawn like the spheres'/boxes' and then scaled per
               axis by U[0.5, 2] and uniformly rotated; albedo U[0.1, 0.9]^3;
               1 % emitters with grey radiance U[5, 20]; background
               (0.18, 0.22, 0.30).
  clustered    64 Gaussian clusters (sigma 0.05), 50 % affine, 50 % spheres,
               a ground box, 0.5 % emitters, BVH leaf 4.
The room's BVH is the weighted SAH (lambda 1) with the reference's calibrated
weights; the seeded fields (het, clustered, batch, diff) use the weighted SAH
(lambda 1) with the gfx950 weights of this library's kernels
(scene.GFX950_WEIGHTS); leaf 8 unless stated.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .scene import BvhBuildConfig, GpuScene
from .tracer import GpuCamera

SCENE_SEED_BASE = 0x4D49333535580000


def splitmix64(x: int) -> int:
    m = (1 << 64) - 1
    z = (x + 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return z ^ (z >> 31)


def config_seed(k: int) -> int:
    return splitmix64(SCENE_SEED_BASE + k)


@dataclass
class Workload:
    name: str
    scene: GpuScene
    camera: GpuCamera
    width: int
    height: int
    spp: int
    max_depth: int
    rr_start: int = -1
    frame_seed: int = 0x5EED
    extra: dict = field(default_factory=dict)

    @property
    def paths(self) -> int:
        return self.width * self.height * self.spp


def _mat(diffuse, emit=None):
    n = len(diffuse)
    m = np.zeros((n, 13), np.float64)
    m[:, 0:3] = diffuse
    if emit is not None:
        m[:, 9:12] = emit
    m[:, 12] = 1.5
    return m


def _rotations(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform random rotations (Shoemake's quaternion method), shape (n, 3, 3)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1 - u1), np.sqrt(u1)
    q = np.stack([a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2), b * np.sin(2 * np.pi * u3),
                  b * np.cos(2 * np.pi * u3)], axis=1)
    x, y, z, w = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], 1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], 1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], 1)], 1)


def _transforms(rot: np.ndarray, scale: np.ndarray, centre: np.ndarray) -> np.ndarray:
    """Forward 3x4 transforms  x -> R diag(scale) x + centre."""
    lin = rot * scale[:, None, :]
    t = np.zeros((len(rot), 3, 4))
    t[:, :, :3] = lin
    t[:, :, 3] = centre
    return t.reshape(-1, 12)


def field_scene(seed: int, n: int, *, clustered: bool = False, leaf_size: int = 8, emissive_fraction: float = 0.01,
                ground: bool = False, affine_fraction: float = 0.10, box_fraction: float = 0.30,
                bvh: BvhBuildConfig | None = None) -> GpuScene:
    rng = np.random.default_rng(seed)
    n_aff = int(round(n * affine_fraction))
    n_box = int(round(n * box_fraction))
    n_sph = n - n_aff - n_box
    n_as = n_aff // 2
    n_ab = n_aff - n_as
    if clustered:
        centres64 = rng.uniform(-1, 1, (64, 3))
        pick = rng.integers(0, 64, n)
        centres = centres64[pick] + rng.normal(0, 0.05, (n, 3))
    else:
        centres = rng.uniform(-1, 1, (n, 3))
    radius = 0.004 * np.exp(0.5 * rng.standard_normal(n))
    edges = rng.uniform(0.002, 0.01, (n, 3))
    albedo = rng.uniform(0.1, 0.9, (n, 3))
    emit = np.zeros((n, 3))
    n_emit = max(1, int(round(n * emissive_fraction))) if emissive_fraction > 0 else 0
    if n_emit:
        idx = rng.choice(n, n_emit, replace=False)
        emit[idx] = rng.uniform(5, 20, n_emit)[:, None]
    mats = _mat(albedo, emit)
    o = 0
    sc = GpuScene(bvh or BvhBuildConfig.gfx950(leaf_size))
    sc.set_background((0.18, 0.22, 0.30))
    sl = slice(o, o + n_sph); o += n_sph
    sc.add_spheres(np.concatenate([centres[sl], radius[sl, None]], 1), mats[sl])
    bl = slice(o, o + n_box); o += n_box
    boxes = np.concatenate([centres[bl] - edges[bl] / 2, centres[bl] + edges[bl] / 2], 1)
    if ground:
        boxes = np.concatenate([boxes, [[-4.0, -1.35, -4.0, 4.0, -1.25, 4.0]]], 0)
        gm = _mat(np.array([[0.5, 0.5, 0.5]]))
        sc.add_boxes(boxes, np.concatenate([mats[bl], gm], 0))
    else:
        sc.add_boxes(boxes, mats[bl])
    if n_as:
        al = slice(o, o + n_as); o += n_as
        scale = radius[al, None] * rng.uniform(0.5, 2.0, (n_as, 3))
        sc.add_affine_transforms("sphere", _transforms(_rotations(rng, n_as), scale, centres[al]), mats[al])
    if n_ab:
        al = slice(o, o + n_ab); o += n_ab
        scale = edges[al] / 2 * rng.uniform(0.5, 2.0, (n_ab, 3))
        sc.add_affine_transforms("box", _transforms(_rotations(rng, n_ab), scale, centres[al]), mats[al])
    return sc


def field_camera() -> GpuCamera:
    return GpuCamera.look_at((0.0, 0.0, -3.2), (0.0, 0.0, 0.0), 40.0)


def room_scene(leaf_size: int = 8) -> GpuScene:
    sc = GpuScene(BvhBuildConfig.weighted(leaf_size))
    sc.set_background((0.0, 0.0, 0.0))
    t = 0.05
    walls = np.array([
        [-t, -t, -t, 1 + t, 0.0, 1 + t],      # floor
        [-t, 1.0, -t, 1 + t, 1 + t, 1 + t],   # ceiling
        [-t, -t, 1.0, 1 + t, 1 + t, 1 + t],   # back
        [-t, -t, -t, 0.0, 1 + t, 1 + t],      # left
        [1.0, -t, -t, 1 + t, 1 + t, 1 + t],   # right
        [0.35, 0.985, 0.35, 0.65, 0.995, 0.65],  # light 0.3 x 0.01 x 0.3
    ])
    alb = np.array([[0.7, 0.7, 0.7], [0.7, 0.7, 0.7], [0.7, 0.7, 0.7], [0.75, 0.25, 0.25], [0.25, 0.75, 0.25],
                    [0.0, 0.0, 0.0]])
    emit = np.zeros((6, 3))
    emit[5] = 10.0
    sc.add_boxes(walls, _mat(alb, emit))
    sc.add_spheres([[0.72, 0.2, 0.40, 0.2], [0.33, 0.2, 0.28, 0.2]], _mat(np.array([[0.8, 0.6, 0.2], [0.2, 0.4, 0.8]])))
    a = math.radians(30.0)
    rot = np.array([[[math.cos(a), 0.0, math.sin(a)], [0.0, 1.0, 0.0], [-math.sin(a), 0.0, math.cos(a)]]])
    tr = _transforms(rot, np.array([[0.15, 0.3, 0.15]]), np.array([[0.42, 0.3, 0.72]]))
    sc.add_affine_transforms("box", tr, _mat(np.array([[0.6, 0.6, 0.6]])))
    return sc


def room_camera() -> GpuCamera:
    return GpuCamera.look_at((0.5, 0.5, -1.45), (0.5, 0.5, 0.5), 40.0)


def room() -> Workload:
    return Workload("room-256", room_scene(), room_camera(), 256, 256, 16, 4)


def het(n: int = 100_000, width: int = 1920, height: int = 1080, spp: int = 64, depth: int = 8,
        rr_start: int = 3) -> Workload:
    return Workload(f"het-{n}", field_scene(config_seed(2), n), field_camera(), width, height, spp, depth, rr_start)


def clustered(n: int = 1_000_000, width: int = 1920, height: int = 1080, spp: int = 16, depth: int = 6) -> Workload:
    sc = field_scene(config_seed(3), n, clustered=True, leaf_size=4, emissive_fraction=0.005, ground=True,
                     affine_fraction=0.5, box_fraction=0.0)
    return Workload(f"clustered-{n}", sc, field_camera(), width, height, spp, depth, -1)


RESOLUTIONS = [(256, 256), (512, 512), (1024, 1024), (1920, 1080)]
SPPS = [4, 16, 64]


def batch_specs(count: int = 64, max_prims: int = 1_000_000) -> list[dict]:
    """Batch scene i: N_i = round(exp(U[ln 1e3, ln 1e6])), resolution and spp drawn uniformly."""
    base = config_seed(4)
    specs = []
    for i in range(count):
        seed = splitmix64(base + i)
        rng = np.random.default_rng(seed)
        n = int(round(math.exp(rng.uniform(math.log(1e3), math.log(1e6)))))
        w, h = RESOLUTIONS[int(rng.integers(0, len(RESOLUTIONS)))]
        spp = SPPS[int(rng.integers(0, len(SPPS)))]
        specs.append({"index": i, "seed": seed, "prims": min(n, max_prims), "width": w, "height": h, "spp": spp,
                      "max_depth": 6, "rr_start": 3})
    return specs


def batch_scene(spec: dict, bvh: BvhBuildConfig | None = None) -> Workload:
    sc = field_scene(spec["seed"], spec["prims"], bvh=bvh)
    sc.bvh_stats()          # packs and builds the BVH (the native build releases the GIL)
    return Workload(f"batch-{spec['index']}", sc, field_camera(), spec["width"], spec["height"], spec["spp"],
                    spec["max_depth"], spec["rr_start"], frame_seed=spec["seed"] & 0xFFFFFFFF)


def batch(count: int = 64, threads: int = 16, max_prims: int = 1_000_000,
          bvh: BvhBuildConfig | None = None) -> list[Workload]:
    """The 64-scene batch, scenes generated and their BVHs built on a thread pool
    (`bvh` replaces the default weighted SAH, e.g. with calibrated weights)."""
    from concurrent.futures import ThreadPoolExecutor
    specs = batch_specs(count, max_prims)
    order = sorted(range(count), key=lambda i: -specs[i]["prims"])   # largest first for balance
    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        built = dict(zip(order, ex.map(lambda i: batch_scene(specs[i], bvh), order)))
    return [built[i] for i in range(count)]


def diff(n: int = 10_000) -> Workload:
    return Workload(f"diff-{n}", field_scene(config_seed(5), n), field_camera(), 512, 512, 32, 4, -1)
