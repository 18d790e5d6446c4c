"""Scene packing and BVH build: Python face of the C++ GpuScene mirror.

Mirrors src/xyz/marsavic/gfxlab/gpu/GpuScene.java (packing, material dedup,
object/weighted SAH, DFS flattening) and BvhBuildConfig.java /
PrimitiveCostModel.java (weights and defaults).  The work is done in
libmrad (master-rad_amd/csrc/scene_builder.cpp); this module only marshals
numpy arrays.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N

# GpuScene.java:19-21: the largest affine spectral condition an upload accepts, and the smallest sphere radius or
# box extent that survives float packing (sqrt of the smallest normal float).
MAX_AFFINE_SPECTRAL_CONDITION = 65_536.0
MIN_PACKED_PRIMITIVE_EXTENT = float(np.sqrt(np.float32(np.finfo(np.float32).tiny), dtype=np.float32))

# PrimitiveCostModel.java:11-19 (RTX 3080 calibration of the reference).
SPHERE = 1.000000000
BOX = 1.288951704
AFFINE_SPHERE = 2.552743938
AFFINE_BOX = 2.728115282
PLANE = 2.860319327
NODE_AABB = 0.507189194
INTERIOR_TRAVERSAL = 1.275576713

# gfx950 (MI355X) weights of this library's kernels, from `python -m
# mrad.calibrate --in-situ`: primitive classes by the reference's isolated
# protocol (5 contexts), node AABB and interior traversal fitted in situ on
# the trace kernel of the 64-scene batch (an AABB test costs 1/27 of a
# primitive test there).
GFX950_WEIGHTS = {"sphere": 1.000000000, "box": 1.296939838, "affineSphere": 1.633064197,
                  "affineBox": 1.661005931, "plane": 1.311160810, "nodeAabb": 0.041857181,
                  "interiorTraversal": 0.083714363}


@dataclass(frozen=True)
class BvhBuildConfig:
    """BvhBuildConfig record (BvhBuildConfig.java:6-15); defaults of fromProperties (:52-64)."""
    mode: str = "UNIFORM_SAH"
    leaf_size: int = 8
    traversal_weight: float = INTERIOR_TRAVERSAL
    sphere_weight: float = SPHERE
    box_weight: float = BOX
    affine_sphere_weight: float = AFFINE_SPHERE
    affine_box_weight: float = AFFINE_BOX
    lam: float = 0.0

    @staticmethod
    def weighted(leaf_size: int = 8, lam: float = 1.0) -> "BvhBuildConfig":
        return BvhBuildConfig("WEIGHTED_SAH", leaf_size, lam=lam)

    @staticmethod
    def uniform(leaf_size: int = 8) -> "BvhBuildConfig":
        return BvhBuildConfig("UNIFORM_SAH", leaf_size)

    @staticmethod
    def gfx950(leaf_size: int = 8, lam: float = 1.0) -> "BvhBuildConfig":
        """Weighted SAH with the gfx950 weights (GFX950_WEIGHTS)."""
        w = GFX950_WEIGHTS
        return BvhBuildConfig("WEIGHTED_SAH", leaf_size, traversal_weight=w["interiorTraversal"],
                              sphere_weight=w["sphere"], box_weight=w["box"],
                              affine_sphere_weight=w["affineSphere"], affine_box_weight=w["affineBox"], lam=lam)

    def summary(self) -> str:
        """BvhBuildConfig.summary (BvhBuildConfig.java:71-76), the recorder's
        accelerationConfig."""
        from .recorder import java_fixed
        w = (self.traversal_weight, self.sphere_weight, self.box_weight, self.affine_sphere_weight,
             self.affine_box_weight, self.lam)
        return ("mode={},leafSize={},interiorTraversal={},sphere={},box={},affineSphere={},affineBox={},lambda={}"
                .format(self.mode, self.leaf_size, *(java_fixed(x, 6) for x in w)))

    def native(self) -> N.BvhConfig:
        modes = {"UNIFORM_SAH": N.MRAD_BVH_UNIFORM_SAH, "WEIGHTED_SAH": N.MRAD_BVH_WEIGHTED_SAH}
        if self.mode not in modes:
            raise ValueError(f"BVH mode {self.mode} is not built here (UNIFORM_SAH, WEIGHTED_SAH)")
        return N.BvhConfig(modes[self.mode], self.leaf_size, self.traversal_weight, self.sphere_weight,
                           self.box_weight, self.affine_sphere_weight, self.affine_box_weight, self.lam)


def material_array(diffuse, reflective=(0, 0, 0), refractive=(0, 0, 0), emittance=(0, 0, 0), ior=1.5, n=None):
    """Materials as (n, 13) doubles in packed order (GpuScene.java:951-957)."""
    cols = [np.broadcast_to(np.asarray(v, dtype=np.float64), (n, 3)) if n is not None else np.asarray(v, np.float64)
            for v in (diffuse, reflective, refractive, emittance)]
    if n is None:
        return np.concatenate([*cols, [float(ior)]]).reshape(1, 13)
    iorc = np.broadcast_to(np.asarray(ior, np.float64).reshape(-1, 1), (n, 1))
    return np.ascontiguousarray(np.concatenate([*cols, iorc], axis=1))


def _f64(a, width):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1, width))
    return a


class GpuScene:
    """GpuScene (GpuScene.java:18-700) backed by the native builder."""

    def __init__(self, config: BvhBuildConfig | None = None):
        self.config = config or BvhBuildConfig()
        h = C.c_void_p()
        N.check(N.lib.mrad_scene_create(C.byref(self.config.native()), C.byref(h)))
        self._h = h
        self._view = None
        self.key = id(self)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            N.lib.mrad_scene_destroy(h)
            self._h = None

    def _touch(self):
        self._view = None

    def set_background(self, rgb) -> "GpuScene":
        N.check(N.lib.mrad_scene_set_background(self._h, *map(float, rgb)))
        self._touch()
        return self

    def add_spheres(self, center_radius, materials) -> "GpuScene":
        cr, m = _f64(center_radius, 4), _f64(materials, 13)
        assert len(cr) == len(m)
        N.check(N.lib.mrad_scene_add_spheres(self._h, len(cr), N.ptr(cr, N.dp), N.ptr(m, N.dp)))
        self._touch()
        return self

    def add_boxes(self, min_max, materials) -> "GpuScene":
        b, m = _f64(min_max, 6), _f64(materials, 13)
        assert len(b) == len(m)
        N.check(N.lib.mrad_scene_add_boxes(self._h, len(b), N.ptr(b, N.dp), N.ptr(m, N.dp)))
        self._touch()
        return self

    def add_planes(self, point_normal, materials) -> "GpuScene":
        p, m = _f64(point_normal, 6), _f64(materials, 13)
        assert len(p) == len(m)
        N.check(N.lib.mrad_scene_add_planes(self._h, len(p), N.ptr(p, N.dp), N.ptr(m, N.dp)))
        self._touch()
        return self

    def add_affine_transforms(self, kind: str, transforms, materials) -> "GpuScene":
        """Affine SceneSphere/SceneBox by forward 3x4 transform (GpuSceneBuilder.java:37-78)."""
        k = {"sphere": N.MRAD_PRIM_AFFINE_SPHERE, "box": N.MRAD_PRIM_AFFINE_BOX}[kind]
        t, m = _f64(transforms, 12), _f64(materials, 13)
        assert len(t) == len(m)
        N.check(N.lib.mrad_scene_add_affine_transforms(self._h, k, len(t), N.ptr(t, N.dp), N.ptr(m, N.dp)))
        self._touch()
        return self

    def add_affine(self, kind: str, inverse12, inverse_transpose12, center3, bounds6, materials) -> "GpuScene":
        """GpuScene.addAffineSphere/addAffineBox with explicit payload (GpuScene.java:457-479)."""
        k = {"sphere": N.MRAD_PRIM_AFFINE_SPHERE, "box": N.MRAD_PRIM_AFFINE_BOX}[kind]
        a, b, c, d, m = (_f64(inverse12, 12), _f64(inverse_transpose12, 12), _f64(center3, 3), _f64(bounds6, 6),
                         _f64(materials, 13))
        N.check(N.lib.mrad_scene_add_affine(self._h, k, len(a), *(N.ptr(x, N.dp) for x in (a, b, c, d, m))))
        self._touch()
        return self

    # ---- packed data (GpuScene.java:507-607)
    def arrays(self) -> N.SceneArrays:
        if self._view is None:
            v = N.SceneArrays()
            N.check(N.lib.mrad_scene_view(self._h, C.byref(v)))
            self._view = v
        return self._view

    def _np(self, field, count, width, dtype):
        v = self.arrays()
        n = getattr(v, count) if isinstance(count, str) else count
        if n == 0:
            return np.zeros((0, width), dtype)
        p = getattr(v, field)
        return np.ctypeslib.as_array(p, shape=(n * width,)).view(dtype).reshape(n, width).copy()

    def sphere_data(self): return self._np("spheres", "sphere_count", 5, np.float32)
    def box_data(self): return self._np("boxes", "box_count", 7, np.float32)
    def plane_data(self): return self._np("planes", "plane_count", 26, np.float32)
    def affine_sphere_data(self): return self._np("affine_spheres", "affine_sphere_count", 22, np.float32)
    def affine_box_data(self): return self._np("affine_boxes", "affine_box_count", 22, np.float32)
    def material_data(self): return self._np("materials", "material_count", 13, np.float32)
    def bvh_node_bounds(self): return self._np("node_bounds", "node_count", 6, np.float32)
    def bvh_node_data(self): return self._np("node_data", "node_count", 4, np.int32)
    def bvh_primitive_refs(self): return self._np("primitive_refs", "primitive_ref_count", 1, np.int32).reshape(-1)
    def bvh_root_indices(self): return self._np("root_indices", "root_count", 1, np.int32).reshape(-1)

    def background(self):
        return np.array(self.arrays().background[:], dtype=np.float32)

    @property
    def revision(self) -> int:
        return int(self.arrays().revision)

    def counts(self) -> dict:
        v = self.arrays()
        return {"spheres": v.sphere_count, "boxes": v.box_count, "planes": v.plane_count,
                "affine_spheres": v.affine_sphere_count, "affine_boxes": v.affine_box_count,
                "materials": v.material_count, "nodes": v.node_count, "refs": v.primitive_ref_count}

    def primitive_count(self) -> int:
        c = self.counts()
        return c["spheres"] + c["boxes"] + c["planes"] + c["affine_spheres"] + c["affine_boxes"]

    def material_count(self) -> int:
        """materialCount (GpuScene.java:493-495): the distinct materials registered so far."""
        return self.counts()["materials"]

    def material_index_of(self, material) -> int:
        """materialIndexOf (:497-500): the packed index of a material given as its 13 values in packed order
        (diffuse, reflective, refractive, emittance, ior), or -1 if the scene has no such material."""
        row = np.asarray(material, np.float64).reshape(13).astype(np.float32)
        hits = np.flatnonzero(np.all(self.material_data() == row[None, :], axis=1))
        return int(hits[0]) if len(hits) else -1

    def bvh_stats(self) -> N.BvhStats:
        s = N.BvhStats()
        N.check(N.lib.mrad_scene_bvh_stats(self._h, C.byref(s)))
        return s

    def wire(self) -> dict:
        """All packed arrays as numpy copies (the reference's upload set)."""
        return {"spheres": self.sphere_data(), "boxes": self.box_data(), "planes": self.plane_data(),
                "affine_spheres": self.affine_sphere_data(), "affine_boxes": self.affine_box_data(),
                "node_bounds": self.bvh_node_bounds(), "node_data": self.bvh_node_data(),
                "refs": self.bvh_primitive_refs(), "roots": self.bvh_root_indices(),
                "materials": self.material_data(), "background": self.background()}
