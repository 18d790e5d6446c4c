"""Shared inputs and comparisons for the parity tests."""
from __future__ import annotations

import hashlib
import struct

import numpy as np


def camera_grid_rays(camera: np.ndarray, width: int, height: int, phase=(0.5, 0.5)) -> np.ndarray:
    """Pinhole ray grid from the packed 12-float camera (BenchmarkRays.primaryGrid, pixel phase)."""
    cam = np.asarray(camera, np.float64)
    pos, fwd, right, up = cam[0:3], cam[3:6], cam[6:9], cam[9:12]
    aspect = width / height
    ys, xs = np.mgrid[0:height, 0:width]
    u = ((xs + phase[0]) * 2.0 / width - 1.0) * aspect
    v = 1.0 - (ys + phase[1]) * 2.0 / height
    d = fwd[None, None] + right[None, None] * u[..., None] + up[None, None] * v[..., None]
    o = np.broadcast_to(pos, d.shape)
    return np.concatenate([o, d], -1).reshape(-1, 6).astype(np.float32)


def random_rays(rng: np.random.Generator, n: int, box=1.2, inside_fraction=0.3) -> np.ndarray:
    """Rays from outside and inside the scene cube with random, partly axis-aligned directions."""
    o = rng.uniform(-box, box, (n, 3))
    outside = rng.random(n) >= inside_fraction
    o[outside] *= 2.5
    target = rng.uniform(-1, 1, (n, 3))
    d = target - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d *= rng.uniform(0.5, 2.0, (n, 1))           # unnormalized directions, like the integrator's
    zero = rng.random((n, 3)) < 0.05             # exact zero components exercise the d == 0 branches
    d[zero] = 0.0
    d[np.all(d == 0, axis=1), 2] = 1.0
    return np.concatenate([o, d], 1).astype(np.float32)


def topology_sha256(scene) -> str:
    """MethodCatalog.topologySha256 (benchmark/MethodCatalog.java:164-172, :208-223)."""
    h = hashlib.sha256()
    for arr, is_float in ((scene.bvh_node_bounds().reshape(-1), True), (scene.bvh_node_data().reshape(-1), False),
                          (scene.bvh_primitive_refs(), False), (scene.bvh_root_indices(), False)):
        h.update(struct.pack(">i", len(arr)))
        h.update(arr.astype(">f4" if is_float else ">i4").tobytes())
    return h.hexdigest()


def assert_replay_equal(gpu, ref, check_counters=True):
    """Bit-exact comparison of a GPU TraceReplayResult with the oracle tuple."""
    flags, dist, order, mat, normals, per = ref
    np.testing.assert_array_equal(gpu.hit_flags, flags)
    np.testing.assert_array_equal(gpu.hit_primitive_orders, order)
    np.testing.assert_array_equal(gpu.hit_material_indices, mat)
    np.testing.assert_array_equal(gpu.hit_distances.view(np.uint32), dist.view(np.uint32))
    np.testing.assert_array_equal(gpu.hit_normals.view(np.uint32), normals.view(np.uint32))
    if check_counters and gpu.per_ray_counters is not None:
        np.testing.assert_array_equal(gpu.per_ray_counters, per)


def corrupt_view(scene, **arrays):
    """A copy of the scene's wire view (mrad_scene_arrays) with some arrays replaced, e.g.
    corrupt_view(sc, primitive_refs=refs, spheres=sph).  The replacement numpy arrays are returned with the
    view so that they outlive it."""
    import ctypes as C

    from mrad import _native as N
    v = N.SceneArrays()
    C.pointer(v)[0] = scene.arrays()
    keep = []
    for name, arr in arrays.items():
        is_int = name in ("primitive_refs", "node_data", "root_indices")
        keep.append(np.ascontiguousarray(arr, np.int32 if is_int else np.float32))
        setattr(v, name, keep[-1].ctypes.data_as(N.ip if is_int else N.fp))
    v.revision = v.revision + 1000
    return v, keep


def bad_material_variants(scene):
    """Wire arrays of `scene` with one material index out of range (past the table, below -1, non-integral
    or NaN) in each primitive class the scene has; each item is (label, {array name: array})."""
    out = []
    mc = scene.counts()["materials"]
    for name, data, col in (("spheres", scene.sphere_data(), 4), ("boxes", scene.box_data(), 6),
                            ("affine_spheres", scene.affine_sphere_data(), 21),
                            ("affine_boxes", scene.affine_box_data(), 21), ("planes", scene.plane_data(), 24)):
        if not len(data):
            continue
        for bad in (float(mc), -2.0, 0.5, float("nan")):
            arr = np.array(data, np.float32, copy=True)
            arr[len(arr) // 2, col] = bad
            out.append((f"{name}={bad}", {name: arr}))
    return out


class WireScene:
    """A scene given by its wire arrays, as a caller that builds its own BVH would pass them to the boundary (for
    example the per-type forests of the reference's PER_TYPE_SAH modes, up to four roots).  It provides what the
    tracer and the oracle use of GpuScene: arrays(), key and wire()."""
    _next_key = 1 << 40

    def __init__(self, wire: dict):
        import ctypes as C

        from mrad import _native as N
        self._wire = dict(wire)
        self.key = WireScene._next_key
        WireScene._next_key += 1
        v = N.SceneArrays()
        self._keep = []

        def put(field, count_field, name, ptr, dtype, width):
            a = np.ascontiguousarray(wire[name], dtype).reshape(-1)
            self._keep.append(a)
            setattr(v, field, a.ctypes.data_as(ptr))
            if count_field:
                setattr(v, count_field, len(a) // width)

        put("spheres", "sphere_count", "spheres", N.fp, np.float32, 5)
        put("boxes", "box_count", "boxes", N.fp, np.float32, 7)
        put("planes", "plane_count", "planes", N.fp, np.float32, 26)
        put("affine_spheres", "affine_sphere_count", "affine_spheres", N.fp, np.float32, 22)
        put("affine_boxes", "affine_box_count", "affine_boxes", N.fp, np.float32, 22)
        put("node_bounds", "node_count", "node_bounds", N.fp, np.float32, 6)
        put("node_data", None, "node_data", N.ip, np.int32, 4)
        put("primitive_refs", "primitive_ref_count", "refs", N.ip, np.int32, 1)
        put("root_indices", "root_count", "roots", N.ip, np.int32, 1)
        put("materials", "material_count", "materials", N.fp, np.float32, 13)
        v.background = (C.c_float * 3)(*[float(x) for x in wire["background"]])
        v.revision = 1
        self._view = v

    def arrays(self):
        return self._view

    def wire(self) -> dict:
        return self._wire


def per_type_forest(full, parts) -> WireScene:
    """The wire arrays of `full` with its BVH replaced by a forest of one tree per primitive class: the trees of
    `parts` (scenes holding one class each, built from the same primitives in the same order as `full`),
    concatenated in order with their child and ref indices rebased; one root per tree."""
    bounds, data, refs, roots = [], [], [], []
    node_base = ref_base = 0
    for p in parts:
        nd = p.bvh_node_data().copy()
        interior = nd[:, 0] >= 0
        nd[interior, 0] += node_base
        nd[interior, 1] += node_base
        nd[~interior, 2] += ref_base
        bounds.append(p.bvh_node_bounds())
        data.append(nd)
        refs.append(p.bvh_primitive_refs())
        roots.extend(int(r) + node_base for r in p.bvh_root_indices())
        node_base += len(nd)
        ref_base += len(refs[-1])
    wire = full.wire()
    wire.update(node_bounds=np.concatenate(bounds), node_data=np.concatenate(data), refs=np.concatenate(refs),
                roots=np.array(roots, np.int32))
    return WireScene(wire)
