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
