"""ctypes face of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product (master-rad_amd/)
never imports it.

Scenes are passed in the reference's packed wire format (GpuScene.java:24-31)
as produced by mrad.GpuScene.wire().
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
STAT_COUNT = 14

fp = C.POINTER(C.c_float)
ip = C.POINTER(C.c_int32)


class OrScene(C.Structure):
    _fields_ = [("spheres", fp), ("sphereCount", C.c_int), ("boxes", fp), ("boxCount", C.c_int),
                ("planes", fp), ("planeCount", C.c_int), ("affineSpheres", fp), ("affineSphereCount", C.c_int),
                ("affineBoxes", fp), ("affineBoxCount", C.c_int), ("nodeBounds", fp), ("nodeCount", C.c_int),
                ("nodeData", ip), ("refs", ip), ("refCount", C.c_int), ("roots", ip), ("rootCount", C.c_int),
                ("useBvh", C.c_int), ("materials", fp), ("materialCount", C.c_int), ("bg", C.c_float * 3),
                ("stackSize", C.c_int)]


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def _lib() -> C.CDLL:
    if not LIB_PATH.exists():
        build()
    lib = C.CDLL(str(LIB_PATH))
    lib.or_hash_u32.restype = C.c_uint32
    lib.or_hash_u32.argtypes = [C.c_uint32]
    lib.or_rand01.restype = C.c_float
    lib.or_rand01.argtypes = [C.POINTER(C.c_uint32)]
    lib.or_sincosf.restype = None
    lib.or_sincosf.argtypes = [C.c_float, fp, fp]
    lib.or_path_seed.restype = C.c_uint32
    lib.or_path_seed.argtypes = [C.c_uint64, C.c_int, C.c_uint32, C.c_int]
    lib.or_camera_ray.restype = None
    lib.or_camera_ray.argtypes = [fp, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_uint32), fp]
    lib.or_render_pixels.restype = C.c_int
    lib.or_render_pixels.argtypes = [C.POINTER(OrScene), fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_uint64, C.c_int, ip, C.c_int, fp, C.POINTER(C.c_uint64), C.c_int]
    lib.or_trace_replay.restype = C.c_int
    lib.or_trace_replay.argtypes = [C.POINTER(OrScene), fp, C.c_int, ip, fp, ip, ip, fp, C.POINTER(C.c_uint32)]
    lib.or64_render.restype = C.c_int
    lib.or64_render.argtypes = [C.POINTER(OrScene), fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64,
                                C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int]
    lib.or64_render_window.restype = C.c_int
    lib.or64_render_window.argtypes = [C.POINTER(OrScene), fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.c_int, C.c_int, C.c_int, C.c_uint64, C.POINTER(C.c_double),
                                       C.POINTER(C.c_double), C.c_int]
    lib.or_adjoint.restype = C.c_int
    lib.or_adjoint.argtypes = [C.POINTER(OrScene), fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64,
                               C.c_int, fp, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int]
    lib.or_cost_accumulator.restype = C.c_float
    lib.or_cost_accumulator.argtypes = [fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    lib.or_record_paths.restype = C.c_int
    lib.or_record_paths.argtypes = [C.POINTER(OrScene), fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_int,
                                    C.c_int, ip, C.POINTER(C.c_uint8), fp, C.POINTER(C.c_uint8),
                                    C.POINTER(C.c_uint8)]
    lib.or_render_fixed_paths64.restype = C.c_int
    lib.or_render_fixed_paths64.argtypes = [C.POINTER(OrScene), fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64,
                                            C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.or64_closest.restype = C.c_int
    lib.or64_closest.argtypes = [C.POINTER(OrScene), fp, C.POINTER(C.c_double), ip, ip, C.POINTER(C.c_double)]
    return lib


lib = _lib()


class Scene:
    """Holds numpy copies of the wire arrays and the C view of them."""

    def __init__(self, wire: dict, use_bvh: bool = True, stack_size: int = 32):
        self.w = {k: np.ascontiguousarray(v) for k, v in wire.items()}
        w = self.w

        def f(name):
            a = w[name]
            return a.ctypes.data_as(fp) if a.size else None

        def i(name):
            a = w[name]
            return a.ctypes.data_as(ip) if a.size else None

        s = OrScene()
        s.spheres, s.sphereCount = f("spheres"), len(w["spheres"])
        s.boxes, s.boxCount = f("boxes"), len(w["boxes"])
        s.planes, s.planeCount = f("planes"), len(w["planes"])
        s.affineSpheres, s.affineSphereCount = f("affine_spheres"), len(w["affine_spheres"])
        s.affineBoxes, s.affineBoxCount = f("affine_boxes"), len(w["affine_boxes"])
        s.nodeBounds, s.nodeCount = f("node_bounds"), (len(w["node_bounds"]) if use_bvh else 0)
        s.nodeData = i("node_data")
        s.refs, s.refCount = i("refs"), (len(w["refs"]) if use_bvh else 0)
        s.roots, s.rootCount = i("roots"), (len(w["roots"]) if use_bvh else 0)
        s.useBvh = 1 if use_bvh else 0
        s.materials, s.materialCount = f("materials"), len(w["materials"])
        for k in range(3):
            s.bg[k] = float(w["background"][k])
        s.stackSize = stack_size
        self.c = s


def hash_u32(x: int) -> int:
    return lib.or_hash_u32(x)


def sincosf(x: float) -> tuple[float, float]:
    s, c = C.c_float(), C.c_float()
    lib.or_sincosf(x, C.byref(s), C.byref(c))
    return s.value, c.value


def path_seed(frame_seed: int, frame_index: int, pixel: int, sample: int) -> int:
    return lib.or_path_seed(frame_seed, frame_index, pixel, sample)


def camera_ray(camera: np.ndarray, width: int, height: int, x: int, y: int, rng: int):
    """renderKernel's camera ray of pixel (x, y) from RNG state rng: (ray6 float32, rng after the two jitter draws)."""
    cam = np.ascontiguousarray(camera, np.float32)
    state = C.c_uint32(rng & 0xFFFFFFFF)
    out = np.zeros(6, np.float32)
    lib.or_camera_ray(cam.ctypes.data_as(fp), width, height, x, y, C.byref(state), out.ctypes.data_as(fp))
    return out, state.value


def trace_replay(scene: Scene, rays: np.ndarray):
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1)
    n = len(rays) // 6
    flags = np.zeros(n, np.int32)
    dist = np.zeros(n, np.float32)
    order = np.zeros(n, np.int32)
    mat = np.zeros(n, np.int32)
    normals = np.zeros(3 * n, np.float32)
    per = np.zeros((n, STAT_COUNT), np.uint32)
    rc = lib.or_trace_replay(C.byref(scene.c), rays.ctypes.data_as(fp), n, flags.ctypes.data_as(ip),
                             dist.ctypes.data_as(fp), order.ctypes.data_as(ip), mat.ctypes.data_as(ip),
                             normals.ctypes.data_as(fp), per.ctypes.data_as(C.POINTER(C.c_uint32)))
    assert rc == 0
    return flags, dist, order, mat, normals, per


def render_pixels(scene: Scene, camera: np.ndarray, width: int, height: int, max_depth: int, frame_index: int,
                  frame_seed: int, spp: int, pixels=None, rr_start: int = -1, threads: int | None = None):
    """Float image (n, 3) of the listed pixels (all when None) and the 14 frame counters."""
    cam = np.ascontiguousarray(camera, np.float32)
    if pixels is not None:
        pixels = np.ascontiguousarray(pixels, np.int32)
        n = len(pixels)
    else:
        n = width * height
    out = np.zeros((n, 3), np.float32)
    stats = (C.c_uint64 * STAT_COUNT)()
    threads = threads or min(16, os.cpu_count() or 1)
    rc = lib.or_render_pixels(C.byref(scene.c), cam.ctypes.data_as(fp), width, height, max_depth, rr_start,
                              frame_index, frame_seed & ((1 << 64) - 1), spp,
                              pixels.ctypes.data_as(ip) if pixels is not None else None, n,
                              out.ctypes.data_as(fp), stats, threads)
    assert rc == 0
    return out, np.array(stats[:], dtype=np.uint64)


def render64(scene: Scene, camera: np.ndarray, width: int, height: int, max_depth: int, spp: int, seed: int,
             rr_start: int = -1, threads: int | None = None, variance: bool = False):
    cam = np.ascontiguousarray(camera, np.float32)
    mean = np.zeros((height * width, 3), np.float64)
    var = np.zeros((height * width, 3), np.float64) if variance else None
    threads = threads or min(16, os.cpu_count() or 1)
    rc = lib.or64_render(C.byref(scene.c), cam.ctypes.data_as(fp), width, height, max_depth, rr_start, spp,
                         seed & ((1 << 64) - 1), mean.ctypes.data_as(C.POINTER(C.c_double)),
                         var.ctypes.data_as(C.POINTER(C.c_double)) if variance else None, threads)
    assert rc == 0
    return (mean, var) if variance else mean


def render64_window(scene: Scene, camera: np.ndarray, width: int, height: int, window, max_depth: int, spp: int,
                    seed: int, rr_start: int = -1, threads: int | None = None):
    """FP64 mean and per-pixel sample variance ((ch*cw, 3) each) of the window (x0, y0, cw, ch) of a
    width x height frame: the same pixels, camera mapping and random numbers as render64's full frame."""
    x0, y0, cw, ch = window
    cam = np.ascontiguousarray(camera, np.float32)
    mean = np.zeros((ch * cw, 3), np.float64)
    var = np.zeros((ch * cw, 3), np.float64)
    threads = threads or min(16, os.cpu_count() or 1)
    rc = lib.or64_render_window(C.byref(scene.c), cam.ctypes.data_as(fp), width, height, x0, y0, cw, ch, max_depth,
                                rr_start, spp, seed & ((1 << 64) - 1), mean.ctypes.data_as(C.POINTER(C.c_double)),
                                var.ctypes.data_as(C.POINTER(C.c_double)), threads)
    assert rc == 0
    return mean, var


def adjoint(scene: Scene, camera: np.ndarray, width: int, height: int, max_depth: int, frame_index: int,
            frame_seed: int, spp: int, grad_image: np.ndarray, primitive_count: int, rr_start: int = -1,
            threads: int | None = None):
    """Path-replay adjoint with FP32 paths and FP64 sums: (grad_albedo, grad_emission) per primitive."""
    cam = np.ascontiguousarray(camera, np.float32)
    g = np.ascontiguousarray(grad_image, np.float32).reshape(-1)
    ga = np.zeros(primitive_count * 3, np.float64)
    ge = np.zeros(primitive_count * 3, np.float64)
    threads = threads or min(16, os.cpu_count() or 1)
    rc = lib.or_adjoint(C.byref(scene.c), cam.ctypes.data_as(fp), width, height, max_depth, rr_start, frame_index,
                        frame_seed & ((1 << 64) - 1), spp, g.ctypes.data_as(fp),
                        ga.ctypes.data_as(C.POINTER(C.c_double)), ge.ctypes.data_as(C.POINTER(C.c_double)), threads)
    assert rc == 0
    return ga.reshape(-1, 3), ge.reshape(-1, 3)


class RecordedPaths:
    """The vertices of a frame's paths (or_record_paths): for the FP64 re-evaluation of the frame's estimator
    under other material parameters, an exact polynomial in albedo and emission for Lambertian materials."""

    def __init__(self, scene: Scene, camera: np.ndarray, width: int, height: int, max_depth: int, frame_index: int,
                 frame_seed: int, spp: int, threads: int | None = None):
        n = width * height * spp
        self.width, self.height, self.spp, self.depth = width, height, spp, max(1, max_depth)
        self.mat = np.full((n, self.depth), -1, np.int32)
        self.lobe = np.zeros((n, self.depth), np.uint8)
        self.prob = np.zeros((n, self.depth), np.float32)
        self.count = np.zeros(n, np.uint8)
        self.end = np.zeros(n, np.uint8)
        self.background = np.asarray(scene.w["background"], np.float64).reshape(3)
        cam = np.ascontiguousarray(camera, np.float32)
        threads = threads or min(16, os.cpu_count() or 1)
        u8 = C.POINTER(C.c_uint8)
        rc = lib.or_record_paths(C.byref(scene.c), cam.ctypes.data_as(fp), width, height, max_depth, frame_index,
                                 frame_seed & ((1 << 64) - 1), spp, threads, self.mat.ctypes.data_as(ip),
                                 self.lobe.ctypes.data_as(u8), self.prob.ctypes.data_as(fp),
                                 self.count.ctypes.data_as(u8), self.end.ctypes.data_as(u8))
        assert rc == 0
        self._index = None

    def radiance(self, materials64: np.ndarray, paths=None) -> np.ndarray:
        """FP64 radiance (k, 3) of the given paths (all when None) under the material table materials64 (m x 13):
        sum over vertices of T_i Le_i with T_{i+1} = T_i c_i[lobe] / p_i, plus T_n times the background when the
        last ray missed (or_render_fixed_paths64's arithmetic)."""
        sel = slice(None) if paths is None else paths
        mat, lobe, prob = self.mat[sel], self.lobe[sel], self.prob[sel]
        count, end = self.count[sel], self.end[sel]
        flat = np.ascontiguousarray(materials64, np.float64).reshape(-1)
        k = len(count)
        T = np.ones((k, 3))
        L = np.zeros((k, 3))
        for i in range(self.depth):
            live = np.flatnonzero(count > i)
            if len(live) == 0:
                break
            mi = mat[live, i].astype(np.int64)
            ok = mi >= 0
            base = np.where(ok, mi, 0) * 13
            Tl = T[live]
            le = np.stack([flat[base + 9], flat[base + 10], flat[base + 11]], 1)
            L[live] += np.where(ok[:, None], Tl * le, 0.0)
            # every live vertex scatters but one at which the path was absorbed
            scat = ok & ~((end[live] == 2) & (count[live] == i + 1))
            col = base + 3 * lobe[live, i].astype(np.int64)
            p = prob[live, i].astype(np.float64)
            inv = np.where(p > 0, 1.0 / np.where(p > 0, p, 1.0), 0.0)
            f = np.stack([flat[col], flat[col + 1], flat[col + 2]], 1) * inv[:, None]
            T[live] = np.where(scat[:, None], Tl * f, np.where(ok[:, None], Tl, 0.0))
        L += np.where((end == 1)[:, None], T * self.background[None, :], 0.0)
        return L

    def image(self, materials64: np.ndarray) -> np.ndarray:
        """The frame (W*H*3 doubles): each pixel the mean of its spp path radiances."""
        return self.radiance(materials64).reshape(-1, self.spp, 3).mean(axis=1).reshape(-1)

    def paths_of_material(self, m: int) -> np.ndarray:
        """Indices of the paths with a vertex on material m."""
        if self._index is None:
            flat = self.mat.reshape(-1)
            order = np.argsort(flat, kind="stable")
            self._index = (flat[order], order // self.depth)
        keys, paths = self._index
        lo, hi = np.searchsorted(keys, m, "left"), np.searchsorted(keys, m, "right")
        return np.unique(paths[lo:hi])


def render_fixed_paths64(scene: Scene, camera: np.ndarray, width: int, height: int, max_depth: int, frame_index: int,
                         frame_seed: int, spp: int, materials64: np.ndarray) -> np.ndarray:
    """The frame's FP32 paths with radiance re-evaluated in double from materials64 (n x 13)."""
    cam = np.ascontiguousarray(camera, np.float32)
    m = np.ascontiguousarray(materials64, np.float64).reshape(-1)
    out = np.zeros(width * height * 3, np.float64)
    rc = lib.or_render_fixed_paths64(C.byref(scene.c), cam.ctypes.data_as(fp), width, height, max_depth, frame_index,
                                     frame_seed & ((1 << 64) - 1), spp, m.ctypes.data_as(C.POINTER(C.c_double)),
                                     out.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == 0
    return out


def closest64(scene: Scene, ray6):
    r = np.ascontiguousarray(ray6, np.float32)
    t = C.c_double()
    kind, index = C.c_int32(), C.c_int32()
    n = (C.c_double * 3)()
    f = lib.or64_closest(C.byref(scene.c), r.ctypes.data_as(fp), C.byref(t), C.byref(kind), C.byref(index), n)
    return bool(f), t.value, kind.value, index.value, np.array(n[:])


def cost_accumulators(workload: np.ndarray, records: int, profiles: int, op: int, profile: int, offset: int,
                      iterations: int, copies: int, threads: int) -> np.ndarray:
    """or_cost_accumulator of threads 0..threads-1: the calibration kernel's per-thread outputs."""
    w = np.ascontiguousarray(workload, np.float32)
    p = w.ctypes.data_as(fp)
    return np.array([lib.or_cost_accumulator(p, records, profiles, op, profile, offset, iterations, copies, t)
                     for t in range(threads)], np.float32)
