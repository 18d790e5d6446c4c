"""The reference's fixed benchmark workloads: Layered Hero scenes, their camera
and the metrics-frame seed, so that whole frames of the reference's own
evaluation can be rendered here and their per-frame counters compared with
the published ones.

  LayeredHeroFamily.create/camera   graphics3d/scene/v1/LayeredHeroFamily.java:55-136
  SceneSupport.sphere/box/camera    graphics3d/scene/v1/SceneSupport.java:42-89
  Affine composition and rotations  graphics3d/Affine.java:35-41, :114-136
  BenchmarkRays.gpuCamera           benchmark/BenchmarkRays.java:28-66
  BenchmarkProtocol.metricsSeed     benchmark/BenchmarkProtocol.java:295-297, :385-400
  BenchmarkWorkloads.fixed          benchmark/BenchmarkWorkloads.java:48-70

All arithmetic is IEEE double in the reference's operation order (Python
floats are doubles and evaluate left to right like Java), so the packed FP32
scene is the reference's bit for bit.  Angles of the reference's Affine
rotations are in turns; their sine and cosine are taken as sin(2 pi t) and
cos(2 pi t) of the C library.
"""
from __future__ import annotations

import math

import numpy as np

from .scene import BvhBuildConfig, GpuScene, material_array
from .tracer import GpuCamera

# ---- BenchmarkProtocol (benchmark/BenchmarkProtocol.java:39-46, :62-63)
WIDTH, HEIGHT = 1920, 1080
SAMPLES_PER_PIXEL = 1
MAXIMUM_PATH_DEPTH = 12
STACK_CAPACITY = 32
LEAF_SIZE = 8
WEIGHTED_LAMBDA = 1.0
HERO_SCALES = (96, 1_000, 10_000, 100_000, 1_000_000)
HERO_LEAF_SIZES = (1, 2, 4, 8, 16, 32, 64)
HERO_LAMBDAS = (0.0, 0.25, 0.5, 0.75, 1.0, 2.0, 5.0, 10.0, 20.0, 50.0, 100.0)
PUBLICATION_ROW_IDS = ("hero-96", "hero-1k", "hero-10k", "hero-100k", "hero-1m", "overlap-10k")

_M64 = (1 << 64) - 1
_MIX = 0x9E3779B97F4A7C15


def _rotl(v: int, r: int) -> int:
    return ((v << r) | (v >> (64 - r))) & _M64


def seed_domain(label: str) -> int:
    """BenchmarkProtocol.seedDomain (:402-408): Java String-hash style, 64-bit wrap."""
    v = 1_125_899_906_842_597
    for ch in label:
        v = (31 * v + ord(ch)) & _M64
    return v


def _seed(namespace: int, first: int, second: int) -> int:
    """BenchmarkProtocol.seed (:385-393); returns the unsigned 64-bit value."""
    v = namespace ^ _rotl(((first + 1) * _MIX) & _M64, 17) ^ _rotl(((second + 1) * _MIX) & _M64, 41)
    v = ((v ^ (v >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    v = ((v ^ (v >> 27)) * 0x94D049BB133111EB) & _M64
    v ^= v >> 31
    return namespace if v == 0 else v


METRICS_SEED_DOMAIN = seed_domain("evaluation metrics")
TIMING_SEED_DOMAIN = seed_domain("evaluation timing")


def workload_key(row_id: str) -> int:
    """BenchmarkProtocol.workloadKey (:279-289) of a fixed (Hero or overlap) row: its ordinal."""
    return PUBLICATION_ROW_IDS.index(row_id)


def metrics_seed(key: int) -> int:
    """BenchmarkProtocol.metricsSeed (:295-297): the frame seed of the counter frame."""
    return _seed(METRICS_SEED_DOMAIN, key, 0)


def timing_seed(key: int, comparison: int) -> int:
    return _seed(TIMING_SEED_DOMAIN, key, comparison)


# ---- Affine (graphics3d/Affine.java), rows (m00..m03, m10..m13, m20..m23)
IDENTITY = (1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0)


def then(a, t):
    """Affine.then (:35-41): first a, then t (t * a), in the reference's summation order."""
    m00, m01, m02, m03, m10, m11, m12, m13, m20, m21, m22, m23 = a
    t00, t01, t02, t03, t10, t11, t12, t13, t20, t21, t22, t23 = t
    return (t00 * m00 + t01 * m10 + t02 * m20, t00 * m01 + t01 * m11 + t02 * m21,
            t00 * m02 + t01 * m12 + t02 * m22, t00 * m03 + t01 * m13 + t02 * m23 + t03,
            t10 * m00 + t11 * m10 + t12 * m20, t10 * m01 + t11 * m11 + t12 * m21,
            t10 * m02 + t11 * m12 + t12 * m22, t10 * m03 + t11 * m13 + t12 * m23 + t13,
            t20 * m00 + t21 * m10 + t22 * m20, t20 * m01 + t21 * m11 + t22 * m21,
            t20 * m02 + t21 * m12 + t22 * m22, t20 * m03 + t21 * m13 + t22 * m23 + t23)


def sin_t(turns: float) -> float:
    return math.sin(turns * 2.0 * math.pi)


def cos_t(turns: float) -> float:
    return math.cos(turns * 2.0 * math.pi)


def translation(x, y, z):
    return (1.0, 0.0, 0.0, x, 0.0, 1.0, 0.0, y, 0.0, 0.0, 1.0, z)


def scaling(x, y, z):
    return (x, 0.0, 0.0, 0.0, 0.0, y, 0.0, 0.0, 0.0, 0.0, z, 0.0)


def rotation_about_x(t):
    c, s = cos_t(t), sin_t(t)
    return (1.0, 0.0, 0.0, 0.0, 0.0, c, -s, 0.0, 0.0, s, c, 0.0)


def rotation_about_y(t):
    c, s = cos_t(t), sin_t(t)
    return (c, 0.0, s, 0.0, 0.0, 1.0, 0.0, 0.0, -s, 0.0, c, 0.0)


def rotation_about_z(t):
    c, s = cos_t(t), sin_t(t)
    return (c, -s, 0.0, 0.0, s, c, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0)


def _at(a, v):
    m00, m01, m02, m03, m10, m11, m12, m13, m20, m21, m22, m23 = a
    x, y, z = v
    return (m00 * x + m01 * y + m02 * z + m03, m10 * x + m11 * y + m12 * z + m13, m20 * x + m21 * y + m22 * z + m23)


def _linear(a, v):
    m00, m01, m02, _, m10, m11, m12, _, m20, m21, m22, _ = a
    x, y, z = v
    return (m00 * x + m01 * y + m02 * z, m10 * x + m11 * y + m12 * z, m20 * x + m21 * y + m22 * z)


# ---- cameras (cameras/Perspective.java, cameras/TransformedCamera.java, SceneSupport.camera :64-77)
class Camera:
    """TransformedCamera(Perspective(k), transform): the sensor point (sx, sy)
    leaves the origin along (sx, sy, 1/k), then the transform is applied."""

    def __init__(self, k: float, transform=IDENTITY):
        self.k = k
        self.transform = transform

    def ray(self, sx: float, sy: float):
        d = (sx, sy, 1 / self.k)
        return _at(self.transform, (0.0, 0.0, 0.0)), _linear(self.transform, d)


def support_camera(position, pitch: float, yaw: float, roll: float = 0.0) -> Camera:
    t = then(then(then(then(IDENTITY, translation(*position)), rotation_about_x(pitch)), rotation_about_y(yaw)),
             rotation_about_z(roll))
    return Camera(1.0 / 3.0, t)


def _sensor(x: float, y: float, width: int, height: int):
    aspect = width / height
    return ((-1.0 + 2.0 * x / width) * aspect, 1.0 - 2.0 * y / height)


def gpu_camera(camera: Camera, width: int, height: int) -> GpuCamera:
    """BenchmarkRays.gpuCamera (:28-46): the linear camera through the centre pixel
    and its right and upper neighbours."""
    cx, cy = (width - 1) / 2.0, (height - 1) / 2.0
    c = _sensor(cx + 0.5, cy + 0.5, width, height)
    r = _sensor(min(width - 1, cx + 1.0) + 0.5, cy + 0.5, width, height)
    u = _sensor(cx + 0.5, min(height - 1, cy + 1.0) + 0.5, width, height)
    cp, cd = camera.ray(*c)
    _, rd = camera.ray(*r)
    _, ud = camera.ray(*u)
    right = tuple((a - b) / (r[0] - c[0]) for a, b in zip(rd, cd))
    up = tuple((a - b) / (u[1] - c[1]) for a, b in zip(ud, cd))
    fwd = tuple(d - rr * c[0] - uu * c[1] for d, rr, uu in zip(cd, right, up))
    return GpuCamera(cp, fwd, right, up)


def primary_grid(camera: Camera, width: int, height: int, phase_x: float = 0.5, phase_y: float = 0.5) -> np.ndarray:
    """BenchmarkRays.primaryGrid (:14-26): one ray per grid cell (x + phase_x, y + phase_y) of a width x height
    sensor, rows from the top, as float32 (origin, direction) records."""
    out = np.empty((width * height, 6), np.float32)
    k = 0
    for y in range(height):
        for x in range(width):
            o, d = camera.ray(*_sensor(x + phase_x, y + phase_y, width, height))
            out[k] = (*o, *d)
            k += 1
    return out


# ---- Layered Hero (graphics3d/scene/v1/LayeredHeroFamily.java)
PRIMITIVES_PER_MODULE = 8
_DEPTH_SEPARATION, _BOX_HALF_LENGTH, _BOX_DEPTH_TILT_TURNS = 0.750, 2.000, 0.150
_ASPECT = 16.0 / 9.0
_SENSOR_HALF_HEIGHT = 0.92
_SENSOR_HALF_WIDTH = _ASPECT * _SENSOR_HALF_HEIGHT
_PERSPECTIVE_Z = 3.0
_LAYER_OFFSET = 0.28
_MODULE_SCALE = 0.35
_DEPTH_STEP = 2.35
_OUTER_X, _OUTER_Y, _MIDDLE_Y = 0.72, 0.72, 0.76
_SPHERE_RADIUS = 0.150
_BOX_HALF_WIDTH, _BOX_HALF_DEPTH = 0.065, 0.110
_DIAGONAL_TURNS = 0.125
_BOX_NORMAL_OFFSET = 0.140


def _surface(emission, diffuse=(0.0, 0.0, 0.0), reflective=(0.0, 0.0, 0.0), refractive=(0.0, 0.0, 0.0), ior=1.5):
    """SceneSupport.surface -> Material.surface (Material.java:86-88) in packed order."""
    return material_array(diffuse, reflective, refractive, emission, ior)[0]


# Palette(green, cyan, blue, red, orange) (:44-49)
_GREEN = _surface((0.015, 0.56, 0.36))
_CYAN = _surface((0.015, 0.34, 0.78))
_BLUE = _surface((0.13, 0.48, 0.98))
_RED = _surface((1.52, 0.045, 0.008))
_ORANGE = _surface((1.28, 0.25, 0.010))


class _Grid:
    """LayeredHeroFamily.Grid (:226-273)."""

    def __init__(self, modules: int):
        nominal = max(1, math.ceil(math.sqrt(modules * _ASPECT)))
        self.rows = max(1, (modules + nominal - 1) // nominal)
        self.base_columns = modules // self.rows
        self.extra_rows = modules % self.rows
        self.cell_w = 2.0 * _SENSOR_HALF_WIDTH / nominal
        self.cell_h = 2.0 * _SENSOR_HALF_HEIGHT / self.rows

    def _centred_extra(self, row: int) -> bool:
        rows, extra = self.rows, self.extra_rows
        if extra <= 0:
            return False
        if rows & 1:
            dist = abs(row - rows // 2)
            return dist <= extra // 2 if extra & 1 else 0 < dist <= extra // 2
        pairs = extra // 2
        lower, upper = rows // 2 - pairs, rows // 2 + pairs
        return lower <= row < upper if (extra & 1) == 0 else lower <= row <= upper

    def columns_in_row(self, row: int) -> int:
        return self.base_columns + (1 if self._centred_extra(row) else 0)

    def sensor_x(self, column: int, in_row: int) -> float:
        return (column - (in_row - 1) * 0.5) * self.cell_w

    def sensor_y(self, row: int) -> float:
        return _SENSOR_HALF_HEIGHT - (row + 0.5) * self.cell_h


def _world(sx, sy, depth):
    return (sx * depth / _PERSPECTIVE_Z, sy * depth / _PERSPECTIVE_Z, depth)


def hero_primitives(primitive_count: int):
    """LayeredHeroFamily.create (:55-86) as typed lists in the reference's
    insertion order: spheres (centre, radius, material), axis-aligned boxes
    (min, max, material), affine boxes (forward 3x4 transform, material) and
    the background."""
    if primitive_count <= 0 or primitive_count % PRIMITIVES_PER_MODULE:
        raise ValueError(f"primitiveCount must be a positive multiple of 8: {primitive_count}")
    modules = primitive_count // PRIMITIVES_PER_MODULE
    grid = _Grid((modules + 2) // 3)
    spheres, sphere_mats, aboxes, abox_mats = [], [], [], []
    layer_cos = [math.cos((layer / 3.0 - 0.25) * 2.0 * math.pi) for layer in range(3)]
    layer_sin = [math.sin((layer / 3.0 - 0.25) * 2.0 * math.pi) for layer in range(3)]
    rot = {}
    for sign in (1, -1):
        base = sign * _DIAGONAL_TURNS
        radians = base * 2.0 * math.pi
        axis_x, axis_y = -math.sin(radians), math.cos(radians)
        rot[sign] = (base, axis_y, -axis_x)        # baseRotation, normalX, normalY

    def sphere(sx, sy, scale, depth, mat):
        spheres.append((*_world(sx, sy, depth), _SPHERE_RADIUS * scale * (depth / _PERSPECTIVE_Z)))
        sphere_mats.append(mat)

    module = 0
    for row in range(grid.rows):
        if module >= modules:
            break
        in_row = grid.columns_in_row(row)
        for column in range(in_row):
            if module >= modules:
                break
            cell_x, cell_y = grid.sensor_x(column, in_row), grid.sensor_y(row)
            cell_scale = min(grid.cell_w, grid.cell_h)
            for layer in range(3):
                if module >= modules:
                    break
                sx = cell_x + _LAYER_OFFSET * cell_scale * layer_cos[layer]
                sy = cell_y + _LAYER_OFFSET * cell_scale * layer_sin[layer]
                scale = _MODULE_SCALE * cell_scale
                depth = 4.8 + layer * _DEPTH_STEP
                sign = 1 if ((row + column + layer) & 1) == 0 else -1
                # addModule (:138-167)
                half_sep = 0.5 * _DEPTH_SEPARATION * (scale * depth / _PERSPECTIVE_Z)
                cheap, box_depth = depth - half_sep, depth + half_sep
                sphere(sx - _OUTER_X * scale, sy + _OUTER_Y * scale, scale, cheap, _GREEN)
                sphere(sx, sy + _MIDDLE_Y * scale, scale, cheap, _BLUE)
                sphere(sx + _OUTER_X * scale, sy + _OUTER_Y * scale, scale, cheap, _CYAN)
                sphere(sx - _OUTER_X * scale, sy - _OUTER_Y * scale, scale, cheap, _CYAN)
                sphere(sx, sy - _MIDDLE_Y * scale, scale, cheap, _GREEN)
                sphere(sx + _OUTER_X * scale, sy - _OUTER_Y * scale, scale, cheap, _BLUE)
                # addBoxPair (:169-204)
                base, nx, ny = rot[sign]
                for box in range(2):
                    s = -1.0 if box == 0 else 1.0
                    bx = sx + s * scale * _BOX_NORMAL_OFFSET * nx
                    by = sy + s * scale * _BOX_NORMAL_OFFSET * ny
                    screen = base + (0.0 if box == 0 else 0.5)
                    unit = scale * box_depth / _PERSPECTIVE_Z
                    t = then(IDENTITY, scaling(_BOX_HALF_WIDTH * unit, _BOX_HALF_LENGTH * unit, _BOX_HALF_DEPTH * unit))
                    t = then(t, rotation_about_x(s * _BOX_DEPTH_TILT_TURNS))
                    t = then(t, rotation_about_z(screen))
                    t = then(t, translation(*_world(bx, by, box_depth)))
                    aboxes.append(t)
                    abox_mats.append(_RED if box == 0 else _ORANGE)
                module += 1
    if len(spheres) + len(aboxes) != primitive_count:
        raise RuntimeError("Layered module count mismatch")
    # addRayTracingEnvironment(solids, 12.8) (:88-118)
    rear = 12.8
    mirror = _surface((0.002, 0.003, 0.006), reflective=(0.82, 0.88, 0.98))
    lens = _surface((0.004, 0.010, 0.018), reflective=(0.06, 0.08, 0.10), refractive=(0.88, 0.96, 1.00), ior=1.33)
    frame = _surface((0.025, 0.22, 0.72), diffuse=(0.010, 0.035, 0.10), reflective=(0.22, 0.30, 0.46))
    fz, fs = rear - 0.10, rear / 8.8
    boxes, box_mats = [], []
    for centre, radii, mat in (
            ((0.0, 0.0, rear + 0.045), (2.66 * fs, 1.57 * fs, 0.045), mirror),
            ((-2.72 * fs, 0.0, fz), (0.055 * fs, 1.68 * fs, 0.045), frame),
            ((2.72 * fs, 0.0, fz), (0.055 * fs, 1.68 * fs, 0.045), frame),
            ((0.0, 1.68 * fs, fz), (2.78 * fs, 0.055 * fs, 0.045), frame),
            ((0.0, -1.68 * fs, fz), (2.78 * fs, 0.055 * fs, 0.045), frame)):
        # SceneBox.axisAligned -> Box.fromCenterAndRadii: p = c - r, q = c + r
        boxes.append((*(c - r for c, r in zip(centre, radii)), *(c + r for c, r in zip(centre, radii))))
        box_mats.append(mat)
    for x in (-2.15, 2.15):
        for y in (-1.18, 1.18):
            spheres.append((x, y, 4.20, 0.22))
            sphere_mats.append(lens)
    return {"spheres": (np.array(spheres, np.float64), np.array(sphere_mats)),
            "boxes": (np.array(boxes, np.float64), np.array(box_mats)),
            "affine_boxes": (np.array(aboxes, np.float64), np.array(abox_mats)),
            "background": (0.002, 0.004, 0.014)}


def hero_scene(primitive_count: int, config: BvhBuildConfig | None = None) -> GpuScene:
    """GpuSceneBuilder.from(LayeredHeroFamily.create(n), config) (gpu/GpuSceneBuilder.java:27-77)."""
    p = hero_primitives(primitive_count)
    sc = GpuScene(config or BvhBuildConfig.uniform(LEAF_SIZE))
    sc.set_background(p["background"])
    sc.add_spheres(*p["spheres"])
    sc.add_boxes(*p["boxes"])
    sc.add_affine_transforms("box", *p["affine_boxes"])
    return sc


def hero_camera() -> Camera:
    """LayeredHeroFamily.camera (:120-122)."""
    return support_camera((0.0, 0.0, 0.0), 0.0, 0.0)


def hero_gpu_camera(width: int = WIDTH, height: int = HEIGHT) -> GpuCamera:
    return gpu_camera(hero_camera(), width, height)


# ---- GI Test (graphics3d/scene/v1/GITestPlain.java:15-49) and its catalog camera
GI_TEST_WIDTH, GI_TEST_HEIGHT = 1280, 720       # the platform comparison's frame (921,600 camera samples)


def _hsb(h: float, s: float, b: float):
    """Color.hsb (Color.java:57-74) for the hues GI Test uses."""
    h = h - math.floor(h)
    base = int(h * 6.0)
    f = h * 6.0 - base
    p, q, t = b * (1.0 - s), b * (1.0 - s * f), b * (1.0 - s * (1.0 - f))
    return [(b, t, p), (q, b, p), (p, b, t), (p, q, b), (t, p, b), (b, p, q)][base]


def gi_test_scene(config: BvhBuildConfig | None = None) -> GpuScene:
    """GITestPlain: five half-spaces (matte side and back walls, an emissive ceiling), a glass sphere, a mirror
    sphere, a matte sphere and a partly reflective sphere; black background.  Materials are
    SceneSupport.matte / emissive / surface (SceneSupport.java:16-37, Material.java:86-92)."""
    def m(emission=(0.0, 0.0, 0.0), diffuse=(0.0, 0.0, 0.0), reflective=(0.0, 0.0, 0.0), refractive=(0.0, 0.0, 0.0),
          ior=1.5):
        return material_array(diffuse, reflective, refractive, emission, ior)[0]

    grey = (0.7, 0.7, 0.7)
    sc = GpuScene(config or BvhBuildConfig.uniform(LEAF_SIZE))
    sc.set_background((0.0, 0.0, 0.0))
    spheres = [[-0.2, -0.5, 0.0, 0.3], [0.5, -0.5, -0.3, 0.3], [0.0, 0.2, 0.0, 0.2], [-0.4, 0.5, 0.1, 0.2]]
    smats = [m(reflective=(0.05,) * 3, refractive=(0.95,) * 3, ior=1.4), m(reflective=(1.0,) * 3),
             m(diffuse=grey), m(reflective=(0.9,) * 3)]
    sc.add_spheres(spheres, np.array(smats))
    planes = [[-1, 0, 0, 1, 0, 0], [1, 0, 0, -1, 0, 0], [0, -1, 0, 0, 1, 0], [0, 1, 0, 0, -1, 0], [0, 0, 1, 0, 0, -1]]
    pmats = [m(diffuse=_hsb(0.0, 0.5, 0.7)), m(diffuse=_hsb(1.0 / 3.0, 0.5, 0.7)), m(diffuse=grey),
             m(emission=(1.0, 1.0, 1.0)), m(diffuse=grey)]
    sc.add_planes(planes, np.array(pmats))
    return sc


def gi_test_camera() -> Camera:
    """SceneCatalog.create(GI_TEST) (playground/SceneCatalog.java:35-37): SceneSupport.camera((0, 0, -4), 0, 0)."""
    return support_camera((0.0, 0.0, -4.0), 0.0, 0.0)


def gi_test_gpu_camera(width: int = GI_TEST_WIDTH, height: int = GI_TEST_HEIGHT) -> GpuCamera:
    return gpu_camera(gi_test_camera(), width, height)


# ---- City of Night (graphics3d/scene/v1/CityOfNightV1.java) and Signal Chamber (SignalChamber.java), the
# catalog's other introductory scenes (playground/SceneCatalog.java:39-46)
_CITY_SEED = 0x1F234AB9C6D8EE07
_CITY_EPS, _CITY_PANEL = 0.0025, 0.004
_CITY_MOON = 4.0
_CITY_WINDOWS = [(30.0, 14.0, 6.0), (10.0, 22.0, 40.0), (24.0, 12.0, 30.0), (16.0, 28.0, 18.0)]


def _mul(c, k):
    return (c[0] * k, c[1] * k, c[2] * k)


def _clamp(v, lo, hi):
    return max(lo, min(hi, v))


class _City:
    """CityOfNightV1's construction (:29-81, :83-389) in the reference's draw order.  Boxes are
    SceneSupport.box(centre, radii) = axis-aligned [centre - radii, centre + radii]."""

    def __init__(self):
        from .random_scenes import SplittableRandom
        self.r = SplittableRandom(_CITY_SEED)
        self.boxes, self.mats = [], []

    def box(self, c, radii, mat):
        self.boxes.append((c[0] - radii[0], c[1] - radii[1], c[2] - radii[2],
                           c[0] + radii[0], c[1] + radii[1], c[2] + radii[2]))
        self.mats.append(mat)

    def light(self, c, radii, glow):
        self.box(c, radii, _surface(glow))

    def front_back(self, x, y, z, hw, hh, rz, ht, glow, front):
        zc = z + rz + ht + _CITY_EPS if front else z - rz - ht - _CITY_EPS
        self.light((x, y, zc), (hw, hh, ht), glow)

    def side(self, x, y, z, rx, ht, hh, hd, glow, right):
        xc = x + rx + ht + _CITY_EPS if right else x - rx - ht - _CITY_EPS
        self.light((xc, y, z), (ht, hh, hd), glow)

    def build(self):
        r = self.r
        for gz in range(32):
            z_base = -4.45 + gz * 0.60
            for gx in range(-16, 17):
                keep = 0.67 if gz < 6 else (0.63 if gz < 12 else 0.57)
                if abs(gx) > 13:
                    keep -= 0.08
                if r.next_double() > keep:
                    continue
                x = gx * 0.64 + r.next_double_in(-0.10, 0.10)
                z = z_base + r.next_double_in(-0.10, 0.10)
                rx = r.next_double_in(0.10, 0.20)
                rz = r.next_double_in(0.10, 0.20)
                height = 0.10 + math.pow(r.next_double(), 1.5) * 0.18 + math.pow(r.next_double(), 4.2) * 1.10
                if gz > 12 and abs(gx) < 8 and r.next_double() < 0.065:
                    height *= r.next_double_in(2.2, 3.5)
                    rx *= r.next_double_in(0.76, 0.92)
                    rz *= r.next_double_in(0.76, 0.92)
                self.tower(x, z, rx, rz, height, gz)
        for _ in range(26):                                            # addForegroundMassing (:83-106)
            x = r.next_double_in(-6.8, 6.8)
            z = r.next_double_in(-5.4, -0.9)
            rx = r.next_double_in(0.14, 0.32)
            rz = r.next_double_in(0.16, 0.34)
            height = r.next_double_in(0.12, 0.34)
            diffuse = _hsb(r.next_double(), r.next_double_in(0.12, 0.32), r.next_double_in(0.16, 0.28))
            self.box((x, height * 0.5, z), (rx, height * 0.5, rz),
                     _surface((0.0, 0.0, 0.0), diffuse, (0.006, 0.007, 0.008)))
        self.box((0.0, 5.55, 5.1), (13.5, 0.045, 15.0), _surface(_mul((0.22, 0.24, 0.27), _CITY_MOON)))
        return self

    def tower(self, x, z, rx, rz, height, band):                       # addTower (:108-152)
        r = self.r
        y = height * 0.5
        mirror = r.next_double() < 0.04
        colorful = r.next_double() < 0.62
        if colorful:
            diffuse = _hsb(r.next_double(), r.next_double_in(0.42, 0.82), r.next_double_in(0.34, 0.60))
        else:
            diffuse = _hsb(r.next_double(), r.next_double_in(0.16, 0.32), r.next_double_in(0.22, 0.36))
        reflective = (0.44, 0.48, 0.54) if mirror else (0.026, 0.030, 0.034)
        self.box((x, y, z), (rx, y, rz), _surface((0.0, 0.0, 0.0), diffuse, reflective))
        if height > 0.78 and r.next_double() < 0.38:
            self.box((x, 0.08, z), (rx * 1.08, 0.08, rz * 1.08),
                     _surface((0.0, 0.0, 0.0), _mul(diffuse, 0.90), _mul(reflective, 0.55)))
        glow = _CITY_WINDOWS[(r.next_int_bounded(len(_CITY_WINDOWS)) + band) % len(_CITY_WINDOWS)]
        self.accents(x, z, rx, rz, height, glow)
        pick = r.next_double()
        if pick < 0.20:
            self.dense_bands(x, z, rx, rz, height, glow)
        elif pick < 0.32:
            self.grid(x, z, rx, rz, height, glow, 0.64)
        elif pick < 0.42:
            self.grid(x, z, rx, rz, height, _mul(glow, 0.98), 0.40)
        elif pick < 0.52:
            self.corners(x, z, rx, rz, height, glow)

    def accents(self, x, z, rx, rz, height, glow):                     # addFacadeAccents (:154-189)
        r = self.r
        if height > 0.48 and r.next_double() < 0.42:
            self.strip_fb(x, z, rx, rz, height, True, _mul(glow, r.next_double_in(1.20, 1.66)))
        if height > 0.54 and r.next_double() < 0.18:
            self.strip_fb(x, z, rx, rz, height, False, _mul(glow, r.next_double_in(1.00, 1.30)))
        if height > 0.48 and r.next_double() < 0.30:
            self.strip_side(x, z, rx, rz, height, True, _mul(glow, r.next_double_in(1.12, 1.54)))
        if height > 0.54 and r.next_double() < 0.18:
            self.strip_side(x, z, rx, rz, height, False, _mul(glow, r.next_double_in(1.00, 1.34)))
        if r.next_double() < 0.22:
            self.billboard(x, z, rx, rz, height, True, _mul(glow, r.next_double_in(1.14, 1.44)))
        if r.next_double() < 0.06:
            h = height * r.next_double_in(0.72, 0.96)
            self.billboard(x, z, rx, rz, h, False, _mul(glow, r.next_double_in(1.00, 1.18)))
        if r.next_double() < 0.12:
            self.side_billboard(x, z, rx, rz, height, True, _mul(glow, r.next_double_in(1.08, 1.42)))
        if r.next_double() < 0.10:
            self.side_billboard(x, z, rx, rz, height, False, _mul(glow, r.next_double_in(1.00, 1.28)))

    def strip_fb(self, x, z, rx, rz, height, front, glow):            # addFrontBackStrip (:191-204)
        hw = min(rx * 0.08, 0.010)
        hh = min(height * 0.38, max(0.15, height * 0.28))
        cy = _clamp(height * self.r.next_double_in(0.42, 0.62), hh + 0.04, height - hh - 0.04)
        self.front_back(x, cy, z, hw, hh, rz, _CITY_PANEL, glow, front)

    def strip_side(self, x, z, rx, rz, height, right, glow):          # addSideStrip (:206-220)
        hd = min(rz * 0.08, 0.010)
        hh = min(height * 0.36, max(0.15, height * 0.26))
        cy = _clamp(height * self.r.next_double_in(0.40, 0.60), hh + 0.04, height - hh - 0.04)
        self.side(x, cy, z, rx, _CITY_PANEL, hh, hd, glow, right)

    def side_billboard(self, x, z, rx, rz, height, right, glow):      # addSideBillboard (:222-236)
        r = self.r
        hh = min(height * 0.24, 0.14 + r.next_double() * 0.08)
        cy = _clamp(height * r.next_double_in(0.32, 0.72), hh + 0.05, height - hh - 0.05)
        hd = min(rz * 0.70, 0.24 + r.next_double() * 0.10)
        self.side(x, cy, z, rx, _CITY_PANEL, hh, hd, glow, right)

    def dense_bands(self, x, z, rx, rz, height, glow):                # addDenseBands (:238-254)
        r = self.r
        bands = max(3, int(math.floor(height / 0.26)))
        for i in range(bands):
            y = 0.08 + i * (height - 0.16) / bands
            if r.next_double() < 0.08:
                continue
            ws = 0.42 + r.next_double() * 0.34
            g = _mul(glow, r.next_double_in(1.16, 1.52))
            self.front_back(x, y, z, rx * ws, 0.016, rz, 0.010, g, True)           # addFrontBackWindow (:353-356)
            self.front_back(x, y, z, rx * ws, 0.016, rz, 0.010, _mul(g, 0.86), False)

    def grid(self, x, z, rx, rz, height, glow, lit_chance):            # addGrid (:278-303)
        r = self.r
        rows = max(4, int(math.floor(height / 0.26)))
        cols = max(2, int(math.floor((rx * 2.0) / 0.11)))
        hw = min(0.040, rx * 0.24)
        for row in range(rows):
            y = 0.10 + row * (height - 0.20) / rows
            for col in range(cols):
                if r.next_double() > lit_chance:
                    continue
                xo = 0.0 if cols == 1 else (-rx * 0.68 + col * (rx * 1.36 / (cols - 1)))
                lit = _mul(glow, r.next_double_in(1.12, 1.64))
                self.front_back(x + xo, y, z, hw, 0.015, rz, 0.010, lit, True)
                if r.next_double() < 0.64:
                    self.front_back(x + xo, y, z, hw, 0.015, rz, 0.010, _mul(lit, 0.88), False)

    def corners(self, x, z, rx, rz, height, glow):                    # addCornerPattern (:305-321)
        r = self.r
        edge = height * 0.30
        y = height * 0.50
        lit = _mul(glow, r.next_double_in(1.12, 1.48))
        self.front_back(x - rx * 0.72, y, z, 0.014, edge, rz, 0.010, lit, True)
        self.front_back(x + rx * 0.72, y, z, 0.014, edge, rz, 0.010, _mul(lit, 0.94), True)
        if r.next_double() < 0.54:
            self.side(x, y, z - rz * 0.52, rx, 0.010, edge * 0.78, 0.015, _mul(lit, 0.84), False)

    def billboard(self, x, z, rx, rz, height, front, glow):           # addBillboard (:323-337)
        r = self.r
        hh = min(height * 0.22, 0.14 + r.next_double() * 0.08)
        cy = _clamp(height * r.next_double_in(0.30, 0.70), hh + 0.05, height - hh - 0.05)
        hw = min(rx * 0.76, 0.28 + r.next_double() * 0.12)
        lit = _mul(glow, r.next_double_in(1.24, 1.70))
        self.front_back(x, cy, z, hw, hh, rz, _CITY_PANEL, lit, front)


def city_of_night_scene(config: BvhBuildConfig | None = None) -> GpuScene:
    """CityOfNightV1: a ground half-space (a dark glossy surface), towers of axis-aligned boxes on a jittered grid
    with emissive window panels, strips and billboards, 26 foreground blocks and an emissive moon slab, all drawn
    from SplittableRandom(0x1F234AB9C6D8EE07); background (0.021, 0.023, 0.031)."""
    c = _City().build()
    sc = GpuScene(config or BvhBuildConfig.uniform(LEAF_SIZE))
    sc.set_background((0.021, 0.023, 0.031))
    sc.add_boxes(np.array(c.boxes, np.float64), np.array(c.mats))
    ground = _surface((0.0, 0.0, 0.0), (0.150, 0.154, 0.160), (0.024, 0.026, 0.028))
    sc.add_planes([[0.0, 0.0, 0.0, 0.0, 1.0, 0.0]], np.array([ground]))
    return sc


def city_of_night_camera() -> Camera:
    """SceneCatalog.create(CITY_OF_NIGHT_V1) (:39-42): SceneSupport.camera((0, 0.34, -6.6), 0.08, -0.01)."""
    return support_camera((0.0, 0.34, -6.6), 0.08, -0.01)


_SIGNAL_SEED = 0x5C82D0A4B7319E6F


def _signal_color(r):
    """SignalChamber.randomPixelColor (:91-101)."""
    if r.next_double() < 0.08:
        hot = r.next_double_in(10.0, 18.0)
        return (hot, hot * r.next_double_in(0.95, 1.05), hot * r.next_double_in(1.00, 1.12))
    base = _hsb(r.next_double(), r.next_double_in(0.55, 0.92), r.next_double_in(0.60, 1.00))
    return _mul(base, r.next_double_in(9.0, 24.0))


def signal_chamber_scene(config: BvhBuildConfig | None = None) -> GpuScene:
    """SignalChamber: six half-spaces (reflective floor, ceiling, front and back, matte side walls) and two walls of
    101 x 151 emissive spheres of radius 0.014 (30,502), colours from SplittableRandom(0x5C82D0A4B7319E6F);
    black background."""
    from .random_scenes import SplittableRandom
    refl = _surface((0.0, 0.0, 0.0), (0.055, 0.060, 0.070), (0.44, 0.47, 0.52))
    side = _surface((0.0, 0.0, 0.0), (0.05, 0.055, 0.065))
    planes = [[0.0, -1.0, 0.0, 0.0, 1.0, 0.0], [0.0, 1.96, 0.0, 0.0, -1.0, 0.0], [-1.26, 0.0, 0.0, 1.0, 0.0, 0.0],
              [1.26, 0.0, 0.0, -1.0, 0.0, 0.0], [0.0, 0.0, 6.50, 0.0, 0.0, -1.0], [0.0, 0.0, -2.60, 0.0, 0.0, 1.0]]
    pmats = [refl, refl, side, side, refl, refl]
    rows, cols = 101, 151
    y0, z0 = -0.80, -0.10
    ys, zs = 2.60 / (rows - 1), 6.35 / (cols - 1)
    r = SplittableRandom(_SIGNAL_SEED)
    spheres, mats = [], []
    for row in range(rows):
        for col in range(cols):
            left, right = _signal_color(r), _signal_color(r)
            y, z = y0 + row * ys, z0 + col * zs
            spheres.append((-1.268, y, z, 0.014))
            mats.append(_surface(left))
            spheres.append((1.268, y, z, 0.014))
            mats.append(_surface(right))
    sc = GpuScene(config or BvhBuildConfig.uniform(LEAF_SIZE))
    sc.set_background((0.0, 0.0, 0.0))
    sc.add_spheres(np.array(spheres, np.float64), np.array(mats))
    sc.add_planes(planes, np.array(pmats))
    return sc


def signal_chamber_camera() -> Camera:
    """SceneCatalog.create(SIGNAL_CHAMBER) (:43-46): SceneSupport.camera((0, 0.20, -1.40), 0.016, 0)."""
    return support_camera((0.0, 0.20, -1.40), 0.016, 0.0)


def scale_id(scale: int) -> str:
    """BenchmarkProtocol.compact (:410-417) as used in the row ids."""
    return "hero-" + {1_000: "1k", 10_000: "10k", 100_000: "100k", 1_000_000: "1m"}.get(scale, str(scale))


# ---- Gallery overlap (graphics3d/scene/v1/GalleryOverlapSource.java, GalleryOverlapScene.java)
OVERLAP_OBJECTS = 10_000
_OVERLAP_SPHERES, _OVERLAP_LARGE, _OVERLAP_FOREGROUND = 3_000, 1_000, 128
_GOLDEN_ANGLE_TURNS = 0.3819660112501051


def _unit_hash(index: int, salt: int) -> float:
    """GalleryOverlapSource.unit / GalleryOverlapScene.unit (Java long arithmetic)."""
    v = (index * 0x9E3779B97F4A7C15 + salt) & _M64
    v = ((v ^ (v >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    v = ((v ^ (v >> 27)) * 0x94D049BB133111EB) & _M64
    v ^= v >> 31
    return (v >> 11) * 2.0 ** -53


def _evenly(index: int, population: int, selected: int) -> bool:
    return (index + 1) * selected // population != index * selected // population


_OVL_MATS = [_surface((0.46, 0.018, 0.006)), _surface((0.006, 0.17, 0.62)), _surface((0.014, 0.42, 0.22)),
             _surface((0.70, 0.16, 0.010))]
_OVL_SPHERE_MATS = [_surface((0.10, 0.56, 0.88)), _surface((0.92, 0.23, 0.035)), _surface((0.055, 0.72, 0.34)),
                    _surface((0.90, 0.50, 0.040))]
_OVL_MIRROR = _surface((0.0, 0.0, 0.0), reflective=(0.92, 0.92, 0.92), ior=1.5)
_OVL_GLASS = _surface((0.0, 0.0, 0.0), reflective=(0.05, 0.05, 0.05), refractive=(0.95, 0.95, 0.95), ior=1.45)


def _v_sub(a, b): return (a[0] - b[0], a[1] - b[1], a[2] - b[2])
def _v_add(a, b): return (a[0] + b[0], a[1] + b[1], a[2] + b[2])
def _v_mul(a, k): return (a[0] * k, a[1] * k, a[2] * k)
def _v_dot(a, b): return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]
def _v_len(a): return math.sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2])
def _v_cross(a, b): return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])
def _v_rejection(a, d): return _v_sub(a, _v_mul(d, _v_dot(a, d) / _v_dot(d, d)))
def _v_normalized_to(a, length): return _v_mul(a, length / _v_len(a))


def overlap_primitives():
    """GalleryOverlapScene.create (:47-126) over GalleryOverlapSource.overlapGallery
    (:62-151): 3,000 spheres and 7,000 thin ribbons (affine boxes) in the
    reference's insertion order."""
    spheres, sphere_mats, aboxes, abox_mats = [], [], [], []
    sphere_n = large_n = fg_n = inter_n = 0
    for index in range(OVERLAP_OBJECTS):
        turn = index * _GOLDEN_ANGLE_TURNS
        turn = turn - math.floor(turn)
        angle = turn * 2.0 * math.pi
        progress = (index + 0.5) / OVERLAP_OBJECTS
        radius = 0.12 + 1.05 * math.sqrt(progress)
        phase = angle + 5.0 * math.pi * progress
        sx, sy = 1.34 * radius * math.cos(phase), 0.72 * radius * math.sin(phase)
        rot_turns = turn + 0.25 + 0.020 * math.sin(phase * 3.0)
        sphere = _evenly(index, OVERLAP_OBJECTS, _OVERLAP_SPHERES)
        large = sphere and _evenly(sphere_n, _OVERLAP_SPHERES, _OVERLAP_LARGE)
        fg = large and _evenly(large_n, _OVERLAP_LARGE, _OVERLAP_FOREGROUND)
        depth = 5.03 + 0.07 * _unit_hash(index, 0x4B1D5A77) if fg else 5.2 + 7.8 * _unit_hash(index, 0x4B1D5A77)
        mat = _OVL_MATS[index & 3]
        if sphere:
            u = _unit_hash(index, 0x23A4D159)
            r_sensor = (0.018 + 0.006 * u if fg else 0.013 + 0.005 * u) if large else 0.0040 + 0.0020 * u
            if large:
                if fg:
                    cls = (fg_n * 53) % 128
                    fg_n += 1
                    mat = _OVL_GLASS if cls < 48 else _OVL_MIRROR if cls < 96 else _OVL_SPHERE_MATS[large_n & 3]
                else:
                    cls = (inter_n * 361) % 872
                    inter_n += 1
                    mat = _OVL_GLASS if cls < 80 else _OVL_MIRROR if cls < 160 else _OVL_SPHERE_MATS[large_n & 3]
                large_n += 1
            spheres.append((*_world(sx, sy, depth), r_sensor * depth / _PERSPECTIVE_Z))
            sphere_mats.append(mat)
            sphere_n += 1
            continue
        # the source's thin axis-aligned box (GalleryOverlapSource :104-118) ...
        half_len_s = 1.03 + 0.34 * _unit_hash(index, 0x6F23E991)
        half_w_s = 0.000045 + 0.000025 * _unit_hash(index, 0x19C5B7D3)
        radians = rot_turns * 2.0 * math.pi
        horizontal = abs(math.sin(radians)) >= abs(math.cos(radians))
        hl, hw = half_len_s * depth / _PERSPECTIVE_Z, half_w_s * depth / _PERSPECTIVE_Z
        half = (hl, hw, 0.008) if horizontal else (hw, hl, 0.008)
        c0 = _world(sx, sy, depth)
        p, q = _v_sub(c0, half), _v_add(c0, half)
        # ... turned into a ribbon across depth (GalleryOverlapScene :69-110)
        centre = _v_mul(_v_add(p, q), 0.5)
        radii = _v_mul(_v_sub(q, p), 0.5)
        horiz = radii[0] > radii[1]
        old_len = radii[0] if horiz else radii[1]
        old_w = radii[1] if horiz else radii[0]
        scx = _PERSPECTIVE_Z * centre[0] / centre[2]
        scy = _PERSPECTIVE_Z * centre[1] / centre[2]
        s_len = _PERSPECTIVE_Z * old_len / centre[2]
        s_w = _PERSPECTIVE_Z * old_w / centre[2]
        near = 5.04 + 0.14 * _unit_hash(index, 0x754A32D192ED03)
        far = 12.82 + 0.14 * _unit_hash(index, 0x16B4CE29890A1F)
        reverse = _unit_hash(index, 0x6D2B79F5AA010D) < 0.5
        d_minus, d_plus = (far, near) if reverse else (near, far)
        m_s = (scx - s_len, scy, 0.0) if horiz else (scx, scy - s_len, 0.0)
        p_s = (scx + s_len, scy, 0.0) if horiz else (scx, scy + s_len, 0.0)
        m_e, p_e = _world(m_s[0], m_s[1], d_minus), _world(p_s[0], p_s[1], d_plus)
        r_centre = _v_mul(_v_add(m_e, p_e), 0.5)
        long_axis = _v_mul(_v_sub(p_e, m_e), 0.5)
        mean_depth = 0.5 * (near + far)
        new_w = s_w * mean_depth / _PERSPECTIVE_Z
        want = (0.0, 1.0, 0.0) if horiz else (-1.0, -0.0, -0.0)
        w_axis = _v_normalized_to(_v_rejection(want, long_axis), new_w)
        t_axis = _v_normalized_to(_v_cross(long_axis, w_axis), 0.008)
        # Affine.unitVectors(long, width, thickness).then(translation(centre))
        t = (long_axis[0], w_axis[0], t_axis[0], 0.0, long_axis[1], w_axis[1], t_axis[1], 0.0,
             long_axis[2], w_axis[2], t_axis[2], 0.0)
        aboxes.append(then(t, translation(*r_centre)))
        abox_mats.append(mat)
    if (sphere_n, large_n, fg_n, len(aboxes)) != (_OVERLAP_SPHERES, _OVERLAP_LARGE, _OVERLAP_FOREGROUND, 7_000):
        raise RuntimeError("overlap population mismatch")
    return {"spheres": (np.array(spheres, np.float64), np.array(sphere_mats)),
            "affine_boxes": (np.array(aboxes, np.float64), np.array(abox_mats)),
            "background": (0.003, 0.006, 0.014)}


def overlap_scene(config: BvhBuildConfig | None = None) -> GpuScene:
    p = overlap_primitives()
    sc = GpuScene(config or BvhBuildConfig.uniform(LEAF_SIZE))
    sc.set_background(p["background"])
    sc.add_spheres(*p["spheres"])
    sc.add_affine_transforms("box", *p["affine_boxes"])
    return sc


def overlap_camera() -> Camera:
    """GalleryOverlapScene.create's front camera: SceneSupport.camera(0, 0, 0, 0.125)."""
    return support_camera((0.0, 0.0, 0.0), 0.0, 0.0, 0.125)
