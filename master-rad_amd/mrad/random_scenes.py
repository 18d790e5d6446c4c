"""The random first-hit families F1-F4 of the benchmark study
(benchmark/RandomSingleRaySceneFactory.java, benchmark/BenchmarkWorkloads.java:74-100).

A layout places n envelope centres in [-1, 1]^3 by rejection sampling; the four
populations put spheres, axis-aligned boxes, affine spheres and affine boxes on
the same centres:

    F1  half spheres, half affine boxes      (SPHERE_AFFINE_BOX_50_50)
    F2  four primitive kinds in equal shares (FOUR_TYPES_EQUAL)
    F3  spheres only                         (ALL_SPHERE)
    F4  affine boxes only                    (ALL_AFFINE_BOX)

Random numbers come from java.util.SplittableRandom, which the reference takes
from the JDK.  It is restated here from its published algorithm, SplitMix64
(G. L. Steele, D. Lea, C. H. Flood, "Fast splittable pseudorandom number
generators", OOPSLA 2014), with the JDK's derived draws (nextInt(bound) by
rejection of over-represented values, nextDouble as the top 53 bits,
nextDouble(origin, bound) as r * (bound - origin) + origin clamped below bound,
nextBoolean as the sign of nextInt).  Java evaluates each double operation
without fused multiply-add, as Python does.  Math.cbrt, Math.sin and Math.cos
come from the C library here (numpy.cbrt, math.sin, math.cos); Java's may differ
in the last bit, which moves an envelope or an axis by at most one ulp.
"""
from __future__ import annotations

import math

import numpy as np

from .scene import BvhBuildConfig, GpuScene, material_array

_M64 = (1 << 64) - 1
GOLDEN_GAMMA = 0x9E3779B97F4A7C15

FAMILIES = ("F1", "F2", "F3", "F4")
OBJECT_COUNTS = (100, 1_000, 10_000)
LAYOUT_IDS = range(100)                          # BenchmarkProtocol FIRST..LAST_RANDOM_LAYOUT_ID
WARMUP_LAYOUT_ID = 100                           # BenchmarkProtocol.RANDOM_WARMUP_LAYOUT_ID (:74)
LAYOUT_EXPERIMENT_SEED = 5_641_122_812_212_105_804   # BenchmarkProtocol.RANDOM_LAYOUT_EXPERIMENT_SEED (:75)

BASE_ENVELOPE_SCALE = 0.58                       # RandomSingleRaySceneFactory.java:50-51
GAP_FRACTION = 0.08
MAX_LAYOUT_ATTEMPTS = 64
MAX_OBJECT_PLACEMENT_ATTEMPTS = 20_000
LAYOUT_SEED_DOMAIN = 0x4C41594F55544C00
POSITION_SEED_DOMAIN = 0x16B00B5
VALIDATION_EPSILON = 1.0e-12

SPHERE, BOX, AFFINE_SPHERE, AFFINE_BOX = range(4)               # PrimitiveKind ordinals
ALL_SPHERE, SPHERE_AFFINE_BOX_50_50, ALL_AFFINE_BOX, FOUR_TYPES_EQUAL = range(4)   # Population ordinals
POPULATION = {"F1": SPHERE_AFFINE_BOX_50_50, "F2": FOUR_TYPES_EQUAL, "F3": ALL_SPHERE, "F4": ALL_AFFINE_BOX}

# Material.matte(TYPE_COLORS[kind]) (:58-63): diffuse only, ior 1.5
TYPE_MATERIALS = material_array(np.array([(0.10, 0.32, 0.92), (0.12, 0.72, 0.30), (0.92, 0.12, 0.04),
                                          (0.35, 0.58, 0.96)]), n=4)
ENVIRONMENT = (0.18, 0.22, 0.30)                 # BenchmarkWorkloads.RANDOM_ENVIRONMENT (:15)


def _rotl(v: int, r: int) -> int:
    v &= _M64
    return ((v << r) | (v >> (64 - r))) & _M64


def mix(seed: int, a: int, b: int) -> int:
    """RandomSingleRaySceneFactory.mix (:427-432) as an unsigned 64-bit value."""
    v = (seed ^ _rotl(a * GOLDEN_GAMMA, 21) ^ b) & _M64
    v = ((v ^ (v >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    v = ((v ^ (v >> 27)) * 0x94D049BB133111EB) & _M64
    return v ^ (v >> 31)


def layout_seed(experiment_seed: int, object_count: int, layout_id: int) -> int:
    """RandomSingleRaySceneFactory.layoutSeed (:71-82); unsigned 64-bit."""
    if layout_id < 0:
        raise ValueError("Layout identifier must be nonnegative.")
    return mix(experiment_seed, ((object_count << 32) ^ (layout_id & 0xFFFFFFFF)) & _M64, LAYOUT_SEED_DOMAIN)


class SplittableRandom:
    """java.util.SplittableRandom(seed): SplitMix64 with the golden gamma."""

    __slots__ = ("seed",)

    def __init__(self, seed: int):
        self.seed = seed & _M64

    def _next_seed(self) -> int:
        self.seed = (self.seed + GOLDEN_GAMMA) & _M64
        return self.seed

    def next_long(self) -> int:
        """mix64 of the next seed (Stafford's variant 13), unsigned."""
        z = self._next_seed()
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def next_int(self) -> int:
        """mix32 of the next seed: the high half of the second product, as a signed int."""
        z = self._next_seed()
        z = ((z ^ (z >> 33)) * 0x62A9D9ED799705F5) & _M64
        r = (((z ^ (z >> 28)) * 0xCB24D0A5C88C35B3) & _M64) >> 32
        return r - (1 << 32) if r >> 31 else r

    def next_int_bounded(self, bound: int) -> int:
        """nextInt(bound): a power of two masks; otherwise u = r >>> 1 is
        redrawn while u + (bound - 1) - u % bound overflows int."""
        if bound <= 0:
            raise ValueError("bound must be positive")
        r = self.next_int()
        m = bound - 1
        if bound & m == 0:
            return r & m
        u = (r & 0xFFFFFFFF) >> 1
        while True:
            r = u % bound
            if u + m - r < (1 << 31):
                return r
            u = (self.next_int() & 0xFFFFFFFF) >> 1

    def next_double(self) -> float:
        return (self.next_long() >> 11) * 1.1102230246251565e-16          # 0x1.0p-53

    def next_double_in(self, origin: float, bound: float) -> float:
        r = self.next_double()
        if origin < bound:
            r = r * (bound - origin) + origin
            if r >= bound:
                r = math.nextafter(bound, -math.inf)
        return r

    def next_boolean(self) -> bool:
        return self.next_int() < 0


def _cell(p, size: float):
    return (math.floor(p[0] / size), math.floor(p[1] / size), math.floor(p[2] / size))


def _try_positions(count: int, minimum_distance: float, limit: float, seed: int):
    """tryPositions (:245-276): rejection sampling against a hash grid of
    cell size minimum_distance."""
    rnd = SplittableRandom(seed)
    grid: dict = {}
    out = []
    d2 = minimum_distance * minimum_distance
    for _ in range(count):
        accepted = None
        for _attempt in range(MAX_OBJECT_PLACEMENT_ATTEMPTS):
            c = (rnd.next_double_in(-limit, limit), rnd.next_double_in(-limit, limit), rnd.next_double_in(-limit, limit))
            cx, cy, cz = _cell(c, minimum_distance)
            ok = True
            for x in (cx - 1, cx, cx + 1):
                for y in (cy - 1, cy, cy + 1):
                    for z in (cz - 1, cz, cz + 1):
                        for e in grid.get((x, y, z), ()):
                            dx, dy, dz = c[0] - e[0], c[1] - e[1], c[2] - e[2]
                            if dx * dx + dy * dy + dz * dz < d2:
                                ok = False
                                break
                        if not ok:
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if ok:
                accepted = c
                break
        if accepted is None:
            return None
        out.append(accepted)
        grid.setdefault(_cell(accepted, minimum_distance), []).append(accepted)
    return out


def envelope_radius(object_count: int) -> float:
    return BASE_ENVELOPE_SCALE / float(np.cbrt(float(object_count)))


def layout(object_count: int, seed: int):
    """RandomSingleRaySceneFactory.layout (:84-101): (centres, envelope radius)."""
    if object_count < 2:
        raise ValueError("The random study requires at least two objects.")
    radius = envelope_radius(object_count)
    minimum_distance = radius * (2.0 + GAP_FRACTION)
    limit = 1.0 - radius
    position_seed = mix(seed, object_count, POSITION_SEED_DOMAIN)
    for attempt in range(MAX_LAYOUT_ATTEMPTS):
        centres = _try_positions(object_count, minimum_distance, limit, mix(position_seed, attempt, 0x1A7A77E5))
        if centres is not None:
            _validate(centres, radius)
            return centres, radius
    raise RuntimeError(f"Could not place {object_count} objects after {MAX_LAYOUT_ATTEMPTS} "
                       "deterministic layout attempts.")


def _validate(centres, radius: float) -> None:
    """validate (:364-394): unit-box containment and the envelope gap."""
    minimum_distance = radius * (2.0 + GAP_FRACTION)
    grid: dict = {}
    for i, c in enumerate(centres):
        if max(abs(c[0]), abs(c[1]), abs(c[2])) + radius > 1.0 + VALIDATION_EPSILON:
            raise RuntimeError(f"Unit-box containment failed at object {i}.")
        cx, cy, cz = _cell(c, minimum_distance)
        for x in (cx - 1, cx, cx + 1):
            for y in (cy - 1, cy, cy + 1):
                for z in (cz - 1, cz, cz + 1):
                    for j, e in grid.get((x, y, z), ()):
                        dx, dy, dz = c[0] - e[0], c[1] - e[1], c[2] - e[2]
                        if math.sqrt(dx * dx + dy * dy + dz * dz) + VALIDATION_EPSILON < minimum_distance:
                            raise RuntimeError(f"Envelope gap failed for objects {j} and {i}.")
        grid.setdefault((cx, cy, cz), []).append((i, c))


def population_kinds(object_count: int, seed: int, population: int):
    """populationKinds (:134-171): per-kind quotas in kind order, then a
    Fisher-Yates shuffle."""
    if population == ALL_SPHERE:
        quotas = [object_count, 0, 0, 0]
    elif population == ALL_AFFINE_BOX:
        quotas = [0, 0, 0, object_count]
    elif population == SPHERE_AFFINE_BOX_50_50:
        spheres = object_count // 2
        if object_count & 1 and mix(seed, object_count, 0xF1) & 1 == 0:
            spheres += 1
        quotas = [spheres, 0, 0, object_count - spheres]
    else:                                                   # equalFourWayQuotas (:173-183)
        quotas = [object_count // 4] * 4
        offset = mix(seed, object_count, 0xF2) & 3           # floorMod(signed, 4)
        for i in range(object_count % 4):
            quotas[(offset + i) % 4] += 1
    kinds = [k for k in range(4) for _ in range(quotas[k])]
    rnd = SplittableRandom(mix(seed, object_count, 0x504F50554C415449 + population))
    for i in range(len(kinds) - 1, 0, -1):
        j = rnd.next_int_bounded(i + 1)
        kinds[i], kinds[j] = kinds[j], kinds[i]
    return kinds


def _random_rotation(rnd: SplittableRandom):
    """randomRotation (:318-339): a uniform axis ey, an orthonormal ex from a
    helper axis, and a roll about ey; returns the 3x3 columns as rows m."""
    y = rnd.next_double_in(-1.0, 1.0)
    phi = rnd.next_double_in(0.0, 2.0 * math.pi)
    radial = math.sqrt(1.0 - y * y)
    ey = (radial * math.cos(phi), y, radial * math.sin(phi))
    helper = (0.0, 1.0, 0.0) if abs(ey[1]) < 0.9 else (1.0, 0.0, 0.0)
    k = helper[0] * ey[0] + helper[1] * ey[1] + helper[2] * ey[2]
    v = (helper[0] - ey[0] * k, helper[1] - ey[1] * k, helper[2] - ey[2] * k)
    n = math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
    ex = (v[0] / n, v[1] / n, v[2] / n)
    ez = (ex[1] * ey[2] - ex[2] * ey[1], ex[2] * ey[0] - ex[0] * ey[2], ex[0] * ey[1] - ex[1] * ey[0])
    roll = rnd.next_double_in(0.0, 2.0 * math.pi)
    c, s = math.cos(roll), math.sin(roll)
    a = tuple(ex[i] * c + ez[i] * s for i in range(3))
    b = tuple(ex[i] * -s + ez[i] * c for i in range(3))
    # Affine.unitVectors(a, ey, b): columns a, ey, b
    return (a[0], ey[0], b[0], 0.0, a[1], ey[1], b[1], 0.0, a[2], ey[2], b[2], 0.0)


def _transform(center, axes, rnd: SplittableRandom):
    """scaling(axes).then(rotation).then(translation(center)) (:311-316)."""
    from .workloads import scaling, then, translation
    t = then(scaling(*axes), _random_rotation(rnd))
    return then(t, translation(*center))


def primitives(centres, radius: float, seed: int, population: int):
    """generate (:103-132) + primitive (:278-309): typed lists in index order."""
    kinds = population_kinds(len(centres), seed, population)
    out = {SPHERE: [], BOX: [], AFFINE_SPHERE: [], AFFINE_BOX: []}
    half = radius / math.sqrt(3.0)
    for index, (kind, c) in enumerate(zip(kinds, centres)):
        rnd = SplittableRandom(mix(seed, index, 0x5A9E5 + 101 * kind))
        if kind == SPHERE:
            out[SPHERE].append((c[0], c[1], c[2], radius))
        elif kind == BOX:
            out[BOX].append((c[0] - half, c[1] - half, c[2] - half, c[0] + half, c[1] + half, c[2] + half))
        elif kind == AFFINE_SPHERE:
            if rnd.next_boolean():
                axes = (radius, radius * rnd.next_double_in(0.30, 0.44), radius * rnd.next_double_in(0.30, 0.44))
            else:
                axes = (radius, radius * rnd.next_double_in(0.82, 0.96), radius * rnd.next_double_in(0.26, 0.40))
            out[AFFINE_SPHERE].append(_transform(c, axes, rnd))
        else:
            p = (1.0, rnd.next_double_in(0.28, 0.42), rnd.next_double_in(0.28, 0.42))
            k = radius / math.sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2])
            out[AFFINE_BOX].append(_transform(c, (p[0] * k, p[1] * k, p[2] * k), rnd))
    return kinds, out


def row_id(family: str, object_count: int) -> str:
    return f"{family}-" + {100: "100", 1_000: "1k", 10_000: "10k"}[object_count]


# BenchmarkProtocol.publicationRows (:183-193): the random rows follow the six fixed ones
RANDOM_ROW_IDS = tuple(row_id(f, n) for f in FAMILIES for n in OBJECT_COUNTS)


def workload_key(row: str, layout_id: int) -> int:
    """BenchmarkProtocol.workloadKey (:279-289) of a random row's layout."""
    if layout_id not in LAYOUT_IDS:
        raise ValueError("Unknown benchmark random layout")
    return 6 + RANDOM_ROW_IDS.index(row) * len(LAYOUT_IDS) + layout_id


def metrics_seed(row: str, layout_id: int) -> int:
    """The frame seed of a random layout's counter frame (BenchmarkProtocol.metricsSeed)."""
    from .workloads import metrics_seed as seed_of
    return seed_of(workload_key(row, layout_id))


def random_scene(family: str, object_count: int, layout_id: int, config: BvhBuildConfig | None = None,
                 layout_cache: dict | None = None) -> GpuScene:
    """BenchmarkWorkloads.random (:74-90) as a GpuScene: the layout of
    (object count, layout id) under the study's experiment seed, populated as
    the family says, over the random environment colour.  ``layout_cache``
    shares layouts between families."""
    seed = layout_seed(LAYOUT_EXPERIMENT_SEED, object_count, layout_id)
    key = (object_count, layout_id)
    if layout_cache is not None and key in layout_cache:
        centres, radius = layout_cache[key]
    else:
        centres, radius = layout(object_count, seed)
        if layout_cache is not None:
            layout_cache[key] = (centres, radius)
    _, prims = primitives(centres, radius, seed, POPULATION[family])
    sc = GpuScene(config or BvhBuildConfig.uniform(8))
    sc.set_background(ENVIRONMENT)
    if prims[SPHERE]:
        sc.add_spheres(np.array(prims[SPHERE]), np.repeat(TYPE_MATERIALS[SPHERE:SPHERE + 1], len(prims[SPHERE]), 0))
    if prims[BOX]:
        sc.add_boxes(np.array(prims[BOX]), np.repeat(TYPE_MATERIALS[BOX:BOX + 1], len(prims[BOX]), 0))
    for kind, name in ((AFFINE_SPHERE, "sphere"), (AFFINE_BOX, "box")):
        if prims[kind]:
            sc.add_affine_transforms(name, np.array(prims[kind]),
                                     np.repeat(TYPE_MATERIALS[kind:kind + 1], len(prims[kind]), 0))
    return sc


def _normalized(v):
    n = math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
    return (v[0] / n, v[1] / n, v[2] / n)


def _cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def random_camera():
    """BenchmarkWorkloads.RANDOM_CAMERA (:16-17): SceneSupport.cameraLookAt
    (SceneSupport.java:93-103) from (-3, 0, 0) towards the origin, perspective 0.30."""
    from .workloads import IDENTITY, Camera, then, translation
    position, target = (-3.0, 0.0, 0.0), (0.0, 0.0, 0.0)
    forward = _normalized(tuple(t - p for t, p in zip(target, position)))
    right = _normalized(_cross((0.0, 1.0, 0.0), forward))
    up = _normalized(_cross(forward, right))
    basis = (right[0], up[0], forward[0], 0.0, right[1], up[1], forward[1], 0.0,
             right[2], up[2], forward[2], 0.0)                 # Affine.unitVectors(right, up, forward)
    return Camera(0.30, then(then(IDENTITY, basis), translation(*position)))


def random_gpu_camera(width: int, height: int):
    from .workloads import gpu_camera
    return gpu_camera(random_camera(), width, height)
