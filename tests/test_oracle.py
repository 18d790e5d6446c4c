"""Pins of the CPU oracle itself (CPU only).

Each check is chosen so a plausible slip in the restatement (dropped term,
wrong sign or index, swapped operands) fails: closed-form intersections,
FP64 brute force, BVH-versus-linear equality, exact counter identities of the
reference, an invertibility check of the hash and the furnace closed form of
the integrator.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import mrad
from mrad import configs
from oracle import oracle as O

from helpers import camera_grid_rays, random_rays
from scenes import gi_camera, gi_test_scene, small_field

M32 = 0xFFFFFFFF


# ------------------------------------------------------------------ RNG / math

def _inv_xorshift(y, k):
    x = y
    for _ in range(32 // k + 1):
        x = y ^ (x >> k)
    return x & M32


def _inv_hash(h):
    """Inverse of each step of triple32 (GpuKernelSources.java:165-174), built from the reference constants."""
    x = _inv_xorshift(h, 14)
    x = (x * pow(0x31848bab, -1, 1 << 32)) & M32
    x = _inv_xorshift(x, 15)
    x = (x * pow(0xac4c1b51, -1, 1 << 32)) & M32
    x = _inv_xorshift(x, 11)
    x = (x * pow(0xed5ad4bb, -1, 1 << 32)) & M32
    return _inv_xorshift(x, 17)


def test_hash_is_triple32():
    assert O.hash_u32(0) == 0
    rng = np.random.default_rng(0)
    for x in rng.integers(0, 1 << 32, 2000, dtype=np.uint64).tolist() + [1, 0x9e3779b9, M32]:
        h = O.hash_u32(int(x))
        assert _inv_hash(h) == x
        assert h != x or x == 0


def test_rand01_grid_and_range():
    import ctypes as C
    s = C.c_uint32(12345)
    vals = [O.lib.or_rand01(C.byref(s)) for _ in range(10000)]
    arr = np.array(vals, np.float64)
    assert arr.min() >= 0.0 and arr.max() < 1.0
    assert np.all(arr * 2 ** 24 == np.floor(arr * 2 ** 24))
    assert abs(arr.mean() - 0.5) < 0.01


def test_sincos_accuracy_over_sampling_range():
    xs = np.concatenate([np.linspace(0, 2 * np.pi, 20001, dtype=np.float32),
                         np.float32([0.0, np.pi / 4, np.pi / 2, np.pi, 3 * np.pi / 2, 6.2831855])])
    for x in xs.tolist():
        s, c = O.sincosf(x)
        assert abs(s - math.sin(x)) <= 3e-7 and abs(c - math.cos(x)) <= 3e-7, x
    s, c = O.sincosf(-1.0)
    assert abs(s + math.sin(1.0)) < 3e-7 and abs(c - math.cos(1.0)) < 3e-7


def test_path_seed_distinct_and_mixing():
    seeds = {O.path_seed(0xABCDEF0123456789, 3, p, s) for p in range(300) for s in range(16)}
    assert len(seeds) == 300 * 16
    # every input word changes the seed (frame seed halves, frame index, pixel, sample)
    base = O.path_seed(1, 2, 3, 4)
    assert len({base, O.path_seed(1 << 32 | 1, 2, 3, 4), O.path_seed(1, 3, 3, 4), O.path_seed(1, 2, 4, 4),
                O.path_seed(1, 2, 3, 5), O.path_seed(2, 2, 3, 4)}) == 6


GOLDEN = 0x9E3779B9


def _rand01_ref(state: int) -> tuple[float, int]:
    """rand01 (GpuKernelSources.java:176-179) from the pinned hash: state = h(state + 0x9e3779b9)."""
    state = O.hash_u32((state + 0x9E3779B9) & M32)
    return (state >> 8) * 2.0 ** -24, state


def _seed_base(frame_seed: int, frame: int, pixel: int, sample: int) -> int:
    """Peels the sample layers off a path seed: h^-1(seed) ^ h(s) ^ h(s * 0x9E3779B9 mod 2^32)."""
    inner = _inv_hash(O.path_seed(frame_seed, frame, pixel, sample))
    return inner ^ O.hash_u32(sample & M32) ^ O.hash_u32((sample * GOLDEN) & M32)


def test_path_seed_sample_layer_is_invariant():
    """seed = h(base ^ h(s) ^ h(s * 0x9E3779B9)) (GpuKernelSources.java:1250-1253): peeling the final hash and
    the two sample terms leaves one base per (frame seed, frame, pixel), whatever the sample.  A dropped final
    hash, a missing or misplaced golden-ratio product or a sample term hashed twice changes the base per s."""
    rng = np.random.default_rng(11)
    for _ in range(40):
        fs = int(rng.integers(0, 1 << 63)) | (int(rng.integers(0, 2)) << 63)
        frame, pixel = int(rng.integers(0, 1 << 20)), int(rng.integers(0, 1 << 31))
        bases = {_seed_base(fs, frame, pixel, s) for s in list(range(70)) + [1 << 20, (1 << 31) - 1]}
        assert len(bases) == 1
    # all-zero inputs: every hash of 0 is 0 (triple32 fixes 0), so the seed is 0
    assert O.path_seed(0, 0, 0, 0) == 0 and _seed_base(0, 0, 0, 0) == 0


def test_path_seed_base_layer_xor_structure():
    """base = h(h(lo) ^ h(hi) ^ h(frame) ^ h(pixel)) (GpuKernelSources.java:1231-1249): h^-1(base) changes by
    exactly h(a) ^ h(b) when one input word changes from a to b, for each of the four words, and is 0 for all-zero
    inputs, which fixes h^-1(base) = h(lo) ^ h(hi) ^ h(frame) ^ h(pixel) everywhere."""
    H = O.hash_u32
    rng = np.random.default_rng(12)

    def core(fs, frame, pixel):
        return _inv_hash(_seed_base(fs, frame, pixel, int(rng.integers(0, 64))))

    assert core(0, 0, 0) == 0
    for _ in range(60):
        lo1, lo2, hi1, hi2 = (int(v) for v in rng.integers(0, 1 << 32, 4))
        f1, f2 = (int(v) for v in rng.integers(0, 1 << 31, 2))
        p1, p2 = (int(v) for v in rng.integers(0, 1 << 32, 2))
        fs = hi1 << 32 | lo1
        assert core(fs, f1, p1) ^ core(fs, f1, p2) == H(p1) ^ H(p2)                 # pixel word
        assert core(fs, f1, p1) ^ core(fs, f2, p1) == H(f1) ^ H(f2)                 # frame index
        assert core(hi1 << 32 | lo1, f1, p1) ^ core(hi1 << 32 | lo2, f1, p1) == H(lo1) ^ H(lo2)   # low half
        assert core(hi1 << 32 | lo1, f1, p1) ^ core(hi2 << 32 | lo1, f1, p1) == H(hi1) ^ H(hi2)   # high half
    # a negative frame index enters as its 32-bit two's complement
    assert core(5, -3, 7) ^ core(5, 0, 7) == H((-3) & M32)


def _cam(pos=(0.25, -0.5, 2.0), fwd=(0.0, 0.0, 1.0), right=(1.0, 0.0, 0.0), up=(0.0, 1.0, 0.0)):
    return np.array([*pos, *fwd, *right, *up], np.float32)


@pytest.mark.parametrize("w,h", [(1920, 1080), (256, 256), (37, 91)])
def test_camera_ray_lies_in_its_pixel_footprint(w, h):
    """renderKernel's camera ray (GpuKernelSources.java:1256-1264) with fwd = z, right = x, up = y has direction
    (u, v, 1): u / aspect lies in the pixel's footprint [x, x + 1) * 2 / W - 1 and v in 1 - [y, y + 1) * 2 / H,
    at the offset of the first and second jitter draws of the path's RNG, and the state advances by exactly those
    two draws.  A missing +0.5, a flipped v, a wrong aspect or a swapped draw fails."""
    rng = np.random.default_rng(w * 7 + h)
    aspect = np.float32(w) / np.float32(h)
    fracx, fracy = [], []
    for _ in range(1500):
        x, y, st = int(rng.integers(0, w)), int(rng.integers(0, h)), int(rng.integers(0, 1 << 32))
        ray, after = O.camera_ray(_cam(), w, h, x, y, st)
        assert ray[0:3].tolist() == [0.25, -0.5, 2.0] and ray[5] == 1.0
        jx, s1 = _rand01_ref(st)
        jy, s2 = _rand01_ref(s1)
        assert after == s2
        px = (float(ray[3]) / float(aspect) + 1.0) * w / 2.0
        py = (1.0 - float(ray[4])) * h / 2.0
        slack = 4e-6 * max(w, h)
        assert x - slack <= px <= x + 1 + slack and y - slack <= py <= y + 1 + slack, (x, y, px, py)
        assert abs((px - x) - jx) <= slack and abs((py - y) - jy) <= slack
        fracx.append(px - x)
        fracy.append(py - y)
    for f in (np.array(fracx), np.array(fracy)):     # jitter uniform over the footprint
        assert abs(f.mean() - 0.5) < 0.03 and abs(f.var() - 1.0 / 12.0) < 0.01
        hist, _ = np.histogram(f, bins=10, range=(0, 1))
        assert hist.min() > 100


def test_camera_ray_is_fwd_plus_right_u_plus_up_v():
    """For a rotated, scaled camera basis the direction is fwd + right u + up v with the (u, v) of the axis
    camera for the same pixel and state: right carries u, up carries v, fwd enters once and the origin is the
    camera position."""
    rng = np.random.default_rng(5)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    fwd, right, up = q[:, 0] * 1.7, q[:, 1] * 0.6, q[:, 2] * 0.9
    cam = _cam(pos=(1.5, 2.5, -3.0), fwd=fwd, right=right, up=up)
    for _ in range(300):
        x, y, st = int(rng.integers(0, 640)), int(rng.integers(0, 480)), int(rng.integers(0, 1 << 32))
        axis, _ = O.camera_ray(_cam(), 640, 480, x, y, st)
        ray, _ = O.camera_ray(cam, 640, 480, x, y, st)
        np.testing.assert_array_equal(ray[0:3], cam[0:3])
        d = ray[3:6].astype(np.float64) - cam[3:6].astype(np.float64)
        r64, u64 = cam[6:9].astype(np.float64), cam[9:12].astype(np.float64)
        u = d @ r64 / (r64 @ r64)
        v = d @ u64 / (u64 @ u64)
        assert abs(u - float(axis[3])) < 2e-6 and abs(v - float(axis[4])) < 2e-6


def test_frame_seed_is_splitmix64():
    # SplitMix64 of 0 is the published first output of the generator seeded with 0
    assert mrad.mix64(0) == 0xE220A8397B1DCDAF
    assert mrad.frame_seed(0, 0) == 0xE220A8397B1DCDAF


# ---------------------------------------------------------- closed-form hits

def _scene(spheres=(), boxes=(), affine=(), leaf=8):
    sc = mrad.GpuScene(mrad.BvhBuildConfig.uniform(leaf))
    if len(spheres):
        sc.add_spheres(spheres, mrad.material_array((0.5, 0.5, 0.5), n=len(spheres)))
    if len(boxes):
        sc.add_boxes(boxes, mrad.material_array((0.5, 0.5, 0.5), n=len(boxes)))
    for kind, t in affine:
        sc.add_affine_transforms(kind, [t], mrad.material_array((0.5, 0.5, 0.5)))
    return sc


def _replay(sc, rays, use_bvh=True, stack=32):
    return O.trace_replay(O.Scene(sc.wire(), use_bvh=use_bvh, stack_size=stack), np.asarray(rays, np.float32))


def test_sphere_closed_forms():
    sc = _scene(spheres=[[0, 0, 0, 1.0]])
    rays = [[0, 0, -5, 0, 0, 1], [0, 0, -5, 0, 0, 2], [0, 0, 0, 0, 0, 1], [0, 1.0, -5, 0, 0, 1],
            [0, 0, 5, 0, 0, 1], [0.6, 0, -5, 0, 0, 1]]
    f, t, o, m, n, per = _replay(sc, rays)
    assert f.tolist() == [1, 1, 1, 0, 0, 1]
    assert t[0] == 4.0 and t[1] == 2.0 and t[2] == 1.0          # unnormalized dir: t in units of dir
    np.testing.assert_array_equal(n.reshape(-1, 3)[0], [0, 0, -1])
    np.testing.assert_array_equal(n.reshape(-1, 3)[2], [0, 0, 1])  # inside origin: exit, outward normal
    assert abs(t[5] - (5 - math.sqrt(1 - 0.36))) < 1e-6
    assert o[0] == 0 and o[3] == -1 and m[3] == -1 and np.isinf(t[3])


def test_box_closed_forms_and_zero_direction_branches():
    sc = _scene(boxes=[[-1, -1, -1, 1, 1, 1]])
    rays = [[0, 0, -5, 0, 0, 1], [0, 0, 0, 0, 0, 1], [2, 0, -5, 0, 0, 1], [0.999, 0.5, -5, 0, 0, 1],
            [0, 0, -5, 0, 0, -1], [-3, -3, -3, 1, 1, 1], [0.5, -5, 0.5, 0, 3, 0]]
    f, t, o, m, n, _ = _replay(sc, rays)
    n = n.reshape(-1, 3)
    assert f.tolist() == [1, 1, 0, 1, 0, 1, 1]
    assert t[0] == 4.0 and t[1] == 1.0 and t[3] == 4.0 and t[5] == 2.0
    assert abs(t[6] - 4.0 / 3.0) < 1e-7
    np.testing.assert_array_equal(n[0], [0, 0, -1])
    np.testing.assert_array_equal(n[1], [0, 0, 1])
    np.testing.assert_array_equal(n[6], [0, -1, 0])
    assert n[5].tolist() in ([-1, 0, 0],)  # equal entry on all axes keeps the x axis


def test_affine_equals_plain_primitive_under_similarity():
    """An affine sphere with a uniform scale is a sphere; an affine box rotated 90 deg is an axis-aligned box."""
    c = (0.3, -0.2, 0.4)
    aff_s = _scene(affine=[("sphere", [0.5, 0, 0, c[0], 0, 0.5, 0, c[1], 0, 0, 0.5, c[2]])])
    plain_s = _scene(spheres=[[*c, 0.5]])
    rot = [0, 0, 3.0, c[0], 0, 2.0, 0, c[1], -1.0, 0, 0, c[2]]  # x->-z*1, y->y*2, z->x*3
    aff_b = _scene(affine=[("box", rot)])
    plain_b = _scene(boxes=[[c[0] - 3, c[1] - 2, c[2] - 1, c[0] + 3, c[1] + 2, c[2] + 1]])
    rng = np.random.default_rng(5)
    for a, p, half in ((aff_s, plain_s, (0.6, 0.6, 0.6)), (aff_b, plain_b, (3.5, 2.5, 1.5))):
        target = np.asarray(c) + rng.uniform(-1, 1, (3000, 3)) * half
        origin = np.asarray(c) + rng.normal(0, 1, (3000, 3)) * 8.0
        rays = np.concatenate([origin, (target - origin) * rng.uniform(0.05, 0.3, (3000, 1))], 1).astype(np.float32)
        fa, ta, _, _, na, _ = _replay(a, rays)
        fp, tp, _, _, np_, _ = _replay(p, rays)
        both = (fa == 1) & (fp == 1)
        assert (fa != fp).sum() <= 3          # grazing rays only
        assert both.sum() > 500
        np.testing.assert_allclose(ta[both], tp[both], rtol=2e-5, atol=2e-5)
        np.testing.assert_allclose(na.reshape(-1, 3)[both], np_.reshape(-1, 3)[both], atol=2e-3)


def test_equal_distance_tie_takes_lower_order():
    sc = _scene(spheres=[[0, 0, 0, 1.0], [0, 0, 0, 1.0], [0, 0, 0, 1.0]])
    for use_bvh in (True, False):
        f, t, o, _, _, _ = _replay(sc, [[0, 0, -5, 0, 0, 1], [0.3, 0.1, -5, 0, 0, 1]], use_bvh=use_bvh)
        assert o.tolist() == [0, 0]


# --------------------------------------------------------- traversal checks

FIELD_SCENES = [lambda: small_field(3000), lambda: small_field(4000, 5, clustered=True), lambda: configs.room_scene(),
                lambda: gi_test_scene()]


@pytest.mark.parametrize("make", FIELD_SCENES)
def test_bvh_equals_linear_scan(make):
    sc = make()
    rng = np.random.default_rng(17)
    rays = np.concatenate([random_rays(rng, 3000), camera_grid_rays(configs.field_camera().packed(), 32, 18)])
    b = _replay(sc, rays, use_bvh=True)
    lin = _replay(sc, rays, use_bvh=False)
    np.testing.assert_array_equal(b[0], lin[0])
    np.testing.assert_array_equal(b[2], lin[2])
    np.testing.assert_array_equal(b[1].view(np.uint32), lin[1].view(np.uint32))
    np.testing.assert_array_equal(b[4].view(np.uint32), lin[4].view(np.uint32))


@pytest.mark.parametrize("make", FIELD_SCENES)
def test_counter_identities(make):
    sc = make()
    rays = random_rays(np.random.default_rng(4), 2000)
    per = _replay(sc, rays)[5].astype(np.int64)
    roots = len(sc.bvh_root_indices())
    # thesis eq. 4.4: N_AABB = roots + 2 N_internal (per ray)
    np.testing.assert_array_equal(per[:, 1], roots + 2 * per[:, 10])
    np.testing.assert_array_equal(per[:, 11], per[:, 12] + per[:, 13])           # leaf-kind partition
    np.testing.assert_array_equal(per[:, 2], per[:, 3:8].sum(1))                  # tests by class
    assert np.all(per[:, 0] == 1) and np.all(per[:, 8] == 0)


def test_stack_overflow_falls_back_to_linear():
    sc = small_field(4000, 5, clustered=True)
    rays = random_rays(np.random.default_rng(8), 3000)
    small = _replay(sc, rays, stack=1)
    lin = _replay(sc, rays, use_bvh=False)
    per = small[5]
    assert per[:, 8].sum() > 100                       # overflows happened
    np.testing.assert_array_equal(small[2], lin[2])    # same closest hits as the linear scan
    ov = per[:, 8] == 1
    assert np.all(per[ov, 2] >= len(sc.bvh_primitive_refs()))  # the fallback tests every ref


def test_fp32_matches_fp64_closest_hits():
    """GpuBvhCorrectnessCheck CPU-vs-GPU tolerances (GpuBvhCorrectnessCheck.java:31-32, :206-238)."""
    sc = small_field(3000)
    osc = O.Scene(sc.wire())
    rays = np.concatenate([random_rays(np.random.default_rng(9), 1500),
                           camera_grid_rays(configs.field_camera().packed(), 32, 18)])
    f, t, o, m, n, _ = O.trace_replay(osc, rays)
    n = n.reshape(-1, 3)
    offsets = {0: 0, 1: len(sc.sphere_data()), 2: len(sc.sphere_data()) + len(sc.box_data()),
               3: len(sc.sphere_data()) + len(sc.box_data()) + len(sc.affine_sphere_data())}
    mismatch = 0
    for i, r in enumerate(rays):
        hit, t64, kind, idx, n64 = O.closest64(osc, r)
        if bool(f[i]) != hit or (hit and offsets[kind] + idx != o[i]):
            mismatch += 1
            continue
        if hit:
            assert abs(t[i] - t64) <= 2e-3 + 2e-4 * max(1.0, abs(t64))
            assert np.all(np.abs(n[i] - n64) <= 2e-3)
    assert mismatch <= 2   # grazing-angle ties only


# --------------------------------------------------------------- integrator

def _furnace_scene(albedo, emission):
    """Closed room of six wall boxes (solids' normals point into the room), all emitting and reflecting."""
    sc = mrad.GpuScene()
    t = 0.1
    walls = [[-1 - t, -1 - t, -1 - t, 1 + t, -1, 1 + t], [-1 - t, 1, -1 - t, 1 + t, 1 + t, 1 + t],
             [-1 - t, -1 - t, -1 - t, -1, 1 + t, 1 + t], [1, -1 - t, -1 - t, 1 + t, 1 + t, 1 + t],
             [-1 - t, -1 - t, -1 - t, 1 + t, 1 + t, -1], [-1 - t, -1 - t, 1, 1 + t, 1 + t, 1 + t]]
    sc.add_boxes(walls, mrad.material_array((albedo,) * 3, emittance=(emission,) * 3, n=6))
    return sc


@pytest.mark.parametrize("depth", [1, 2, 4, 8])
def test_furnace_closed_form(depth):
    """Camera inside a closed diffuse emitter: every path sums E * rho^k, k < depth."""
    rho, E = 0.6, 2.0
    sc = _furnace_scene(rho, E)
    cam = mrad.GpuCamera.look_at((0.1, 0.2, -0.3), (0.3, 0.1, 0.5), 60.0).packed()
    img, st = O.render_pixels(O.Scene(sc.wire()), cam, 16, 16, depth, 0, 5, 3)
    expect = E * (1 - rho ** depth) / (1 - rho)
    np.testing.assert_allclose(img, expect, rtol=2e-6)
    assert st[0] == 16 * 16 * 3 * depth
    m64 = O.render64(O.Scene(sc.wire()), cam, 8, 8, depth, 4, 1)
    np.testing.assert_allclose(m64, expect, rtol=1e-7)   # materials are packed to FP32


def test_furnace_roulette_unbiased():
    rho, E, depth = 0.6, 2.0, 12
    sc = _furnace_scene(rho, E)
    cam = mrad.GpuCamera.look_at((0.1, 0.2, -0.3), (0.3, 0.1, 0.5), 60.0).packed()
    img, st = O.render_pixels(O.Scene(sc.wire()), cam, 64, 64, depth, 0, 5, 16, rr_start=3)
    expect = E * (1 - rho ** depth) / (1 - rho)
    assert abs(img.mean() - expect) < 0.01 * expect
    assert st[0] < 64 * 64 * 16 * depth  # roulette ended paths early


def test_emitter_seen_directly():
    sc = mrad.GpuScene()
    sc.set_background((0.1, 0.2, 0.3))
    sc.add_boxes([[-0.5, -0.5, 1, 0.5, 0.5, 1.2]], mrad.material_array((0, 0, 0), emittance=(3, 4, 5)))
    cam = mrad.GpuCamera.look_at((0, 0, -2), (0, 0, 1), 20.0).packed()
    img, _ = O.render_pixels(O.Scene(sc.wire()), cam, 8, 8, 1, 0, 1, 1)
    assert np.all(img[27] == np.float32([3, 4, 5]))          # centre pixel sees the emitter
    img, _ = O.render_pixels(O.Scene(sc.wire()), cam, 8, 8, 0, 0, 1, 1)
    assert np.all(img == 0)                                     # depth 0 renders black


def test_fp32_render_agrees_with_fp64_within_monte_carlo_band():
    """The FP32 path's tolerance on the room at 32x32: each pixel of the FP32
    estimate lies within 3.29 sigma of the FP64 reference's mean."""
    w = configs.room()
    osc = O.Scene(w.scene.wire())
    cam = w.camera.packed()
    W = H = 24
    spp32 = 256
    img32, _ = O.render_pixels(osc, cam, W, H, w.max_depth, 0, 123, spp32)
    ref, var = O.render64(osc, cam, W, H, w.max_depth, 4096, 77, variance=True)
    var32 = np.maximum(var, 1e-12)
    sigma = np.sqrt(var32 / spp32 + var32 / 4096)
    z = np.abs(img32 - ref) / np.maximum(sigma, 1e-9)
    assert (z <= 3.29).mean() >= 0.99
    relmse = np.mean((img32 - ref) ** 2 / (ref ** 2 + 1e-2))
    expected = np.mean(sigma ** 2 / (ref ** 2 + 1e-2))    # pure Monte-Carlo noise of both estimates
    assert relmse <= 1.5 * expected
