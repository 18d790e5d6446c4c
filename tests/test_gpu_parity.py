"""GPU parity: the HIP kernels against the FP32 oracle, bit for bit.

Closest-hit replay compares hit flag, distance bits, primitive order,
material index, normal bits and every per-ray counter (node visits and tests
per class).  Rendering compares whole images bit for bit and the 14 frame
counters exactly.  Full-size configurations are checked on sampled pixels.
"""
from __future__ import annotations

import numpy as np
import pytest

import mrad
from mrad import configs
from oracle import oracle as O

from helpers import assert_replay_equal, camera_grid_rays, random_rays
from scenes import gi_camera, gi_test_scene, small_field

pytestmark = pytest.mark.gpu

SCENES = {
    "room": (lambda: configs.room_scene(), configs.room_camera),
    "gi_test": (lambda: gi_test_scene(), gi_camera),
    "field3k": (lambda: small_field(3000), configs.field_camera),
    "clustered4k": (lambda: small_field(4000, 5, clustered=True), configs.field_camera),
}


def _rays(scene_name, camera):
    rng = np.random.default_rng(1234)
    grid = camera_grid_rays(camera.packed(), 48, 27)
    if scene_name in ("room",):
        rnd = random_rays(rng, 2048) * np.float32(0.5) + np.float32(0.5) * np.array([1, 1, 1, 0, 0, 0], np.float32)
    else:
        rnd = random_rays(rng, 2048)
    return np.concatenate([grid, rnd])


@pytest.mark.parametrize("name", list(SCENES))
@pytest.mark.parametrize("use_bvh,stack", [(True, 32), (True, 8), (False, 32)])
def test_replay_bit_exact(require_gpu, name, use_bvh, stack):
    make_scene, make_cam = SCENES[name]
    scene, cam = make_scene(), make_cam()
    rays = _rays(name, cam)
    print(f"replay {name} bvh={use_bvh} stack={stack} rays={len(rays)}", flush=True)
    ref = O.trace_replay(O.Scene(scene.wire(), use_bvh=use_bvh, stack_size=stack), rays)
    with mrad.GpuRayTracer(8, 8, 1, use_bvh, collect_metrics=True, bvh_stack_size=stack) as t:
        res = t.trace_replay(scene, rays, collect_metrics=True, per_ray_counters=True)
    assert_replay_equal(res, ref)
    agg = ref[5].astype(np.uint64)
    assert res.stats["rays"] == len(rays)
    assert res.stats["aabb_tests"] == int(agg[:, 1].sum())
    assert res.stats["max_stack_size"] == int(agg[:, 9].max())
    if stack == 8 and name.startswith("clustered") and use_bvh:
        assert res.stats["stack_overflows"] > 0   # the capacity-8 run exercises the linear fallback


def test_replay_launch_plan(require_gpu):
    """launchTraceReplayLocked's chunks (GpuRayTracer.java:624-673): rays in launches of at most
    replay_rays_per_launch give the results and counters of one launch; a linear replay's launches carry at most
    2^20 primitive-work units per ray (8 per sphere, box or plane, 64 per affine primitive), and a plan beyond that
    budget, or beyond 8192 launches, is an invalid argument (:743-769)."""
    scene, cam = small_field(3000), configs.field_camera()
    rays = _rays("field3k", cam)
    ref = O.trace_replay(O.Scene(scene.wire()), rays)
    with mrad.GpuRayTracer(8, 8, 1, True, collect_metrics=True, replay_rays_per_launch=1000) as t:
        res = t.trace_replay(scene, rays, collect_metrics=True, per_ray_counters=True)
        assert t.last_frame_stats()["physical_launches"] == -(-len(rays) // 1000)
    assert_replay_equal(res, ref)
    ns, nb, na = len(scene.sphere_data()), len(scene.box_data()), len(scene.affine_sphere_data())
    work = 8 * (ns + nb + len(scene.plane_data())) + 64 * (na + len(scene.affine_box_data()))
    per = (1 << 20) // work
    lin = O.trace_replay(O.Scene(scene.wire(), use_bvh=False), rays)
    with mrad.GpuRayTracer(8, 8, 1, False, collect_metrics=True) as t:
        res = t.trace_replay(scene, rays, collect_metrics=True, per_ray_counters=True)
        assert t.last_frame_stats()["physical_launches"] == -(-len(rays) // per)
    assert_replay_equal(res, lin)
    big = mrad.GpuScene()                      # 2^17 + 1 spheres: 8 (2^17 + 1) > 2^20 units per ray
    big.add_spheres([[i * 1e-3, 0.0, 5.0, 1e-4] for i in range((1 << 17) + 1)],
                    mrad.material_array((0.5, 0.5, 0.5), n=(1 << 17) + 1))
    with mrad.GpuRayTracer(8, 8, 1, False) as t:
        with pytest.raises(mrad.InvalidArgument, match="exceeding the watchdog-safe budget of 1048576"):
            t.trace_replay(big, rays[:16])
    edge = mrad.GpuScene()                     # exactly 2^20 units: one ray per launch, so 8193 rays need 8193
    edge.add_spheres([[i * 1e-3, 0.0, 5.0, 1e-4] for i in range(1 << 17)], mrad.material_array((0.5, 0.5, 0.5), n=1 << 17))
    many = np.tile(rays[:1], (8193, 1))
    with mrad.GpuRayTracer(8, 8, 1, False) as t:
        with pytest.raises(mrad.InvalidArgument, match="8193 watchdog-safe chunks, exceeding the limit of 8192"):
            t.trace_replay(edge, many)


def test_replay_uninstrumented_matches(require_gpu):
    scene, cam = small_field(3000), configs.field_camera()
    rays = _rays("field3k", cam)
    ref = O.trace_replay(O.Scene(scene.wire()), rays)
    with mrad.GpuRayTracer(8, 8, 1, True) as t:
        res = t.trace_replay(scene, rays)
    assert_replay_equal(res, ref, check_counters=False)


RENDER_CASES = [
    # name, width, height, spp, depth, rr, frame index
    ("room", 64, 64, 4, 4, -1, 0),
    ("room", 48, 40, 3, 4, -1, 5),
    ("gi_test", 64, 48, 2, 12, -1, 1),
    ("field3k", 96, 54, 2, 8, 3, 0),
    ("clustered4k", 80, 45, 2, 6, -1, 2),
]


@pytest.mark.parametrize("name,w,h,spp,depth,rr,frame", RENDER_CASES)
def test_render_bit_exact(require_gpu, name, w, h, spp, depth, rr, frame):
    make_scene, make_cam = SCENES[name]
    scene, cam = make_scene(), make_cam()
    seed = mrad.frame_seed(0x1234ABCD, frame)
    print(f"render {name} {w}x{h} spp={spp} depth={depth} rr={rr}", flush=True)
    ref, ref_stats = O.render_pixels(O.Scene(scene.wire()), cam.packed(), w, h, depth, frame, seed, spp, rr_start=rr)
    for metrics, lanes in ((False, 0), (True, 0), (True, 16)):
        with mrad.GpuRayTracer(w, h, spp, True, collect_metrics=metrics) as t:
            if rr >= 0:
                t.set_russian_roulette(rr)
            t.set_leaf_batch(lanes)
            img = np.zeros(w * h * 3, np.float32)
            t.render_sample(img, scene, cam, depth, frame, seed)
            stats = t.last_frame_stats()
        np.testing.assert_array_equal(img.view(np.uint32), ref.reshape(-1).view(np.uint32))
        if metrics:
            got = np.array([stats[k] for k in mrad.STAT_NAMES], dtype=np.uint64)
            np.testing.assert_array_equal(got, ref_stats)


@pytest.mark.parametrize("lanes", [0, 16])
def test_render_overflow_fallback_bit_exact(require_gpu, lanes):
    """A capacity-8 stack overflows on the clustered field: the linear fallback
    of traceBvh, in the plain and in the leaf-batched traversal loop."""
    scene, cam = SCENES["clustered4k"][0](), mrad.GpuCamera.look_at((0.2, 0.1, -0.3), (0, 0, 1), 90)
    w, h, spp, depth, frame = 160, 90, 2, 6, 3
    seed = mrad.frame_seed(0x5EED, frame)
    print(f"render clustered4k stack 8 leaf batch {lanes}", flush=True)
    ref, ref_stats = O.render_pixels(O.Scene(scene.wire(), stack_size=8), cam.packed(), w, h, depth, frame, seed, spp)
    assert int(ref_stats[mrad.STAT_NAMES.index("stack_overflows")]) > 0
    with mrad.GpuRayTracer(w, h, spp, True, collect_metrics=True, bvh_stack_size=8) as t:
        t.set_leaf_batch(lanes)
        img = np.zeros(w * h * 3, np.float32)
        t.render_sample(img, scene, cam, depth, frame, seed)
        stats = t.last_frame_stats()
    np.testing.assert_array_equal(img.view(np.uint32), ref.reshape(-1).view(np.uint32))
    np.testing.assert_array_equal(np.array([stats[k] for k in mrad.STAT_NAMES], dtype=np.uint64), ref_stats)


def test_render_split_launches_equal(require_gpu):
    """Physical launches of sample ranges accumulate exactly like one launch."""
    scene, cam = configs.room_scene(), configs.room_camera()
    out = []
    for per in (0, 1, 3):
        with mrad.GpuRayTracer(40, 30, 7, True, samples_per_launch=per) as t:
            img = np.zeros(40 * 30 * 3, np.float32)
            t.render_sample(img, scene, cam, 4, 2, 99)
            out.append(img)
            assert t.last_frame_stats()["physical_launches"] == (1 if per == 0 else -(-7 // per))
    np.testing.assert_array_equal(out[0], out[1])
    np.testing.assert_array_equal(out[0], out[2])


def test_render_pixel_tiles_equal_one_launch(require_gpu):
    """The reference's launch plan (pixel tiles of at most renderPixelsPerLaunch pixels, one sample per launch,
    GpuRayTracer.java:317-432) renders the same frame and counters as one launch, for tile sizes that do and do
    not divide the frame; a per-launch duration limit below the launch time fails the frame with IllegalState."""
    scene, cam = configs.room_scene(), configs.room_camera()
    w, h, spp = 70, 45, 3
    ref, ref_stats = None, None
    for ppl, per in ((0, 0), (1000, 1), (70 * 9, 2), (1, 3)):
        with mrad.GpuRayTracer(w, h, spp, True, collect_metrics=True, samples_per_launch=per,
                               pixels_per_launch=ppl) as t:
            img = np.zeros(w * h * 3, np.float32)
            t.render_sample(img, scene, cam, 4, 0, 77)
            st = t.last_frame_stats()
        if ref is None:
            ref, ref_stats = img, st
            continue
        np.testing.assert_array_equal(img.view(np.uint32), ref.view(np.uint32))
        assert all(st[k] == ref_stats[k] for k in mrad.STAT_NAMES)
        tiles = {1000: 4, 70 * 9: 5, 1: w * h}[ppl]
        assert st["physical_launches"] == tiles * (-(-spp // (per or spp)))
    with mrad.GpuRayTracer(w, h, spp, True, max_physical_kernel_nanos=1) as t:
        with pytest.raises(mrad.IllegalState, match="less than 1 ns"):
            t.render_sample(np.zeros(w * h * 3, np.float32), scene, cam, 4, 0, 77)


def test_full_size_field_sampled_pixels(require_gpu):
    """het-100k at 1920x1080: sampled pixels of the full GPU frame equal the oracle's."""
    w = configs.het(100_000, spp=2)
    seed = mrad.frame_seed(7, 0)
    with mrad.GpuRayTracer(w.width, w.height, w.spp, True) as t:
        t.set_russian_roulette(w.rr_start)
        img = np.zeros(w.width * w.height * 3, np.float32)
        t.render_sample(img, w.scene, w.camera, w.max_depth, 0, seed)
    pix = np.random.default_rng(3).choice(w.width * w.height, 1500, replace=False).astype(np.int32)
    ref, _ = O.render_pixels(O.Scene(w.scene.wire()), w.camera.packed(), w.width, w.height, w.max_depth, 0, seed,
                             w.spp, pixels=pix, rr_start=w.rr_start)
    got = img.reshape(-1, 3)[pix]
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_steady_state_samples(require_gpu):
    """conditionSteadyStateSample uploads and reports valid diagnostics; a steady-state sample of the resident
    scene renders the same frame; one that would upload another scene raises IllegalState."""
    scene, cam = configs.room_scene(), configs.room_camera()
    other = small_field(500)
    with mrad.GpuRayTracer(32, 24, 4, True, samples_per_launch=2) as t:
        cs = t.condition_steady_state_sample(scene, cam, 4, 0, 11)
        assert isinstance(cs, mrad.ConditioningStats)
        assert cs.physical_launches == 2 == t.last_physical_kernel_launch_count()
        assert cs.upload_nanos > 0
        a = np.zeros(32 * 24 * 3, np.float32)
        t.render_steady_state_sample(a, scene, cam, 4, 0, 11)
        b = np.zeros_like(a)
        t.render_sample(b, scene, cam, 4, 0, 11)
        np.testing.assert_array_equal(a, b)
        with pytest.raises(mrad.IllegalState):
            t.render_steady_state_sample(a, other, cam, 4, 0, 11)


def _quotient_edge_rays(scene, rng, n=6000):
    """Rays whose slab quotients leave the fast reciprocal path's range or hit its edge cases: direction
    components from 1e-40 (subnormal) to 1e30 and exact zeros, origins on node bound planes (zero numerators)
    and origins with tiny coordinates (subnormal-scale numerators)."""
    o = rng.uniform(-1.2, 1.2, (n, 3))
    mag = 10.0 ** rng.uniform(-40, 30, (n, 3))
    d = np.where(rng.random((n, 3)) < 0.5, -mag, mag)
    d[rng.random((n, 3)) < 0.1] = 0.0
    d[np.all(d == 0, axis=1), 0] = 1.0
    bounds = scene.bvh_node_bounds().reshape(-1)
    on_plane = rng.random((n, 3)) < 0.2
    o[on_plane] = rng.choice(bounds, on_plane.sum())
    tiny = rng.random((n, 3)) < 0.1
    o[tiny] = rng.uniform(-1, 1, tiny.sum()) * 10.0 ** rng.uniform(-38, -15, tiny.sum())
    # aim half of the rays at sphere centres so that many traverse deeply and hit
    aim = rng.random(n) < 0.5
    centres = scene.sphere_data()[:, 0:3].astype(np.float64)
    tgt = centres[rng.integers(0, len(centres), aim.sum())]
    d[aim] = (tgt - o[aim]) * 10.0 ** rng.uniform(-25, 25, (aim.sum(), 1))
    return np.concatenate([o, d], 1).astype(np.float32)


@pytest.mark.parametrize("name", ["field3k", "clustered4k"])
def test_replay_quotient_edge_cases(require_gpu, name):
    """The reciprocal-and-correction quotients of the slab and sphere tests equal IEEE division on every
    ray, including those that take the division fallback (bit-exact hits and per-ray counters)."""
    make_scene, _ = SCENES[name]
    scene = make_scene()
    rays = _quotient_edge_rays(scene, np.random.default_rng(77))
    assert np.all(np.isfinite(rays))
    print(f"quotient edge rays {name} n={len(rays)}", flush=True)
    ref = O.trace_replay(O.Scene(scene.wire()), rays)
    with mrad.GpuRayTracer(8, 8, 1, True, collect_metrics=True) as t:
        res = t.trace_replay(scene, rays, collect_metrics=True, per_ray_counters=True)
    assert_replay_equal(res, ref)
    assert int(ref[0].sum()) > 1000     # enough hits to exercise the primitive tests
