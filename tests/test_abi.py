"""The C-ABI library: symbols, error mapping and host-side validation (CPU only;
no compute call needs a GPU here)."""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import mrad
from mrad import _native as N
from mrad import configs

from helpers import bad_material_variants, corrupt_view
from scenes import gi_test_scene, small_field

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "mrad.h").read_text()
    return sorted(set(re.findall(r"\b(mrad_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(N.LIB_PATH))
    names = declared_symbols()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(N.SIGNATURES), "binding and header disagree"


def test_validate_full_path_depth():
    mrad.validate_full_path_depth(0)
    mrad.validate_full_path_depth(32)
    for bad in (-1, 33):
        with pytest.raises(mrad.InvalidArgument, match=r"\[0, 32\]"):
            mrad.validate_full_path_depth(bad)


def test_code_object_hash_is_sha256():
    h = mrad.GpuRayTracer.compiled_code_object_sha256()
    assert re.fullmatch(r"[0-9a-f]{64}", h)


def test_invalid_stack_size_rejected():
    for bad in (4, 257):
        with pytest.raises(mrad.InvalidArgument):
            mrad.GpuRayTracer(8, 8, 1, True, bvh_stack_size=bad)


def test_tracer_without_device_is_unavailable_and_errors_are_illegal_state(monkeypatch):
    if mrad.GpuRayTracer(1, 1, 1, True).is_available():
        pytest.skip("a HIP device is present; the unavailable path is exercised on CPU hosts")
    t = mrad.GpuRayTracer(4, 4, 1, True)
    assert not t.is_available()
    sc = mrad.GpuScene()
    sc.add_spheres([[0, 0, 0, 1]], mrad.material_array((0.5, 0.5, 0.5)))
    cam = mrad.GpuCamera.look_at((0, 0, -3), (0, 0, 0), 40)
    with pytest.raises(mrad.IllegalState, match="unavailable"):
        t.render_sample(np.zeros(48, np.float32), sc, cam, 4, 0, 1)
    with pytest.raises(mrad.IllegalState):
        t.trace_replay(sc, np.zeros(6, np.float32))


def test_argument_checks_precede_device_checks():
    t = mrad.GpuRayTracer(4, 4, 1, True)
    sc = mrad.GpuScene()
    sc.add_spheres([[0, 0, 0, 1]], mrad.material_array((0.5, 0.5, 0.5)))
    cam = mrad.GpuCamera.look_at((0, 0, -3), (0, 0, 0), 40)
    with pytest.raises(mrad.InvalidArgument):
        t.render_sample(np.zeros(48, np.float32), sc, cam, 33, 0, 1)
    with pytest.raises(mrad.InvalidArgument):
        t.render_sample(np.zeros(10, np.float32), sc, cam, 4, 0, 1)
    with pytest.raises(mrad.InvalidArgument):
        t.trace_replay(sc, np.zeros(5, np.float32))
    with pytest.raises(mrad.InvalidArgument, match="GPU scene must not be null"):
        t.trace_replay(None, np.zeros(6, np.float32))
    with pytest.raises(mrad.InvalidArgument, match="GPU scene and camera must not be null"):
        t.render_sample(np.zeros(48, np.float32), None, cam, 4, 0, 1)
    rc = N.lib.mrad_upload_scene(None, None, 0)
    assert rc == N.MRAD_ERR_INVALID_ARGUMENT
    assert b"null" in N.lib.mrad_last_error()


def test_leaf_batch_setter_checks_arguments():
    """mrad_set_leaf_batch: 0 or 16 on any tracer handle (no device needed); a
    null handle and other widths are invalid arguments."""
    t = mrad.GpuRayTracer(4, 4, 1, True)
    t.set_leaf_batch(16)
    t.set_leaf_batch(0)
    for bad in (-1, 1, 8, 32):
        with pytest.raises(mrad.InvalidArgument):
            t.set_leaf_batch(bad)
    assert N.lib.mrad_set_leaf_batch(None, 0) == N.MRAD_ERR_INVALID_ARGUMENT
    assert b"null" in N.lib.mrad_last_error()


def test_trace_pass_settings_are_checked_at_creation():
    """trace_passes 1..3, pass_pops in [1, 2^20], suspend_capacity in [64, queue_capacity] and
    pass_dense_lanes -1 or in [1, 64] (0 picks the default); checked before any device work, so no GPU is
    needed."""
    for ok in (dict(trace_passes=2), dict(trace_passes=3, pass_pops=1, suspend_capacity=64, pass_dense_lanes=-1),
               dict(trace_passes=2, pass_pops=1 << 20, suspend_capacity=4096, pass_dense_lanes=64),
               dict(pass_dense_lanes=1)):
        mrad.BatchRenderer(mrad.WavefrontConfig(queue_capacity=4096, **ok)).close()
    for bad, msg in ((dict(trace_passes=4), "trace passes"), (dict(trace_passes=-1), "trace passes"),
                     (dict(pass_pops=-5), "pass pops"), (dict(pass_pops=(1 << 20) + 1), "pass pops"),
                     (dict(suspend_capacity=32), "suspend capacity"),
                     (dict(suspend_capacity=4097), "suspend capacity"),
                     (dict(pass_dense_lanes=65), "pass dense lanes"), (dict(pass_dense_lanes=-2), "pass dense lanes")):
        with pytest.raises(mrad.InvalidArgument, match=msg):
            mrad.BatchRenderer(mrad.WavefrontConfig(queue_capacity=4096, **bad))


def test_batch_config_defaults_are_resolved():
    """mrad_batch_get_config reports the parameters in use: two capped trace passes of 80 pops (dense-wave
    threshold 16, continuation lists of half the queue) for a queue of 2^25 paths or more, one pass below, the
    queue rounded up to whole waves; explicit values are kept.  No device work, so no GPU is needed."""
    with mrad.BatchRenderer(mrad.WavefrontConfig(queue_capacity=1 << 22)) as b:
        c = b.effective_config()
        assert (c["trace_passes"], c["pass_pops"], c["pass_dense_lanes"]) == (1, 80, 16)
        assert c["suspend_capacity"] == 1 << 21 and c["waves_per_cu"] == 28 and c["lds_stack_entries"] == 10
    with mrad.BatchRenderer(mrad.WavefrontConfig(queue_capacity=1 << 25)) as b:
        c = b.effective_config()
        assert (c["trace_passes"], c["pass_pops"], c["suspend_capacity"]) == (2, 80, 1 << 24)
    with mrad.BatchRenderer(mrad.WavefrontConfig(queue_capacity=(1 << 25) - 100, trace_passes=3, pass_pops=7,
                                                 pass_dense_lanes=-1)) as b:
        c = b.effective_config()
        assert c["queue_capacity"] == (1 << 25) - 64
        assert (c["trace_passes"], c["pass_pops"], c["pass_dense_lanes"]) == (3, 7, -1)
    assert N.lib.mrad_batch_get_config(None, None) == N.MRAD_ERR_INVALID_ARGUMENT


def test_destroy_null_is_noop():
    N.lib.mrad_destroy(None)
    N.lib.mrad_scene_destroy(None)


def test_conditioning_stats_validation():
    """ConditioningStats rejects the diagnostics GpuRayTracer.java:108-116 rejects."""
    ok = mrad.ConditioningStats(2, 0, 100, 60, 500)
    assert ok.physical_launches == 2
    for bad in [(0, 0, 100, 60, 500), (1, -1, 100, 60, 500), (1, 0, 0, 0, 500), (1, 0, 100, 0, 500),
                (1, 0, 100, 101, 500), (1, 0, 100, 60, 0)]:
        with pytest.raises(mrad.InvalidArgument):
            mrad.ConditioningStats(*bad)


def test_calibration_entry_points_check_arguments_first():
    """mrad_cost_create / mrad_cost_time reject bad arguments before touching a device."""
    h = C.c_void_p()
    cus = C.c_int()
    w = np.zeros(16, np.float32)
    for records, profiles in ((100, 3), (0, 3), (64, 0)):
        rc = N.lib.mrad_cost_create(0, w.ctypes.data_as(N.fp), records, profiles, C.byref(cus), C.byref(h))
        assert rc == N.MRAD_ERR_INVALID_ARGUMENT
    assert N.lib.mrad_cost_create(0, None, 64, 3, C.byref(cus), C.byref(h)) == N.MRAD_ERR_INVALID_ARGUMENT
    ms = C.c_float()
    assert N.lib.mrad_cost_time(None, 1, 0, 0, 1, 1, 1, 64, C.byref(ms)) == N.MRAD_ERR_INVALID_ARGUMENT
    N.lib.mrad_cost_destroy(None)


@pytest.mark.parametrize("make", [configs.room_scene, gi_test_scene, lambda: small_field(400, 2)])
def test_scene_with_bad_material_index_is_rejected_on_the_host(make):
    """A material index that is not -1 (no material, GpuScene.materialIndexOf) or inside the scene's table
    (past it, below -1, non-integral, NaN) is an invalid argument, found before any value reaches a kernel;
    mrad_batch_add_scene validates on the host, so no device is needed here."""
    sc = make()
    variants = bad_material_variants(sc)
    assert len(variants) >= 8
    cam = configs.field_camera()
    with mrad.BatchRenderer(mrad.WavefrontConfig(queue_capacity=4096)) as b:
        for label, arrays in variants:
            v, keep = corrupt_view(sc, **arrays)
            view = mrad.batch.make_view(cam, 8, 8, 1, 4, 0, 1)
            rc = N.lib.mrad_batch_add_scene(b._h, C.byref(v), C.byref(view), None)
            assert rc == N.MRAD_ERR_INVALID_ARGUMENT, label
            assert b"material index" in N.lib.mrad_last_error(), label
        assert b.add_scene(sc, cam, 8, 8, 1, 4) == 0    # the valid scene is still accepted, as the first


def test_batch_accepts_more_than_2_32_paths_and_checks_view_limits():
    """Paths of a batch are numbered with 64 bits: views totalling more than 2^32 paths are accepted (host logic,
    no device needed).  A view of more than 2^32 pixels or 2^30 samples is an invalid argument."""
    sc, cam = small_field(200, 4), configs.field_camera()
    with mrad.BatchRenderer(mrad.WavefrontConfig()) as b:
        for _ in range(3):
            b.add_scene(sc, cam, 2048, 2048, 512, 4)          # 2.1e9 paths each
        for w, h, spp in ((1 << 17, 1 << 16, 1), (8, 8, (1 << 30) + 1)):
            view = mrad.batch.make_view(cam, w, h, spp, 4, 0, 1)
            rc = N.lib.mrad_batch_add_scene(b._h, C.byref(sc.arrays()), C.byref(view), None)
            assert rc == N.MRAD_ERR_INVALID_ARGUMENT
    for bad in (-1, (1 << 30) + 1):
        with pytest.raises(mrad.InvalidArgument):
            mrad.BatchRenderer(mrad.WavefrontConfig(result_chunk_paths=bad))


def test_render_launch_plan_limits_are_checked_before_the_device():
    """renderLaunchCount (GpuRayTracer.java:550-564): a plan of more than 8192 pixel-tile x sample launches is an
    invalid argument even where no device is present; pixels per launch outside [0, 262144] and a negative
    watchdog limit are rejected at creation."""
    sc = mrad.GpuScene()
    sc.add_spheres([[0, 0, 0, 1]], mrad.material_array((0.5, 0.5, 0.5)))
    cam = mrad.GpuCamera.look_at((0, 0, -3), (0, 0, 0), 40)
    t = mrad.GpuRayTracer(128, 128, 1, True, pixels_per_launch=1)          # 16384 one-pixel tiles
    with pytest.raises(mrad.InvalidArgument, match="require 16384 watchdog-safe .* exceeding the limit of 8192"):
        t.render_sample(np.zeros(128 * 128 * 3, np.float32), sc, cam, 4, 0, 1)
    with pytest.raises(mrad.InvalidArgument, match="exceeding the limit of 8192"):
        t.render_launch_count()
    t.close()
    # 8192 launches exactly is allowed: 64 x 128 two-pixel tiles of one sample each
    t = mrad.GpuRayTracer(128, 128, 1, True, pixels_per_launch=2)
    assert t.render_launch_count() == 8192
    t.close()
    # tiles x sample launches, against the provenance layout rule
    from mrad.benchmark_runner import render_tile_count
    for w, h, spp, per, ppl in ((1920, 1080, 4, 1, 262_144), (1920, 1080, 4, 2, 262_144), (640, 480, 3, 0, 0),
                                (1000, 7, 5, 1, 333), (7, 1000, 2, 1, 5)):
        t = mrad.GpuRayTracer(w, h, spp, True, samples_per_launch=per, pixels_per_launch=ppl)
        tiles = render_tile_count(w, h, ppl) if ppl else 1
        ranges = 1 if per == 0 else -(-spp // per)
        assert t.render_launch_count() == tiles * ranges, (w, h, spp, per, ppl)
        t.close()
    for kw in (dict(pixels_per_launch=262_145), dict(pixels_per_launch=-1), dict(max_physical_kernel_nanos=-5),
               dict(replay_rays_per_launch=262_145), dict(replay_rays_per_launch=-1)):
        with pytest.raises(mrad.InvalidArgument):
            mrad.GpuRayTracer(8, 8, 1, True, **kw)


def test_render_form_checks_arguments_first():
    """mrad_set_render_form: 0 or 1, checked before the device; the wavefront form skips the megakernel's pixel-tile
    plan (a plan of 16384 launches is then no error) and still needs a device to render."""
    t = mrad.GpuRayTracer(128, 128, 1, True, pixels_per_launch=1)
    for bad in (2, -1, "fast"):
        with pytest.raises(mrad.InvalidArgument, match="render form"):
            t.set_render_form(bad)
    t.set_render_form("wavefront")
    sc = mrad.GpuScene()
    sc.add_spheres([[0, 0, 0, 1]], mrad.material_array((0.5, 0.5, 0.5)))
    cam = mrad.GpuCamera.look_at((0, 0, -3), (0, 0, 0), 40)
    with pytest.raises(mrad.IllegalState):
        t.render_sample(np.zeros(128 * 128 * 3, np.float32), sc, cam, 4, 0, 1)
    t.set_render_form("megakernel")
    with pytest.raises(mrad.InvalidArgument, match="exceeding the limit of 8192"):
        t.render_sample(np.zeros(128 * 128 * 3, np.float32), sc, cam, 4, 0, 1)
    t.close()
