"""GPU: mrad_trace_replay_noalloc, the replay entry point that allocates no
device memory (its staging lives in the library's static scratch array),
returns exactly what mrad_trace_replay returns: two calls in a row, the second
with fewer rays, with and without per-ray counters."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import mrad
from mrad import _native as N
from mrad import configs

from helpers import camera_grid_rays, random_rays
from scenes import small_field

pytestmark = pytest.mark.gpu

MAX_RAYS = 65536


def _replay(entry, t, rays, collect, per_ray):
    n = len(rays)
    out = dict(flags=np.zeros(n, np.int32), dist=np.zeros(n, np.float32), order=np.zeros(n, np.int32),
               mat=np.zeros(n, np.int32), normals=np.zeros(3 * n, np.float32),
               per=np.zeros((n, N.STAT_COUNT), np.uint32) if per_ray else None)
    r = np.ascontiguousarray(rays, np.float32).reshape(-1)
    rc = entry(t._handle(), N.ptr(r, N.fp), n, 1 if collect else 0, N.ptr(out["flags"], N.ip),
               N.ptr(out["dist"], N.fp), N.ptr(out["order"], N.ip), N.ptr(out["mat"], N.ip),
               N.ptr(out["normals"], N.fp), N.ptr(out["per"], N.up))
    assert rc == N.MRAD_OK, N.lib.mrad_last_error()
    out["stats"] = t.last_frame_stats()
    return out


def _assert_same(a, b):
    for k in ("flags", "order", "mat"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    for k in ("dist", "normals"):
        np.testing.assert_array_equal(a[k].view(np.uint32), b[k].view(np.uint32), err_msg=k)
    if a["per"] is not None:
        np.testing.assert_array_equal(a["per"], b["per"])
    for k in mrad.STAT_NAMES:
        assert a["stats"][k] == b["stats"][k], k


@pytest.mark.parametrize("collect,per_ray", [(True, True), (False, False)])
def test_noalloc_replay_equals_replay_twice(require_gpu, collect, per_ray):
    scene, cam = small_field(3000), configs.field_camera()
    rng = np.random.default_rng(31)
    big = np.concatenate([camera_grid_rays(cam.packed(), 160, 90), random_rays(rng, MAX_RAYS - 160 * 90)])
    small = random_rays(np.random.default_rng(32), 777)
    assert len(big) == MAX_RAYS
    print(f"noalloc replay collect={collect} per_ray={per_ray}: {len(big)} then {len(small)} rays", flush=True)
    with mrad.GpuRayTracer(8, 8, 1, True, collect_metrics=collect) as t:
        t.ensure_scene_resident(scene)
        for rays in (big, small):
            got = _replay(N.lib.mrad_trace_replay_noalloc, t, rays, collect, per_ray)
            ref = _replay(N.lib.mrad_trace_replay, t, rays, collect, per_ray)
            _assert_same(got, ref)
            assert int(got["flags"].sum()) > 0


def test_noalloc_replay_rejects_more_rays_than_its_scratch(require_gpu):
    scene = small_field(500)
    with mrad.GpuRayTracer(8, 8, 1, True) as t:
        t.ensure_scene_resident(scene)
        n = MAX_RAYS + 1
        rays = np.zeros(6 * n, np.float32)
        outs = [np.zeros(n, np.int32), np.zeros(n, np.float32), np.zeros(n, np.int32), np.zeros(n, np.int32),
                np.zeros(3 * n, np.float32)]
        rc = N.lib.mrad_trace_replay_noalloc(t._handle(), N.ptr(rays, N.fp), n, 0, N.ptr(outs[0], N.ip),
                                             N.ptr(outs[1], N.fp), N.ptr(outs[2], N.ip), N.ptr(outs[3], N.ip),
                                             N.ptr(outs[4], N.fp), None)
        assert rc == N.MRAD_ERR_INVALID_ARGUMENT
        assert b"at most 65536 rays" in N.lib.mrad_last_error()
