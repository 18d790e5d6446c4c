"""GPU: scenes whose wire arrays hold out-of-range refs or leaves are rejected
on upload with an invalid-argument error (IllegalArgumentException), and a
rejected scene leaves a batch's arena as it was."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import mrad
from mrad import _native as N
from mrad import configs

from helpers import bad_material_variants, corrupt_view
from scenes import gi_test_scene, small_field

pytestmark = pytest.mark.gpu


def _corrupt(scene, refs=None, node_data=None):
    """A copy of the scene's wire view with some arrays replaced; the numpy
    arrays are returned so they outlive the view."""
    v = N.SceneArrays()
    C.pointer(v)[0] = scene.arrays()
    keep = []
    if refs is not None:
        keep.append(np.ascontiguousarray(refs, np.int32))
        v.primitive_refs = keep[-1].ctypes.data_as(N.ip)
    if node_data is not None:
        keep.append(np.ascontiguousarray(node_data, np.int32))
        v.node_data = keep[-1].ctypes.data_as(N.ip)
    v.revision = v.revision + 1000
    return v, keep


def _bad_variants(scene):
    refs = scene.bvh_primitive_refs().copy()
    counts = scene.counts()
    bad_ref = refs.copy()
    bad_ref[0] = (bad_ref[0] & np.int32(-0x40000000)) | np.int32(counts["spheres"] + counts["boxes"] + 5)
    nd = scene.bvh_node_data().copy()
    leaf = int(np.nonzero(nd[:, 0] < 0)[0][0])
    nd_bad = nd.copy()
    nd_bad[leaf, 2] = counts["refs"]        # first ref past the end
    return [dict(refs=bad_ref), dict(node_data=nd_bad)]


def test_tracer_rejects_out_of_range_refs_and_leaves(require_gpu):
    sc = small_field(1000, 3)
    for bad in _bad_variants(sc):
        v, keep = _corrupt(sc, **bad)
        with mrad.GpuRayTracer(8, 8, 1, True) as t:
            rc = N.lib.mrad_upload_scene(t._h, C.byref(v), C.c_uint64(99))
            assert rc == N.MRAD_ERR_INVALID_ARGUMENT, N.lib.mrad_last_error()
            assert b"out of range" in N.lib.mrad_last_error()
            # the tracer still works with a valid scene afterwards
            img = np.zeros(8 * 8 * 3, np.float32)
            t.render_sample(img, sc, configs.field_camera(), 4, 0, 1)
            assert np.isfinite(img).all()


def test_batch_arena_unchanged_after_rejected_scene(require_gpu):
    good, cam = small_field(2000, 4), configs.field_camera()
    seed = mrad.frame_seed(21, 0)
    with mrad.GpuRayTracer(40, 30, 2, True) as t:
        ref = np.zeros(40 * 30 * 3, np.float32)
        t.render_sample(ref, good, cam, 6, 0, seed)
    bad_scene = small_field(1000, 3)
    with mrad.BatchRenderer(mrad.WavefrontConfig(queue_capacity=4096)) as b:
        for bad in _bad_variants(bad_scene):
            v, keep = _corrupt(bad_scene, **bad)
            view = mrad.batch.make_view(cam, 40, 30, 2, 6, 0, seed)
            idx = C.c_int()
            rc = N.lib.mrad_batch_add_scene(b._h, C.byref(v), C.byref(view), C.byref(idx))
            assert rc == N.MRAD_ERR_INVALID_ARGUMENT
        b.add_scene(good, cam, 40, 30, 2, 6, 0, seed)
        b.render()
        np.testing.assert_array_equal(b.image(0).view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("make", [configs.room_scene, gi_test_scene])
def test_tracer_rejects_bad_material_indices(require_gpu, make):
    """mrad_upload_scene rejects a material index outside -1 / [0, materials) before any launch, and the
    tracer renders a valid scene afterwards."""
    sc = make()
    with mrad.GpuRayTracer(8, 8, 1, True) as t:
        for label, arrays in bad_material_variants(sc):
            v, keep = corrupt_view(sc, **arrays)
            rc = N.lib.mrad_upload_scene(t._h, C.byref(v), C.c_uint64(77))
            assert rc == N.MRAD_ERR_INVALID_ARGUMENT, label
            assert b"material index" in N.lib.mrad_last_error()
        img = np.zeros(8 * 8 * 3, np.float32)
        t.render_sample(img, sc, configs.room_camera(), 4, 0, 1)
        assert np.isfinite(img).all()
