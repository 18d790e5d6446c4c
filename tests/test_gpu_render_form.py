"""GPU: mrad_render_sample in the wavefront form (mrad_set_render_form) against the megakernel and the oracle.

Every frame (image bits and, with metrics, all 14 counters) equals the megakernel's for the same view. The frames
are rendered in a row with new frame indices and seeds, which reuses the form's graph. A new resident scene makes a
new graph. The full-size het-100k frame is compared pixel for pixel and timed in both forms."""
from __future__ import annotations

import numpy as np
import pytest

import mrad
from mrad import configs
from oracle import oracle as O

from scenes import gi_camera, gi_test_scene, small_field

pytestmark = pytest.mark.gpu

CASES = [
    # scene, camera, width, height, spp, depth, roulette
    (configs.room_scene, configs.room_camera, 64, 64, 4, 4, -1),
    (lambda: gi_test_scene(), gi_camera, 64, 48, 2, 12, -1),
    (lambda: small_field(3000), configs.field_camera, 96, 54, 2, 8, 3),
    (lambda: small_field(4000, 5, clustered=True), configs.field_camera, 80, 45, 3, 6, -1),
]


def _frame(t, scene, cam, depth, frame, seed, w, h):
    img = np.zeros(w * h * 3, np.float32)
    t.render_sample(img, scene, cam, depth, frame, seed)
    st = t.last_frame_stats()
    return img, np.array([st[k] for k in mrad.STAT_NAMES], np.uint64), st


def test_wavefront_form_equals_megakernel_and_oracle(require_gpu):
    w0 = None
    for make, make_cam, w, h, spp, depth, rr in CASES:
        scene, cam = make(), make_cam()
        mk = mrad.GpuRayTracer(w, h, spp, True, collect_metrics=True)
        if w0 is None or w0[0] != (w, h, spp):
            if w0 is not None:
                w0[1].close()
            wf = mrad.GpuRayTracer(w, h, spp, True, collect_metrics=True)
            wf.set_render_form("wavefront")
            w0 = ((w, h, spp), wf)
        wf = w0[1]
        for t in (mk, wf):
            t.set_russian_roulette(rr)
        for frame in range(3):
            seed = mrad.frame_seed(0xF0F0 + frame, frame)
            print(f"{w}x{h} {spp} spp depth {depth} frame {frame}", flush=True)
            a, ca, _ = _frame(mk, scene, cam, depth, frame, seed, w, h)
            b, cb, st = _frame(wf, scene, cam, depth, frame, seed, w, h)
            np.testing.assert_array_equal(b.view(np.uint32), a.view(np.uint32))
            np.testing.assert_array_equal(cb, ca)
            assert st["physical_launches"] >= 2 and 0 < st["maximum_physical_kernel_nanos"] <= st["kernel_nanos"]
            if frame == 0:
                ref, ref_stats = O.render_pixels(O.Scene(scene.wire()), cam.packed(), w, h, depth, frame, seed, spp,
                                                 rr_start=rr)
                np.testing.assert_array_equal(b.view(np.uint32), ref.reshape(-1).view(np.uint32))
                np.testing.assert_array_equal(cb, ref_stats)
        mk.close()
    w0[1].close()


def test_wavefront_form_full_size_het100k(require_gpu):
    w = configs.het(100_000, spp=16)
    seed = mrad.frame_seed(7, 0)
    out, ms = {}, {}
    for form in ("megakernel", "wavefront"):
        with mrad.GpuRayTracer(w.width, w.height, w.spp, True) as t:
            t.set_russian_roulette(w.rr_start)
            t.set_render_form(form)
            img = np.zeros(w.width * w.height * 3, np.float32)
            t.render_sample(img, w.scene, w.camera, w.max_depth, 0, seed)        # first frame builds the graph
            t.render_sample(img, w.scene, w.camera, w.max_depth, 0, seed)
            out[form], ms[form] = img, t.last_frame_stats()["kernel_nanos"] / 1e6
    print(f"het-100k 1080p 16 spp: megakernel {ms['megakernel']:.1f} ms, wavefront form {ms['wavefront']:.1f} ms",
          flush=True)
    np.testing.assert_array_equal(out["wavefront"].view(np.uint32), out["megakernel"].view(np.uint32))
