"""GPU: the batched wavefront form renders every scene bit-identically to the
megakernel (and so to the oracle), with the same frame counters."""
from __future__ import annotations

import numpy as np
import pytest

import mrad
from mrad import configs

from scenes import gi_camera, gi_test_scene, small_field

pytestmark = pytest.mark.gpu


def _mega(scene, cam, w, h, spp, depth, frame, seed, rr=-1, metrics=False, linear=False):
    with mrad.GpuRayTracer(w, h, spp, not linear, collect_metrics=metrics) as t:
        if rr >= 0:
            t.set_russian_roulette(rr)
        img = np.zeros(w * h * 3, np.float32)
        t.render_sample(img, scene, cam, depth, frame, seed)
        return img, t.last_frame_stats()


CASES = [
    # scene factory, camera, w, h, spp, depth, rr, frame
    (configs.room_scene, configs.room_camera, 64, 48, 5, 4, -1, 0),
    (gi_test_scene, gi_camera, 40, 40, 3, 12, -1, 2),
    (lambda: small_field(3000), configs.field_camera, 96, 54, 4, 8, 3, 1),
    (lambda: small_field(4000, 5, clustered=True), configs.field_camera, 80, 45, 2, 6, -1, 0),
]


@pytest.mark.parametrize("use_graph", [True, False])
@pytest.mark.parametrize("capacity", [1 << 20, 4096])
def test_batch_equals_megakernel(require_gpu, use_graph, capacity):
    cfg = mrad.WavefrontConfig(queue_capacity=capacity, use_graph=use_graph, waves_per_cu=8)
    refs = []
    with mrad.BatchRenderer(cfg) as b:
        for k, (mk_scene, mk_cam, w, h, spp, depth, rr, frame) in enumerate(CASES):
            sc, cam = mk_scene(), mk_cam()
            seed = mrad.frame_seed(0xBA7C4 + k, frame)
            b.add_scene(sc, cam, w, h, spp, depth, frame, seed, rr)
            refs.append(_mega(sc, cam, w, h, spp, depth, frame, seed, rr)[0])
        print(f"batch capacity={capacity} graph={use_graph}", flush=True)
        st = b.render()
        assert st["steps"] >= 2
        for k, ref in enumerate(refs):
            np.testing.assert_array_equal(b.image(k).view(np.uint32), ref.view(np.uint32))


def test_batch_counters_equal_megakernel(require_gpu):
    sc, cam = small_field(3000), configs.field_camera()
    seed = mrad.frame_seed(5, 0)
    ref, ref_stats = _mega(sc, cam, 64, 36, 4, 8, 0, seed, rr=3, metrics=True)
    with mrad.BatchRenderer(mrad.WavefrontConfig(queue_capacity=2048, collect_metrics=True)) as b:
        b.add_scene(sc, cam, 64, 36, 4, 8, 0, seed, 3)
        st = b.render()
        np.testing.assert_array_equal(b.image(0).view(np.uint32), ref.view(np.uint32))
    for k in mrad.STAT_NAMES:
        assert st[k] == ref_stats[k], k


@pytest.mark.parametrize("regroup,lds", [(True, 10), (False, 4)])
def test_regroup_and_lds_split_bit_exact_with_counters(require_gpu, regroup, lds):
    """Primitive-class regrouping of leaf tests and the 4-entry LDS split change neither images nor counters."""
    cases = [(lambda: small_field(4000, 5, clustered=True), configs.field_camera, 80, 45, 3, 6, -1),
             (gi_test_scene, gi_camera, 40, 40, 2, 12, -1),
             (lambda: small_field(3000), configs.field_camera, 64, 36, 4, 8, 3)]
    for k, (mk_scene, mk_cam, w, h, spp, depth, rr) in enumerate(cases):
        sc, cam = mk_scene(), mk_cam()
        seed = mrad.frame_seed(77 + k, 0)
        ref, ref_stats = _mega(sc, cam, w, h, spp, depth, 0, seed, rr, metrics=True)
        for metrics in (False, True):
            cfg = mrad.WavefrontConfig(queue_capacity=5000, leaf_regroup=regroup, lds_stack_entries=lds,
                                       collect_metrics=metrics)
            with mrad.BatchRenderer(cfg) as b:
                b.add_scene(sc, cam, w, h, spp, depth, 0, seed, rr)
                st = b.render()
                np.testing.assert_array_equal(b.image(0).view(np.uint32), ref.view(np.uint32))
            if metrics:
                for name in mrad.STAT_NAMES:
                    assert st[name] == ref_stats[name], (k, name)


def test_configure_keeps_scenes_and_images(require_gpu):
    from mrad.autotune import tune
    sc, cam = small_field(3000), configs.field_camera()
    with mrad.BatchRenderer(mrad.WavefrontConfig(queue_capacity=2048)) as b:
        b.add_scene(sc, cam, 48, 27, 3, 6, 0, 11, 3)
        b.render()
        first = b.image(0)
        best, table = tune(b, grid={"queue_capacity": [1024, 8192], "waves_per_cu": [4, 8],
                                    "lds_stack_entries": [4, 10]})
        assert len(table) >= 4
        b.render()
        np.testing.assert_array_equal(first, b.image(0))


def test_batch_linear_scan_and_zero_depth(require_gpu):
    sc, cam = configs.room_scene(), configs.room_camera()
    with mrad.BatchRenderer(mrad.WavefrontConfig(queue_capacity=1000, linear_scan=True)) as b:
        b.add_scene(sc, cam, 32, 32, 2, 4, 0, 77)
        b.add_scene(sc, cam, 16, 8, 3, 0, 0, 78)
        b.render()
        ref = _mega(sc, cam, 32, 32, 2, 4, 0, 77, linear=True)[0]
        np.testing.assert_array_equal(b.image(0).view(np.uint32), ref.view(np.uint32))
        assert np.all(b.image(1) == 0)


def test_batch_rerender_is_deterministic(require_gpu):
    sc, cam = small_field(2000), configs.field_camera()
    with mrad.BatchRenderer(mrad.WavefrontConfig(queue_capacity=3000)) as b:
        b.add_scene(sc, cam, 48, 27, 3, 6, 0, 9, 3)
        b.render()
        first = b.image(0)
        b.render()
        np.testing.assert_array_equal(first, b.image(0))


@pytest.mark.parametrize("chunk,capacity", [(4096, 2048), (2999, 1 << 20)])
def test_batch_small_result_chunks_equal_megakernel(require_gpu, chunk, capacity):
    """Result chunks of a few thousand paths (dozens of chunks of whole pixels, the two result buffers taking
    turns, generation held back until a buffer is resolved) give the megakernel's images bit for bit, also with
    a zero-depth scene in the batch and chunk sizes that are no multiple of the sample counts."""
    cfg = mrad.WavefrontConfig(queue_capacity=capacity, result_chunk_paths=chunk, waves_per_cu=8)
    refs = []
    with mrad.BatchRenderer(cfg) as b:
        for k, (mk_scene, mk_cam, w, h, spp, depth, rr, frame) in enumerate(CASES):
            sc, cam = mk_scene(), mk_cam()
            seed = mrad.frame_seed(0xC4A + k, frame)
            b.add_scene(sc, cam, w, h, spp, depth, frame, seed, rr)
            refs.append(_mega(sc, cam, w, h, spp, depth, frame, seed, rr)[0])
        sc, cam = configs.room_scene(), configs.room_camera()
        b.add_scene(sc, cam, 20, 10, 3, 0, 0, 5)
        print(f"batch result chunks {chunk}, pool {capacity}", flush=True)
        st = b.render()
        assert st["paths"] == sum(c[2] * c[3] * c[4] for c in CASES) + 600
        for k, ref in enumerate(refs):
            np.testing.assert_array_equal(b.image(k).view(np.uint32), ref.view(np.uint32))
        assert np.all(b.image(len(refs)) == 0)
        # a different chunk size on the same batch re-plans and renders the same images
        b.configure(mrad.WavefrontConfig(queue_capacity=capacity, result_chunk_paths=0, waves_per_cu=8))
        b.render()
        for k, ref in enumerate(refs):
            np.testing.assert_array_equal(b.image(k).view(np.uint32), ref.view(np.uint32))


def test_batch_of_more_than_2_32_paths(require_gpu):
    """Two 1024x1024 views of 2100 samples each (4.4e9 paths, more than 2^32) render in one batch, each image
    equal to the megakernel's frame bit for bit."""
    sc, cam = small_field(1000, 9), configs.field_camera()
    w, h, spp, depth = 1024, 1024, 2100, 2
    seeds = [mrad.frame_seed(0xB16, k) for k in range(2)]
    with mrad.BatchRenderer(mrad.WavefrontConfig(queue_capacity=1 << 24)) as b:
        for k in range(2):
            b.add_scene(sc, cam, w, h, spp, depth, k, seeds[k])
        print(f"batch of {2 * w * h * spp} paths", flush=True)
        st = b.render()
        assert st["paths"] == 2 * w * h * spp > 1 << 32
        for k in range(2):
            ref = _mega(sc, cam, w, h, spp, depth, k, seeds[k])[0]
            np.testing.assert_array_equal(b.image(k).view(np.uint32), ref.view(np.uint32))


PASS_CASES = [(lambda: small_field(4000, 5, clustered=True), configs.field_camera, 80, 45, 3, 6, -1),
              (gi_test_scene, gi_camera, 40, 40, 2, 12, -1),
              (lambda: small_field(3000), configs.field_camera, 64, 36, 4, 8, 3)]


@pytest.mark.parametrize("passes,pops,cap,dense", [(2, 4, 0, -1), (3, 2, 0, -1), (3, 8, 64, -1), (2, 1, 128, -1),
                                                  (2, 128, 0, -1), (2, 4, 0, 0), (3, 2, 64, 8), (2, 16, 0, 64)])
def test_capped_trace_passes_equal_single_pass(require_gpu, passes, pops, cap, dense):
    """Closest-hit queries suspended after `pops` pops and resumed by the next
    pass give the single pass's images and all 14 counters; tiny budgets
    suspend most queries several times, and a 64- or 128-item list overflows
    (those queries finish in their own pass).  With `dense` -1 every wave is
    suspended at its budget; otherwise a wave with that many lanes still
    traversing keeps going and is checked again (0: the library's threshold)."""
    for k, (mk_scene, mk_cam, w, h, spp, depth, rr) in enumerate(PASS_CASES):
        sc, cam = mk_scene(), mk_cam()
        seed = mrad.frame_seed(91 + k, 0)
        out = {}
        for name, extra in (("single", dict(trace_passes=1)),
                            ("passes", dict(trace_passes=passes, pass_pops=pops, suspend_capacity=cap,
                                            pass_dense_lanes=dense))):
            cfg = mrad.WavefrontConfig(queue_capacity=5056, collect_metrics=True, **extra)
            with mrad.BatchRenderer(cfg) as b:
                b.add_scene(sc, cam, w, h, spp, depth, 0, seed, rr)
                st = b.render()
                out[name] = (b.image(0).copy(), st)
        print(f"case {k} passes={passes} pops={pops} cap={cap} dense={dense}: rays {out['single'][1]['rays']}",
              flush=True)
        np.testing.assert_array_equal(out["passes"][0].view(np.uint32), out["single"][0].view(np.uint32))
        for n in mrad.STAT_NAMES:
            assert out["passes"][1][n] == out["single"][1][n], n


@pytest.mark.parametrize("passes", [1, 2])
def test_static_super_chunks_equal_megakernel(require_gpu, passes):
    """A grid of one wave per CU (64 blocks) and a 256x256, 8-spp frame (524,288 paths in a 2^20 pool) give the
    trace kernel whole static rounds of 2048-path super-chunks (60 % of the pool, interleaved over the blocks)
    and a global tail; images and all 14 counters equal the megakernel's, in one pass and in two."""
    sc, cam = small_field(3000), configs.field_camera()
    seed = mrad.frame_seed(0x5C, 0)
    ref, ref_stats = _mega(sc, cam, 256, 256, 8, 6, 0, seed, rr=3, metrics=True)
    for metrics in (False, True):
        cfg = mrad.WavefrontConfig(queue_capacity=1 << 20, waves_per_cu=1, collect_metrics=metrics,
                                   trace_passes=passes, pass_pops=24)
        with mrad.BatchRenderer(cfg) as b:
            b.add_scene(sc, cam, 256, 256, 8, 6, 0, seed, 3)
            st = b.render()
            np.testing.assert_array_equal(b.image(0).view(np.uint32), ref.view(np.uint32))
        if metrics:
            for k in mrad.STAT_NAMES:
                assert st[k] == ref_stats[k], k


def test_configure_resolves_pass_defaults_for_the_new_queue(require_gpu):
    """mrad_batch_configure resolves the zero fields for the new launch parameters: a 2^22-path batch traces in
    one pass, and reconfigured to a 2^25-path queue in two; the image is the same either way."""
    sc, cam = small_field(3000), configs.field_camera()
    seed = mrad.frame_seed(0x77, 0)
    with mrad.BatchRenderer(mrad.WavefrontConfig(queue_capacity=1 << 22)) as b:
        b.add_scene(sc, cam, 96, 64, 4, 6, 0, seed, 3)
        assert b.effective_config()["trace_passes"] == 1
        b.render()
        one = b.image(0).copy()
        b.configure(mrad.WavefrontConfig(queue_capacity=1 << 25))
        c = b.effective_config()
        assert (c["trace_passes"], c["queue_capacity"], c["suspend_capacity"]) == (2, 1 << 25, 1 << 24)
        b.render()
        np.testing.assert_array_equal(b.image(0).view(np.uint32), one.view(np.uint32))
