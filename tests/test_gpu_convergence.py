"""GPU: the megakernel's image of the heterogeneous 100k field (1920x1080, 8
bounces, roulette after bounce 3) converges to the reference CPU renderer's
FP64 image (oracle_fp64.c, the semantics of graphics3d/raytracers/
PathTracer.java:37-71 with its own random numbers) within the Monte-Carlo
confidence band: on a 48x48 crop at the frame's centre, where the field is
densest, at least 99 % of pixel channels satisfy
|I - R| <= 3.29 sqrt(s^2 (1/spp + 1/spp_ref)) + 4 spp 2^-24 |R|
(s^2 the FP64 run's per-pixel sample variance; the second term is the FP32
rounding of a spp-sample float sum), and relMSE = mean (I - R)^2 / (R^2 + 1e-2)
is within 20 % of its expectation mean(s^2 (1/spp + 1/spp_ref) / (R^2 + 1e-2)).

At 1024 GPU samples the 1 % emitters (radiance 5-20) still make the pixel
estimates heavier-tailed than a normal law (a pixel that catches one rare
emitter path lands outside 3.29 sigma), so the band holds at ~99.3 %, not
99.9 %; at 64 samples it is ~98.4 % with relMSE within 6 % of expectation,
the signature of noise rather than bias."""
from __future__ import annotations

import numpy as np
import pytest

import mrad
from mrad import configs
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_het100k_converges_to_fp64_reference_within_band(require_gpu):
    w = configs.het(100_000)
    spp, ref_spp = 1024, 8192
    x0, y0, cw, ch = 936, 516, 48, 48
    seed = mrad.frame_seed(7, 0)
    print(f"het-100k 1920x1080 {spp} spp on the GPU", flush=True)
    with mrad.GpuRayTracer(w.width, w.height, spp, True) as t:
        t.set_russian_roulette(w.rr_start)
        img = np.zeros(w.width * w.height * 3, np.float32)
        t.render_sample(img, w.scene, w.camera, w.max_depth, 0, seed)
    crop = img.reshape(w.height, w.width, 3)[y0:y0 + ch, x0:x0 + cw].reshape(-1, 3).astype(np.float64)
    print(f"FP64 reference, {cw}x{ch} crop at {ref_spp} spp", flush=True)
    ref, var = O.render64_window(O.Scene(w.scene.wire()), w.camera.packed(), w.width, w.height, (x0, y0, cw, ch),
                                 w.max_depth, ref_spp, 0x5EED, rr_start=w.rr_start)
    v = np.maximum(var, 0.0) * (1.0 / spp + 1.0 / ref_spp)
    band = np.abs(crop - ref) <= 3.29 * np.sqrt(v) + 4.0 * spp * 2.0 ** -24 * np.abs(ref)
    den = ref ** 2 + 1e-2
    relmse, expected = float(np.mean((crop - ref) ** 2 / den)), float(np.mean(v / den))
    print(f"band {band.mean():.4f}, relMSE {relmse:.5f} (expected {expected:.5f})", flush=True)
    assert band.mean() >= 0.99
    assert 0.8 <= relmse / expected <= 1.2
