"""GPU: the path-replay adjoint against the oracle adjoint (same paths, FP64
sums) and against central finite differences of the GPU forward pass with
common random numbers (h = 1e-3)."""
from __future__ import annotations

import numpy as np
import pytest

import mrad
from mrad import configs
from mrad.adjoint import PathReplayRenderer, l2_loss
from oracle import oracle as O

from scenes import small_field

pytestmark = pytest.mark.gpu


def _target(r: PathReplayRenderer, seed=3):
    rng = np.random.default_rng(seed)
    m = r.base_materials.copy()
    m[:, 0:3] *= rng.uniform(0.8, 1.2, (len(m), 3)).astype(np.float32)
    m[:, 9:12] *= rng.uniform(0.8, 1.2, (len(m), 1)).astype(np.float32)
    img = r.forward(m)
    r.forward(r.base_materials)
    return img


@pytest.mark.parametrize("case", ["room", "field", "field_rr"])
def test_gpu_adjoint_matches_oracle(require_gpu, case):
    if case == "room":
        sc, cam, w, h, spp, depth, rr = configs.room_scene(), configs.room_camera(), 48, 40, 4, 4, -1
    else:
        sc, cam, w, h, spp, depth, rr = small_field(3000), configs.field_camera(), 64, 48, 4, 4, (2 if case == "field_rr" else -1)
    r = PathReplayRenderer(sc, cam, w, h, spp, depth, frame_seed=0xA11CE, rr_start=rr)
    target = _target(r)
    img = r.forward()
    _, gimg = l2_loss(img, target)
    ga, ge = r.backward(gimg)
    oa, oe = O.adjoint(O.Scene(sc.wire()), cam.packed(), w, h, depth, 0, 0xA11CE, spp, gimg, r.primitive_count,
                       rr_start=rr)
    for got, ref in ((ga, oa), (ge, oe)):
        scale = np.abs(ref).max()
        assert scale > 0
        np.testing.assert_allclose(got, ref, rtol=1e-4, atol=1e-5 * scale)
    r.close()


def _fd_rounding_bound(img_p, img_m, delta, spp, span):
    """Largest loss-difference error FP32 rounding of the two images can cause, over the step span:
    each changed pixel is a sum of spp path radiances, each rounded in FP32."""
    changed = img_p != img_m
    ulp = np.spacing(np.maximum(np.abs(img_p), np.abs(img_m)).astype(np.float32)).astype(np.float64)
    return float(np.sum(np.abs(delta[changed]) * 4.0 * spp * ulp[changed])) / span


def test_gpu_adjoint_central_differences_diff_field(require_gpu):
    """Heterogeneous field at 10k primitives (4 bounces, no roulette) at 128x128, 8 spp: the 56 strongest
    albedo channels and 8 emitter radiances, central differences h = 1e-3 of the GPU forward pass.
    The agreement bar is 1e-3 relative plus the FP32 rounding bound of the two images (a bright pixel,
    e.g. one with an emitter sample, resolves a 1e-3 step of a dim path only to ~1e-3); most parameters
    meet 1e-3 outright."""
    w = configs.diff(10_000)
    W, H, SPP = 128, 128, 8
    r = PathReplayRenderer(w.scene, w.camera, W, H, SPP, w.max_depth, frame_seed=0xD1FF)
    target = _target(r, 9)
    base = r.base_materials
    img = r.forward(base)
    loss0, gimg = l2_loss(img, target)
    ga, ge = r.backward(gimg)
    gam, gem = r.per_material(ga), r.per_material(ge)
    # 64 parameters with a visible effect on the image: the 56 largest albedo gradients and 8 emitters
    order = np.argsort(-np.abs(gam).reshape(-1))[:56]
    picks = [("albedo", int(i // 3), int(i % 3)) for i in order]
    emitters = np.where(base[:, 9] > 0)[0]
    em_order = emitters[np.argsort(-np.abs(gem[emitters].sum(1)))][:8]
    picks += [("emission", int(m), -1) for m in em_order]
    h = 1e-3
    bad, strict = [], 0
    delta = gimg.astype(np.float64)
    for kind, m, c in picks:
        plus, minus = base.copy(), base.copy()
        if kind == "albedo":
            plus[m, c] += np.float32(h); minus[m, c] -= np.float32(h)
            analytic = gam[m, c]
            span = np.float64(plus[m, c]) - np.float64(minus[m, c])
        else:   # grey radiance: all three channels scaled together
            plus[m, 9:12] += np.float32(h); minus[m, 9:12] -= np.float32(h)
            analytic = gem[m].sum()
            span = np.float64(plus[m, 9]) - np.float64(minus[m, 9])
        ip, im = r.forward(plus), r.forward(minus)
        fd = (l2_loss(ip, target)[0] - l2_loss(im, target)[0]) / span
        err = abs(fd - analytic)
        strict += err <= 1e-3 * abs(fd)
        if err > 1e-3 * abs(fd) + _fd_rounding_bound(ip, im, delta, SPP, span):
            bad.append((kind, m, c, fd, analytic))
    r.forward(base)
    r.close()
    assert not bad, bad
    assert strict >= len(picks) // 2, strict


def test_gpu_adjoint_matches_oracle_per_parameter_diff_field(require_gpu):
    """Every parameter with a non-negligible gradient agrees with the FP64-summed oracle to 1e-4."""
    w = configs.diff(10_000)
    W, H, SPP = 96, 96, 4
    r = PathReplayRenderer(w.scene, w.camera, W, H, SPP, w.max_depth, frame_seed=0xD1FF)
    target = _target(r, 9)
    img = r.forward()
    _, gimg = l2_loss(img, target)
    ga, ge = r.backward(gimg)
    oa, oe = O.adjoint(O.Scene(w.scene.wire()), w.camera.packed(), W, H, w.max_depth, 0, 0xD1FF, SPP, gimg,
                       r.primitive_count)
    for got, ref in ((ga, oa), (ge, oe)):
        big = np.abs(ref) > 1e-3 * np.abs(ref).max()
        assert big.sum() > 100
        np.testing.assert_allclose(got[big], ref[big], rtol=1e-4)
    r.close()


def test_torch_function_backward(require_gpu):
    torch = pytest.importorskip("torch")
    from mrad.adjoint import PathReplayFunction
    sc, cam = configs.room_scene(), configs.room_camera()
    r = PathReplayRenderer(sc, cam, 32, 32, 2, 4, frame_seed=5)
    alb = torch.tensor(r.base_materials[:, 0:3], requires_grad=True)
    emi = torch.tensor(r.base_materials[:, 9:12], requires_grad=True)
    img = PathReplayFunction.apply(alb, emi, r)
    loss = (img ** 2).mean()
    loss.backward()
    _, g = l2_loss(img.detach().numpy().reshape(-1), np.zeros(32 * 32 * 3, np.float32))
    ga, ge = r.backward(g)
    np.testing.assert_allclose(alb.grad.numpy(), r.per_material(ga), rtol=1e-5, atol=1e-9)
    np.testing.assert_allclose(emi.grad.numpy(), r.per_material(ge), rtol=1e-5, atol=1e-9)
    r.close()
