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


def test_gpu_adjoint_central_differences_diff_field(require_gpu):
    """Heterogeneous field at 10k primitives (4 bounces, no roulette) at 128x128, 8 spp: 32 sampled
    albedo channels and the radiance of emitters, central differences h = 1e-3."""
    w = configs.diff(10_000)
    W, H, SPP = 128, 128, 8
    r = PathReplayRenderer(w.scene, w.camera, W, H, SPP, w.max_depth, frame_seed=0xD1FF)
    target = _target(r, 9)
    base = r.base_materials
    img = r.forward(base)
    loss0, gimg = l2_loss(img, target)
    ga, ge = r.backward(gimg)
    gam, gem = r.per_material(ga), r.per_material(ge)
    rng = np.random.default_rng(4)
    # parameters with a visible effect: the largest albedo gradients and the emitters
    order = np.argsort(-np.abs(gam).reshape(-1))[:200]
    picks = [("albedo", int(i // 3), int(i % 3)) for i in rng.choice(order, 24, replace=False)]
    emitters = np.where(base[:, 9] > 0)[0]
    em_order = emitters[np.argsort(-np.abs(gem[emitters].sum(1)))][:8]
    picks += [("emission", int(m), -1) for m in em_order]
    h = 1e-3
    bad = []
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
        fd = (l2_loss(r.forward(plus), target)[0] - l2_loss(r.forward(minus), target)[0]) / span
        if abs(fd - analytic) > 1e-3 * max(abs(fd), abs(analytic)) + 1e-12:
            bad.append((kind, m, c, fd, analytic))
    r.forward(base)
    r.close()
    assert not bad, bad


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
