"""GPU: the path-replay adjoint against the oracle adjoint (same paths, FP64
sums) and against central finite differences (h = 1e-3) with common random
numbers on the differentiable configuration."""
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


def test_gpu_adjoint_central_differences_diff_config(require_gpu):
    """The differentiable configuration itself: the 10k heterogeneous field at 512x512, 32 spp, 4 bounces, loss
    the pixel L2 to the frame under seed-perturbed albedos and emissions (each value scaled by U[0.8, 1.2]).
    The GPU path-replay gradients of 64 parameters drawn at random (seeded) among those with a non-zero
    gradient, 56 albedo channels and 8 emitter radiances, are within 1e-3 of central differences (h = 1e-3,
    common random numbers: the frame's own paths, re-evaluated in double as an exact polynomial in the
    parameters), and every gradient of weight equals the FP64-summed oracle adjoint of the same paths to 1e-4."""
    from fd_paths import PathFD, draw_parameters, per_material
    w = configs.diff(10_000)
    seed = 0xD1FF
    osc = O.Scene(w.scene.wire())
    mats = w.scene.material_data().astype(np.float64)
    print("recording the diff-10k frame's paths (512x512, 32 spp)", flush=True)
    rec = O.RecordedPaths(osc, w.camera.packed(), w.width, w.height, w.max_depth, 0, seed, w.spp)
    rng = np.random.default_rng(9)
    tm = mats.copy()
    tm[:, 0:3] *= rng.uniform(0.8, 1.2, (len(mats), 3))
    tm[:, 9:12] *= rng.uniform(0.8, 1.2, (len(mats), 1))
    fd = PathFD(rec, mats, rec.image(tm))
    gimg = fd.grad_image()
    r = PathReplayRenderer(w.scene, w.camera, w.width, w.height, w.spp, w.max_depth, frame_seed=seed)
    print("GPU backward", flush=True)
    ga, ge = r.backward(gimg)
    prim_mat, n_prim = r.prim_material, r.primitive_count
    r.close()
    oa, oe = O.adjoint(osc, w.camera.packed(), w.width, w.height, w.max_depth, 0, seed, w.spp, gimg, n_prim)
    for got, ref in ((ga, oa), (ge, oe)):
        big = np.abs(ref) > 1e-3 * np.abs(ref).max()
        assert big.sum() > 1000
        np.testing.assert_allclose(got[big], ref[big], rtol=1e-4)
    gam, gem = per_material(ga, prim_mat, len(mats)), per_material(ge, prim_mat, len(mats))
    picks = draw_parameters(np.random.default_rng(64), gam, gem, np.flatnonzero(mats[:, 9] > 0))
    assert len(picks) == 64
    print("central differences of 64 parameters", flush=True)
    bad = []
    for kind, m, c in picks:
        f = fd.central_difference(kind, m, c, 1e-3)
        g = gam[m, c] if kind == "albedo" else gem[m].sum()
        if not abs(f - g) <= 1e-3 * abs(f):
            bad.append((kind, m, c, f, g))
    assert not bad, bad


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
