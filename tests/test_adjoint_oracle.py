"""CPU: the oracle's path-replay adjoint against central finite differences of
its own forward renderer (common random numbers), and closed forms."""
from __future__ import annotations

import numpy as np
import pytest

import mrad
from mrad import configs
from mrad.adjoint import l2_loss, primitive_material_indices
from oracle import oracle as O

W, H, SPP, DEPTH = 16, 16, 4, 4
SEED = 0xADC0DE


def _render(wire, cam, materials, rr=-1):
    w = dict(wire)
    w["materials"] = materials
    img, _ = O.render_pixels(O.Scene(w), cam, W, H, DEPTH, 0, SEED, SPP, rr_start=rr)
    return img.reshape(-1)


def _per_material(grad_prim, prim_mat, n_mat):
    out = np.zeros((n_mat, 3))
    np.add.at(out, prim_mat, grad_prim)
    return out


@pytest.mark.parametrize("rr", [-1, 2])
def test_oracle_adjoint_matches_central_differences_room(rr):
    w = configs.room()
    wire, cam = w.scene.wire(), w.camera.packed()
    mats = wire["materials"].copy()
    rng = np.random.default_rng(1)
    target_mats = mats.copy()
    target_mats[:, 0:3] *= rng.uniform(0.8, 1.2, (len(mats), 3)).astype(np.float32)
    target_mats[:, 9:12] *= np.float32(1.1)
    target = _render(wire, cam, target_mats, rr)
    img = _render(wire, cam, mats, rr)
    loss, gimg = l2_loss(img, target)
    prim_mat = primitive_material_indices(w.scene)
    ga, ge = O.adjoint(O.Scene(wire), cam, W, H, DEPTH, 0, SEED, SPP, gimg, len(prim_mat), rr_start=rr)
    gam, gem = _per_material(ga, prim_mat, len(mats)), _per_material(ge, prim_mat, len(mats))
    h = 1e-3
    checked = 0
    for m in range(len(mats)):
        for c in range(3):
            for slot, analytic in ((c, gam[m, c]), (9 + c, gem[m, c])):
                if slot >= 9 and mats[m, slot] == 0 and m != 5:
                    continue
                if slot < 9 and not np.any(mats[m, 0:3] > 0):
                    continue   # a black surface absorbs: albedo 0 is the edge of the lobe-selection domain
                plus, minus = mats.copy(), mats.copy()
                plus[m, slot] += np.float32(h)
                minus[m, slot] -= np.float32(h)
                hp = np.float64(plus[m, slot]) - np.float64(mats[m, slot])
                hm = np.float64(mats[m, slot]) - np.float64(minus[m, slot])
                fd = (l2_loss(_render(wire, cam, plus, rr), target)[0] -
                      l2_loss(_render(wire, cam, minus, rr), target)[0]) / (hp + hm)
                scale = max(abs(fd), abs(analytic), 1e-7)
                if rr < 0:
                    assert abs(fd - analytic) <= 1e-3 * scale + 1e-9, (m, slot, fd, analytic)
                checked += 1
    assert checked >= 21


def test_recorded_paths_reproduce_the_fixed_path_renderer():
    """The recorded vertices of a frame (or_record_paths), re-evaluated in numpy, give the FP64 fixed-path
    frame of or_render_fixed_paths64 exactly, for the scene's materials and for perturbed ones."""
    w = configs.diff(10_000)
    wire, cam = w.scene.wire(), w.camera.packed()
    osc = O.Scene(wire)
    mats = wire["materials"].astype(np.float64)
    rec = O.RecordedPaths(osc, cam, 40, 30, w.max_depth, 0, SEED, 3)
    pert = mats * np.random.default_rng(4).uniform(0.8, 1.2, mats.shape)
    for m in (mats, pert):
        ref = O.render_fixed_paths64(osc, cam, 40, 30, w.max_depth, 0, SEED, 3, m)
        np.testing.assert_array_equal(rec.image(m), ref)
    # the room: boxes, spheres, an affine box, absorbing black surfaces and an emitter with zero albedo
    r = configs.room()
    osc = O.Scene(r.scene.wire())
    rm = r.scene.wire()["materials"].astype(np.float64)
    rec = O.RecordedPaths(osc, r.camera.packed(), 24, 24, r.max_depth, 0, SEED, 4)
    assert (rec.end == 2).any() and (rec.end == 0).any()
    np.testing.assert_array_equal(rec.image(rm), O.render_fixed_paths64(osc, r.camera.packed(), 24, 24, r.max_depth,
                                                                        0, SEED, 4, rm))


def test_adjoint_central_differences_diff_field_fixed_paths():
    """Heterogeneous 10k field, 4 bounces: the oracle adjoint of the frame's paths against central differences
    (h = 1e-3) of the same FP32 paths re-evaluated in double, on 64 parameters drawn at random among those with
    a non-zero gradient (56 albedo channels and 8 emitter radiances), all within 1e-3 relative."""
    from fd_paths import PathFD, draw_parameters, per_material
    w = configs.diff(10_000)
    wire, cam = w.scene.wire(), w.camera.packed()
    osc = O.Scene(wire)
    mats = wire["materials"].astype(np.float64)
    rec = O.RecordedPaths(osc, cam, 96, 96, w.max_depth, 0, SEED, 4)
    rng = np.random.default_rng(9)
    tm = mats.copy()
    tm[:, 0:3] *= rng.uniform(0.8, 1.2, (len(mats), 3))
    tm[:, 9:12] *= rng.uniform(0.8, 1.2, (len(mats), 1))
    fd = PathFD(rec, mats, rec.image(tm))
    prim_mat = primitive_material_indices(w.scene)
    ga, ge = O.adjoint(osc, cam, 96, 96, w.max_depth, 0, SEED, 4, fd.grad_image(), len(prim_mat))
    gam, gem = per_material(ga, prim_mat, len(mats)), per_material(ge, prim_mat, len(mats))
    picks = draw_parameters(np.random.default_rng(64), gam, gem, np.flatnonzero(mats[:, 9] > 0))
    assert len(picks) == 64 and sum(k == "emission" for k, _, _ in picks) == 8
    for kind, m, c in picks:
        f = fd.central_difference(kind, m, c, 1e-3)
        g = gam[m, c] if kind == "albedo" else gem[m].sum()
        assert abs(f - g) <= 1e-3 * abs(f), (kind, m, c, f, g)


def test_emission_gradient_is_linear_response():
    """Images are linear in emission: dLoss/dscale of all emitters == sum(Le * dLoss/dLe)."""
    w = configs.room()
    wire, cam = w.scene.wire(), w.camera.packed()
    mats = wire["materials"]
    img = _render(wire, cam, mats)
    target = img * np.float32(0.9)
    _, gimg = l2_loss(img, target)
    prim_mat = primitive_material_indices(w.scene)
    _, ge = O.adjoint(O.Scene(wire), cam, W, H, DEPTH, 0, SEED, SPP, gimg, len(prim_mat))
    le = mats[prim_mat, 9:12].astype(np.float64)
    # d/ds loss(I(s)) at s=1 with Le -> s Le and no background: sum over primitives of Le . gradient
    analytic = float(np.sum(le * ge))
    emit_img = img.astype(np.float64)          # the room's background is black: every pixel is emitted light
    direct = float(np.sum(gimg.astype(np.float64) * emit_img))
    assert abs(analytic - direct) <= 1e-5 * abs(direct)
