"""Primitive-cost calibration (CPU): the workload generator's layout and the
measurement protocol's arithmetic, driven by a synthetic timer whose costs
are known, so the recovered weights are known exactly."""
from __future__ import annotations

import numpy as np
import pytest

from mrad import calibrate as K


def _sections(w, n=64, profiles=3):
    o = 0
    out = {}
    for name, width in (("rays", 6 * profiles), ("spheres", 5), ("boxes", 7), ("affines", 22), ("planes", 26),
                        ("aabbs", 6), ("pairs", 12), ("materials", 13)):
        out[name] = w[o:o + n * width].reshape(-1, 6 if name == "rays" else width)
        o += n * width
    assert o == len(w)
    return out


def test_workload_layout_follows_reference_records():
    """PrimitiveCostBenchmark.buildWorkload (:488-715): ranges of every record kind."""
    n = 64
    s = _sections(K.build_workload(n), n)
    d = s["rays"][:, 3:6].astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, rtol=1e-6)
    assert np.all(s["rays"][:n, :3] == np.array([0, 0, -3], np.float32))     # profile 0: pinhole at z = -3
    assert np.all(np.abs(s["rays"][2 * n:, 0]) <= 0.68)                         # profile 2: spread origins
    sp = s["spheres"]
    assert np.all((sp[:, 3] >= 0.58) & (sp[:, 3] <= 0.92)) and np.array_equal(sp[:, 4], np.arange(n))
    b = s["boxes"]
    centre = (b[:, 0:3] + b[:, 3:6]) / 2
    np.testing.assert_allclose(centre[:, :2], sp[:, :2], atol=1e-6)
    half = (b[:, 3:6] - b[:, 0:3]) / 2
    assert np.all((half >= 0.52 - 1e-6) & (half <= 0.87 + 1e-6))
    assert np.array_equal(s["aabbs"], b[:, :6])
    p = s["pairs"]
    assert np.array_equal(p[:, 0:3], b[:, 0:3]) and np.array_equal(p[:, 9:12], b[:, 3:6])
    assert np.array_equal(p[:, 3], p[:, 6])                                     # shared split plane
    a = s["affines"]
    assert np.array_equal(a[:, 3], sp[:, 0]) and np.array_equal(a[:, 7], sp[:, 1])
    lin = a[:, [0, 1, 4, 5]].reshape(-1, 2, 2).astype(np.float64)
    det = np.linalg.det(lin)
    assert np.all(det > 0)                                                      # rotation over positive scales
    pl = s["planes"]
    np.testing.assert_allclose(np.linalg.norm(pl[:, 0:3].astype(np.float64), axis=1), 1.0, rtol=1e-6)
    assert np.all(pl[:, 24] == -1) and np.all(pl[:, 25] == -1)


def test_workload_rejects_non_power_of_two():
    with pytest.raises(ValueError):
        K.build_workload(100)


def test_operation_order_rotates_and_reverses():
    """operationOrder (:267-278): every context is a permutation of 1..7."""
    assert K.operation_order(1) == [1, 2, 3, 4, 5, 6, 7]
    for ctx in range(1, 6):
        order = K.operation_order(ctx)
        assert sorted(order) == list(range(1, 8))
        assert order[0] == 1 + (ctx - 1) % 7
        if ctx % 2 == 0:
            assert order[1] == 1 + (ctx - 2) % 7          # walks backwards


class _FakeTimer:
    """Launch time = tests * (setup + cost[op]) ns, with a per-launch jitter
    that the pairing cancels (identical for the two launches of a pair)."""

    def __init__(self, costs):
        self.costs = costs
        self.calls = 0

    def __call__(self, op, profile, offset, iterations, copies, blocks, threads):
        self.calls += 1
        tests = blocks * threads * iterations * copies
        per = 2.0 + (0.0 if op == 0 else self.costs[op] * (1.0 + 0.1 * profile))
        return tests * per * 1e-6


def test_protocol_recovers_known_relative_costs():
    costs = {1: 3.0, 2: 3.9, 3: 7.5, 4: 8.1, 5: 1.5, 6: 3.8, 7: 8.6}
    timer = _FakeTimer(costs)
    p = K.Protocol(contexts=2, warmups=1, repeats=3, iterations=8, sublaunches=2)
    cal = K.run(p, timer=timer, compute_units=2)
    for op, label in enumerate(K.LABELS[1:], start=1):
        assert cal.weights[label] == pytest.approx(costs[op] / costs[1], rel=1e-9), label
    # 2 contexts x 3 profiles x 7 operations x (1 + 3) relative samples x 2 pairs x 2 sub-launches x 2 launches
    assert timer.calls == 2 * 3 * 7 * 4 * 2 * 2 * 2
    assert len(cal.per_context) == 2 and cal.weights["sphere"] == 1.0


def test_protocol_rejects_non_positive_costs():
    timer = _FakeTimer({1: 0.0, 2: 1.0, 3: 1.0, 4: 1.0, 5: 1.0, 6: 1.0, 7: 1.0})
    with pytest.raises(RuntimeError):
        K.run(K.Protocol(contexts=1, warmups=0, repeats=1, iterations=2, sublaunches=1), timer=timer, compute_units=1)


def test_bvh_config_uses_calibrated_weights():
    w = {"sphere": 1.0, "box": 1.1, "affineSphere": 2.0, "affineBox": 2.2, "nodeAabb": 0.4,
         "interiorTraversal": 1.3, "plane": 2.5}
    cfg = K.bvh_config(w, leaf_size=4, lam=0.5)
    assert (cfg.mode, cfg.leaf_size, cfg.lam) == ("WEIGHTED_SAH", 4, 0.5)
    assert (cfg.traversal_weight, cfg.box_weight, cfg.affine_box_weight) == (1.3, 1.1, 2.2)
    cfg.native()


def test_in_situ_fit_recovers_linear_costs():
    """Synthetic renders whose time is exactly c0 N + a N_aabb + b N_prim."""
    c0, a, b = 0.2, 0.005, 0.07
    mixes = [(92.2, 1.43), (92.7, 1.18), (88.2, 4.60), (93.4, 1.42)]
    n = 1_000_000
    samples = [{"segments": n, "aabb_tests": int(x * n), "primitive_tests": int(y * n),
                "trace_ns": c0 * n + a * int(x * n) + b * int(y * n)} for x, y in mixes]
    fit = K.fit_in_situ(samples)
    assert fit["segment_ns"] == pytest.approx(c0, rel=1e-6)
    assert fit["aabb_ns"] == pytest.approx(a, rel=1e-6)
    assert fit["primitive_ns"] == pytest.approx(b, rel=1e-6)
    assert fit["rms_residual_ns"] < 1e-9
    with pytest.raises(ValueError):
        K.fit_in_situ(samples[:2])


def test_in_situ_weights_in_sphere_units():
    iso = {"sphere": 1.0, "box": 1.5, "affineSphere": 2.0, "affineBox": 2.0, "nodeAabb": 1.0,
           "interiorTraversal": 2.0, "plane": 1.0}
    counts = {"sphere_tests": 50, "box_tests": 50, "affine_sphere_tests": 0, "affine_box_tests": 0}
    # the mean tested primitive costs 1.25 spheres, so 0.1 ns per primitive test is 0.08 ns per sphere test
    w = K.in_situ_weights({"aabb_ns": 0.008, "primitive_ns": 0.1}, iso, counts)
    assert w["nodeAabb"] == pytest.approx(0.1)
    assert w["interiorTraversal"] == pytest.approx(0.2)
    assert w["box"] == 1.5 and w["sphere"] == 1.0
