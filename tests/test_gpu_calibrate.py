"""GPU: a short primitive-cost calibration run on this device through the
C-ABI (mrad_cost_*): every weight positive and finite, the sphere the unit."""
from __future__ import annotations

import math

import pytest

from mrad import calibrate as K

pytestmark = pytest.mark.gpu


def test_short_calibration_gives_positive_weights(require_gpu):
    p = K.Protocol(contexts=1, warmups=1, repeats=2, iterations=64, sublaunches=2)
    print("calibration: 1 context, 2 repeats, 64 iterations", flush=True)
    cal = K.run(p)
    assert cal.weights["sphere"] == 1.0
    for label, w in cal.weights.items():
        assert math.isfinite(w) and 0.05 < w < 50.0, (label, w)
    assert all(s > 0 for s in cal.sphere_net_ns)


def test_calibration_operations_are_their_intersectors(require_gpu):
    """Every operation of the calibration kernel, on every ray profile of the reference's calibration workload,
    folds the result of the intended routine: each thread's accumulator equals the oracle's restatement with the
    oracle's sphere, box, affine-sphere, affine-box, entry-distance, node-pair and plane routines, bit for bit.
    The operations' accumulators all differ from one another, so two swapped operation indices would fail."""
    import numpy as np

    from oracle import oracle as O
    records, profiles = K.RECORDS, K.PROFILES
    w = K.build_workload(records, profiles)
    timer = K._Timer(w, records, profiles)
    blocks, threads, iterations, copies = 2, 256, 6, 3
    n = blocks * threads
    seen = {}
    try:
        for op in range(8):
            for profile in range(profiles):
                timer(op, profile, 17, iterations, copies, blocks, threads)
                got = timer.output(n)
                ref = O.cost_accumulators(w, records, profiles, op, profile, 17, iterations, copies, n)
                np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32), err_msg=f"op {op}")
                seen[(op, profile)] = got
    finally:
        timer.close()
    for profile in range(profiles):
        outs = [seen[(op, profile)] for op in range(8)]
        for a in range(8):
            for b in range(a + 1, 8):
                assert not np.array_equal(outs[a], outs[b]), (profile, a, b)
