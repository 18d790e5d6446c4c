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
