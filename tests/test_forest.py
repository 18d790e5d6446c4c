"""Per-type forests at the boundary (CPU).  The reference's PER_TYPE_SAH modes pack one tree per primitive class
and traverse up to four roots, inserting them farthest first so that the nearest is popped first and equal entry
distances keep their order (GpuKernelSources.java:863-885).  This build's builder does not make forests (the
per-type modes are a construction study), but the boundary accepts them from a caller: here the oracle's
traversal of such a forest equals its linear scan bit for bit, and its per-ray counters obey thesis eq. 4.4 with
four roots (N_aabb = roots + 2 N_internal)."""
from __future__ import annotations

import numpy as np

from mrad import configs
from oracle import oracle as O

from helpers import camera_grid_rays, per_type_forest, random_rays
from scenes import per_type_parts


def test_forest_traversal_equals_linear_scan():
    full, parts = per_type_parts()
    forest = per_type_forest(full, parts)
    assert len(forest.wire()["roots"]) == 4
    rays = np.concatenate([camera_grid_rays(configs.field_camera().packed(), 64, 36),
                           random_rays(np.random.default_rng(3), 2048)])
    lin = O.trace_replay(O.Scene(full.wire(), use_bvh=False), rays)
    got = O.trace_replay(O.Scene(forest.wire()), rays)
    assert lin[0].sum() > 500
    for a, b in zip(lin[:5], got[:5]):
        np.testing.assert_array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))
    per = got[5].astype(np.int64)
    np.testing.assert_array_equal(per[:, 1], 4 + 2 * per[:, 10])
    single = O.trace_replay(O.Scene(full.wire()), rays)[5]
    assert not np.array_equal(single, got[5])            # a different traversal, the same hits
