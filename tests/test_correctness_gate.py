"""The CPU half of the reference's BVH correctness gate (benchmark/GpuBvhCorrectnessCheck.java:49-132) on its own
scenario list: for every scenario the FP32 oracle's linear closest hits agree with an FP64 brute force within the
gate's CPU tolerances (:31-32, :206-238: same hit flag and primitive, |dt| <= 2e-3 + 2e-4 max(1, |t|), normals within
2e-3); and for every build method the oracle's BVH traversal equals its linear scan bit for bit (the gate asks
|dt| <= 1e-4 + 1e-5 max(1, |t|), equal primitive order and material, normals within 2e-5, :241-272), with no stack
overflow and a maximum stack below the capacity of 32 (:103-106)."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O

from scenes import gate_methods, gate_rays, gate_scenarios

SCENARIOS = gate_scenarios()


def _offsets(sc):
    ns, nb, na = len(sc.sphere_data()), len(sc.box_data()), len(sc.affine_sphere_data())
    return {0: 0, 1: ns, 2: ns + nb, 3: ns + nb + na}


@pytest.mark.parametrize("name,make,camera", SCENARIOS, ids=[s[0] for s in SCENARIOS])
def test_gate_linear_fp32_agrees_with_fp64(name, make, camera):
    sc = make(gate_methods()[0][1])
    rays = gate_rays(camera)
    osc = O.Scene(sc.wire(), use_bvh=False)
    f, t, order, mat, nrm, _ = O.trace_replay(osc, rays)
    nrm = nrm.reshape(-1, 3)
    off = _offsets(sc)
    n_planes = len(sc.plane_data())
    assert f.sum() > 0, name
    mismatch = []
    for i, r in enumerate(rays):
        hit, t64, kind, idx, n64 = O.closest64(osc, r)
        if bool(f[i]) != hit:
            mismatch.append(i)
            continue
        if not hit:
            continue
        want = off[kind] + idx if kind in off else None
        if want is not None and want != order[i]:
            mismatch.append(i)
            continue
        if abs(float(t[i]) - t64) > 2e-3 + 2e-4 * max(1.0, abs(t64)) or np.any(np.abs(nrm[i] - n64) > 2e-3):
            mismatch.append(i)
    assert not mismatch, (name, mismatch[:8], n_planes)


@pytest.mark.parametrize("name,make,camera", SCENARIOS, ids=[s[0] for s in SCENARIOS])
def test_gate_bvh_equals_linear(name, make, camera):
    rays = gate_rays(camera)
    for method, cfg in gate_methods():
        sc = make(cfg)
        lin = O.trace_replay(O.Scene(sc.wire(), use_bvh=False), rays)
        bvh = O.trace_replay(O.Scene(sc.wire(), use_bvh=True, stack_size=32), rays)
        for a, b in zip(lin[:5], bvh[:5]):
            np.testing.assert_array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32),
                                          err_msg=f"{name} {method}")
        per = bvh[5]
        assert int(per[:, 8].sum()) == 0, (name, method, "stack overflow")
        assert int(per[:, 9].max()) < 32, (name, method, "maximum stack")
