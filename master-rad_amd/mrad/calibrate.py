"""Primitive-cost calibration on this GPU: per-operation costs relative to a
sphere test, the weights of the weighted SAH (gpu/PrimitiveCostModel.java:11-19).

Follows the reference's calibration campaign (gpu/PrimitiveCostBenchmark.java,
benchmarks/config/primitive-cost-protocol.json):

  * workload: 4096 records of each kind and 3 ray profiles, generated by the
    reference's hashes and formulas (buildWorkload, :488-715);
  * observation: one operation paired with the setup (wrapper-only) kernel,
    32 alternating sub-launches of 64 iterations x 8 operation copies per
    thread; cost = (operation - setup) ns per test (runPair, :201-242);
  * each operation is paired with a local sphere observation, the pair order
    alternating by repeat (runRelativeSample, :244-265); 4 warm-up and 12
    measured pairs per operation and profile;
  * aggregation: per profile, median operation net / median local sphere net;
    per context, the median over the 3 profiles; final, the median over the
    contexts (:352-410); operation order rotated and reversed by context
    (operationOrder, :267-278).

What is measured are this library's traversal-time operations (mrad_calibrate.hip):
the distance-and-accept tests of mrad_device.hpp, one node entry distance,
and a 64-byte node record's expansion.  Two choices differ from the
reference: the grid is 7 blocks of 256 threads per CU (the trace kernel's
occupancy; the reference's fixed 256 blocks would put one wave on each SIMD
of this GPU), and contexts are rounds in one process rather than fresh
processes.
"""
from __future__ import annotations

import ctypes as C
import math
import statistics
from dataclasses import dataclass, field

import numpy as np

from . import _native as N

LABELS = ("setup", "sphere", "box", "affineSphere", "affineBox", "nodeAabb", "interiorTraversal", "plane")
RECORDS = 4096
PROFILES = 3
F32 = np.float32


def _hash32(v: int) -> int:
    """PrimitiveCostBenchmark.hash32 (:698-707) on 32-bit two's complement."""
    v &= 0xFFFFFFFF
    v ^= v >> 17
    v = (v * 0xED5AD4BB) & 0xFFFFFFFF
    v ^= v >> 11
    v = (v * 0xAC4C1B51) & 0xFFFFFFFF
    v ^= v >> 15
    v = (v * 0x31848BAB) & 0xFFFFFFFF
    v ^= v >> 14
    return v


def _unit(v: int) -> np.float32:
    return F32((v & 0xFFFFFFFF) >> 8) * F32(2.0 ** -24)


def _signed_unit(v: int) -> np.float32:
    return F32(2.0) * _unit(v) - F32(1.0)


def build_workload(records: int = RECORDS, profiles: int = PROFILES) -> np.ndarray:
    """The calibration workload in the reference's float layout (buildWorkload, :488-580)."""
    if records <= 0 or records & (records - 1):
        raise ValueError("calibration workload size must be a power of two")
    n = records
    ray_off = 0
    sph_off = ray_off + profiles * n * 6
    box_off = sph_off + n * 5
    aff_off = box_off + n * 7
    pla_off = aff_off + n * 22
    aabb_off = pla_off + n * 26
    pair_off = aabb_off + n * 6
    mat_off = pair_off + n * 12
    data = np.zeros(mat_off + n * 13, np.float32)
    for p in range(profiles):
        for r in range(n):
            s = _hash32(r ^ ((p * 0x6D2B79F5) & 0xFFFFFFFF))
            x, y, z = _signed_unit(s), _signed_unit(_hash32(s ^ 0x51ED270B)), _signed_unit(_hash32(s ^ 0xA54FF53A))
            ox, oy, oz, dz = 0.0, 0.0, -3.0, 1.0
            if p == 0:
                dx, dy = float(x) * 0.00075, float(y) * 0.00075
            elif p == 1:
                dx, dy = float(x) * 0.36, float(y) * 0.24 - 0.12
            else:
                ox, oy = float(x) * 0.68, float(y) * 0.68
                oz += float(z) * 0.20
                dx, dy = float(x) * 0.055, float(y) * 0.055
            inv = 1.0 / math.sqrt(dx * dx + dy * dy + dz * dz)
            o = ray_off + (p * n + r) * 6
            data[o:o + 6] = (ox, oy, oz, dx * inv, dy * inv, dz * inv)
    for r in range(n):
        s = _hash32((r * 0x9E3779B9) & 0xFFFFFFFF)
        x, y, z = _signed_unit(s), _signed_unit(_hash32(s ^ 0x243F6A88)), _signed_unit(_hash32(s ^ 0xB7E15162))
        cx, cy, cz = x * F32(0.28), y * F32(0.22), z * F32(0.18) + F32(0.35)
        o = sph_off + r * 5
        data[o:o + 5] = (cx, cy, cz, F32(0.58) + F32(0.34) * _unit(_hash32(s ^ 0x13198A2E)), r)
        hx = F32(0.52) + F32(0.35) * _unit(_hash32(s ^ 0x03707344))
        hy = F32(0.52) + F32(0.35) * _unit(_hash32(s ^ 0xA4093822))
        hz = F32(0.52) + F32(0.35) * _unit(_hash32(s ^ 0x299F31D0))
        o = box_off + r * 7
        data[o:o + 7] = (cx - hx, cy - hy, cz - hz, cx + hx, cy + hy, cz + hz, r)
        data[aabb_off + r * 6:aabb_off + r * 6 + 6] = data[o:o + 6]
        # writeAffineRecord (:582-624)
        angle = float(_signed_unit(_hash32(s ^ 0x452821E6))) * 0.85
        c, sn = math.cos(angle), math.sin(angle)
        sx = 0.62 + 0.78 * float(_unit(_hash32(s ^ 0x38D01377)))
        sy = 0.62 + 0.78 * float(_unit(_hash32(s ^ 0xBE5466CF)))
        sz = 0.62 + 0.78 * float(_unit(_hash32(s ^ 0x34E90C6C)))
        i00, i01, i10, i11, i22 = F32(c / sx), F32(sn / sx), F32(-sn / sy), F32(c / sy), F32(1.0 / sz)
        o = aff_off + r * 22
        data[o:o + 22] = (i00, i01, 0, cx, i10, i11, 0, cy, 0, 0, i22, cz, i00, i10, 0, i01, i11, 0, 0, 0, i22, r)
        # writePlaneRecord (:626-653)
        nx = float(_signed_unit(_hash32(s ^ 0xC0AC29B7))) * 0.12
        ny = float(_signed_unit(_hash32(s ^ 0xC97C50DD))) * 0.12
        nz = -1.0
        il = 1.0 / math.sqrt(nx * nx + ny * ny + nz * nz)
        nX, nY, nZ = F32(nx * il), F32(ny * il), F32(nz * il)
        pz = F32(0.15) * _signed_unit(_hash32(s ^ 0x3F84D5B5))
        o = pla_off + r * 26
        data[o:o + 7] = (nX, nY, nZ, -(nX * cx + nY * cy + nZ * pz), cx, cy, pz)
        data[o + 23], data[o + 24], data[o + 25] = r, -1.0, -1.0
        # writeNodePair (:655-680)
        spread = F32(0.04) * _signed_unit(_hash32(s ^ 0x9216D5D9))
        split = max(F32(-0.75) * hx, min(F32(0.75) * hx, spread))
        o = pair_off + r * 12
        data[o:o + 12] = (cx - hx, cy - hy, cz - hz, cx + split, cy + hy, cz + hz,
                          cx + split, cy - hy, cz - hz, cx + hx, cy + hy, cz + hz)
        # writeMaterialRecord (:682-696)
        o = mat_off + r * 13
        keys = (0xD1310BA6, 0x98DFB5AC, 0x2FFD72DB, 0xD01ADFB7, 0xB8E1AFED, 0x6A267E96, 0xBA7C9045, 0xF12C7F99,
                0x24A19947, 0xB3916CF7, 0x0801F2E2, 0x858EFC16, 0x636920D8)
        lo = (0.35,) * 3 + (0.02,) * 3 + (0.01,) * 3 + (0.0,) * 3 + (1.1,)
        span = (0.55,) * 3 + (0.18,) * 3 + (0.09,) * 3 + (0.01,) * 3 + (0.7,)
        data[o:o + 13] = [F32(a) + F32(b) * _unit(_hash32(s ^ k)) for a, b, k in zip(lo, span, keys)]
    return data


def operation_order(context: int, operations: int = len(LABELS) - 1) -> list[int]:
    """PrimitiveCostBenchmark.operationOrder (:267-278): rotated, reversed on even contexts."""
    offset = (context - 1) % operations
    reverse = (context & 1) == 0
    return [1 + (offset + (-k if reverse else k)) % operations for k in range(operations)]


@dataclass
class Protocol:
    contexts: int = 5
    warmups: int = 4
    repeats: int = 12
    iterations: int = 2048
    sublaunches: int = 32
    copies: int = 8
    threads: int = 256
    blocks_per_cu: int = 7
    profiles: int = PROFILES


@dataclass
class Calibration:
    weights: dict
    per_context: list = field(default_factory=list)
    per_profile: list = field(default_factory=list)   # (context, profile, {label: ratio})
    sphere_net_ns: list = field(default_factory=list)


class _Timer:
    """One launch of the calibration kernel per call (mrad_cost_time)."""

    def __init__(self, workload: np.ndarray, records: int, profiles: int, device: int = 0):
        h = C.c_void_p()
        cus = C.c_int()
        w = np.ascontiguousarray(workload, np.float32)
        N.check(N.lib.mrad_cost_create(device, N.ptr(w, N.fp), records, profiles, C.byref(cus), C.byref(h)))
        self._h = h
        self.compute_units = int(cus.value)

    def __call__(self, op, profile, offset, iterations, copies, blocks, threads) -> float:
        ms = C.c_float()
        N.check(N.lib.mrad_cost_time(self._h, op, profile, offset, iterations, copies, blocks, threads,
                                     C.byref(ms)))
        return float(ms.value)

    def output(self, n: int) -> np.ndarray:
        """The per-thread accumulators of the last launch (mrad_cost_output)."""
        out = np.zeros(n, np.float32)
        N.check(N.lib.mrad_cost_output(self._h, N.ptr(out, N.fp), n))
        return out

    def close(self):
        if self._h is not None and self._h.value:
            N.lib.mrad_cost_destroy(self._h)
        self._h = None


def _pair(timer, p: Protocol, blocks: int, op: int, profile: int, repeat: int) -> float:
    """runPair (:201-242): net ns per test of `op` against the setup kernel."""
    setup_first_start = (repeat & 1) == 0
    per = p.iterations // p.sublaunches
    setup_ms = op_ms = 0.0
    for k in range(p.sublaunches):
        off = k * per
        if ((k & 1) == 0) == setup_first_start:
            setup_ms += timer(0, profile, off, per, p.copies, blocks, p.threads)
            op_ms += timer(op, profile, off, per, p.copies, blocks, p.threads)
        else:
            op_ms += timer(op, profile, off, per, p.copies, blocks, p.threads)
            setup_ms += timer(0, profile, off, per, p.copies, blocks, p.threads)
    tests = float(blocks) * p.threads * p.iterations * p.copies
    return (op_ms - setup_ms) * 1e6 / tests


def _relative(timer, p, blocks, op, profile, repeat):
    """runRelativeSample (:244-265): (operation net, local sphere net)."""
    if (repeat & 1) == 0:
        a = _pair(timer, p, blocks, op, profile, repeat)
        s = _pair(timer, p, blocks, 1, profile, repeat + 1)
    else:
        s = _pair(timer, p, blocks, 1, profile, repeat + 1)
        a = _pair(timer, p, blocks, op, profile, repeat)
    return a, s


def run(protocol: Protocol | None = None, timer=None, compute_units: int | None = None, log=None) -> Calibration:
    """Runs the campaign; `timer` defaults to the HIP kernel on device 0."""
    p = protocol or Protocol()
    own = timer is None
    if own:
        timer = _Timer(build_workload(RECORDS, p.profiles), RECORDS, p.profiles)
    if compute_units is None:
        compute_units = getattr(timer, "compute_units", 256)
    blocks = max(1, compute_units * p.blocks_per_cu)
    result = Calibration(weights={})
    try:
        for ctx in range(1, p.contexts + 1):
            runs = {}
            for profile in range(p.profiles):
                for op in operation_order(ctx):
                    for w in range(p.warmups):
                        _relative(timer, p, blocks, op, profile, w)
                    runs[(profile, op)] = [_relative(timer, p, blocks, op, profile, r) for r in range(p.repeats)]
            ratios = {op: [] for op in range(1, len(LABELS))}
            for profile in range(p.profiles):
                sphere_net = statistics.median(s for _, s in runs[(profile, 1)])
                if not (sphere_net > 0 and math.isfinite(sphere_net)):
                    raise RuntimeError(f"profile {profile}: sphere net cost {sphere_net} ns is not positive")
                result.sphere_net_ns.append(sphere_net)
                row = {}
                for op in range(1, len(LABELS)):
                    net = statistics.median(a for a, _ in runs[(profile, op)])
                    local = statistics.median(s for _, s in runs[(profile, op)])
                    if not (net > 0 and local > 0):
                        raise RuntimeError(f"profile {profile} {LABELS[op]}: non-positive net cost")
                    ratio = 1.0 if op == 1 else net / local
                    ratios[op].append(ratio)
                    row[LABELS[op]] = ratio
                result.per_profile.append((ctx, profile, row))
            weights = {LABELS[op]: statistics.median(v) for op, v in ratios.items()}
            result.per_context.append(weights)
            if log:
                log(f"context {ctx}: " + ", ".join(f"{k} {v:.4f}" for k, v in weights.items()))
        result.weights = {k: statistics.median(c[k] for c in result.per_context) for k in LABELS[1:]}
    finally:
        if own:
            timer.close()
    return result


# ------------------------------------------------------------------ in situ
#
# The isolated costs above price each operation alone.  Inside the trace
# kernel an operation's cost also depends on divergence and on the latency of
# the fetches around it, so the traversal weight is also fitted in situ: the
# same frames are rendered with BVHs built under several cost settings, and
# the trace kernel's time is regressed on the exact counters of each render,
#   time = c0 N_segments + a N_aabb + b N_primitive_tests,
# all renders having the same segments (the hits do not depend on the BVH).
# a / b prices an AABB test in primitive tests; with the isolated ratios
# between primitive classes, b is converted to sphere units.

def fit_in_situ(samples: list[dict]) -> dict:
    """Least-squares fit of (c0, a, b) from samples {segments, aabb_tests,
    primitive_tests, trace_ns}; returns ns per segment, AABB test and
    primitive test."""
    if len(samples) < 3:
        raise ValueError("the in-situ fit needs at least three BVH variants")
    A = np.array([[1.0, s["aabb_tests"] / s["segments"], s["primitive_tests"] / s["segments"]] for s in samples])
    y = np.array([s["trace_ns"] / s["segments"] for s in samples])
    (c0, a, b), *_ = np.linalg.lstsq(A, y, rcond=None)
    resid = y - A @ np.array([c0, a, b])
    return {"segment_ns": float(c0), "aabb_ns": float(a), "primitive_ns": float(b),
            "rms_residual_ns": float(np.sqrt(np.mean(resid ** 2)))}


def in_situ_weights(fit: dict, isolated: dict, class_counts: dict) -> dict:
    """Weights in sphere-test units: the fitted AABB and interior costs, the
    isolated ratios between primitive classes.  class_counts are the sphere,
    box, affine sphere and affine box tests of the fitted frames."""
    cls = {"sphere": "sphere_tests", "box": "box_tests", "affineSphere": "affine_sphere_tests",
           "affineBox": "affine_box_tests"}
    total = sum(class_counts[v] for v in cls.values())
    mean_class = sum(isolated[k] * class_counts[v] for k, v in cls.items()) / total
    sphere_ns = fit["primitive_ns"] / mean_class
    w = dict(isolated)
    w["nodeAabb"] = fit["aabb_ns"] / sphere_ns
    w["interiorTraversal"] = 2.0 * fit["aabb_ns"] / sphere_ns
    return w


def in_situ(make_workloads, isolated: dict, variants: dict | None = None, log=None) -> dict:
    """Renders the workloads (a list of configs.Workload built by
    make_workloads(bvh_config)) in the wavefront form under each BVH variant
    and fits the in-situ costs."""
    from .batch import BatchRenderer, WavefrontConfig
    from .scene import BvhBuildConfig
    variants = variants or {
        "reference weights, lambda 1": BvhBuildConfig.weighted(8, 1.0),
        "reference weights, lambda 5": BvhBuildConfig.weighted(8, 5.0),
        "isolated gfx950 weights, lambda 1": bvh_config(isolated, 8, 1.0),
        "uniform SAH": BvhBuildConfig.uniform(8),
    }
    samples = []
    for name, cfg in variants.items():
        ws = make_workloads(cfg)

        def batch(**kw):
            b = BatchRenderer(WavefrontConfig(**kw))
            for x in ws:
                b.add_scene(x.scene, x.camera, x.width, x.height, x.spp, x.max_depth, 0, x.frame_seed, x.rr_start)
            return b

        with batch(time_kernels=True) as b:
            b.render()
            trace_ns = min(b.render()["trace_nanos"] for _ in range(2))
        with batch(collect_metrics=True) as b:
            st = b.render()
        samples.append({"variant": name, "segments": st["rays"], "aabb_tests": st["aabb_tests"],
                        "primitive_tests": st["primitive_tests"], "trace_ns": trace_ns,
                        **{k: st[k] for k in ("sphere_tests", "box_tests", "affine_sphere_tests", "affine_box_tests")}})
        if log:
            log(f"in situ {name}: trace {trace_ns / 1e6:.1f} ms, {st['aabb_tests'] / st['rays']:.2f} AABB and "
                f"{st['primitive_tests'] / st['rays']:.2f} primitive tests per segment")
        del ws
    fit = fit_in_situ(samples)
    weights = in_situ_weights(fit, isolated, samples[0])
    return {"fit": fit, "weights": weights, "samples": samples}


def bvh_config(weights: dict, leaf_size: int = 8, lam: float = 1.0):
    """A weighted-SAH BvhBuildConfig with calibrated weights (BvhBuildConfig.java:6-15)."""
    from .scene import BvhBuildConfig
    return BvhBuildConfig("WEIGHTED_SAH", leaf_size, traversal_weight=weights["interiorTraversal"],
                          sphere_weight=1.0, box_weight=weights["box"], affine_sphere_weight=weights["affineSphere"],
                          affine_box_weight=weights["affineBox"], lam=lam)


def main(argv=None) -> None:
    import argparse
    import json
    import sys
    import time
    ap = argparse.ArgumentParser(description="primitive-cost calibration of the weighted SAH on this GPU")
    ap.add_argument("--contexts", type=int, default=5)
    ap.add_argument("--repeats", type=int, default=12)
    ap.add_argument("--warmups", type=int, default=4)
    ap.add_argument("--iterations", type=int, default=2048)
    ap.add_argument("--sublaunches", type=int, default=32)
    ap.add_argument("--in-situ", action="store_true",
                    help="also fit the traversal weight in the trace kernel on the 64-scene batch")
    a = ap.parse_args(argv)
    p = Protocol(contexts=a.contexts, repeats=a.repeats, warmups=a.warmups, iterations=a.iterations,
                 sublaunches=a.sublaunches)
    t0 = time.perf_counter()
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    cal = run(p, log=log)
    out = {"weights": cal.weights, "per_context": cal.per_context, "sphere_net_ns_per_test": cal.sphere_net_ns,
           "protocol": p.__dict__}
    if a.in_situ:
        from . import configs
        out["in_situ"] = in_situ(lambda cfg: configs.batch(64, threads=16, bvh=cfg), cal.weights, log=log)
    out["seconds"] = time.perf_counter() - t0
    print(json.dumps(out))


if __name__ == "__main__":
    main()
