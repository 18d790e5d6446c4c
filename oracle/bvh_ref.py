"""Pure-Python restatement of the reference BVH builder, for small inputs.

TEST INFRASTRUCTURE ONLY (the oracle; imported by tests/ only).

Follows src/xyz/marsavic/gfxlab/gpu/GpuScene.java literally: primitive
bounds and centroids (:626-649, :79, :877-903), full-sweep object SAH that
re-sorts the node's primitives per axis by (centroid, order) at every node
(:1453-1488), the leaf rule (:1424-1447) and DFS pre-order flattening
(:772-808).  It is quadratic-ish and meant for a few hundred primitives: it
checks the C++ builder, which sorts once and partitions.
"""
from __future__ import annotations

import math
import struct

import numpy as np

REF_SPHERE, REF_BOX, REF_ASPHERE, REF_ABOX = 0x00000000, 0x40000000, 0x80000000, 0xC0000000


def _f32(x: float) -> float:
    return struct.unpack("f", struct.pack("f", x))[0]


def _next_down2(x: float) -> float:
    return float(np.nextafter(np.nextafter(np.float32(x), np.float32(-np.inf)), np.float32(-np.inf)))


def _next_up2(x: float) -> float:
    return float(np.nextafter(np.nextafter(np.float32(x), np.float32(np.inf)), np.float32(np.inf)))


def _java_min(a: float, b: float) -> float:
    if a != a:
        return a
    if a == 0.0 and b == 0.0:
        return a if math.copysign(1.0, a) < 0 else b
    return a if a <= b else b


def _java_max(a: float, b: float) -> float:
    if a != a:
        return a
    if a == 0.0 and b == 0.0:
        return b if math.copysign(1.0, a) < 0 else a
    return a if a >= b else b


def _java_compare_key(v: float):
    # Double.compare total order: -0.0 < +0.0
    return (v, math.copysign(1.0, v))


def _union(a, b):
    return ([_java_min(a[0][i], b[0][i]) for i in range(3)], [_java_max(a[1][i], b[1][i]) for i in range(3)])


def _area(b) -> float:
    dx, dy, dz = b[1][0] - b[0][0], b[1][1] - b[0][1], b[1][2] - b[0][2]
    return 2.0 * (dx * dy + dx * dz + dy * dz)


class Prim:
    __slots__ = ("ref", "order", "bounds", "weight", "centroid")

    def __init__(self, ref, order, bounds, weight):
        self.ref, self.order, self.bounds, self.weight = ref, order, bounds, weight
        self.centroid = [(bounds[0][i] + bounds[1][i]) * 0.5 for i in range(3)]


def primitives(spheres, boxes, affine_sphere_bounds, affine_box_bounds, weights):
    """BvhPrimitive list in global order (GpuScene.java:626-649). weights = (sphere, box, asphere, abox)
    construction weights q."""
    out = []
    order = 0
    for i, (cx, cy, cz, r) in enumerate(spheres):
        cx, cy, cz, r = _f32(cx), _f32(cy), _f32(cz), _f32(r)
        lo = [_next_down2(_f32(cx - r)), _next_down2(_f32(cy - r)), _next_down2(_f32(cz - r))]
        hi = [_next_up2(_f32(cx + r)), _next_up2(_f32(cy + r)), _next_up2(_f32(cz + r))]
        out.append(Prim(REF_SPHERE | i, order, (lo, hi), weights[0])); order += 1
    for i, b in enumerate(boxes):
        lo = [_next_down2(_f32(v)) for v in b[:3]]
        hi = [_next_up2(_f32(v)) for v in b[3:]]
        out.append(Prim(REF_BOX | i, order, (lo, hi), weights[1])); order += 1
    for kind, lst, w in ((REF_ASPHERE, affine_sphere_bounds, weights[2]), (REF_ABOX, affine_box_bounds, weights[3])):
        for i, b in enumerate(lst):
            lo = [_next_down2(_f32(v)) for v in b[:3]]
            hi = [_next_up2(_f32(v)) for v in b[3:]]
            out.append(Prim(kind | i, order, (lo, hi), w)); order += 1
    return out


def build(prims, leaf_size: int, traversal_weight: float):
    """Returns (node_bounds f32 (n,6), node_data i32 (n,4), refs i32, roots i32)."""
    arr = list(prims)

    def node(start, end):
        bounds = arr[start].bounds
        cb = (list(arr[start].centroid), list(arr[start].centroid))
        for i in range(start + 1, end):
            bounds = _union(bounds, arr[i].bounds)
            c = arr[i].centroid
            cb = ([_java_min(cb[0][k], c[k]) for k in range(3)], [_java_max(cb[1][k], c[k]) for k in range(3)])
        count = end - start
        if count == 1:
            return ("leaf", bounds, start, end)
        best = None
        node_area = _area(bounds)
        for axis in range(3):
            if not (cb[1][axis] - cb[0][axis] > 1.0e-12):
                continue
            s = sorted(arr[start:end], key=lambda p: (_java_compare_key(p.centroid[axis]), p.order))
            pre, prew = [None] * count, [0.0] * count
            suf, sufw = [None] * count, [0.0] * count
            pre[0], prew[0] = s[0].bounds, s[0].weight
            for i in range(1, count):
                pre[i] = _union(pre[i - 1], s[i].bounds)
                prew[i] = prew[i - 1] + s[i].weight
            suf[count - 1], sufw[count - 1] = s[count - 1].bounds, s[count - 1].weight
            for i in range(count - 2, -1, -1):
                suf[i] = _union(suf[i + 1], s[i].bounds)
                sufw[i] = sufw[i + 1] + s[i].weight
            for i in range(1, count):
                cost = node_area * traversal_weight + _area(pre[i - 1]) * prew[i - 1] + _area(suf[i]) * sufw[i]
                if best is None or cost < best[2]:
                    best = (s, i, cost)
        if best is None:
            return ("leaf", bounds, start, end)
        wsum = 0.0
        for i in range(start, end):
            wsum += arr[i].weight
        if count <= leaf_size and not (best[2] < _area(bounds) * wsum):
            return ("leaf", bounds, start, end)
        arr[start:end] = best[0]
        mid = start + best[1]
        left = node(start, mid)
        right = node(mid, end)
        return ("inner", bounds, left, right)

    root = node(0, len(arr))
    nb, nd, refs = [], [], []

    def flatten(n):
        idx = len(nb)
        nb.append(list(n[1][0]) + list(n[1][1]))
        nd.append([0, 0, 0, 0])
        if n[0] == "leaf":
            start, end = n[2], n[3]
            t0 = arr[start].ref & 0xC0000000
            kind = -1 if any((arr[i].ref & 0xC0000000) != t0 for i in range(start, end)) else \
                {REF_SPHERE: -2, REF_BOX: -3, REF_ASPHERE: -4, REF_ABOX: -5}[t0]
            nd[idx] = [kind, -1, len(refs), end - start]
            refs.extend(arr[i].ref for i in range(start, end))
            return idx
        left = flatten(n[2])
        right = flatten(n[3])
        nd[idx] = [left, right, -1, 0]
        return idx

    flatten(root)
    refs_i32 = np.array(refs, dtype=np.uint32).view(np.int32)
    return (np.array(nb, np.float32).reshape(-1, 6), np.array(nd, np.int32).reshape(-1, 4), refs_i32,
            np.array([0], np.int32))
