"""Central differences of a frame's L2 loss over its recorded paths.

The frame's paths are recorded once by the oracle (or_record_paths: FP32
geometry and sampling decisions); under other albedos and emissions the same
paths give an exact polynomial, evaluated here in double.  A parameter of
material m changes only the paths with a vertex on m and the pixels they fall
in, so each central difference re-evaluates just those.
"""
from __future__ import annotations

import numpy as np

from oracle import oracle as O


class PathFD:
    def __init__(self, rec: O.RecordedPaths, materials64: np.ndarray, target: np.ndarray):
        self.rec = rec
        self.mats = np.asarray(materials64, np.float64)
        self.target = np.asarray(target, np.float64).reshape(-1, 3)
        self.radiance = rec.radiance(self.mats)
        self.image = self.radiance.reshape(-1, rec.spp, 3).mean(axis=1)
        self.n = self.image.size

    def loss(self) -> float:
        return float(np.mean((self.image - self.target) ** 2))

    def grad_image(self) -> np.ndarray:
        """dLoss/dImage of the FP64 frame, W*H*3 floats."""
        return (2.0 * (self.image - self.target) / self.n).reshape(-1).astype(np.float32)

    def _delta_loss(self, mats: np.ndarray, paths: np.ndarray) -> float:
        spp = self.rec.spp
        change = (self.rec.radiance(mats, paths) - self.radiance[paths]) / spp
        pix, inv = np.unique(paths // spp, return_inverse=True)
        d = np.zeros((len(pix), 3))
        np.add.at(d, inv, change)
        old = self.image[pix] - self.target[pix]
        return float(np.sum((old + d) ** 2 - old ** 2)) / self.n

    def central_difference(self, kind: str, m: int, c: int, h: float = 1e-3) -> float:
        """dLoss/d(albedo[m, c]) or, for kind "emission", d/d(grey radiance of m) (all three channels)."""
        paths = self.rec.paths_of_material(m)
        if len(paths) == 0:
            return 0.0
        plus, minus = self.mats.copy(), self.mats.copy()
        cols = [c] if kind == "albedo" else [9, 10, 11]
        plus[m, cols] += h
        minus[m, cols] -= h
        span = plus[m, cols[0]] - minus[m, cols[0]]
        return (self._delta_loss(plus, paths) - self._delta_loss(minus, paths)) / span


def per_material(grad_prim: np.ndarray, prim_material: np.ndarray, n_materials: int) -> np.ndarray:
    out = np.zeros((n_materials, 3))
    ok = prim_material >= 0
    np.add.at(out, prim_material[ok], np.asarray(grad_prim, np.float64)[ok])
    return out


def draw_parameters(rng: np.random.Generator, g_albedo: np.ndarray, g_emission: np.ndarray, emitters: np.ndarray,
                    n_albedo: int = 56, n_emission: int = 8) -> list[tuple[str, int, int]]:
    """Parameters drawn uniformly (seeded) among those with a non-zero gradient: n_albedo albedo channels and
    n_emission emitter radiances."""
    cand = np.flatnonzero(g_albedo.reshape(-1) != 0)
    picks = [("albedo", int(i // 3), int(i % 3)) for i in rng.choice(cand, n_albedo, replace=False)]
    em = emitters[np.abs(g_emission[emitters]).sum(1) != 0]
    picks += [("emission", int(m), -1) for m in rng.choice(em, min(n_emission, len(em)), replace=False)]
    return picks
