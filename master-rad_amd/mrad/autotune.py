"""Autotuner of the wavefront form's launch parameters.

Coordinate descent over queue capacity (paths in flight), waves per CU of
the persistent kernels, the LDS split of the traversal stack (entries per
lane in LDS), the stack depth (entries per lane of the ahead-of-time stack
storage), the closest-hit passes per bounce step with their pop budget
(capped passes regroup long traversals into full waves) and the lanes still
traversing that keep a wave from being suspended, on a loaded batch: each candidate is applied with
mrad_batch_configure (scenes stay resident), rendered once to warm and then
timed.  Images do not depend on these parameters (every path runs the same
arithmetic), so only time is compared.  A time budget bounds the search: once
it is spent, the best configuration so far is kept.
"""
from __future__ import annotations

import dataclasses
import time

from .batch import BatchRenderer, WavefrontConfig

GRID = {
    "queue_capacity": [1 << 22, 1 << 23, 1 << 24, 1 << 25, 1 << 26, 1 << 27],
    "waves_per_cu": [16, 24, 28, 32],
    "lds_stack_entries": [4, 10],
    "stack_storage": [32, 64],
    "trace_passes": [1, 2],
    "pass_pops": [64, 80, 96],
    "pass_dense_lanes": [-1, 16, 32],
}


def _time(batch: BatchRenderer, cfg: WavefrontConfig, repeats: int) -> float:
    batch.configure(cfg)
    batch.render()
    best = float("inf")
    for _ in range(repeats):
        t0 = time.perf_counter()
        batch.render()
        best = min(best, time.perf_counter() - t0)
    return best


def tune(batch: BatchRenderer, start: WavefrontConfig | None = None, grid: dict | None = None, passes: int = 1,
         repeats: int = 1, log=None, budget_s: float | None = None) -> tuple[WavefrontConfig, list[dict]]:
    """Returns the fastest configuration found and the table of every trial.
    With `budget_s`, no new candidate is started once that many seconds have
    passed."""
    grid = grid or GRID
    best = dataclasses.replace(start or batch.config, use_graph=True, time_kernels=False, collect_metrics=False)
    seen: dict[tuple, float] = {}
    table = []
    t0 = time.perf_counter()

    def trial(cfg):
        key = tuple(getattr(cfg, k) for k in grid)
        if key not in seen:
            seen[key] = _time(batch, cfg, repeats)
            table.append({**{k: getattr(cfg, k) for k in grid}, "seconds": seen[key]})
            if log:
                log(f"autotune {table[-1]}")
        return seen[key]

    best_t = trial(best)
    for _ in range(passes):
        for name, values in grid.items():
            for v in values:
                if budget_s is not None and time.perf_counter() - t0 > budget_s:
                    batch.configure(best)
                    return best, table
                cand = dataclasses.replace(best, **{name: v})
                t = trial(cand)
                if t < best_t:
                    best, best_t = cand, t
    batch.configure(best)
    return best, table
