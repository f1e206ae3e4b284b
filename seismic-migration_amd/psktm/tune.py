"""Autotuner for the migration kernels' tile shape and stage size.

A candidate is (tile_x, tile_y, traces_per_stage): the bins of an output tile
(8 or 16, which sets the register tile of every lane) and the traces staged in
LDS per barrier pair. The tuner times each candidate on the caller's own
workload (a user-supplied timing function), keeps the fastest, and stores the
choice with all timings in a JSON cache keyed by a workload signature. With a
cache entry present it returns the cached choice without timing anything, so
the choice is deterministic for a fixed cache. Ties go to the earlier
candidate.

Trials on part of a workload can rank candidates within their run-to-run
spread (on the 64-gather batch, nine candidates timed on a quarter of the
gathers came within 17%, and the quickest there ran 15% slower on the whole
batch than the runner-up). ``refine_fn`` then times the ``refine_top``
fastest trial candidates on the whole workload and the choice is the
quickest of those.
"""
from __future__ import annotations

import json
import math
import os
import time
from typing import Callable, Iterable, Optional, Sequence, Tuple

Candidate = Tuple[int, int, int]

# 3-D patches: square-ish tiles first; 2-D lines use (16,1) / (8,1).
CANDIDATES_3D: Sequence[Candidate] = (
    (4, 4, 4), (4, 4, 2), (4, 4, 8), (8, 2, 4), (2, 8, 4), (4, 2, 4), (2, 4, 4),
    (8, 2, 8), (2, 8, 8),
)
CANDIDATES_2D: Sequence[Candidate] = ((16, 1, 4), (16, 1, 2), (16, 1, 8), (8, 1, 4))


def default_cache_path() -> str:
    return os.environ.get("PSKTM_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "psktm", "tune.json"))


def signature(kind: str, **fields) -> str:
    """Stable cache key: kind plus sorted field=value pairs."""
    parts = [kind] + [f"{k}={fields[k]}" for k in sorted(fields)]
    return ";".join(parts)


class Autotuner:
    def __init__(self, cache_path: Optional[str] = None):
        self.cache_path = cache_path or default_cache_path()
        self.cache = {}
        if os.path.exists(self.cache_path):
            with open(self.cache_path) as f:
                self.cache = json.load(f)

    def lookup(self, key: str) -> Optional[Candidate]:
        e = self.cache.get(key)
        return tuple(e["choice"]) if e else None

    def choose(self, key: str, candidates: Iterable[Candidate],
               time_fn: Callable[[Candidate], Optional[float]], retune: bool = False,
               refine_fn: Optional[Callable[[Candidate], Optional[float]]] = None,
               refine_top: int = 3) -> Candidate:
        """Fastest candidate for ``key``; ``time_fn`` returns seconds, or None
        when a candidate is not supported for the workload. With ``refine_fn``
        (seconds on the whole workload), the ``refine_top`` fastest candidates
        of ``time_fn`` are timed again by it and the fastest there is chosen."""
        if not retune:
            hit = self.lookup(key)
            if hit is not None:
                return hit
        timings = {}
        best, best_t = None, math.inf
        for c in candidates:
            c = tuple(int(v) for v in c)
            t = time_fn(c)
            timings[",".join(map(str, c))] = t
            if t is not None and t < best_t:
                best, best_t = c, t
        if best is None:
            raise RuntimeError(f"no candidate ran for {key}")
        entry = {"choice": list(best), "seconds": best_t, "timings": timings}
        if refine_fn is not None and refine_top > 1:
            ranked = sorted((t, i, c) for i, (c, t) in
                            enumerate((tuple(int(v) for v in k.split(",")), t)
                                      for k, t in timings.items()) if t is not None)
            refined = {}
            best, best_t = None, math.inf
            for _, _, c in ranked[:refine_top]:
                t = refine_fn(c)
                refined[",".join(map(str, c))] = t
                if t is not None and t < best_t:
                    best, best_t = c, t
            if best is None:
                raise RuntimeError(f"no candidate ran on the whole workload for {key}")
            entry = {"choice": list(best), "seconds": best_t, "timings": timings,
                     "refined": refined}
        entry["tuned_at"] = time.strftime("%Y-%m-%dT%H:%M:%S")
        self.cache[key] = entry
        self.save()
        return best

    def save(self):
        d = os.path.dirname(self.cache_path)
        if d:
            os.makedirs(d, exist_ok=True)
        tmp = self.cache_path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.cache, f, indent=1, sort_keys=True)
        os.replace(tmp, self.cache_path)


def time_gpu(run: Callable[[], None], repeats: int = 2) -> float:
    """Median of ``repeats`` timed calls of ``run`` after one warm-up (GPU)."""
    import torch
    run()
    torch.cuda.synchronize()
    ts = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        run()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    ts.sort()
    return ts[len(ts) // 2]
