"""Autotuner of the engine's tier options (bucket bounds, chunk and partition
sizes, LDS merge limits, the column-blocked modularity pass).

The options change only how work is mapped onto kernels, never a result
(every tier computes the same exact FP64 sums and the same packed argmax), so
the tuner may pick any of them by time alone.  It times one sweep from the
singleton state on the device for each candidate (lv_set_options on one
resident handle, median of `repeats`), keeps the fastest, and caches the
choice in a JSON file keyed by a signature of the graph's shape (vertex and
edge counts rounded to powers of two and the degree histogram by bucket).
No counterpart exists in the reference.
"""
from __future__ import annotations

import itertools
import json
import math
import os
import tempfile
import time

import numpy as np

import louvain_amd as L

DEFAULT_CACHE = os.path.join(tempfile.gettempdir(), "gpulouvain_tune.json")

# candidate values per option (0 = the library default)
SPACE = {
    "heavy_wg_max": [512, 1024, 2048, 4096],
    "chunk_edges": [256, 512, 1024],
    "part_target": [256, 512, 1024],
    "merge_limit": [1024, 1536],
    "xcd_rows": [0, 1, 2, -1],
}
_BOUNDS = [4, 8, 16, 32, 84, 319, 2048]


def signature(g: L.CSR) -> str:
    """Shape key: log2 of V and E, and log2 of the vertex count per degree range."""
    deg = g.degrees()
    edges = np.searchsorted(np.array(_BOUNDS), deg, side="left")
    hist = np.bincount(edges[deg > 0], minlength=len(_BOUNDS) + 1)
    parts = [int(round(math.log2(max(g.V, 1)))), int(round(math.log2(max(g.E, 1))))]
    parts += [int(round(math.log2(h))) if h > 0 else -1 for h in hist]
    return "v1:" + ",".join(str(p) for p in parts)


def load_cache(path: str = DEFAULT_CACHE) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save_cache(cache: dict, path: str = DEFAULT_CACHE) -> None:
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def candidates(space: dict = SPACE):
    keys = sorted(space)
    for vals in itertools.product(*(space[k] for k in keys)):
        yield dict(zip(keys, vals))


def time_sweep(G: L.Graph, repeats: int = 3) -> float:
    """Median wall time (ms) of reset + one sweep."""
    ts = []
    for _ in range(repeats):
        G.reset()
        t = time.perf_counter()
        G.sweep()
        ts.append((time.perf_counter() - t) * 1e3)
    return float(np.median(ts))


def time_level_sweep(G: L.Graph, start: np.ndarray, repeats: int = 3) -> float:
    """Median wall time (ms) of one sweep of the current level from `start`."""
    ts = []
    for _ in range(repeats):
        G.set_communities(start)
        t = time.perf_counter()
        G.sweep()
        ts.append((time.perf_counter() - t) * 1e3)
    return float(np.median(ts))


def tune(g: L.CSR, device: int = 0, space: dict = SPACE, repeats: int = 3,
         cache_path: str | None = DEFAULT_CACHE, timer=time_sweep, second_level: bool = True) -> dict:
    """Best options for graphs shaped like g: {"options": {...}, "ms": t,
    "default_ms": t0, "tried": n, "signature": key, "cached": bool}.

    The objective is the first sweep from singletons at level 0 plus, with
    second_level, the first sweep from singletons at level 1 (where a run
    spends most of its sweeps on graphs whose first phase merges little)."""
    key = signature(g) + (":l2" if second_level else "")
    cache = load_cache(cache_path) if cache_path else {}
    if key in cache:
        out = dict(cache[key])
        out["cached"] = True
        return out
    results = []
    with L.Graph(g, device=device) as G:
        G.sweep()  # warm the kernels
        default_ms = timer(G, repeats)
        for opt in candidates(space):
            G.set_options(**opt)
            results.append([timer(G, repeats), opt])
        if second_level:
            G.set_options()
            G.reset()
            if G.optimise(1e-6):
                G.aggregate()
                V1, _ = G.level_size()
                start = np.arange(V1, dtype=np.int32)
                default_ms += time_level_sweep(G, start, repeats)
                for r in results:
                    G.set_options(**r[1])
                    r[0] += time_level_sweep(G, start, repeats)
                G.set_options()
    best_ms, best = min(results, key=lambda r: r[0])
    out = {"options": best, "ms": best_ms, "default_ms": default_ms, "tried": len(results),
           "signature": key, "cached": False}
    if cache_path:
        cache[key] = {k: v for k, v in out.items() if k != "cached"}
        save_cache(cache, cache_path)
    return out
