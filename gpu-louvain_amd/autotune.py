"""Autotuner of the engine's tier options: the lanes per vertex of the
sub-wave buckets (degrees up to 32), the degree cuts inside degree groups 5
and 6, the lanes per row and the LDS table sizes of the
one-workgroup tiers, the chunked tier's chunk, partition and merge sizes, and
the column-blocked passes.

The options change only how work is mapped onto kernels, never a result
(every tier computes the same exact FP64 sums and the same packed argmax), so
the tuner may pick any of them by time alone.  Single graphs: it times a
sweep from the singleton state at level 0 and at level 1 on one resident
handle (lv_set_options, median of `repeats`).  Batches: it times whole
batched runs of a sample of the batch (lv_batch_set_options).  Small spaces
are searched exhaustively; larger ones by coordinate descent (one option at
a time over its values, the others held at the best so far, repeated).  The
choice is cached in a JSON file keyed by a signature of the graph's shape
(vertex and edge counts rounded to powers of two and the degree histogram by
bucket).  The reference's launch shapes are compile-time tables
(modularity_optimisation.cuh:7-19, community_aggregation.cu:286-293).
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

# candidate values per option; a value of the library's default is listed
# first, so that the search starts at the defaults
SPACE = {
    "sub_lanes": [(4, 4, 4), (4, 4, 8), (4, 8, 8), (8, 16, 32), (8, 8, 16), (4, 8, 16)],
    "heavy_wg_max": [4096, 2048, 1024],
    "wg_bound_a": [512, 384, 640],
    "wg_bound_b": [1024, 768, 1536],
    "wg_bound_c": [2048, 1536, 3072],
    "wave_split": [128, 96, 192, 255],
    "wg_lanes": [(128, 256, 256, 1024), (256, 256, 256, 1024), (128, 128, 256, 1024),
                 (128, 256, 512, 1024), (128, 256, 256, 512), (128, 512, 512, 1024)],
    "table_pct": [0, 125, 150, 200],
    "chunk_edges": [1024, 512, 256],
    "part_target": [1024, 512, 256],
    "merge_limit": [1536, 1024],
    "xcd_rows": [0, 1, 2, -1],
}
# the options that shape a batch's launches (its graphs are small: the
# chunked tier and the column-blocked passes rarely apply)
BATCH_SPACE = {k: SPACE[k] for k in ("sub_lanes", "heavy_wg_max", "wg_bound_a", "wg_bound_b",
                                     "wave_split", "wg_lanes", "table_pct")}
EXHAUSTIVE_MAX = 64
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
    """Every combination of the space's values (exhaustive search)."""
    keys = sorted(space)
    for vals in itertools.product(*(space[k] for k in keys)):
        yield dict(zip(keys, vals))


def space_size(space: dict) -> int:
    return int(np.prod([len(v) for v in space.values()]))


def _jsonable(opt: dict) -> dict:
    return {k: (list(v) if isinstance(v, tuple) else v) for k, v in opt.items()}


def search(evaluate, space: dict, passes: int = 2) -> tuple[float, dict, list]:
    """Minimises evaluate(options) over the space: exhaustively when it has at
    most EXHAUSTIVE_MAX points, else by coordinate descent from the first
    value of every option.  Returns (best time, best options, [(time, options)
    for every evaluation])."""
    tried = []

    def ev(opt):
        for t, o in tried:
            if o == opt:
                return t
        t = evaluate(opt)
        tried.append((t, dict(opt)))
        return t

    if space_size(space) <= EXHAUSTIVE_MAX:
        for opt in candidates(space):
            ev(opt)
        t, best = min(tried, key=lambda r: r[0])
        return t, best, tried
    best = {k: v[0] for k, v in space.items()}
    best_t = ev(best)
    for _ in range(passes):
        improved = False
        for k in sorted(space):
            for val in space[k]:
                if val == best[k]:
                    continue
                opt = dict(best)
                opt[k] = val
                t = ev(opt)
                if t < best_t:
                    best_t, best, improved = t, opt, True
        if not improved:
            break
    return best_t, best, tried


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
    spends most of its sweeps)."""
    key = signature(g) + (":l2" if second_level else "") + ":v2"
    cache = load_cache(cache_path) if cache_path else {}
    if key in cache:
        out = dict(cache[key])
        out["cached"] = True
        return out
    # one handle at level 0; with second_level another one at level 1
    G = L.Graph(g, device=device)
    G1, start1 = None, None
    try:
        if second_level:
            G1 = L.Graph(g, device=device)
            if G1.optimise(1e-6):
                G1.aggregate()
                start1 = np.arange(G1.level_size()[0], dtype=np.int32)
                G1.sweep()  # warm
        G.sweep()  # warm the kernels

        def evaluate(opt):
            G.set_options(**opt)
            t = timer(G, repeats)
            if start1 is not None:
                G1.set_options(**opt)
                t += time_level_sweep(G1, start1, repeats)
            return t

        default_ms = evaluate({})
        best_ms, best, tried = search(evaluate, space)
    finally:
        G.close()
        if G1 is not None:
            G1.close()
    out = {"options": _jsonable(best), "ms": best_ms, "default_ms": default_ms, "tried": len(tried),
           "gain": 1.0 - best_ms / default_ms if default_ms > 0 else 0.0, "signature": key,
           "cached": False}
    if cache_path:
        cache[key] = {k: v for k, v in out.items() if k != "cached"}
        save_cache(cache, cache_path)
    return out


def tune_batch(graphs, device: int = 0, space: dict = BATCH_SPACE, repeats: int = 2,
               min_gain: float = 1e-6, cache_path: str | None = DEFAULT_CACHE) -> dict:
    """Best options for batches shaped like `graphs` (a sample of the batch),
    by the time of whole batched runs (lv_batch_run), as tune()."""
    vptr, u = L.union(graphs)
    key = "batch:" + signature(u) + f":G{int(round(math.log2(max(len(graphs), 1))))}:v2"
    cache = load_cache(cache_path) if cache_path else {}
    if key in cache:
        out = dict(cache[key])
        out["cached"] = True
        return out
    with L.Batch(graphs, device=device) as B:
        B.run(min_gain)  # warm

        def evaluate(opt):
            B.set_options(**opt)
            ts = []
            for _ in range(repeats):
                ts.append(B.run(min_gain)[1]["algo_ms"])
            return float(np.median(ts))

        default_ms = evaluate({})
        best_ms, best, tried = search(evaluate, space)
        B.set_options()
    out = {"options": _jsonable(best), "ms": best_ms, "default_ms": default_ms, "tried": len(tried),
           "gain": 1.0 - best_ms / default_ms if default_ms > 0 else 0.0, "signature": key,
           "cached": False}
    if cache_path:
        cache[key] = {k: v for k, v in out.items() if k != "cached"}
        save_cache(cache, cache_path)
    return out
