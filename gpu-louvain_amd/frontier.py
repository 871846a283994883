"""Training of the learned frontier (move predictor) on plain-run traces.

The model (8 -> 32 -> 32 -> 1 MLP, see louvain_amd.Graph.set_model) predicts
whether a vertex moves in the next sweep from features of the state after
the last sweep.  Traces come from plain phases on the device
(lv_trace_phase): features before each sweep, label = moved in that sweep.
Training runs on the device with the library's forward/backward kernels and
Adam (lv_mlp_train).  No counterpart exists in the reference.
"""
from __future__ import annotations

import time

import numpy as np

import louvain_amd as L


def collect_traces(graphs, min_gain: float = 1e-6, per_sweep: int = 16384, max_sweeps: int = 48,
                   max_levels: int = 8, seed: int = 0, device: int = 0):
    """Features and labels from plain multi-level runs on `graphs` (CSRs).

    Every traced sweep contributes up to `per_sweep` vertices of positive
    degree, sampled uniformly; returns (X [n, 8] fp32, y [n] fp32, info)."""
    rng = np.random.default_rng(seed)
    xs, ys = [], []
    info = {"graphs": 0, "sweeps": 0, "samples": 0, "positives": 0, "seconds": 0.0}
    t0 = time.perf_counter()
    for csr in graphs:
        with L.Graph(csr, device=device) as G:
            for _level in range(max_levels):
                feat, moved, changed = G.trace_phase(min_gain, max_sweeps)
                # x0 = log2(1 + degree)/16 is positive exactly for degree > 0
                cand = np.nonzero(feat[0, :, 0] > 0)[0] if feat.shape[0] else np.zeros(0, np.int64)
                for s in range(feat.shape[0]):
                    if cand.size == 0:
                        break
                    pick = cand if cand.size <= per_sweep else rng.choice(cand, per_sweep, replace=False)
                    xs.append(feat[s, pick])
                    ys.append(moved[s, pick].astype(np.float32))
                    info["sweeps"] += 1
                if not changed:
                    break
                G.aggregate()
        info["graphs"] += 1
    X = np.concatenate(xs) if xs else np.zeros((0, L.FEATURES), np.float32)
    y = np.concatenate(ys) if ys else np.zeros(0, np.float32)
    info["samples"] = int(y.size)
    info["positives"] = int(y.sum())
    info["seconds"] = time.perf_counter() - t0
    return X, y, info


def train(X: np.ndarray, y: np.ndarray, epochs: int = 12, batch: int = 8192, lr: float = 2e-3,
          pos_weight: float | None = None, seed: int = 0, device: int = 0):
    """Shuffles, then trains on the device; returns (params, loss per epoch, info)."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(y.size)
    X = np.ascontiguousarray(X[perm])
    y = np.ascontiguousarray(y[perm])
    pos = float(y.sum())
    if pos_weight is None:
        pos_weight = float(np.clip((y.size - pos) / max(pos, 1.0), 1.0, 50.0))
    t0 = time.perf_counter()
    params, losses = L.mlp_train(X, y, L.mlp_init(seed), epochs=epochs, batch=batch, lr=lr,
                                 pos_weight=pos_weight, device=device)
    info = {"epochs": epochs, "batch": batch, "lr": lr, "pos_weight": pos_weight,
            "seconds": time.perf_counter() - t0, "loss_first": float(losses[0]) if len(losses) else None,
            "loss_last": float(losses[-1]) if len(losses) else None}
    return params, losses, info


def recall_threshold(params: np.ndarray, X: np.ndarray, y: np.ndarray, recall: float = 0.97,
                     device: int = 0) -> tuple[float, float]:
    """Largest threshold whose recall of movers on (X, y) is at least `recall`;
    returns (threshold, fraction of samples selected)."""
    p = L.mlp_forward(params, X, device=device)
    pos = np.sort(p[y > 0.5])
    if pos.size == 0:
        return 0.5, float(np.mean(p > 0.5))
    k = int(np.floor((1.0 - recall) * pos.size))
    thr = float(np.nextafter(pos[k], -np.inf)) if k < pos.size else 0.0
    thr = max(thr, 0.0)
    return thr, float(np.mean(p > thr))
