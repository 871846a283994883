"""CPU restatement of the learned-frontier arithmetic (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ and the smoke check as the
checker of the library's frontier kernels.  These components have no
counterpart in the reference; what pins each function is stated in its
docstring and exercised in tests/test_learn_cpu.py.

Parameter layout (1377 floats): W1[32][8], b1[32], W2[32][32], b2[32],
w3[32], b3; W[a][j] multiplies input j of unit a.
"""
from __future__ import annotations

import numpy as np

FEAT = 8
HID = 32
N_PARAMS = HID * FEAT + HID + HID * HID + HID + HID + 1


def unpack(params: np.ndarray):
    p = np.asarray(params, dtype=np.float64)
    assert p.size == N_PARAMS
    o = 0
    W1 = p[o:o + HID * FEAT].reshape(HID, FEAT); o += HID * FEAT
    b1 = p[o:o + HID]; o += HID
    W2 = p[o:o + HID * HID].reshape(HID, HID); o += HID * HID
    b2 = p[o:o + HID]; o += HID
    w3 = p[o:o + HID]; o += HID
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def pack(W1, b1, W2, b2, w3, b3) -> np.ndarray:
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, w3, [b3]])


def bf16(x: np.ndarray) -> np.ndarray:
    """Round fp32 values to bfloat16 (round to nearest, ties to even), as fp32.

    Pinned by exact cases: values representable in bf16 are unchanged, a
    halfway value rounds to the even neighbour."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def logits(params: np.ndarray, feat: np.ndarray, bf16_ops: bool = False) -> np.ndarray:
    """z = w3 . relu(W2 relu(W1 x + b1) + b2) + b3 in FP64.

    bf16_ops rounds the MFMA operands the inference kernel rounds (inputs,
    W1, layer-1 activations, W2) to bf16 first; everything else is exact."""
    W1, b1, W2, b2, w3, b3 = unpack(params)
    X = np.asarray(feat, dtype=np.float64).reshape(-1, FEAT)
    if bf16_ops:
        X = bf16(X.astype(np.float32)).astype(np.float64)
        W1 = bf16(W1.astype(np.float32)).astype(np.float64)
        W2 = bf16(W2.astype(np.float32)).astype(np.float64)
    h1 = np.maximum(X @ W1.T + b1, 0.0)
    if bf16_ops:
        h1 = bf16(h1.astype(np.float32)).astype(np.float64)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    return h2 @ w3 + b3


def forward(params: np.ndarray, feat: np.ndarray, bf16_ops: bool = False) -> np.ndarray:
    return 1.0 / (1.0 + np.exp(-logits(params, feat, bf16_ops)))


def loss(params: np.ndarray, feat: np.ndarray, labels: np.ndarray, pos_weight: float = 1.0) -> float:
    """Mean weighted BCE: -(pw*y*log p + (1-y)*log(1-p)), FP64, via softplus."""
    z = logits(params, feat)
    y = np.asarray(labels, dtype=np.float64)
    sp = lambda t: np.maximum(t, 0.0) + np.log1p(np.exp(-np.abs(t)))
    return float(np.mean(pos_weight * y * sp(-z) + (1.0 - y) * sp(z)))


def grad(params: np.ndarray, feat: np.ndarray, labels: np.ndarray, pos_weight: float = 1.0) -> np.ndarray:
    """Backpropagation of loss() in FP64 (pinned against central differences)."""
    W1, b1, W2, b2, w3, b3 = unpack(params)
    X = np.asarray(feat, dtype=np.float64).reshape(-1, FEAT)
    y = np.asarray(labels, dtype=np.float64)
    n = X.shape[0]
    a1 = X @ W1.T + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0.0)
    z = h2 @ w3 + b3
    p = 1.0 / (1.0 + np.exp(-z))
    dz = (pos_weight * y * (p - 1.0) + (1.0 - y) * p) / n
    gw3 = h2.T @ dz
    gb3 = dz.sum()
    d2 = np.outer(dz, w3) * (a2 > 0)
    gW2 = d2.T @ h1
    gb2 = d2.sum(0)
    d1 = (d2 @ W2) * (a1 > 0)
    gW1 = d1.T @ X
    gb1 = d1.sum(0)
    return pack(gW1, gb1, gW2, gb2, gw3, gb3)


def adam(params, g, m, v, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, Algorithm 1), FP64; returns (params, m, v)."""
    m = beta1 * m + (1 - beta1) * g
    v = beta2 * v + (1 - beta2) * g * g
    mh = m / (1 - beta1 ** t)
    vh = v / (1 - beta2 ** t)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v


NMV_MAX_DEGREE = 1024


def neighbor_moves(rowptr: np.ndarray, col: np.ndarray, moved: np.ndarray,
                   max_degree: int = NMV_MAX_DEGREE) -> np.ndarray:
    """nmv[u] = number of row entries (v, u) with v a mover (moved[v] != 0)
    of degree <= max_degree."""
    V = len(rowptr) - 1
    deg = np.diff(rowptr)
    src = np.repeat(np.arange(V), deg)
    sel = (np.asarray(moved, dtype=bool) & (deg <= max_degree))[src]
    return np.bincount(col[: len(src)][sel], minlength=V).astype(np.int32)


def features(rowptr, k, comm, sigma, size, margin, kown, hist, nmv, M) -> np.ndarray:
    """The 8 frontier features per vertex in fp32, from the state after a sweep.

    sigma (FP64 community totals) is rounded to fp32 first, as the library's
    community record holds it; M enters as fp32 and V/(2M) as fp32."""
    V = len(rowptr) - 1
    f32 = np.float32
    deg = np.diff(rowptr).astype(np.int64)
    kv = np.asarray(k, dtype=np.float64).astype(f32)
    cur = np.asarray(comm)
    scur = np.asarray(sigma, dtype=np.float64).astype(f32)[cur]
    sz = np.asarray(size)[cur].astype(f32)
    h = np.asarray(hist, dtype=np.uint32)
    M32 = f32(M)
    vscale = f32(V / (2.0 * M)) if M > 0 else f32(0)
    x = np.zeros((V, FEAT), f32)
    with np.errstate(divide="ignore", invalid="ignore"):
        x[:, 0] = np.log2(f32(1) + deg.astype(f32)) * f32(1 / 16)
        x[:, 1] = np.where(kv > 0, np.asarray(kown, f32) / kv, f32(0))
        x[:, 2] = np.where(kv > 0, np.clip(np.asarray(margin, f32) * M32 / kv, f32(-4), f32(4)) * f32(0.25),
                           f32(0))
        x[:, 3] = np.array([bin(int(t) & 0xF).count("1") for t in h], f32) * f32(0.25)
        x[:, 4] = np.log2(f32(1) + sz) * f32(1 / 24)
        x[:, 5] = np.log2(f32(1) + scur * vscale) * f32(1 / 24)
        x[:, 6] = np.where(deg > 0, np.minimum(f32(1), np.asarray(nmv).astype(f32) / deg.astype(f32)), f32(0))
    since = np.full(V, 8, np.int64)
    for b in range(7, -1, -1):
        since = np.where((h >> b) & 1, b, since)
    x[:, 7] = np.minimum(since, 8).astype(f32) * f32(0.125)
    return x
