"""Pins of the CPU restatement of the learned-frontier arithmetic (oracle/learn.py).

Each check fixes a property a plausible slip would break: bf16 rounding
against exactly representable and halfway values, the parameter layout
against a hand-built network, backpropagation against central differences,
Adam against its closed-form first step, the neighbour-move count against
a loop, and the features against hand-computed values.
"""
import numpy as np
import pytest

import graphs
import learn


def test_bf16_rounding_exact_cases():
    one = np.float32(1.0)
    cases = [
        (1.0, 1.0),
        (-2.5, -2.5),
        (1.0 + 2.0**-8, 1.0),                  # halfway, even neighbour is 1.0
        (1.0 + 3 * 2.0**-8, 1.0 + 2.0**-6),    # halfway, even neighbour is up
        (1.0 + 2.0**-8 + 2.0**-20, 1.0 + 2.0**-7),  # above halfway
        (0.0, 0.0),
    ]
    x = np.array([c[0] for c in cases], np.float32)
    want = np.array([c[1] for c in cases], np.float32)
    assert np.array_equal(learn.bf16(x), want)
    assert learn.bf16(np.array([one]))[0] == one


def test_layout_with_hand_built_network():
    W1 = np.zeros((32, 8)); b1 = np.zeros(32); W2 = np.zeros((32, 32)); b2 = np.zeros(32)
    w3 = np.zeros(32)
    W1[0, 3] = 1.0          # unit 0 reads input 3
    W1[5, 6] = -2.0         # unit 5 reads input 6 negatively
    b1[5] = 1.0
    W2[2, 0] = 1.0          # unit 2 of layer 2 reads unit 0 of layer 1
    W2[7, 5] = 0.5
    w3[2] = 2.0
    w3[7] = -1.0
    p = learn.pack(W1, b1, W2, b2, w3, 0.25)
    assert p.size == learn.N_PARAMS == 1377
    X = np.zeros((3, 8))
    X[0, 3] = 1.5
    X[1, 6] = 0.25
    X[2, 3] = -1.0
    X[2, 6] = 1.0
    # z = 2*relu(x3) - 0.5*relu(1 - 2*x6) + 0.25
    want = np.array([2 * 1.5 - 0.5 * 1.0 + 0.25, 0 - 0.5 * 0.5 + 0.25, 0 - 0 + 0.25])
    assert np.allclose(learn.logits(p, X), want, atol=0, rtol=0)


def test_backprop_matches_central_differences():
    rng = np.random.default_rng(3)
    p = rng.normal(0, 0.4, learn.N_PARAMS)
    X = rng.uniform(-1, 1, (64, 8))
    y = (rng.uniform(size=64) < 0.3).astype(np.float64)
    g = learn.grad(p, X, y, pos_weight=2.5)
    h = 1e-6
    idx = rng.choice(learn.N_PARAMS, 120, replace=False)
    idx = np.concatenate([idx, [0, 255, 256, 287, 288, 1311, 1312, 1344, 1345, 1376]])
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (learn.loss(p + e, X, y, 2.5) - learn.loss(p - e, X, y, 2.5)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_adam_first_step_closed_form():
    p = np.array([1.0, -1.0, 0.5])
    g = np.array([0.3, -2.0, 0.0])
    p1, m, v = learn.adam(p, g, np.zeros(3), np.zeros(3), 1, lr=0.01)
    # first step: m_hat = g, v_hat = g^2 -> step = lr * g / (|g| + eps)
    assert np.allclose(p1, p - 0.01 * np.sign(g), atol=1e-9)


def test_neighbor_moves_against_loop():
    rowptr, col, _ = graphs.random_graph(60, 200, 5, self_loops=4)
    rng = np.random.default_rng(1)
    moved = rng.uniform(size=60) < 0.3
    want = np.zeros(60, np.int32)
    for v in range(60):
        if moved[v]:
            for e in range(rowptr[v], rowptr[v + 1]):
                want[col[e]] += 1
    assert np.array_equal(learn.neighbor_moves(rowptr, col, moved), want)
    # movers above the degree cap do not count
    deg = np.diff(rowptr)
    hub = int(np.argmax(deg))
    moved2 = np.zeros(60, bool)
    moved2[hub] = True
    assert learn.neighbor_moves(rowptr, col, moved2, max_degree=int(deg[hub]) - 1).sum() == 0
    assert learn.neighbor_moves(rowptr, col, moved2, max_degree=int(deg[hub])).sum() == deg[hub]


def test_features_hand_computed_two_triangles():
    rowptr, col, w = graphs.two_triangles()
    V = 6
    k = np.diff(rowptr).astype(np.float64)   # unit weights
    M = k.sum() / 2                           # 7
    comm = np.arange(V)
    sigma = k.copy()
    size = np.ones(V, np.int32)
    margin = np.array([0.0, 0.0, 0.1, -0.2, 0.0, 0.0], np.float32)
    kown = np.zeros(V, np.float32)
    kown[3] = 1.0
    hist = np.array([0, 1, 0b1010, 0, 0, 0x80], np.uint8)
    nmv = np.array([1, 0, 2, 0, 0, 1], np.int32)
    x = learn.features(rowptr, k, comm, sigma, size, margin, kown, hist, nmv, M)
    assert x.shape == (6, 8) and x.dtype == np.float32
    assert np.isclose(x[0, 0], np.log2(3) / 16) and np.isclose(x[2, 0], 2.0 / 16)   # deg 2, 3
    assert x[3, 1] == np.float32(1.0 / 3.0)                                          # kown/k
    assert np.isclose(x[2, 2], 0.1 * 7 / 3 / 4, rtol=1e-6)                           # margin*M/k/4
    assert np.isclose(x[3, 2], -0.2 * 7 / 3 / 4, rtol=1e-6)
    assert x[1, 3] == 0.25 and x[2, 3] == 0.5 and x[5, 3] == 0.0                      # moves in last 4
    assert np.isclose(x[0, 4], 1.0 / 24)                                             # log2(1+1)/24
    assert np.isclose(x[2, 5], np.log2(1 + 3 * 6 / 14) / 24, rtol=1e-6)              # S_cur*V/2M
    assert x[0, 6] == 0.5 and x[2, 6] == np.float32(2 / 3) and x[1, 6] == 0.0          # nmv/deg
    assert x[0, 7] == 1.0 and x[1, 7] == 0.0 and x[2, 7] == 0.125 and x[5, 7] == 7 / 8  # since last move


def test_features_clamp_margin():
    rowptr, col, w = graphs.two_triangles()
    k = np.diff(rowptr).astype(np.float64)
    z = np.zeros(6, np.float32)
    margin = np.full(6, 100.0, np.float32)
    x = learn.features(rowptr, k, np.arange(6), k, np.ones(6, np.int32), margin, z, np.zeros(6, np.uint8),
                       np.zeros(6, np.int32), 7.0)
    assert np.all(x[:, 2] == 1.0)
