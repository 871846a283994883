"""GPU checks of the learned frontier (needs an MI355X).

The MFMA inference is compared with the FP64 forward pass on bf16-rounded
operands (and, loosely, on exact operands); the training gradient with FP64
backpropagation; Adam steps with the FP64 update; features, move margins
and neighbour-move counts with the CPU restatement given the same state;
pruned sweeps with the oracle's masked sweep given the same frontier.
"""
import numpy as np
import pytest

import graphs
import learn

pytestmark = pytest.mark.gpu


def _params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = rng.normal(0.0, 0.35 * scale, learn.N_PARAMS).astype(np.float32)
    return p


def _feat(n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-0.2, 1.0, (n, 8)).astype(np.float32)


@pytest.mark.parametrize("n", [1, 63, 1000, 70001])
def test_mlp_forward_mfma_matches_bf16_forward(lvlib, n):
    print(f"mlp forward n={n}", flush=True)
    p = _params(1)
    X = _feat(n, 2)
    got = lvlib.mlp_forward(p, X)
    want = learn.forward(p, X, bf16_ops=True)
    err = np.abs(got - want)
    # fp32 accumulation against FP64 on identical bf16 operands; a layer-1
    # activation on a bf16 rounding boundary may land on the other side
    assert np.median(err) < 1e-6
    assert err.max() < 5e-3
    assert np.mean(err > 1e-5) < 0.01
    exact = learn.forward(p, X)
    assert np.abs(got - exact).max() < 0.05


def test_mlp_forward_layout_hand_built(lvlib):
    # W[a][j] multiplies input j of unit a; a transposed layout fails this
    W1 = np.zeros((32, 8)); b1 = np.zeros(32); W2 = np.zeros((32, 32)); b2 = np.zeros(32); w3 = np.zeros(32)
    W1[0, 3] = 1.0; W1[17, 6] = -2.0; b1[17] = 1.0
    W2[2, 0] = 1.0; W2[29, 17] = 0.5; w3[2] = 2.0; w3[29] = -1.0
    p = learn.pack(W1, b1, W2, b2, w3, 0.25).astype(np.float32)
    X = np.zeros((5, 8), np.float32)
    X[0, 3] = 1.5; X[1, 6] = 0.25; X[2, 3] = -1.0; X[2, 6] = 1.0; X[3, 3] = 0.5; X[4, 6] = -0.5
    z = learn.logits(p, X)
    got = lvlib.mlp_forward(p, X)
    assert np.allclose(got, 1 / (1 + np.exp(-z)), atol=1e-6)


@pytest.mark.parametrize("n,pw", [(1, 1.0), (300, 2.5), (20000, 4.0)])
def test_mlp_gradient_matches_fp64_backprop(lvlib, n, pw):
    print(f"mlp grad n={n}", flush=True)
    p = _params(5)
    X = _feat(n, 6)
    y = (np.random.default_rng(7).uniform(size=n) < 0.3).astype(np.float32)
    g, loss = lvlib.mlp_grad(p, X, y, pos_weight=pw)
    want = learn.grad(p, X, y, pos_weight=pw)
    scale = np.abs(want).max()
    assert np.abs(g - want).max() <= 2e-5 * scale + 1e-7
    assert abs(loss - learn.loss(p, X, y, pw)) <= 1e-5 * abs(learn.loss(p, X, y, pw))


def test_mlp_adam_steps_match_fp64(lvlib):
    n = 4096
    p0 = _params(9)
    X = _feat(n, 10)
    y = (X[:, 0] + 0.5 * X[:, 3] > 0.6).astype(np.float32)
    lr = 1e-2
    got, losses = lvlib.mlp_train(X, y, p0, epochs=3, batch=n, lr=lr, pos_weight=1.5)
    p = p0.astype(np.float64)
    m = np.zeros_like(p); v = np.zeros_like(p)
    for t in range(1, 4):
        p, m, v = learn.adam(p, learn.grad(p, X, y, 1.5), m, v, t, lr)
    d = np.abs(got - p)
    assert np.quantile(d, 0.99) < 1e-3 * lr
    assert d.max() < 0.5 * lr
    assert losses[2] < losses[0]


def test_mlp_training_learns_a_separable_rule(lvlib):
    rng = np.random.default_rng(11)
    n = 60000
    X = rng.uniform(0, 1, (n, 8)).astype(np.float32)
    y = ((X[:, 1] > 0.5) & (X[:, 6] > 0.3)).astype(np.float32)
    p, losses = lvlib.mlp_train(X, y, lvlib.mlp_init(0), epochs=30, batch=2048, lr=3e-3)
    assert losses[-1] < 0.25 * losses[0]
    pred = lvlib.mlp_forward(p, X) > 0.5
    assert np.mean(pred == (y > 0.5)) > 0.97




@pytest.fixture
def full_frontier(lvlib):
    return lvlib.frontier(kind=lvlib.FRONTIER_ALL)


@pytest.mark.parametrize("case", ["planted", "rmat13", "rmat14_chunked"])
def test_features_and_observations_match_oracle(lvlib, oracle, full_frontier, case):
    print(f"features case {case}", flush=True)
    if case == "planted":
        g, opt = lvlib.gen_planted(4000, 16, 0.02, 0.001, seed=21, unit_weights=True), {}
    elif case == "rmat13":
        g, opt = lvlib.gen_rmat(13, 16, seed=22, unit_weights=True), {}
    else:
        g, opt = lvlib.gen_rmat(14, 16, seed=23), {"heavy_wg_max": 320, "chunk_edges": 64, "part_target": 64}
    og = oracle.Graph.from_csr(g)
    M = oracle.total_weight(og)
    k = oracle.strength(og)
    with lvlib.Graph(g, **opt) as G:
        for s in range(3):
            before = G.communities().copy()
            G.sweep_pruned(full_frontier, track=True)
            after = G.communities()
            st = G.move_state()
            moved = before != after
            assert np.array_equal(st["hist"] & 1, moved.astype(np.uint8)), f"sweep {s}: move flags"
            assert np.array_equal(st["nmv"], learn.neighbor_moves(g.rowptr, g.col, moved)), f"sweep {s}: nmv"
            sigma = np.bincount(after, weights=k, minlength=g.V)
            size = np.bincount(after, minlength=g.V).astype(np.int32)
            want = learn.features(g.rowptr, k, after, sigma, size, st["margin"], st["kown"], st["hist"],
                                  st["nmv"], M)
            got = G.features()
            assert np.abs(got - want).max() <= 2e-6, f"sweep {s}: max feature error {np.abs(got - want).max()}"


@pytest.mark.parametrize("side", [1, -1])
@pytest.mark.parametrize("case", ["planted", "rmat14_chunked"])
def test_neighbor_moves_both_sides_match_oracle(lvlib, full_frontier, case, side):
    # nmv counted from the rows' side (1: every row counts its moved
    # neighbours, heavy rows by chunks) and from the movers' side (-1)
    print(f"nmv case {case} side {side}", flush=True)
    if case == "planted":
        g, opt = lvlib.gen_planted(4000, 16, 0.02, 0.001, seed=24), {}
    else:
        g, opt = lvlib.gen_rmat(14, 16, seed=25), {"heavy_wg_max": 320, "chunk_edges": 64,
                                                   "part_target": 64}
    with lvlib.Graph(g, nmv_pull=side, **opt) as G:
        for s in range(4):
            before = G.communities().copy()
            G.sweep_pruned(full_frontier, track=True)
            moved = before != G.communities()
            st = G.move_state()
            want = learn.neighbor_moves(g.rowptr, g.col, moved)
            assert np.array_equal(st["nmv"], want), f"sweep {s}: nmv"


@pytest.mark.parametrize("case", ["planted", "rmat13_unit", "rmat14_chunked"])
def test_move_margins_match_oracle(lvlib, oracle, case):
    print(f"margins case {case}", flush=True)
    if case == "planted":
        g, opt = lvlib.gen_planted(4000, 16, 0.02, 0.001, seed=31, unit_weights=True), {}
    elif case == "rmat13_unit":
        g, opt = lvlib.gen_rmat(13, 16, seed=32, unit_weights=True), {}
    else:
        g, opt = lvlib.gen_rmat(14, 16, seed=33), {"heavy_wg_max": 320, "chunk_edges": 64, "part_target": 64}
    og = oracle.Graph.from_csr(g)
    with lvlib.Graph(g, **opt) as G:
        G.sweep()   # a mid state: non-singleton communities
        comm = G.communities().copy()
        deg = np.diff(g.rowptr)
        for grp in range(7):
            G.decide(grp)
            st = G.move_state()
            _, margin, kown = oracle.decide_obs(og, comm, grp)
            sel = np.array([oracle.group_of_degree(int(d)) == grp for d in deg])
            if not sel.any():
                continue
            assert np.array_equal(st["kown"][sel], kown[sel]), f"group {grp}: kown"
            assert np.array_equal(st["margin"][sel], margin[sel]), f"group {grp}: margin"


def _mix64(x):
    m = (1 << 64) - 1
    x = (x + 0x9E3779B97F4A7C15) & m
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & m
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & m
    return x ^ (x >> 31)


def test_pruned_all_is_a_plain_sweep(lvlib, full_frontier):
    g = lvlib.gen_rmat(13, 16, seed=40)
    with lvlib.Graph(g) as A, lvlib.Graph(g) as B:
        for s in range(4):
            qa, ma = A.sweep()
            qb, mb, ev = B.sweep_pruned(full_frontier, track=(s % 2 == 0))
            assert ma == mb and qa == qb
            assert np.array_equal(A.communities(), B.communities())
            assert ev == int(np.sum(np.diff(g.rowptr) > 0))


@pytest.mark.parametrize("kind", ["neighbor", "random", "model"])
def test_pruned_sweep_matches_masked_oracle(lvlib, oracle, full_frontier, kind):
    print(f"pruned sweep kind {kind}", flush=True)
    g = lvlib.gen_rmat(14, 16, seed=50)
    opt = {"heavy_wg_max": 320, "chunk_edges": 128, "part_target": 64}
    og = oracle.Graph.from_csr(g)
    deg = np.diff(g.rowptr)
    with lvlib.Graph(g, **opt) as G:
        G.set_model(_params(3, scale=0.6))
        G.sweep_pruned(full_frontier, track=True)
        G.sweep_pruned(full_frontier, track=True)
        for s in range(3):
            comm = G.communities().copy()
            st = G.move_state()
            if kind == "neighbor":
                fr = lvlib.frontier(kind=lvlib.FRONTIER_NEIGHBOR)
                mask = ((st["hist"] & 1) > 0) | (st["nmv"] > 0)
            elif kind == "random":
                fr = lvlib.frontier(kind=lvlib.FRONTIER_RANDOM, random_p=0.3, seed=77)
                # the library's sweep counter: 2 tracked sweeps before, then s
                seed = (77 + 0x9E37 * (2 + s)) & ((1 << 64) - 1)
                mask = np.array([np.float32((_mix64(seed ^ _mix64(v)) >> 40) * 2.0**-24) < np.float32(0.3) for v in range(g.V)])
            else:
                fr = lvlib.frontier(kind=lvlib.FRONTIER_MODEL, threshold=0.5)
                _, act = G.predict(0.5)
                mask = act > 0
            mask = mask & (deg > 0)
            q, moved, ev = G.sweep_pruned(fr, track=True)
            assert ev == int(mask.sum()), f"sweep {s}: evaluated {ev} vs mask {int(mask.sum())}"
            want = comm.copy()
            wm = oracle.sweep_masked(og, want, mask.astype(np.uint8))
            got = G.communities()
            assert wm == moved
            bad = np.nonzero(got != want)[0]
            assert bad.size == 0, f"sweep {s}: {bad.size} differ"
            assert abs(q - oracle.modularity(og, want)) <= 1e-9
