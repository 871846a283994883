"""Batched graphs (needs an MI355X): each graph of a batch must come out
exactly as a run of that graph alone (lv_run), and as the CPU oracle's run."""
import numpy as np
import pytest

import graphs

pytestmark = pytest.mark.gpu


def _mixed(lvlib, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        kind = i % 4
        if kind == 0:
            nv = int(np.exp(rng.uniform(np.log(60), np.log(3000))))
            b = int(rng.integers(4, 16))
            out.append(lvlib.gen_planted(nv, b, min(1.0, 8.0 / max(nv / b - 1, 1)), 2.0 / nv,
                                         seed=int(rng.integers(1 << 30)), unit_weights=True))
        elif kind == 1:
            out.append(lvlib.gen_rmat(int(rng.integers(7, 12)), 8, seed=int(rng.integers(1 << 30))))
        elif kind == 2:
            rp, cl, w = graphs.random_graph(int(rng.integers(20, 300)), int(rng.integers(20, 900)),
                                            int(rng.integers(1 << 30)), True, 3)
            # the union path needs sorted rows for nothing; keep insertion order
            out.append(lvlib.CSR(rp, cl, w))
        else:
            rp, cl, w = graphs.two_triangles()
            out.append(lvlib.CSR(rp, cl, w))
    return out


@pytest.mark.parametrize("n,seed", [(9, 1), (40, 2)])
def test_batch_matches_single_runs(lvlib, n, seed):
    gs = _mixed(lvlib, n, seed)
    print(f"batch of {len(gs)} graphs, V={sum(g.V for g in gs)}", flush=True)
    with lvlib.Batch(gs) as B:
        per, tot = B.run(1e-6)
        comms = B.assignment()
        # a second run from scratch gives the same results
        per2, _ = B.run(1e-6)
        assert per2 == per
    for i, g in enumerate(gs):
        with lvlib.Graph(g) as G:
            st = G.run(1e-6)
            o2c, lv = G.assignment()
        assert per[i]["levels"] == st["levels"], i
        assert per[i]["sweeps"] == st["sweeps"], i
        assert per[i]["final_V"] == lv, i
        assert abs(per[i]["modularity"] - st["modularity"]) <= 1e-12, (i, per[i], st["modularity"])
        assert np.array_equal(comms[i], o2c), i


def test_batch_matches_oracle(lvlib, oracle):
    gs = _mixed(lvlib, 12, 3)
    with lvlib.Batch(gs) as B:
        per, _ = B.run(1e-6)
        comms = B.assignment()
    for i, g in enumerate(gs):
        ost, o2c, lv = oracle.run(oracle.Graph.from_csr(g), 1e-6)
        assert per[i]["final_V"] == lv
        assert np.array_equal(comms[i], o2c)
        assert abs(per[i]["modularity"] - ost["modularity"]) <= 1e-9


@pytest.mark.parametrize("n,seed,thr", [(10, 4, 0.5), (30, 5, 0.35)])
def test_batch_pruned_matches_single_pruned_runs(lvlib, n, seed, thr):
    """Batched learned-pruned runs: every graph follows lv_run_pruned's rule
    on its own sweeps, with features from its own M and vertex count."""
    gs = _mixed(lvlib, n, seed)
    rng = np.random.default_rng(seed)
    params = rng.normal(0.0, 0.3, lvlib.MLP_PARAMS).astype(np.float32)
    fr = lvlib.frontier(kind=lvlib.FRONTIER_MODEL, threshold=thr, full_every=4)
    print(f"pruned batch of {len(gs)} graphs, V={sum(g.V for g in gs)}", flush=True)
    with lvlib.Batch(gs) as B:
        B.set_model(params)
        per, tot = B.run_pruned(1e-6, fr)
        comms = B.assignment()
        per2, _ = B.run_pruned(1e-6, fr)
        assert per2 == per
    assert tot["full_sweeps"] < sum(p["sweeps"] for p in per)  # some sweeps were pruned
    for i, g in enumerate(gs):
        with lvlib.Graph(g) as G:
            G.set_model(params)
            st = G.run_pruned(1e-6, fr)
            o2c, lv = G.assignment()
        assert per[i]["levels"] == st["levels"], i
        assert per[i]["sweeps"] == st["sweeps"], i
        assert per[i]["final_V"] == lv, i
        assert abs(per[i]["modularity"] - st["modularity"]) <= 1e-12, (i, per[i], st["modularity"])
        assert np.array_equal(comms[i], o2c), i


def test_batch_pruned_needs_the_model(lvlib):
    gs = _mixed(lvlib, 4, 6)
    with lvlib.Batch(gs) as B:
        with pytest.raises(lvlib.LouvainError) as ei:
            B.run_pruned(1e-6, lvlib.frontier(kind=lvlib.FRONTIER_MODEL))
        assert ei.value.status == lvlib.LV_ERR_STATE
        B.set_model(np.zeros(lvlib.MLP_PARAMS, np.float32))
        with pytest.raises(lvlib.LouvainError) as ei:
            B.run_pruned(1e-6, lvlib.frontier(kind=lvlib.FRONTIER_NEIGHBOR))
        assert ei.value.status == lvlib.LV_ERR_ARG


def test_batch_with_candidate_bounds_matches(lvlib):
    """Bounds forced on in a batch (per-vertex M in the bound) decide as without."""
    gs = _mixed(lvlib, 12, 7)
    with lvlib.Batch(gs, candidate_bounds=-1) as A, lvlib.Batch(gs, candidate_bounds=1) as B:
        pa, ta = A.run(1e-6)
        pb, tb = B.run(1e-6)
        assert pa == pb
        ca, cb = A.assignment(), B.assignment()
        for i in range(len(gs)):
            assert np.array_equal(ca[i], cb[i]), i
