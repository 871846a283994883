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
