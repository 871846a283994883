"""GPU-vs-oracle parity at the full sizes of the benchmark configurations
(needs an MI355X).

- RMAT-20 (1,048,576 vertices, 31.4M CSR entries): the whole run, bit for
  bit, against the oracle run with several host threads.
- RMAT-22 and RMAT-24 (128.3M and 520.8M entries): given the same mid-run
  community state, every group's decisions and one sweep must equal the
  oracle's; the run's final Q must equal the FP64 modularity of its
  assignment on the original graph; every aggregated level must keep the
  total weight 2M; a second run must give the same assignment.
- 4,096 planted-partition graphs (36.6M vertices) solved as one batch: a
  sample of the graphs against single runs and against the oracle.

The generators' weights lie on the 2^-24 grid (or are unit integers), so all
FP64 sums are exact and decisions compare bit for bit.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

Q_TOL = 1e-9
THREADS = max(1, min(16, os.cpu_count() or 1))


def _say(msg):
    print(msg, flush=True)


@pytest.mark.timeout(600)
def test_rmat20_full_run_matches_oracle(lvlib, oracle):
    g = lvlib.gen_rmat(20, 16, 0.57, 0.19, 0.19, seed=20)
    _say(f"RMAT-20 V={g.V} E={g.E}: oracle run on {THREADS} threads")
    og = oracle.Graph.from_csr(g)
    st_o, o2c_o, lv_o = oracle.run(og, 1e-6, threads=THREADS)
    _say(f"oracle: {st_o['levels']} levels, {st_o['sweeps']} sweeps in {st_o['seconds']:.1f} s")
    with lvlib.Graph(g) as G:
        st = G.run(1e-6)
        o2c, lv = G.assignment()
    assert st["levels"] == st_o["levels"]
    assert st["sweeps"] == st_o["sweeps"]
    assert lv == lv_o
    assert np.array_equal(o2c, o2c_o)
    assert abs(st["modularity"] - st_o["modularity"]) <= Q_TOL


def _mid_state_and_run(lvlib, oracle, g, groups=range(7)):
    og = oracle.Graph.from_csr(g)
    M = oracle.total_weight(og)
    with lvlib.Graph(g) as G:
        # a mid-run state: two sweeps from singletons
        G.sweep()
        G.sweep()
        comm = G.communities().copy()
        for grp in groups:
            got = G.decide(grp)
            want = oracle.decide(og, comm, grp, M, threads=THREADS)
            bad = np.nonzero(got != want)[0]
            assert bad.size == 0, f"group {grp}: {bad.size} decisions differ, first {bad[:5]}"
        _say("decisions equal; one sweep from the same state")
        moved_o = oracle.sweep(og, comm, M, threads=THREADS)
        q_after, moved = G.sweep()
        assert moved == moved_o
        bad = np.nonzero(G.communities() != comm)[0]
        assert bad.size == 0, f"{bad.size} vertices differ after the sweep"
        assert abs(q_after - oracle.modularity(og, comm, M)) <= Q_TOL

        # the whole run, level by level
        G.reset()
        twoM = 2.0 * M
        levels = 0
        while True:
            levels += 1
            if not G.optimise(1e-6):
                break
            G.aggregate()
            h = G.level_graph()
            sw = float(np.sum(h.w, dtype=np.float64))
            # aggregated weights are fp32 sums: each rounds by at most 2^-24 relative
            assert abs(sw - twoM) <= 1e-6 * twoM, (levels, sw, twoM)
        q = G.modularity()
        o2c, lv = G.assignment()
        o2c = o2c.copy()
        G.reset()
        st = G.run(1e-6)
        o2c2, lv2 = G.assignment()
    assert st["levels"] == levels
    assert lv2 == lv and np.array_equal(o2c2, o2c)
    assert st["modularity"] == q
    # Q of the final level against the FP64 modularity of the assignment on
    # the original graph (they differ only by the fp32 rounding of merged weights)
    q0 = oracle.modularity(og, o2c, M)
    _say(f"Q run {q:.9f}, recomputed {q0:.9f}, {levels} levels, {lv} communities")
    assert abs(q - q0) <= 1e-6 * abs(q0)
    # every original vertex maps to a vertex of the final level
    assert o2c.min() >= 0 and o2c.max() < lv


@pytest.mark.timeout(600)
def test_rmat22_mid_state_and_run_properties(lvlib, oracle):
    g = lvlib.gen_rmat(22, 16, 0.57, 0.19, 0.19, seed=22)
    _say(f"RMAT-22 V={g.V} E={g.E}")
    _mid_state_and_run(lvlib, oracle, g)


@pytest.mark.timeout(900)
def test_rmat24_mid_state_and_run_properties(lvlib, oracle):
    g = lvlib.gen_rmat(24, 16, 0.57, 0.19, 0.19, seed=24)
    _say(f"RMAT-24 V={g.V} E={g.E}")
    _mid_state_and_run(lvlib, oracle, g)


@pytest.mark.timeout(600)
def test_batch_4096_planted_graphs(lvlib, oracle):
    rng = np.random.default_rng(4)
    gs = []
    for _ in range(4096):
        n = int(np.exp(rng.uniform(np.log(200), np.log(50000))))
        b = int(rng.integers(4, 65))
        p_in = min(1.0, 12.0 / max(n / b - 1.0, 1.0))
        p_out = 2.0 / max(n - n / b, 1.0)
        gs.append(lvlib.gen_planted(n, b, p_in, p_out, seed=int(rng.integers(1 << 62)), unit_weights=True))
    _say(f"batch of {len(gs)} graphs, V={sum(g.V for g in gs)} E={sum(g.E for g in gs)}")
    with lvlib.Batch(gs) as B:
        per, tot = B.run(1e-6)
        comms = B.assignment()
    assert tot["levels"] == max(p["levels"] for p in per)
    sample = list(range(0, 4096, 256)) + [int(np.argmax([g.V for g in gs]))]
    for i in sample:
        with lvlib.Graph(gs[i]) as G:
            st = G.run(1e-6)
            o2c, lv = G.assignment()
        assert per[i]["levels"] == st["levels"], i
        assert per[i]["sweeps"] == st["sweeps"], i
        assert per[i]["final_V"] == lv, i
        assert abs(per[i]["modularity"] - st["modularity"]) <= 1e-12, i
        assert np.array_equal(comms[i], o2c), i
    for i in sample[:6]:
        st_o, o2c_o, lv_o = oracle.run(oracle.Graph.from_csr(gs[i]), 1e-6, threads=THREADS)
        assert per[i]["final_V"] == lv_o, i
        assert np.array_equal(comms[i], o2c_o), i
        assert abs(per[i]["modularity"] - st_o["modularity"]) <= Q_TOL, i
