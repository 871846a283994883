"""Parity on real (non-dyadic) fp32 weights (needs an MI355X).

The library rounds every weight to a grid of 2^-s (lv_weight_grid) on which
every FP64 sum of weights it forms is exact, so no result depends on the
order in which atomics land.  On inputs whose weights are off the
generators' 2^-24 grid and span many magnitudes these tests check:
  - bit-exact parity with the oracle on the same rounded graph: per-group
    decisions from singletons and from a mid-run state, sweeps, a phase, the
    aggregated CSR and whole runs (Q to 1e-9);
  - against the oracle on the weights as given (no rounding), decisions
    differ only where the oracle's two best options are within 1e-6
    relative of each other (the north-star tie clause), and a run's Q is
    within 1e-6 relative of the FP64 modularity of its assignment;
  - determinism: fresh handles, other kernel tiers (chunked tier forced,
    hash aggregation, incremental modularity, candidate bounds) and a batch
    all give the same assignment and the same Q bit for bit;
  - decimal weights read from a MatrixMarket file by lv_read_mtx.
"""
import os

import numpy as np
import pytest

import graphs

pytestmark = pytest.mark.gpu

Q_TOL = 1e-9
TIE_REL = 1e-6


def _with_weights(lvlib, g, seed, power):
    w = graphs.real_weights(g.rowptr, g.col, seed=seed, power=power)
    return lvlib.CSR(g.rowptr, g.col, w)


def _decimal_graph(lvlib, tmp_dir):
    """A random graph whose weights are decimal strings of mixed magnitude,
    written to MatrixMarket and read back by the library's reader."""
    rng = np.random.default_rng(71)
    n, m = 3000, 24000
    seen, edges = set(), []
    while len(edges) < m:
        i, j = (int(x) for x in rng.integers(0, n, 2))
        key = (min(i, j), max(i, j))
        if i == j or key in seen:
            continue
        seen.add(key)
        edges.append(key)
    pool = ["0.1", "0.3", "0.001", "2.5", "1e-5", "7.25", "0.333333", "0.7", "12.125", "3e-4"]
    weights = [pool[int(x)] for x in rng.integers(0, len(pool), m)]
    # a few self-loops (stored once in their row)
    loops = [(int(v), int(v)) for v in rng.choice(n, 40, replace=False)]
    path = os.path.join(tmp_dir, "decimal_weights.mtx")
    graphs.write_decimal_mtx(path, n, edges + loops, weights + ["0.05"] * len(loops))
    return path


@pytest.fixture(scope="module")
def cases(lvlib, tmp_path_factory):
    out = {}
    out["rmat14_real"] = (_with_weights(lvlib, lvlib.gen_rmat(14, 16, seed=14), 3, 3.0), {})
    out["planted_real"] = (_with_weights(lvlib, lvlib.gen_planted(6000, 30, 0.02, 0.0008, seed=8), 5, 2.0), {})
    out["rmat13_real_tiers"] = (_with_weights(lvlib, lvlib.gen_rmat(13, 16, seed=31), 7, 3.0),
                                {"heavy_wg_max": 320, "chunk_edges": 64, "part_target": 64,
                                 "candidate_bounds": 1, "q_incremental": 1})
    path = _decimal_graph(lvlib, str(tmp_path_factory.mktemp("realw")))
    out["decimal_mtx"] = (lvlib.read_mtx(path), {})
    return out


NAMES = ["rmat14_real", "planted_real", "rmat13_real_tiers", "decimal_mtx"]


@pytest.mark.parametrize("name", NAMES)
def test_weights_are_off_the_dyadic_grid_and_rounded_like_the_oracle(lvlib, oracle, cases, name):
    g, opt = cases[name]
    print(f"case {name} V={g.V} E={g.E}", flush=True)
    og = oracle.Graph.from_csr(g)
    assert og.changed > 0, "the input should need the grid"
    with lvlib.Graph(g, **opt) as G:
        h = G.level_graph()
    rp, cl, w = og.csr_arrays()
    assert np.array_equal(h.rowptr, rp) and np.array_equal(h.col, cl)
    assert np.array_equal(h.w.view(np.uint32), w.view(np.uint32))


@pytest.mark.parametrize("name", NAMES)
def test_decisions_bit_exact_on_the_rounded_graph(lvlib, oracle, cases, name):
    g, opt = cases[name]
    print(f"case {name}", flush=True)
    og = oracle.Graph.from_csr(g)
    mid = np.arange(g.V, dtype=np.int32)
    oracle.sweep(og, mid)
    with lvlib.Graph(g, **opt) as G:
        for state in ("singletons", "mid"):
            comm = np.arange(g.V, dtype=np.int32) if state == "singletons" else mid
            G.set_communities(comm)
            for grp in range(7):
                got = G.decide(grp)
                want = oracle.decide(og, comm, grp)
                bad = np.nonzero(got != want)[0]
                assert bad.size == 0, f"{state} group {grp}: {bad.size} differ, first {bad[:5]}"


@pytest.mark.parametrize("name", NAMES)
def test_decisions_vs_unrounded_oracle_within_tie_clause(lvlib, oracle, cases, name):
    g, opt = cases[name]
    print(f"case {name}", flush=True)
    raw = oracle.Graph.from_csr(g, quantize=False)
    mid = np.arange(g.V, dtype=np.int32)
    oracle.sweep(raw, mid)
    total = differ = 0
    with lvlib.Graph(g, **opt) as G:
        for state in ("singletons", "mid"):
            comm = np.arange(g.V, dtype=np.int32) if state == "singletons" else mid
            G.set_communities(comm)
            for grp in range(7):
                got = G.decide(grp)
                want, gap = oracle.decide_gap(raw, comm, grp)
                bad = np.nonzero(got != want)[0]
                total += int(np.count_nonzero(~np.isnan(gap)))
                differ += bad.size
                assert np.all(gap[bad] < TIE_REL), \
                    f"{state} group {grp}: {bad.size} differ, gaps {gap[bad][:5]}"
    print(f"{name}: {differ} of {total} decisions differ from the unrounded oracle, all within the "
          f"tie clause", flush=True)


@pytest.mark.parametrize("name", NAMES)
def test_sweeps_phase_and_aggregation_match(lvlib, oracle, cases, name):
    g, opt = cases[name]
    print(f"case {name}", flush=True)
    og = oracle.Graph.from_csr(g)
    M = oracle.total_weight(og)
    comm = np.arange(g.V, dtype=np.int32)
    with lvlib.Graph(g, **opt) as G:
        for s in range(3):
            moved_o = oracle.sweep(og, comm, M)
            q, moved = G.sweep()
            assert moved == moved_o, f"sweep {s}"
            assert np.array_equal(G.communities(), comm), f"sweep {s}"
            assert abs(q - oracle.modularity(og, comm, M)) <= Q_TOL
        G.reset()
        comm = np.arange(g.V, dtype=np.int32)
        o2c = np.arange(g.V, dtype=np.int32)
        changed_o, _ = oracle.phase(og, comm, 1e-6, o2c, M)
        assert G.optimise(1e-6) == changed_o
        assert np.array_equal(G.communities(), comm)
        G.aggregate()
        nxt, _ = oracle.aggregate(og, comm, o2c)
        h = G.level_graph()
        rp, cl, w = nxt.csr_arrays()
        assert np.array_equal(h.rowptr, rp) and np.array_equal(h.col, cl)
        assert np.array_equal(h.w.view(np.uint32), w.view(np.uint32))


@pytest.mark.parametrize("name", NAMES)
def test_runs_match_the_oracle_and_repeat_exactly(lvlib, oracle, cases, name):
    g, opt = cases[name]
    print(f"case {name}", flush=True)
    og = oracle.Graph.from_csr(g)
    st_o, o2c_o, lv_o = oracle.run(og, 1e-6)
    runs = []
    variants = [opt, opt, {"heavy_wg_max": 320, "chunk_edges": 128, "agg_mode": 1, "q_incremental": 1,
                           "candidate_bounds": 1, "xcd_rows": 1}]
    for o in variants:
        with lvlib.Graph(g, **o) as G:
            st = G.run(1e-6)
            runs.append((st, *G.assignment()))
    st, o2c, lv = runs[0]
    assert (st["levels"], st["sweeps"], lv) == (st_o["levels"], st_o["sweeps"], lv_o)
    assert np.array_equal(o2c, o2c_o)
    assert abs(st["modularity"] - st_o["modularity"]) <= Q_TOL
    for st2, o2c2, lv2 in runs[1:]:
        assert st2["modularity"] == st["modularity"] and st2["sweeps"] == st["sweeps"] and lv2 == lv
        assert np.array_equal(o2c2, o2c)
    # against the weights as given: the FP64 modularity of the assignment
    raw = oracle.Graph.from_csr(g, quantize=False)
    q_raw = oracle.modularity(raw, o2c.astype(np.int32))
    assert abs(q_raw - st["modularity"]) <= 1e-6 * abs(q_raw)


def test_batch_of_real_weight_graphs_equals_single_runs(lvlib, oracle):
    gs = []
    for i in range(6):
        base = lvlib.gen_planted(600 + 250 * i, 4 + i, 0.05, 0.002, seed=90 + i)
        # graphs of different total weight get different grids
        w = graphs.real_weights(base.rowptr, base.col, seed=i, power=1.0 + i) * np.float32(1.0 + 7 * i)
        gs.append(lvlib.CSR(base.rowptr, base.col, w.astype(np.float32)))
    shifts = [lvlib.weight_grid(g.w)[0] for g in gs]
    assert len(set(shifts)) > 1
    print(f"batch of {len(gs)} graphs, shifts {shifts}", flush=True)
    with lvlib.Batch(gs) as B:
        per, _ = B.run(1e-6)
        comms = B.assignment()
    for i, g in enumerate(gs):
        with lvlib.Graph(g) as G:
            st = G.run(1e-6)
            o2c, lv = G.assignment()
        assert per[i]["modularity"] == st["modularity"] and per[i]["sweeps"] == st["sweeps"]
        assert np.array_equal(comms[i], o2c)
        st_o, o2c_o, _ = oracle.run(oracle.Graph.from_csr(g), 1e-6)
        assert np.array_equal(o2c, o2c_o) and abs(st["modularity"] - st_o["modularity"]) <= Q_TOL


def test_weights_as_given_stay_within_the_tie_clause(lvlib, oracle, cases):
    """exact_sums = -1 keeps the weights unrounded; decisions then follow the
    unrounded oracle up to the tie clause."""
    g, _ = cases["rmat14_real"]
    raw = oracle.Graph.from_csr(g, quantize=False)
    comm = np.arange(g.V, dtype=np.int32)
    with lvlib.Graph(g, exact_sums=-1) as G:
        h = G.level_graph()
        assert np.array_equal(h.w.view(np.uint32), np.asarray(g.w, np.float32).view(np.uint32))
        for grp in range(7):
            got = G.decide(grp)
            want, gap = oracle.decide_gap(raw, comm, grp)
            bad = np.nonzero(got != want)[0]
            assert np.all(gap[bad] < TIE_REL)
        st = G.run(1e-6)
        o2c, _ = G.assignment()
    q_raw = oracle.modularity(raw, o2c.astype(np.int32))
    assert abs(q_raw - st["modularity"]) <= 1e-6 * abs(q_raw)
