"""GPU-vs-oracle parity of the HIP Louvain path (needs an MI355X).

Every check compares the native library (called through the C-ABI) with
the CPU oracle on the same seeded inputs.  The inputs use integer weights or
weights on the 2^-24 grid, where all FP64 sums are exact, so decisions,
sweeps, aggregated CSRs and whole runs must match bit for bit; modularity is
compared to 1e-9 absolute (the bar is 1e-6 relative).
"""
import numpy as np
import pytest

import graphs

pytestmark = pytest.mark.gpu

Q_TOL = 1e-9


def _csr(lvlib, rowptr, col, w):
    return lvlib.CSR(np.asarray(rowptr, np.int32), np.asarray(col, np.int32), np.asarray(w, np.float32))


def _cases(lvlib):
    """(name, CSR, graph options) of the parity inputs."""
    out = []
    out.append(("two_triangles", _csr(lvlib, *graphs.two_triangles()), {}))
    out.append(("planted_c1_unit", lvlib.gen_planted(10000, 20, 0.02, 0.0005, seed=1, unit_weights=True), {}))
    out.append(("rmat12_real", lvlib.gen_rmat(12, 16, seed=12), {}))
    out.append(("rmat14_real", lvlib.gen_rmat(14, 16, seed=14), {}))
    out.append(("rmat13_unit", lvlib.gen_rmat(13, 16, seed=13, unit_weights=True), {}))
    out.append(("random_selfloops_real", _csr(lvlib, *graphs.random_graph(400, 3000, 7, True, 40)), {}))
    out.append(("star3000", _csr(lvlib, *graphs.star(3000)), {}))
    out.append(("star20000", _csr(lvlib, *graphs.star(20000)), {}))
    # force the chunked tier (many chunks and partitions) on a mid-size graph
    out.append(("rmat14_chunked_tier", lvlib.gen_rmat(14, 16, seed=41),
                {"heavy_wg_max": 320, "chunk_edges": 64, "part_target": 64}))
    # modularity carried by the applies' corrections after every sweep
    out.append(("rmat14_incremental_q", lvlib.gen_rmat(14, 16, seed=44),
                {"q_incremental": 1, "heavy_wg_max": 320, "chunk_edges": 256}))
    out.append(("planted_incremental_q", lvlib.gen_planted(5000, 25, 0.03, 0.001, seed=45),
                {"q_incremental": 1}))
    # force the chunked tier's global-table fallback (LDS merge limit 16 pairs)
    out.append(("rmat13_merge_overflow", lvlib.gen_rmat(13, 16, seed=43),
                {"heavy_wg_max": 320, "chunk_edges": 128, "part_target": 100000, "merge_limit": 16}))
    # aggregation by per-community hash merges instead of the radix sort
    out.append(("rmat14_hash_agg", lvlib.gen_rmat(14, 16, seed=47), {"agg_mode": 1}))
    out.append(("planted_hash_agg", lvlib.gen_planted(8000, 40, 0.02, 0.0008, seed=48), {"agg_mode": 1}))
    # the column-blocked (XCD-local) modularity pass forced on a small graph
    out.append(("rmat14_xcd_q", lvlib.gen_rmat(14, 16, seed=46), {"xcd_rows": 1}))
    # the same two with the XCD-aligned chunked tier (column-block chunks,
    # partitions by community-id range)
    out.append(("rmat14_chunked_tier_xcd", lvlib.gen_rmat(14, 16, seed=41),
                {"heavy_wg_max": 320, "chunk_edges": 64, "part_target": 64, "xcd_rows": 2}))
    out.append(("rmat13_merge_overflow_xcd", lvlib.gen_rmat(13, 16, seed=43),
                {"heavy_wg_max": 320, "chunk_edges": 128, "part_target": 100000, "merge_limit": 16,
                 "xcd_rows": 2}))
    # candidate bounds forced on (by default only levels of >= 2^20 vertices)
    out.append(("rmat14_bounds", lvlib.gen_rmat(14, 16, seed=48), {"candidate_bounds": 1}))
    out.append(("planted_bounds", lvlib.gen_planted(8000, 40, 0.02, 0.0008, seed=49),
                {"candidate_bounds": 1}))
    out.append(("rmat14_tiers_bounds", lvlib.gen_rmat(14, 16, seed=41),
                {"heavy_wg_max": 512, "chunk_edges": 128, "candidate_bounds": 1}))
    # bucket-ordered columns for the move kernels' gathers (by default only on
    # levels of >= 2^20 vertices) forced on: every tier, the chunked tier,
    # self-loops, incremental modularity
    out.append(("rmat14_col_order", lvlib.gen_rmat(14, 16, seed=50), {"col_order": 1}))
    out.append(("rmat14_chunked_col_order", lvlib.gen_rmat(14, 16, seed=41),
                {"heavy_wg_max": 320, "chunk_edges": 64, "part_target": 64, "col_order": 1}))
    out.append(("planted_col_order", lvlib.gen_planted(8000, 40, 0.02, 0.0008, seed=51),
                {"col_order": 1}))
    out.append(("random_selfloops_col_order", _csr(lvlib, *graphs.random_graph(400, 3000, 7, True, 40)),
                {"col_order": 1}))
    out.append(("star20000_col_order", _csr(lvlib, *graphs.star(20000)), {"col_order": 1}))
    out.append(("rmat14_incremental_col_order", lvlib.gen_rmat(14, 16, seed=44),
                {"q_incremental": 1, "heavy_wg_max": 320, "chunk_edges": 256, "col_order": 1,
                 "candidate_bounds": 1}))
    return out


@pytest.fixture(scope="module")
def cases(lvlib):
    return _cases(lvlib)


def _names():
    return ["two_triangles", "planted_c1_unit", "rmat12_real", "rmat14_real", "rmat13_unit",
            "random_selfloops_real", "star3000", "star20000", "rmat14_chunked_tier",
            "rmat13_merge_overflow", "rmat14_incremental_q", "planted_incremental_q",
            "rmat14_xcd_q", "rmat14_chunked_tier_xcd", "rmat13_merge_overflow_xcd",
            "rmat14_hash_agg", "planted_hash_agg", "rmat14_bounds", "planted_bounds",
            "rmat14_tiers_bounds", "rmat14_col_order", "rmat14_chunked_col_order",
            "planted_col_order", "random_selfloops_col_order", "star20000_col_order",
            "rmat14_incremental_col_order"]


def _get(cases, name):
    for n, g, opt in cases:
        if n == name:
            return g, opt
    raise KeyError(name)


@pytest.mark.parametrize("name", _names())
def test_decide_singletons_matches_oracle(lvlib, oracle, cases, name):
    g, opt = _get(cases, name)
    print(f"case {name} V={g.V} E={g.E}", flush=True)
    og = oracle.Graph.from_csr(g)
    comm = np.arange(g.V, dtype=np.int32)
    with lvlib.Graph(g, **opt) as G:
        for grp in range(7):
            got = G.decide(grp)
            want = oracle.decide(og, comm, grp)
            bad = np.nonzero(got != want)[0]
            assert bad.size == 0, f"group {grp}: {bad.size} differ, first {bad[:5]} got {got[bad[:5]]} want {want[bad[:5]]}"


@pytest.mark.parametrize("name", _names())
def test_decide_mid_state_matches_oracle(lvlib, oracle, cases, name):
    g, opt = _get(cases, name)
    print(f"case {name}", flush=True)
    og = oracle.Graph.from_csr(g)
    comm = np.arange(g.V, dtype=np.int32)
    oracle.sweep(og, comm)  # a non-trivial state reached by the oracle
    with lvlib.Graph(g, **opt) as G:
        G.set_communities(comm)
        assert np.array_equal(G.communities(), comm)
        for grp in range(7):
            got = G.decide(grp)
            want = oracle.decide(og, comm, grp)
            bad = np.nonzero(got != want)[0]
            assert bad.size == 0, f"group {grp}: {bad.size} differ, first {bad[:5]}"


@pytest.mark.parametrize("name", _names())
def test_sweeps_and_modularity_match_oracle(lvlib, oracle, cases, name):
    g, opt = _get(cases, name)
    print(f"case {name}", flush=True)
    og = oracle.Graph.from_csr(g)
    M = oracle.total_weight(og)
    comm = np.arange(g.V, dtype=np.int32)
    with lvlib.Graph(g, **opt) as G:
        assert abs(G.modularity() - oracle.modularity(og, comm, M)) <= Q_TOL
        for s in range(4):
            moved_o = oracle.sweep(og, comm, M)
            q_after, moved = G.sweep()
            assert moved == moved_o, f"sweep {s}: moved {moved} vs {moved_o}"
            got = G.communities()
            bad = np.nonzero(got != comm)[0]
            assert bad.size == 0, f"sweep {s}: {bad.size} vertices differ"
            q1 = oracle.modularity(og, comm, M)
            assert abs(q_after - q1) <= Q_TOL
            assert abs(G.modularity() - q1) <= Q_TOL


@pytest.mark.parametrize("name", _names())
def test_phase_and_aggregation_match_oracle(lvlib, oracle, cases, name):
    g, opt = _get(cases, name)
    print(f"case {name}", flush=True)
    og = oracle.Graph.from_csr(g)
    M = oracle.total_weight(og)
    comm = np.arange(g.V, dtype=np.int32)
    o2c = np.arange(g.V, dtype=np.int32)
    changed_o, _ = oracle.phase(og, comm, 1e-6, o2c, M)
    with lvlib.Graph(g, **opt) as G:
        changed = G.optimise(1e-6)
        assert changed == changed_o
        assert np.array_equal(G.communities(), comm)
        a, lv = G.assignment()
        assert np.array_equal(a, o2c)
        G.aggregate()
        nxt, nid = oracle.aggregate(og, comm, o2c)
        h = G.level_graph()
        assert h.V == nxt.V and h.E == nxt.E
        rp, cl, w = nxt.csr_arrays()
        assert np.array_equal(h.rowptr, rp)
        assert np.array_equal(h.col, cl)
        assert np.array_equal(h.w.view(np.uint32), w.view(np.uint32))
        a, lv = G.assignment()
        assert lv == nxt.V
        assert np.array_equal(a, o2c)
        # Q is preserved by aggregation up to the fp32 rounding of merged weights
        q_level = G.modularity()
        assert abs(q_level - oracle.modularity(nxt, np.arange(nxt.V, dtype=np.int32), M)) <= Q_TOL


@pytest.mark.parametrize("name", _names())
def test_full_run_matches_oracle(lvlib, oracle, cases, name):
    g, opt = _get(cases, name)
    print(f"case {name}", flush=True)
    og = oracle.Graph.from_csr(g)
    st_o, o2c_o, lv_o = oracle.run(og, 1e-6)
    with lvlib.Graph(g, **opt) as G:
        st = G.run(1e-6)
        o2c, lv = G.assignment()
    assert st["levels"] == st_o["levels"]
    assert st["sweeps"] == st_o["sweeps"]
    assert lv == lv_o
    assert np.array_equal(o2c, o2c_o)
    assert abs(st["modularity"] - st_o["modularity"]) <= Q_TOL


def test_col_order_default_at_scale_is_identical(lvlib):
    """RMAT-21 (about 1.5M vertices with entries: the bucket-ordered columns
    are on by default at levels 0 and 1) against the same run with them off:
    every sweep's modularity and movers, the levels and the assignment agree
    bit for bit."""
    g = lvlib.gen_rmat(21, 16, seed=21)
    print(f"case rmat21 V={g.V} E={g.E}", flush=True)
    runs = []
    for opt in ({}, {"col_order": -1}):
        with lvlib.Graph(g, **opt) as G:
            trace = []
            for _ in range(2):
                q0 = G.modularity()
                for _ in range(6):
                    q, moved = G.sweep()
                    trace.append((q, moved))
                    if moved == 0 or q - q0 < 1e-6:
                        break
                    q0 = q
                G.aggregate()
            G.reset()
            st = G.run(1e-6)
            o2c, lv = G.assignment()
            runs.append((trace, st["sweeps"], st["levels"], st["modularity"], lv, o2c))
    (t1, s1, l1, q1, v1, a1), (t2, s2, l2, q2, v2, a2) = runs
    assert t1 == t2
    assert (s1, l1, q1, v1) == (s2, l2, q2, v2)
    assert np.array_equal(a1, a2)


def test_rerun_after_reset_is_identical(lvlib):
    g = lvlib.gen_rmat(13, 16, seed=5)
    with lvlib.Graph(g) as G:
        a = G.run(1e-6)
        o1, l1 = G.assignment()
        G.reset()
        b = G.run(1e-6)
        o2, l2 = G.assignment()
    assert a["modularity"] == b["modularity"] and l1 == l2
    assert np.array_equal(o1, o2)


@pytest.mark.parametrize("agg_mode", [0, 1])
def test_global_tier_aggregation_with_large_sort(lvlib, oracle, agg_mode):
    """A community with > 8192 distinct neighbour communities: the hash path
    (agg_mode 1) takes the global hash and the merge-path sort; the sort
    path a long run of one row."""
    rowptr, col, w = graphs.star(20000, extra_ring=False)
    g = _csr(lvlib, rowptr, col, w)
    og = oracle.Graph.from_csr(g)
    comm = np.arange(g.V, dtype=np.int32)
    comm[1:50] = 0  # hub community holds 49 leaves; 19951 leaf singletons remain
    o2c = np.arange(g.V, dtype=np.int32)
    with lvlib.Graph(g, agg_mode=agg_mode) as G:
        G.set_communities(comm)
        G.aggregate()
        h = G.level_graph()
    nxt, _ = oracle.aggregate(og, comm, o2c)
    rp, cl, ww = nxt.csr_arrays()
    assert np.array_equal(h.rowptr, rp)
    assert np.array_equal(h.col, cl)
    assert np.array_equal(h.w.view(np.uint32), ww.view(np.uint32))


@pytest.mark.parametrize("kind", ["no_edges", "self_loops_only", "single_edge", "two_components"])
def test_edge_cases(lvlib, oracle, kind):
    if kind == "no_edges":
        rowptr, col, w = np.zeros(6, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32)
    elif kind == "self_loops_only":
        rowptr, col, w = graphs.csr_from_edges(4, [(0, 0), (1, 1), (2, 2)])
    elif kind == "single_edge":
        rowptr, col, w = graphs.csr_from_edges(3, [(0, 2)])
    else:
        rowptr, col, w = graphs.csr_from_edges(8, [(0, 1), (1, 2), (0, 2), (4, 5), (5, 6), (4, 6), (6, 7)])
    g = _csr(lvlib, rowptr, col, w)
    og = oracle.Graph.from_csr(g)
    st_o, o2c_o, lv_o = oracle.run(og, 1e-6)
    with lvlib.Graph(g) as G:
        st = G.run(1e-6)
        o2c, lv = G.assignment()
    assert lv == lv_o and np.array_equal(o2c, o2c_o)
    if g.E > 0:
        assert abs(st["modularity"] - st_o["modularity"]) <= Q_TOL


@pytest.mark.parametrize("scale,seed", [(12, 61), (14, 62)])
def test_xcd_modularity_pass_matches_plain_pass(lvlib, oracle, scale, seed):
    """The column-blocked heavy-row pass gives the same Q as the row pass."""
    g = lvlib.gen_rmat(scale, 16, seed=seed)
    og = oracle.Graph.from_csr(g)
    with lvlib.Graph(g, timing=True, xcd_rows=1) as A, lvlib.Graph(g, xcd_rows=-1) as B:
        for s in range(3):
            qa, ma = A.sweep()
            qb, mb = B.sweep()
            assert ma == mb and qa == qb
            assert abs(A.modularity() - oracle.modularity(og, A.communities())) <= 1e-9
        assert A.kernel_timing_detail(lvlib.DETAIL_NAMES.index("modularity_xcd"))["launches"] > 0
        A.aggregate()
        B.aggregate()
        assert abs(A.modularity() - B.modularity()) == 0.0


@pytest.mark.parametrize("case", ["rmat14_chunked", "self_loops"])
def test_row_pass_through_bucket_ordered_columns(lvlib, oracle, case):
    """The modularity row pass reading the move kernels' bucket-ordered
    columns (col_order) gives the same Q bits as over the original columns;
    every sweep ends with a full pass (q_incremental -1)."""
    if case == "rmat14_chunked":
        g, opt = lvlib.gen_rmat(14, 16, seed=63), {"heavy_wg_max": 320, "chunk_edges": 64,
                                                   "part_target": 64}
    else:
        g, opt = _csr(lvlib, *graphs.random_graph(400, 3000, 7, True, 40)), {}
    print(f"row pass case {case}", flush=True)
    og = oracle.Graph.from_csr(g)
    common = dict(q_incremental=-1, xcd_rows=-1, **opt)
    with lvlib.Graph(g, col_order=1, **common) as A, lvlib.Graph(g, col_order=-1, **common) as B:
        for s in range(4):
            qa, ma = A.sweep()
            qb, mb = B.sweep()
            assert ma == mb and qa == qb, f"sweep {s}"
            assert np.array_equal(A.communities(), B.communities())
            assert abs(qa - oracle.modularity(og, A.communities())) <= 1e-9


@pytest.mark.parametrize("kind", ["rmat", "self_loops"])
def test_singleton_modularity_equals_full_pass(lvlib, oracle, kind):
    """Q of a singleton state (after a level's set-up or a reset) is computed
    from the level's self-loop weight and the sum of S_c^2; it equals the
    full pass's Q (lv_set_communities forces one) bit for bit, at level 0
    and at aggregated levels, whose rows carry self-loops."""
    if kind == "rmat":
        g = lvlib.gen_rmat(14, 16, seed=65)
    else:
        g = _csr(lvlib, *graphs.random_graph(3000, 20000, 66, True, 40))
    og = oracle.Graph.from_csr(g)
    with lvlib.Graph(g) as G:
        for level in range(3):
            V, _ = G.level_size()
            q_fast = G.modularity()
            G.set_communities(np.arange(V, dtype=np.int32))
            q_full = G.modularity()
            assert q_fast == q_full, level
            if level == 0:
                assert abs(q_fast - oracle.modularity(og, np.arange(g.V, dtype=np.int32))) <= 1e-9
                G.reset()
                assert G.modularity() == q_full
            if not G.optimise(1e-6):
                break
            G.aggregate()
