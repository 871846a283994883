"""Aggregated levels stored without their vertices of degree 0 ("compact"
levels, the default) against levels that store every vertex (compact=-1).

The reference keeps every vertex: a vertex without edges stays a singleton
community at every level.  The library drops such vertices from the stored
level (ids stay in the reference's order) and maps every returned array back
to the reference's numbering, so both handles must agree on everything:
level sizes and CSRs, communities, decisions, sweep results, move
observations, features, model outputs, traces and whole runs.
"""
import numpy as np
import pytest

import graphs

pytestmark = pytest.mark.gpu


def _with_isolated(lvlib, g, extra, seed):
    """g with `extra` vertices without edges and a few self-loop-only vertices
    spread over the id range."""
    rng = np.random.default_rng(seed)
    n = g.V + extra
    new_of = np.sort(rng.choice(n, g.V, replace=False)).astype(np.int32)
    deg = np.zeros(n, np.int64)
    deg[new_of] = np.diff(g.rowptr)
    loops = np.setdiff1d(np.arange(n), new_of)[: extra // 10]
    deg[loops] += 1
    rowptr = np.zeros(n + 1, np.int64)
    rowptr[1:] = np.cumsum(deg)
    col = np.zeros(rowptr[-1], np.int32)
    w = np.zeros(rowptr[-1], np.float32)
    for v in range(g.V):
        a, b = g.rowptr[v], g.rowptr[v + 1]
        o = rowptr[new_of[v]]
        col[o:o + b - a] = new_of[g.col[a:b]]
        w[o:o + b - a] = g.w[a:b]
    for x in loops:
        col[rowptr[x + 1] - 1] = x
        w[rowptr[x + 1] - 1] = 1.0
    return lvlib.CSR(rowptr.astype(np.int32), col, w)


def _cases(lvlib):
    return {
        "rmat13": (lvlib.gen_rmat(13, 16, seed=81), {}),
        "rmat13_unit_isolated": (_with_isolated(lvlib, lvlib.gen_rmat(13, 8, seed=82, unit_weights=True), 3000, 5), {}),
        "rmat14_chunked": (lvlib.gen_rmat(14, 16, seed=83), {"heavy_wg_max": 320, "chunk_edges": 64, "part_target": 64}),
        "planted_isolated": (_with_isolated(lvlib, lvlib.gen_planted(3000, 12, 0.02, 0.001, seed=84, unit_weights=True), 500, 6), {}),
    }


@pytest.fixture(scope="module")
def cases(lvlib):
    return _cases(lvlib)


NAMES = ["rmat13", "rmat13_unit_isolated", "rmat14_chunked", "planted_isolated"]


@pytest.mark.parametrize("name", NAMES)
def test_full_runs_agree(lvlib, cases, name):
    print(f"case {name}", flush=True)
    g, opt = cases[name]
    with lvlib.Graph(g, **opt) as A, lvlib.Graph(g, compact=-1, **opt) as B:
        a, b = A.run(1e-6), B.run(1e-6)
        oa, la = A.assignment()
        ob, lb = B.assignment()
    assert a["levels"] == b["levels"] and a["sweeps"] == b["sweeps"]
    assert la == lb and a["final_V"] == b["final_V"]
    assert np.array_equal(oa, ob)
    assert abs(a["modularity"] - b["modularity"]) <= 1e-12


@pytest.mark.parametrize("name", NAMES)
def test_aggregated_level_agrees(lvlib, cases, name):
    print(f"case {name}", flush=True)
    g, opt = cases[name]
    rng = np.random.default_rng(7)
    params = rng.normal(0.0, 0.3, lvlib.MLP_PARAMS).astype(np.float32)
    ALL = lvlib.frontier(kind=lvlib.FRONTIER_ALL)
    with lvlib.Graph(g, **opt) as A, lvlib.Graph(g, compact=-1, **opt) as B:
        for G in (A, B):
            assert G.optimise(1e-6)
            G.aggregate()
            G.set_model(params)
        assert A.level_size() == B.level_size()
        ha, hb = A.level_graph(), B.level_graph()
        assert np.array_equal(ha.rowptr, hb.rowptr) and np.array_equal(ha.col, hb.col)
        assert np.array_equal(ha.w.view(np.uint32), hb.w.view(np.uint32))
        assert np.array_equal(A.assignment()[0], B.assignment()[0])
        assert abs(A.modularity() - B.modularity()) <= 1e-12
        for s in range(3):
            qa, ma, ea = A.sweep_pruned(ALL, track=True)
            qb, mb, eb = B.sweep_pruned(ALL, track=True)
            assert ma == mb and ea == eb and abs(qa - qb) <= 1e-12, f"sweep {s}"
            assert np.array_equal(A.communities(), B.communities()), f"sweep {s}"
            sa, sb = A.move_state(), B.move_state()
            for k in sa:
                assert np.array_equal(sa[k], sb[k]), f"sweep {s}: {k}"
            assert np.array_equal(A.features(), B.features()), f"sweep {s}: features"
            pa, aa = A.predict(0.5)
            pb, ab = B.predict(0.5)
            assert np.array_equal(pa, pb) and np.array_equal(aa, ab), f"sweep {s}: model"
        for grp in range(7):
            assert np.array_equal(A.decide(grp), B.decide(grp)), f"group {grp}"
        fr = lvlib.frontier(kind=lvlib.FRONTIER_NEIGHBOR)
        for s in range(2):
            qa, ma, ea = A.sweep_pruned(fr, track=True)
            qb, mb, eb = B.sweep_pruned(fr, track=True)
            assert ma == mb and ea == eb and abs(qa - qb) <= 1e-12, f"pruned sweep {s}"
            assert np.array_equal(A.communities(), B.communities())
        # a state set from outside, in the reference's numbering
        comm = B.communities()
        A.set_communities(comm)
        B.set_communities(comm)
        assert np.array_equal(A.decide(6), B.decide(6))
        fa, ma_, ca = A.trace_phase(1e-6, 8)
        fb, mb_, cb = B.trace_phase(1e-6, 8)
        assert ca == cb and np.array_equal(fa, fb) and np.array_equal(ma_, mb_)


def test_compact_level_rejects_states_the_reference_cannot_reach(lvlib, cases):
    g, opt = cases["rmat13_unit_isolated"]
    with lvlib.Graph(g, **opt) as A:
        assert A.optimise(1e-6)
        A.aggregate()
        V, _ = A.level_size()
        comm = A.communities()
        h = A.level_graph()
        deg = np.diff(h.rowptr)
        iso = np.nonzero(deg == 0)[0]
        act = np.nonzero(deg > 0)[0]
        assert iso.size > 0 and act.size > 1
        bad = comm.copy()
        bad[iso[0]] = act[0]           # a vertex without edges joins another community
        with pytest.raises(lvlib.LouvainError) as ei:
            A.set_communities(bad)
        assert ei.value.status == lvlib.LV_ERR_ARG
        bad = comm.copy()
        bad[act[0]] = iso[0]           # a vertex joins the community of one without edges
        with pytest.raises(lvlib.LouvainError):
            A.set_communities(bad)
        A.set_communities(np.arange(V, dtype=np.int32))
        assert np.array_equal(A.communities(), np.arange(V))
