"""Candidate bounds in the move kernels (lv_options.candidate_bounds) change
which community records are read, never a decision (needs an MI355X).

With weights >= 0 a candidate's gain is bounded from above without its
record; candidates whose bound cannot change the decision are skipped.
Handles with the bounds forced on (1) and off (-1) must give identical sweeps, runs and
decisions, and identical margins wherever observations are recorded (tracked
sweeps, lv_decide).  A negative weight turns the bounds off by itself."""
import numpy as np
import pytest

import graphs

pytestmark = pytest.mark.gpu


def _cases(lvlib):
    return [
        ("rmat15_real", lvlib.gen_rmat(15, 16, seed=91), {}),
        ("rmat14_unit", lvlib.gen_rmat(14, 16, seed=92, unit_weights=True), {}),
        ("planted_unit", lvlib.gen_planted(6000, 30, 0.02, 0.001, seed=93), {}),
        ("rmat15_tiers", lvlib.gen_rmat(15, 16, seed=94), {"heavy_wg_max": 512, "chunk_edges": 128}),
        ("star", lvlib.CSR(*[np.asarray(x) for x in graphs.star(5000)]), {}),
    ]


@pytest.mark.parametrize("i", range(5))
def test_bounds_change_no_sweep_or_run(lvlib, i):
    name, g, opts = _cases(lvlib)[i]
    print(f"case {name}", flush=True)
    with lvlib.Graph(g, candidate_bounds=-1, **opts) as A, lvlib.Graph(g, candidate_bounds=1, **opts) as B:
        for s in range(6):
            ra, rb = A.sweep(), B.sweep()
            assert ra == rb, (s, ra, rb)
            assert np.array_equal(A.communities(), B.communities()), s
        A.reset()
        B.reset()
        sa, sb = A.run(1e-6), B.run(1e-6)
        for k in ("levels", "sweeps", "final_V", "modularity"):
            assert sa[k] == sb[k], (k, sa[k], sb[k])
        oa, la = A.assignment()
        ob, lb = B.assignment()
        assert la == lb and np.array_equal(oa, ob)


@pytest.mark.parametrize("i", [0, 1, 3])
def test_bounds_keep_observed_margins_exact(lvlib, i):
    name, g, opts = _cases(lvlib)[i]
    print(f"case {name}", flush=True)
    ALL = lvlib.frontier(kind=lvlib.FRONTIER_ALL)
    with lvlib.Graph(g, candidate_bounds=-1, **opts) as A, lvlib.Graph(g, candidate_bounds=1, **opts) as B:
        for s in range(4):
            ra = A.sweep_pruned(ALL, track=True)
            rb = B.sweep_pruned(ALL, track=True)
            assert ra == rb, s
            ma, mb = A.move_state(), B.move_state()
            for k in ma:
                assert np.array_equal(ma[k], mb[k]), (s, k)
        for grp in range(7):
            assert np.array_equal(A.decide(grp), B.decide(grp)), grp
            ma, mb = A.move_state(), B.move_state()
            assert np.array_equal(ma["margin"], mb["margin"]), grp
            assert np.array_equal(ma["kown"], mb["kown"]), grp


def test_negative_weight_turns_bounds_off(lvlib, oracle):
    rowptr, col, w = graphs.random_graph(600, 5000, 11, True, 10)
    w = np.asarray(w, np.float32).copy()
    # one undirected edge made negative (both directions)
    u, v = 0, int(col[rowptr[0]])
    for a_, b_ in ((u, v), (v, u)):
        for e in range(rowptr[a_], rowptr[a_ + 1]):
            if col[e] == b_:
                w[e] = -w[e]
    g = lvlib.CSR(np.asarray(rowptr, np.int32), np.asarray(col, np.int32), w)
    with lvlib.Graph(g, candidate_bounds=-1) as A, lvlib.Graph(g, candidate_bounds=1) as B:
        for s in range(3):
            assert A.sweep() == B.sweep(), s
            assert np.array_equal(A.communities(), B.communities()), s
            ma, mb = A.move_state(), B.move_state()
            assert np.array_equal(ma["margin"], mb["margin"]), s
