"""LDS attempt of the chunked tier's rows (lv_options.heavy_lds) changes where
a heavy row's communities are summed, never a decision (needs an MI355X).

A row above the workgroup tiers is summed in one LDS table when its last
evaluation saw at most 4096 distinct communities; a row that turns out to
have more gives up and is summed by the chunked tier.  Handles with the
attempt off (-1), on (0) and on with a tiny give-up limit (8, so that nearly
every attempt falls back part-way through its row) must give identical
sweeps, runs, decisions and margins, and the same run as the oracle.
"""
import numpy as np
import pytest

import graphs

pytestmark = pytest.mark.gpu

MODES = (-1, 0, 8)


def _cases(lvlib):
    # heavy_wg_max = 320 sends every row above degree 320 to the chunked tier
    return [
        ("planted_dense", lvlib.gen_planted(12000, 12, 0.5, 0.001, seed=95, unit_weights=True),
         {"heavy_wg_max": 320}),
        ("rmat15_tiers", lvlib.gen_rmat(15, 16, seed=96), {"heavy_wg_max": 320, "chunk_edges": 128}),
        ("star", lvlib.CSR(*[np.asarray(x) for x in graphs.star(6000)]), {"heavy_wg_max": 320}),
    ]


def _distinct_per_row(g, comm, rows):
    """Distinct neighbour communities (self-loops excluded) of each listed row."""
    rp = np.asarray(g.rowptr, dtype=np.int64)
    col = np.asarray(g.col)
    out = []
    for v in rows:
        c = col[rp[v]:rp[v + 1]]
        out.append(np.unique(comm[c[c != v]]).size)
    return np.array(out)


def test_heavy_rows_fit_after_one_sweep(lvlib):
    """The dense planted graph does take the LDS path: after one sweep its
    heavy rows see far fewer than 4096 communities, so the next sweeps try
    them in LDS (their hint is that count)."""
    name, g, opts = _cases(lvlib)[0]
    deg = np.diff(np.asarray(g.rowptr))
    heavy = np.nonzero(deg > 320)[0]
    assert heavy.size > 1000
    with lvlib.Graph(g, **opts) as G:
        G.sweep()
        d = _distinct_per_row(g, G.communities(), heavy[:200])
    assert d.max() <= 4096, d.max()


@pytest.mark.parametrize("i", range(3))
def test_heavy_lds_changes_no_sweep_or_run(lvlib, i):
    name, g, opts = _cases(lvlib)[i]
    print(f"case {name}", flush=True)
    hs = [lvlib.Graph(g, heavy_lds=m, **opts) for m in MODES]
    try:
        for s in range(6):
            r = [h.sweep() for h in hs]
            assert r[1] == r[0] and r[2] == r[0], (s, r)
            c0 = hs[0].communities()
            for h in hs[1:]:
                assert np.array_equal(h.communities(), c0), s
        for h in hs:
            h.reset()
        st = [h.run(1e-6) for h in hs]
        for k in ("levels", "sweeps", "final_V", "modularity"):
            assert st[1][k] == st[0][k] and st[2][k] == st[0][k], (k, [x[k] for x in st])
        a0, l0 = hs[0].assignment()
        for h in hs[1:]:
            a, l = h.assignment()
            assert l == l0 and np.array_equal(a, a0)
    finally:
        for h in hs:
            h.close()


@pytest.mark.parametrize("i", [0, 1])
def test_heavy_lds_keeps_margins_and_decisions(lvlib, i):
    name, g, opts = _cases(lvlib)[i]
    print(f"case {name}", flush=True)
    ALL = lvlib.frontier(kind=lvlib.FRONTIER_ALL)
    with lvlib.Graph(g, heavy_lds=-1, **opts) as A, lvlib.Graph(g, heavy_lds=0, **opts) as B:
        # plain sweeps first, so that the heavy rows' hints are set
        for s in range(2):
            assert A.sweep() == B.sweep(), s
        for s in range(3):
            ra = A.sweep_pruned(ALL, track=True)
            rb = B.sweep_pruned(ALL, track=True)
            assert ra == rb, s
        for s in range(2):
            assert A.sweep() == B.sweep(), s
            ma, mb = A.move_state(), B.move_state()
            for k in ma:
                assert np.array_equal(ma[k], mb[k]), (s, k)
        for grp in range(7):
            assert np.array_equal(A.decide(grp), B.decide(grp)), grp


@pytest.mark.parametrize("mode", [0, 8])
def test_heavy_lds_run_matches_oracle(lvlib, oracle, mode):
    name, g, opts = _cases(lvlib)[0]
    print(f"case {name} heavy_lds={mode}", flush=True)
    st_o, o2c_o, lv_o = oracle.run(oracle.Graph.from_csr(g), 1e-6, threads=8)
    with lvlib.Graph(g, heavy_lds=mode, **opts) as G:
        st = G.run(1e-6)
        o2c, lv = G.assignment()
    assert st["levels"] == st_o["levels"] and st["sweeps"] == st_o["sweeps"]
    assert lv == lv_o
    assert np.array_equal(o2c, o2c_o)
    assert abs(st["modularity"] - st_o["modularity"]) <= 1e-9


@pytest.mark.parametrize("i", [0, 1])
def test_heavy_lds_in_pruned_sweeps(lvlib, i):
    """Pruned sweeps: the attempt skips the rows outside the frontier, as the
    chunked tier does; sweeps and pruned runs equal those without it."""
    name, g, opts = _cases(lvlib)[i]
    print(f"case {name}", flush=True)
    hs = [lvlib.Graph(g, heavy_lds=m, **opts) for m in MODES]
    try:
        for h in hs:
            h.sweep()  # plain sweep first: the heavy rows get their hints
        for s in range(4):
            fr = lvlib.frontier(kind=lvlib.FRONTIER_RANDOM, random_p=0.5, seed=100 + s)
            r = [h.sweep_pruned(fr, track=True) for h in hs]
            assert r[1] == r[0] and r[2] == r[0], (s, r)
            c0 = hs[0].communities()
            for h in hs[1:]:
                assert np.array_equal(h.communities(), c0), s
        fr = lvlib.frontier(kind=lvlib.FRONTIER_NEIGHBOR, full_every=4)
        for h in hs:
            h.reset()
        st = [h.run_pruned(1e-6, fr) for h in hs]
        for k in ("levels", "sweeps", "final_V", "modularity"):
            assert st[1][k] == st[0][k] and st[2][k] == st[0][k], (k, [x[k] for x in st])
        a0, _ = hs[0].assignment()
        for h in hs[1:]:
            assert np.array_equal(h.assignment()[0], a0)
    finally:
        for h in hs:
            h.close()
