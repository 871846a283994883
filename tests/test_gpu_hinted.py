"""The small-table attempt of the one-workgroup tiers (lv_options.hinted).

Rows of degree 320 up to the chunked tier whose last evaluation saw at most
256 distinct communities are decided one wave per row in a 512-slot LDS
table; the others (and rows whose communities outgrow the table) go to the
degree-sized tier kernels.  Either way a decision is the reference's
(computeMove, modularity_optimisation.cu:111-228, restated by the oracle):
the tests compare both paths with the oracle on states where the attempt
takes most rows, and whole runs with the attempt on and off.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dense_planted(lvlib):
    # 16 blocks of 1000 vertices, degree about 415: every row lies in the
    # one-workgroup tiers (320..512), and once blocks form a row sees few
    # distinct communities
    return lvlib.gen_planted(16000, 16, 0.4, 0.001, seed=61)


@pytest.mark.parametrize("sweeps,opt", [(1, {}), (3, {}), (3, {"candidate_bounds": 1}),
                                        (3, {"candidate_bounds": -1}),
                                        # every row in the chunked tier: the attempt
                                        # takes them with that tier's own flags
                                        (3, {"heavy_wg_max": 320}),
                                        (3, {"heavy_wg_max": 320, "candidate_bounds": 1})])
def test_attempt_decisions_match_oracle(lvlib, oracle, sweeps, opt):
    g = _dense_planted(lvlib)
    print(f"planted V={g.V} E={g.E} sweeps={sweeps} opt={opt}", flush=True)
    og = oracle.Graph.from_csr(g)
    comm = np.arange(g.V, dtype=np.int32)
    for _ in range(sweeps):
        oracle.sweep(og, comm)
    want = oracle.decide(og, comm, 6)
    with lvlib.Graph(g, **opt) as G:
        G.set_communities(comm)
        first = G.decide(6)   # no hints yet: the tier kernels decide every row
        second = G.decide(6)  # rows with few communities: the small-table attempt
        third = G.decide(6)
    for got in (first, second, third):
        bad = np.nonzero(got != want)[0]
        assert bad.size == 0, f"{bad.size} differ, first {bad[:5]} got {got[bad[:5]]} want {want[bad[:5]]}"


@pytest.mark.parametrize("name", ["planted", "planted_chunked", "rmat15", "rmat14_tiers"])
def test_runs_with_and_without_attempt_match_oracle(lvlib, oracle, name):
    if name == "planted":
        g, opt = _dense_planted(lvlib), {}
    elif name == "planted_chunked":
        g, opt = _dense_planted(lvlib), {"heavy_wg_max": 320}
    elif name == "rmat15":
        g, opt = lvlib.gen_rmat(15, 16, seed=62), {}
    else:  # small workgroup tiers so that more rows take the attempt
        g, opt = lvlib.gen_rmat(14, 16, seed=63), {"wg_bound_a": 400, "wg_bound_b": 600,
                                                    "wg_bound_c": 800, "heavy_wg_max": 1000}
    print(f"case {name} V={g.V} E={g.E}", flush=True)
    with lvlib.Graph(g, **opt) as A:
        sa = A.run(1e-6)
        oa, la = A.assignment()
    with lvlib.Graph(g, hinted=-1, **opt) as B:
        sb = B.run(1e-6)
        ob, lb = B.assignment()
    assert np.array_equal(oa, ob)
    assert la == lb
    assert (sa["levels"], sa["sweeps"]) == (sb["levels"], sb["sweeps"])
    assert sa["modularity"] == sb["modularity"]
    ost, o2c, lv = oracle.run(oracle.Graph.from_csr(g), 1e-6)
    assert la == lv
    assert np.array_equal(oa, o2c)
    assert abs(sa["modularity"] - ost["modularity"]) <= 1e-9
