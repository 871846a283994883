"""Level dump and resume on the device (lv_dump_level, lv_create_from_dump).

A run resumed from a dump of its first aggregated level must finish exactly
as the uninterrupted run: same assignment of the original vertices, same
final modularity, one level fewer.  The dumped MatrixMarket file holds the
level's CSR for the CPU oracle (same edge multiset as lv_level_graph).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _edges(rowptr, col, w):
    src = np.repeat(np.arange(rowptr.size - 1), np.diff(rowptr))
    key = np.stack([src, col, w.view(np.int32)], 1)
    return key[np.lexsort(key.T[::-1])]


@pytest.mark.parametrize("graph", ["rmat13", "planted"])
def test_level_dump_resume_matches_uninterrupted_run(lvlib, oracle, tmp_path, graph):
    print(f"case {graph}", flush=True)
    if graph == "rmat13":
        g = lvlib.gen_rmat(13, 16, seed=9)
    else:
        g = lvlib.gen_planted(6000, 30, 0.02, 0.001, seed=9, unit_weights=True)
    with lvlib.Graph(g) as A:
        st_a = A.run(1e-6)
        o2c_a, lv_a = A.assignment()
    prefix = str(tmp_path / "level1")
    with lvlib.Graph(g) as B:
        assert B.optimise(1e-6)
        B.aggregate()
        B.dump_level(prefix, mtx=True)
        h = B.level_graph()
        o2c_b, lv_b = B.assignment()
    assert st_a["levels"] >= 2
    og = oracle.read_mtx(prefix + ".mtx")
    rp, cl, w = og.csr_arrays()
    assert np.array_equal(_edges(rp, cl, w), _edges(h.rowptr, h.col, h.w))
    with lvlib.Graph.from_dump(prefix) as H:
        assert H.Vorig == g.V
        o2c_h0, lv_h0 = H.assignment()
        assert lv_h0 == lv_b and np.array_equal(o2c_h0, o2c_b)
        st_h = H.run(1e-6)
        o2c_h, lv_h = H.assignment()
        H.reset()
        o2c_r, _ = H.assignment()
    assert lv_h == lv_a and np.array_equal(o2c_h, o2c_a)
    assert st_h["modularity"] == st_a["modularity"]
    assert st_h["levels"] == st_a["levels"] - 1
    assert np.array_equal(o2c_r, o2c_b)  # reset returns to the dumped level
