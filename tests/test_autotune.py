"""Autotuner: cache, signature and search logic on the CPU; on the GPU, that
every candidate (lanes per vertex of the sub-wave buckets, bucket cuts inside
groups 5 and 6, lanes per row and LDS table sizes of the one-workgroup tiers,
chunked-tier sizes, column-blocked passes) gives the same sweep as the
defaults (results do not depend on
tuning), and that tuning single graphs and batches picks one of them."""
import numpy as np
import pytest

import graphs


def test_signature_and_cache_roundtrip(lvlib, tmp_path):
    import autotune as T

    g1 = lvlib.CSR(*graphs.two_triangles())
    g2 = lvlib.CSR(*graphs.star(3000))
    s1, s2 = T.signature(g1), T.signature(g2)
    assert s1 != s2 and s1.startswith("v1:")
    assert T.signature(g1) == s1
    path = str(tmp_path / "tune.json")
    assert T.load_cache(path) == {}
    key = s1 + ":v2"
    T.save_cache({key: {"options": {"chunk_edges": 256}, "ms": 1.0}}, path)
    assert T.load_cache(path)[key]["options"]["chunk_edges"] == 256
    # a cached shape is answered without a device
    out = T.tune(g1, cache_path=path, second_level=False)
    assert out["cached"] and out["options"] == {"chunk_edges": 256}


def test_candidate_space_covers_every_option():
    import autotune as T

    small = {k: T.SPACE[k][:2] for k in ("chunk_edges", "xcd_rows", "wave_split")}
    cands = list(T.candidates(small))
    assert len(cands) == 8 and all(set(c) == set(small) for c in cands)
    assert T.space_size(T.SPACE) > T.EXHAUSTIVE_MAX
    for k in ("wave_split", "wg_bound_a", "wg_bound_b", "wg_bound_c", "wg_lanes", "table_pct",
              "sub_lanes"):
        assert k in T.SPACE
    assert set(T.BATCH_SPACE) <= set(T.SPACE)


def test_coordinate_search_finds_separable_minimum():
    """On a separable cost the coordinate descent reaches the minimum of every
    option, tries far fewer points than the grid, and reports each evaluation."""
    import autotune as T

    space = {"a": [3, 1, 2, 0], "b": [9, 7, 5], "c": [(1, 2), (0, 0)], "d": [4, 8, 6, 2, 1]}
    cost = {"a": {3: 3, 1: 1, 2: 2, 0: 0}, "b": {9: 9, 7: 7, 5: 5}, "c": {(1, 2): 1, (0, 0): 0},
            "d": {4: 4, 8: 8, 6: 6, 2: 2, 1: 1}}
    best_t, best, tried = T.search(lambda o: float(sum(cost[k][v] for k, v in o.items())), space)
    assert best == {"a": 0, "b": 5, "c": (0, 0), "d": 1} and best_t == 6.0
    assert len(tried) < T.space_size(space)
    # exhaustive below the threshold
    small = {"a": [1, 0], "b": [2, 3]}
    t, b, tr = T.search(lambda o: float(o["a"] + o["b"]), small)
    assert b == {"a": 0, "b": 2} and len(tr) == 4


def test_options_struct_fields(lvlib):
    o = lvlib.options_struct(0, False, wg_lanes=(128, 512), table_pct=175, wave_split=200,
                             sub_lanes=(4, 8))
    assert list(o.wg_lanes) == [128, 512, 0, 0] and o.table_pct == 175 and o.wave_split == 200
    assert list(o.sub_lanes) == [4, 8, 0]
    with pytest.raises(TypeError):
        lvlib.options_struct(0, False, no_such_option=1)


SAME_SPACE = [
    {},
    {"heavy_wg_max": 320},
    {"heavy_wg_max": 2048, "chunk_edges": 64, "part_target": 32, "merge_limit": 16},
    {"xcd_rows": 2, "chunk_edges": 64, "heavy_wg_max": 1024},
    {"wave_split": 85},
    {"wave_split": 255},
    {"wg_bound_a": 320, "wg_bound_b": 320, "wg_bound_c": 320},
    {"wg_bound_a": 700, "wg_bound_b": 900, "wg_bound_c": 1100, "heavy_wg_max": 1300},
    {"wg_lanes": (1024, 1024, 1024, 1024)},
    {"wg_lanes": (256, 512, 128, 512), "table_pct": 120},
    {"wg_lanes": (512, 128, 1024, 256), "table_pct": 200},
    {"table_pct": 133, "wave_split": 97, "wg_bound_b": 600},
    {"sub_lanes": (4, 4, 4)},
    {"sub_lanes": (4, 8, 8)},
    {"sub_lanes": (8, 8, 16), "wave_split": 200},
    {"sub_lanes": (4, 16, 16), "candidate_bounds": -1},
]


@pytest.mark.gpu
def test_every_candidate_gives_the_same_sweep(lvlib):
    g = lvlib.gen_rmat(14, 16, seed=71)
    with lvlib.Graph(g) as G:
        G.sweep()
        G.sweep()
        ref_q, ref_m = G.sweep()
        ref_c = G.communities().copy()
        for opt in SAME_SPACE:
            G.set_options(**opt)
            G.reset()
            G.sweep()
            G.sweep()
            q, m = G.sweep()
            assert (q, m) == (ref_q, ref_m), opt
            assert np.array_equal(G.communities(), ref_c), opt


@pytest.mark.gpu
def test_tier_options_give_the_same_run_and_decisions(lvlib, oracle):
    """Whole runs, and decisions against the oracle, under tier options that
    move rows between every workgroup shape."""
    g = lvlib.gen_rmat(13, 16, seed=73)
    og = oracle.Graph.from_csr(g)
    st_o, o2c_o, _ = oracle.run(og, 1e-6)
    comm = np.arange(g.V, dtype=np.int32)
    oracle.sweep(og, comm)
    for opt in SAME_SPACE[4:]:
        with lvlib.Graph(g, **opt) as G:
            st = G.run(1e-6)
            o2c, _ = G.assignment()
            assert st["sweeps"] == st_o["sweeps"] and np.array_equal(o2c, o2c_o), opt
            G.reset()
            G.set_communities(comm)
            for grp in range(1, 7):
                assert np.array_equal(G.decide(grp), oracle.decide(og, comm, grp)), (opt, grp)


@pytest.mark.gpu
def test_tune_picks_a_candidate(lvlib, tmp_path):
    import autotune as T

    g = lvlib.gen_rmat(13, 16, seed=72)
    space = {"heavy_wg_max": [320, 2048], "chunk_edges": [256, 1024], "wg_lanes": [(128, 256, 256, 1024)]}
    out = T.tune(g, space=space, repeats=1, cache_path=str(tmp_path / "t.json"))
    assert out["tried"] == 4 and not out["cached"]
    assert {k: (tuple(v) if isinstance(v, list) else v) for k, v in out["options"].items()} in \
        list(T.candidates(space))
    again = T.tune(g, space=space, cache_path=str(tmp_path / "t.json"))
    assert again["cached"] and again["options"] == out["options"]


@pytest.mark.gpu
def test_tune_batch_and_options_keep_results(lvlib, tmp_path):
    import autotune as T

    gs = [lvlib.gen_planted(300 + 200 * i, 4 + i % 5, 0.06, 0.004, seed=200 + i) for i in range(12)]
    space = {"wave_split": [128, 255], "wg_lanes": [(128, 256, 256, 1024), (256, 512, 512, 1024)]}
    out = T.tune_batch(gs, space=space, repeats=1, cache_path=str(tmp_path / "b.json"))
    assert out["tried"] == 4 and out["signature"].startswith("batch:")
    with lvlib.Batch(gs) as B:
        per0, _ = B.run(1e-6)
        c0 = B.assignment()
        B.set_options(**out["options"])
        per1, _ = B.run(1e-6)
        c1 = B.assignment()
    assert [p["modularity"] for p in per0] == [p["modularity"] for p in per1]
    assert all(np.array_equal(a, b) for a, b in zip(c0, c1))
