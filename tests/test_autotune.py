"""Autotuner: cache and signature logic on the CPU; on the GPU, that every
candidate gives the same sweep as the defaults (results do not depend on
tuning) and that tuning picks one of them."""
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
    T.save_cache({s1: {"options": {"chunk_edges": 256}, "ms": 1.0}}, path)
    assert T.load_cache(path)[s1]["options"]["chunk_edges"] == 256
    # a cached shape is answered without a device
    out = T.tune(g1, cache_path=path, second_level=False)
    assert out["cached"] and out["options"] == {"chunk_edges": 256}


def test_candidate_space_covers_every_option():
    import autotune as T

    cands = list(T.candidates())
    assert len(cands) == int(np.prod([len(v) for v in T.SPACE.values()]))
    assert all(set(c) == set(T.SPACE) for c in cands)


@pytest.mark.gpu
def test_every_candidate_gives_the_same_sweep(lvlib):
    import autotune as T

    g = lvlib.gen_rmat(14, 16, seed=71)
    with lvlib.Graph(g) as G:
        G.sweep()
        G.sweep()
        ref_q, ref_m = G.sweep()
        ref_c = G.communities().copy()
        space = {"heavy_wg_max": [320, 2048], "chunk_edges": [64, 1024], "part_target": [32, 512],
                 "merge_limit": [16, 1536], "xcd_rows": [1, 2, -1]}
        for opt in T.candidates(space):
            G.set_options(**opt)
            G.reset()
            G.sweep()
            G.sweep()
            q, m = G.sweep()
            assert (q, m) == (ref_q, ref_m), opt
            assert np.array_equal(G.communities(), ref_c), opt


@pytest.mark.gpu
def test_tune_picks_a_candidate(lvlib, tmp_path):
    import autotune as T

    g = lvlib.gen_rmat(13, 16, seed=72)
    space = {"heavy_wg_max": [320, 2048], "chunk_edges": [256, 1024], "part_target": [512],
             "merge_limit": [1536], "xcd_rows": [0]}
    out = T.tune(g, space=space, repeats=1, cache_path=str(tmp_path / "t.json"))
    assert out["tried"] == 4 and not out["cached"]
    assert out["options"] in list(T.candidates(space))
    again = T.tune(g, space=space, cache_path=str(tmp_path / "t.json"))
    assert again["cached"] and again["options"] == out["options"]
