"""Host-side checks that need no GPU: the C-ABI library loads and exports
every declared symbol, the MatrixMarket reader follows the reference's
conventions (and agrees with the oracle's reader), the writer round-trips,
the seeded generators have the sizes their definitions imply, and the CLI
front ends print the reference's format on the golden fixtures.
"""
import json
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

GOLD = json.load(open(os.path.join(GOLDEN, "golden.json")))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "louvain.h")).read()
    return sorted(set(re.findall(r"\b(lv_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(lvlib):
    lib = lvlib.load()
    syms = declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_create_without_device_fails_cleanly(lvlib):
    if subprocess.run(["bash", "-c", "ls /dev/kfd"], capture_output=True).returncode == 0:
        pytest.skip("a GPU device node is present")
    g = lvlib.gen_planted(50, 2, 0.5, 0.05, seed=1)
    with pytest.raises(lvlib.LouvainError) as ei:
        lvlib.Graph(g)
    assert ei.value.status in (lvlib.LV_ERR_NODEVICE, lvlib.LV_ERR_HIP)


@pytest.mark.parametrize("name", ["two_triangles.mtx", "ring_of_cliques.mtx", "karate.mtx",
                                  "reader_conventions.mtx"])
def test_reader_matches_oracle_reader(lvlib, oracle, name):
    path = os.path.join(GOLDEN, name)
    a = lvlib.read_mtx(path)
    b = oracle.read_mtx(path)
    rp, cl, w = b.csr_arrays()
    assert np.array_equal(a.rowptr, rp)
    assert np.array_equal(a.col, cl)
    assert np.array_equal(a.w, w)


def test_reader_conventions_hand_computed(lvlib):
    g = lvlib.read_mtx(os.path.join(GOLDEN, "reader_conventions.mtx"))
    gold = GOLD["reader_conventions.mtx"]
    assert g.rowptr.tolist() == gold["rowptr"]
    assert g.col.tolist() == gold["col"]
    assert g.w.tolist() == gold["w"]


def test_reader_errors(lvlib, tmp_path):
    with pytest.raises(lvlib.LouvainError) as ei:
        lvlib.read_mtx(str(tmp_path / "missing.mtx"))
    assert ei.value.status == lvlib.LV_ERR_IO
    bad = tmp_path / "bad.mtx"
    bad.write_text("%%MatrixMarket matrix coordinate real symmetric\n3 3 2\n1 2 1\n4 1 1\n")
    with pytest.raises(lvlib.LouvainError) as ei:
        lvlib.read_mtx(str(bad))
    assert ei.value.status == lvlib.LV_ERR_ARG
    short = tmp_path / "short.mtx"
    short.write_text("%%MatrixMarket matrix coordinate real symmetric\n3 3 5\n1 2 1\n")
    with pytest.raises(lvlib.LouvainError):
        lvlib.read_mtx(str(short))


def _edge_multiset(g):
    src = np.repeat(np.arange(g.V), np.diff(g.rowptr))
    key = np.stack([src, g.col, g.w.view(np.int32)], 1)
    return key[np.lexsort(key.T[::-1])]


def test_writer_round_trip(lvlib, tmp_path):
    g = lvlib.gen_rmat(10, 8, seed=2)
    p = str(tmp_path / "g.mtx")
    lvlib.write_mtx(p, g)
    h = lvlib.read_mtx(p)
    assert h.V == g.V and h.E == g.E
    assert np.array_equal(_edge_multiset(g), _edge_multiset(h))


def test_planted_partition_sizes(lvlib):
    n, B, pin, pout = 10000, 20, 0.02, 0.0005
    g = lvlib.gen_planted(n, B, pin, pout, seed=1)
    s = n // B
    mean = B * s * (s - 1) / 2 * pin + (n * (n - 1) / 2 - B * s * (s - 1) / 2) * pout
    var = B * s * (s - 1) / 2 * pin * (1 - pin) + (n * (n - 1) / 2 - B * s * (s - 1) / 2) * pout * (1 - pout)
    m = g.E // 2
    assert abs(m - mean) < 6 * var ** 0.5, (m, mean)
    _check_simple_symmetric(g)
    assert (g.w == 1.0).all()
    h = lvlib.gen_planted(n, B, pin, pout, seed=1)
    assert np.array_equal(g.col, h.col)


def _check_simple_symmetric(g):
    src = np.repeat(np.arange(g.V), np.diff(g.rowptr))
    assert (src != g.col).all(), "self-loop"
    for v in range(0, g.V, max(1, g.V // 200)):
        row = g.col[g.rowptr[v]:g.rowptr[v + 1]]
        assert (np.diff(row) > 0).all(), "row not strictly ascending (duplicate or unsorted)"
    fwd = set(zip(src.tolist(), g.col.tolist(), g.w.tolist()))
    assert all((b, a, w) in fwd for a, b, w in list(fwd)[:20000])


def test_rmat_properties(lvlib):
    g = lvlib.gen_rmat(14, 16, seed=14)
    assert g.V == 1 << 14
    assert g.E % 2 == 0 and g.E < 2 * 16 * (1 << 14)
    _check_simple_symmetric(g)
    deg = g.degrees()
    assert deg.max() > 20 * deg.mean()  # heavy tail
    grid = g.w.astype(np.float64) * 2 ** 24
    assert (grid == np.round(grid)).all() and (g.w > 0).all() and (g.w <= 1).all()
    h = lvlib.gen_rmat(14, 16, seed=14)
    assert np.array_equal(g.rowptr, h.rowptr) and np.array_equal(g.col, h.col) and np.array_equal(g.w, h.w)
    u = lvlib.gen_rmat(14, 16, seed=15)
    assert not np.array_equal(g.col[:1000], u.col[:1000])


@pytest.mark.parametrize("name", ["two_triangles.mtx", "ring_of_cliques.mtx", "karate.mtx"])
def test_seqlouvain_cli_golden(oracle, name):
    path = os.path.join(GOLDEN, name)
    out = subprocess.run([oracle.SEQ_PATH, "-f", path, "-g", "0.000001", "-v"], capture_output=True,
                         text=True, check=True).stdout.splitlines()
    gold = GOLD[name]
    assert out[0] == f"{gold['modularity']:f}"
    assert len(out[1].split()) == 2
    lv = int(out[2])
    comm = [None] * len(gold["partition"])
    for line in out[3:3 + lv]:
        parts = [int(x) for x in line.split()]
        for x in parts[1:]:
            comm[x - 1] = parts[0] - 1
    assert comm == gold["partition"]


def test_cli_usage(oracle):
    r = subprocess.run([oracle.SEQ_PATH, "-f", "x.mtx"], capture_output=True, text=True)
    assert r.returncode == 1
    assert r.stdout.strip() == "Usage: ./gpulouvain -f mtx-matrix-file -g min-gain [-v]"
    exe = os.path.join(ROOT, "gpu-louvain_amd", "bin", "gpulouvain")
    r = subprocess.run([exe, "-g", "1"], capture_output=True, text=True)
    assert r.returncode == 1
    assert r.stdout.strip() == "Usage: ./gpulouvain -f mtx-matrix-file -g min-gain [-v]"


def test_batch_rejects_edges_across_graphs(lvlib):
    import graphs
    a = lvlib.CSR(*graphs.two_triangles())
    vptr, u = lvlib.union([a, a])
    u.col[0] = 7   # vertex 0 of graph 0 points into graph 1
    import ctypes
    lib = lvlib.load()
    lvlib._batch_sigs(lib)
    h = ctypes.c_void_p()
    st = lib.lv_batch_create(2, lvlib._p(vptr, ctypes.c_int32), lvlib._p(u.rowptr, ctypes.c_int32),
                             lvlib._p(u.col, ctypes.c_int32), lvlib._p(u.w, ctypes.c_float), u.E, None,
                             ctypes.byref(h))
    assert st == lvlib.LV_ERR_ARG
