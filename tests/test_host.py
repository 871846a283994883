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
    b = oracle.read_mtx(path, quantize=False)
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


TRICKY = (
    "%%MatrixMarket matrix coordinate real symmetric\n"
    "% a comment\n"
    "\n"
    "6 6 9\n"
    "1 2 0.5\n"
    "2\t3  1.25\r\n"
    "3 3\n2.0\n"           # an entry across three lines
    "  4 1 1e-3\n"
    "5 6 7\t6 6 0.25\n"     # two entries on one line
    "6\n1\n3.5 2 4 0.125\n"
    "1 1 9\n"
    "4 5 1 99 99 99\n"      # tokens after the last entry are not read
)


@pytest.mark.parametrize("chunk", [None, 1, 2, 5, 13, 64])
def test_reader_chunk_boundaries_match_oracle(lvlib, oracle, tmp_path, monkeypatch, chunk):
    """The parallel reader cuts the entry region into chunks at whitespace
    and gives every token its entry by a prefix over the chunks' token counts;
    tiny chunks put the cuts inside entries and lines."""
    if chunk is not None:
        monkeypatch.setenv("LV_MTX_CHUNK_BYTES", str(chunk))
    tricky = tmp_path / "tricky.mtx"
    tricky.write_bytes(TRICKY.encode())
    pattern = tmp_path / "pattern.mtx"
    pattern.write_text("%%MatrixMarket matrix coordinate pattern symmetric\n4 4 4\n1 2\n2 3 3\n4\n4 1\n")
    g = lvlib.gen_rmat(9, 8, seed=3)
    big = str(tmp_path / "g.mtx")
    lvlib.write_mtx(big, g)
    paths = [str(tricky), str(pattern), big] + [os.path.join(GOLDEN, n) for n in
                                                ("karate.mtx", "reader_conventions.mtx")]
    for path in paths:
        a = lvlib.read_mtx(path)
        b = oracle.read_mtx(path, quantize=False)
        rp, cl, w = b.csr_arrays()
        assert np.array_equal(a.rowptr, rp), path
        assert np.array_equal(a.col, cl), path
        assert np.array_equal(a.w.view(np.uint32), w.view(np.uint32)), path


def test_reader_weights_round_like_strtof(lvlib, oracle, tmp_path):
    """Weights parsed by the reader's fast path equal the C library's
    correctly rounded strtof (through the oracle's fscanf) on decimals of
    every length, exponents, and exact midpoints between floats."""
    rng = np.random.default_rng(11)
    toks = ["1", "0.5", "1e-3", "2.5E+2", "-0.0", "+3.25", "0000.125", "1.000000059604644775390625",
            "1.0000000596046448", "16777217", "16777219", "3.4028235e38", "1e-30", "7.006e-40",
            "0.1", "0.2", "0.3", "123456789012345678901234567890", ".5", "5.", "1e22", "1e23",
            "9007199254740993", "0.000000000000000000001", "0." + "3" * 80, "1" * 70 + "e-69"]
    for _ in range(3000):
        nd = int(rng.integers(1, 22))
        digits = "".join(str(d) for d in rng.integers(0, 10, nd))
        dot = int(rng.integers(0, nd + 1))
        t = digits[:dot] + "." + digits[dot:] if rng.random() < 0.8 else digits
        if rng.random() < 0.3:
            t += "e" + str(int(rng.integers(-25, 25)))
        toks.append(t)
    for _ in range(2000):  # exact midpoints between adjacent floats
        f = np.float32(rng.uniform(1e-3, 1e6))
        lo = float(f)
        hi = float(np.nextafter(f, np.float32(np.inf)))
        toks.append(repr((lo + hi) / 2))
    n = len(toks)
    body = "".join(f"{i % 50 + 1} {(i * 7) % 50 + 1} {t}\n" for i, t in enumerate(toks))
    f = tmp_path / "w.mtx"
    f.write_text(f"%%MatrixMarket matrix coordinate real general\n50 50 {n}\n" + body)
    a = lvlib.read_mtx(str(f))
    b = oracle.read_mtx(str(f), quantize=False)
    rp, cl, w = b.csr_arrays()
    assert np.array_equal(a.col, cl)
    assert np.array_equal(a.w.view(np.uint32), w.view(np.uint32))


@pytest.mark.parametrize("chunk", [None, 3])
def test_reader_error_entries(lvlib, tmp_path, monkeypatch, chunk):
    if chunk is not None:
        monkeypatch.setenv("LV_MTX_CHUNK_BYTES", str(chunk))
    head = "%%MatrixMarket matrix coordinate real symmetric\n3 3 4\n"
    cases = {"1 2 1\n2 3 1\n3 1 1\n": "expected 4 entries, got 3",
             "1 2 1\n2 3 1\n3 1 1\n1\n": "truncated entry 4",
             "1 2 1\n2 3 1\n3 1 1\n1 2\n": "missing weight in entry 4",
             "1 2 1\n2 x 1\n3 1 1\n1 2 1\n": "entry 2 is not a number",
             "1 2 1\n2 3 1\n0 1 1\n1 4 1\n": "entry 3 index out of range"}
    for i, (body, msg) in enumerate(cases.items()):
        f = tmp_path / f"e{i}.mtx"
        f.write_text(head + body)
        with pytest.raises(lvlib.LouvainError) as ei:
            lvlib.read_mtx(str(f))
        assert ei.value.status == lvlib.LV_ERR_ARG
        assert msg in str(ei.value), (msg, str(ei.value))


def test_binary_csr_round_trip_and_errors(lvlib, tmp_path):
    g = lvlib.gen_rmat(11, 8, seed=4)
    p = str(tmp_path / "g.lvcsr")
    lvlib.write_csr(p, g)
    for h in (lvlib.read_csr(p), lvlib.read_graph(p)):
        assert np.array_equal(h.rowptr, g.rowptr) and np.array_equal(h.col, g.col)
        assert np.array_equal(h.w.view(np.uint32), g.w.view(np.uint32))
    m = str(tmp_path / "g.mtx")
    lvlib.write_mtx(m, g)
    a, b = lvlib.read_graph(m), lvlib.read_mtx(m)
    assert np.array_equal(a.rowptr, b.rowptr) and np.array_equal(a.col, b.col)
    raw = open(p, "rb").read()
    (tmp_path / "cut.lvcsr").write_bytes(raw[:-4])
    with pytest.raises(lvlib.LouvainError) as ei:
        lvlib.read_csr(str(tmp_path / "cut.lvcsr"))
    assert ei.value.status == lvlib.LV_ERR_ARG
    with pytest.raises(lvlib.LouvainError) as ei:
        lvlib.read_csr(m)
    assert ei.value.status == lvlib.LV_ERR_ARG


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
