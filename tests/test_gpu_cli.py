"""The product command line (bin/gpulouvain) on the golden fixtures (needs an
MI355X).

gpulouvain keeps the reference's interface (gpulouvain.cu:6-46,
utils.cu:161-186, modularity_optimisation.cu:403-418): line 1 is the final
modularity printed with %f, line 2 the timing line, and with -v the last
level's vertex count followed by one line per community (its id + 1, then
its original vertices + 1).  Each fixture goes through the HIP path; line 1
and the partition are compared with the golden values (networkx-computed
modularity and partitions in tests/golden/golden.json), and the whole output
with the oracle's seqlouvain byte for byte, the timing line excepted.
reader_conventions.mtx (comment lines, a `general` banner listing both
directions, a self-loop, an isolated vertex) has no golden partition; it is
compared with seqlouvain.
"""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
EXE = os.path.join(ROOT, "gpu-louvain_amd", "bin", "gpulouvain")
with open(os.path.join(GOLDEN, "golden.json")) as f:
    GOLD = json.load(f)

FIXTURES = ["two_triangles.mtx", "ring_of_cliques.mtx", "karate.mtx", "reader_conventions.mtx"]


def _run(cmd):
    print("run:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout.splitlines()


def _partition(lines, n):
    lv = int(lines[2])
    assert len(lines) == 3 + lv
    comm = [None] * n
    for line in lines[3:]:
        parts = [int(x) for x in line.split()]
        for x in parts[1:]:
            assert comm[x - 1] is None
            comm[x - 1] = parts[0] - 1
    assert None not in comm
    return comm


@pytest.mark.parametrize("name", FIXTURES)
def test_cli_matches_golden_and_oracle(oracle, name):
    path = os.path.join(GOLDEN, name)
    got = _run([EXE, "-f", path, "-g", "0.000001", "-v"])
    want = _run([oracle.SEQ_PATH, "-f", path, "-g", "0.000001", "-v"])
    # timing line: two numbers (algorithm ms, algorithm + upload ms)
    assert len(got[1].split()) == 2 and all(float(x) >= 0 for x in got[1].split())
    assert got[0] == want[0]
    assert got[2:] == want[2:]
    if "partition" in GOLD.get(name, {}):
        gold = GOLD[name]
        assert got[0] == f"{gold['modularity']:f}"
        assert _partition(got, len(gold["partition"])) == gold["partition"]


def test_cli_without_v_prints_two_lines(oracle):
    got = _run([EXE, "-f", os.path.join(GOLDEN, "karate.mtx"), "-g", "0.000001"])
    assert len(got) == 2 and got[0] == f"{GOLD['karate.mtx']['modularity']:f}"


def test_cli_reads_binary_csr(lvlib, tmp_path):
    g = lvlib.read_mtx(os.path.join(GOLDEN, "karate.mtx"))
    p = str(tmp_path / "karate.lvcsr")
    lvlib.write_csr(p, g)
    a = _run([EXE, "-f", p, "-g", "0.000001", "-v"])
    b = _run([EXE, "-f", os.path.join(GOLDEN, "karate.mtx"), "-g", "0.000001", "-v"])
    assert a[0] == b[0] and a[2:] == b[2:]
