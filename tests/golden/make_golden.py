"""Writes the small MatrixMarket fixtures under tests/golden/ and their
expected results (golden.json).

Expected modularities are computed here with networkx's textbook
`community.modularity` (an independent implementation) for partitions that
are known by construction (two triangles, ring of cliques) or were found by
the CPU oracle (karate club); self-loops are handled by the reference's
convention (weight w counted once in k_v and in 2M), which for networkx means
halving a self-loop's weight.  Run from the repository root:
    python tests/golden/make_golden.py
"""
import json
import os
import subprocess
import sys

import networkx as nx

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def write(name, n, edges, banner="real symmetric", comments=("% fixture",), weights=True):
    path = os.path.join(HERE, name)
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate {banner}\n")
        for c in comments:
            f.write(c + "\n")
        f.write(f"{n} {n} {len(edges)}\n")
        for e in edges:
            if weights:
                f.write(f"{e[0]} {e[1]} {e[2]}\n")
            else:
                f.write(f"{e[0]} {e[1]}\n")
    return path


def nx_graph(n, edges):
    G = nx.Graph()
    G.add_nodes_from(range(n))
    for i, j, w in edges:
        i -= 1
        j -= 1
        w = float(w)
        if i == j:
            w = w / 2.0  # networkx counts a self-loop twice in the degree
        if G.has_edge(i, j):
            G[i][j]["weight"] += w
        else:
            G.add_edge(i, j, weight=w)
    return G


def nx_q(n, edges, comm):
    parts = {}
    for v, c in enumerate(comm):
        parts.setdefault(c, set()).add(v)
    return nx.community.modularity(nx_graph(n, edges), list(parts.values()), weight="weight")


def seq_partition(path):
    out = subprocess.run([os.path.join(ROOT, "oracle", "_build", "seqlouvain"), "-f", path, "-g", "1e-6", "-v"],
                         check=True, capture_output=True, text=True).stdout.splitlines()
    lv = int(out[2])
    comm = {}
    for line in out[3:3 + lv]:
        parts = [int(x) for x in line.split()]
        for x in parts[1:]:
            comm[x - 1] = parts[0] - 1
    return [comm[v] for v in range(len(comm))], out


def main():
    golden = {}
    tri = [(2, 1, 1), (3, 1, 1), (3, 2, 1), (4, 3, 1), (5, 4, 1), (6, 4, 1), (6, 5, 1)]
    write("two_triangles.mtx", 6, tri)
    golden["two_triangles.mtx"] = {"partition": [0, 0, 0, 1, 1, 1],
                                   "modularity": nx_q(6, tri, [0, 0, 0, 1, 1, 1])}

    ring = []
    for c in range(4):
        base = 4 * c + 1
        for a in range(4):
            for b in range(a + 1, 4):
                ring.append((base + b, base + a, 1))
        ring.append((4 * ((c + 1) % 4) + 1, base + 3, 1))
    write("ring_of_cliques.mtx", 16, ring)
    part = [v // 4 for v in range(16)]
    golden["ring_of_cliques.mtx"] = {"partition": part, "modularity": nx_q(16, ring, part)}

    K = nx.karate_club_graph()
    kar = [(max(u, v) + 1, min(u, v) + 1, 1) for u, v in K.edges()]
    p = write("karate.mtx", 34, kar, banner="pattern symmetric", weights=False)
    comm, _ = seq_partition(p)
    golden["karate.mtx"] = {"partition": comm, "modularity": nx_q(34, kar, comm)}

    # reader conventions: comments, self-loops, general banner with both
    # directions listed (stored twice, as the reference does), real weights
    odd = [(1, 1, 2.5), (2, 1, 0.5), (1, 2, 0.5), (3, 2, 1.25), (4, 4, 1), (3, 1, 0.75)]
    write("reader_conventions.mtx", 5, odd, banner="real general", comments=("% a", "%b", "% c d"))
    golden["reader_conventions.mtx"] = {
        # rows in insertion order: triple t puts (j,w) in row i, (i,w) in row j if i != j
        "rowptr": [0, 4, 7, 9, 10, 10],
        "col": [0, 1, 1, 2, 0, 0, 2, 1, 0, 3],
        "w": [2.5, 0.5, 0.5, 0.75, 0.5, 0.5, 1.25, 1.25, 0.75, 1.0],
    }
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(golden, f, indent=1)
    print("wrote", sorted(golden))


if __name__ == "__main__":
    sys.exit(main())
