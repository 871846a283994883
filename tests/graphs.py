"""Small graph builders shared by the tests (plain numpy; no product code)."""
import numpy as np


def csr_from_edges(n, edges, weights=None):
    """Symmetric CSR in the reference reader's layout: for each undirected
    edge (i, j, w) in order, row i gets (j, w) and, if i != j, row j gets (i, w)."""
    rows = [[] for _ in range(n)]
    for t, (i, j) in enumerate(edges):
        w = 1.0 if weights is None else weights[t]
        rows[i].append((j, w))
        if i != j:
            rows[j].append((i, w))
    rowptr = np.zeros(n + 1, np.int32)
    for v in range(n):
        rowptr[v + 1] = rowptr[v] + len(rows[v])
    col = np.array([c for r in rows for c, _ in r], np.int32)
    w = np.array([x for r in rows for _, x in r], np.float32)
    return rowptr, col, w


def two_triangles():
    return csr_from_edges(6, [(0, 1), (0, 2), (1, 2), (2, 3), (3, 4), (3, 5), (4, 5)])


def random_graph(n, m, seed, weighted=False, self_loops=0):
    rng = np.random.default_rng(seed)
    seen = set()
    edges = []
    while len(edges) < m:
        i, j = rng.integers(0, n, 2)
        if i == j:
            continue
        key = (min(i, j), max(i, j))
        if key in seen:
            continue
        seen.add(key)
        edges.append((int(i), int(j)))
    for _ in range(self_loops):
        v = int(rng.integers(0, n))
        if (v, v) not in seen:
            seen.add((v, v))
            edges.append((v, v))
    weights = None
    if weighted:
        weights = [float((1 + rng.integers(0, 1 << 24)) * 2.0**-24) for _ in edges]
    return csr_from_edges(n, edges, weights)


def star(leaves, extra_ring=True):
    """Hub 0 joined to `leaves` vertices; optional ring among the leaves."""
    edges = [(0, i) for i in range(1, leaves + 1)]
    if extra_ring:
        edges += [(i, i + 1) for i in range(1, leaves)]
    return csr_from_edges(leaves + 1, edges)


def _mix64(x):
    """SplitMix64 finaliser over a uint64 numpy array."""
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    x = ((x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    x = ((x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    return x ^ (x >> np.uint64(31))


def real_weights(rowptr, col, seed, low=0.0, power=1.0):
    """Symmetric weights off every dyadic grid the generators use: for each
    undirected pair {u, v}, float32 of x**power with x a float64 uniform in
    (low, 1] drawn from a hash of (min, max, seed), so (u, v) and (v, u)
    agree.  power > 1 skews them toward small values (mixed magnitudes)."""
    rowptr = np.asarray(rowptr, np.int64)
    col = np.asarray(col, np.int64)
    row = np.repeat(np.arange(len(rowptr) - 1, dtype=np.int64), np.diff(rowptr))
    a, b = np.minimum(row, col).astype(np.uint64), np.maximum(row, col).astype(np.uint64)
    with np.errstate(over="ignore"):
        h = _mix64((a << np.uint64(32)) ^ b ^ _mix64(np.full(a.shape, seed, np.uint64)))
    u = (h >> np.uint64(11)).astype(np.float64) * 2.0**-53  # [0, 1)
    return ((1.0 - u * (1.0 - low)) ** power).astype(np.float32)


def write_decimal_mtx(path, n, edges, weights):
    """Symmetric MatrixMarket file whose weights are written as given (strings)."""
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real symmetric\n")
        f.write(f"{n} {n} {len(edges)}\n")
        for (i, j), w in zip(edges, weights):
            f.write(f"{i + 1} {j + 1} {w}\n")


def relabelled(rowptr, col, w, seed):
    """The same graph with its vertices renumbered by a seeded random
    permutation (rows sorted by the new column ids)."""
    rng = np.random.default_rng(seed)
    V = len(rowptr) - 1
    p = rng.permutation(V).astype(np.int64)  # old id -> new id
    rows = np.repeat(np.arange(V), np.diff(rowptr))
    nr, nc = p[rows], p[np.asarray(col, np.int64)]
    o = np.lexsort((nc, nr))
    rp = np.zeros(V + 1, np.int64)
    np.add.at(rp, nr + 1, 1)
    return np.cumsum(rp).astype(np.int32), nc[o].astype(np.int32), np.asarray(w, np.float32)[o]
