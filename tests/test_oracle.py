"""Pins the CPU oracle to the mathematics it restates (no GPU).

The reference ships no tests or golden vectors, so each oracle function is
checked against closed forms, brute force and textbook routines:
  - modularity: closed forms (two triangles 5/14, singletons, one community)
    and the textbook definition recomputed densely;
  - decide(): brute-force FP64 recomputation of Q after every single move
    (argmax, tie rule, singleton rule, move threshold);
  - gain(): identity dQ = g(c) - K_cur/M of the reference's formula;
  - aggregate(): dense contraction C^T A C, compaction order, Q preservation;
  - phase/run: invariants and the known two-triangle answer.
A test that plants a mistake does so in a Python copy of the formula and
checks that the brute force catches it.
"""
import numpy as np
import pytest

import graphs


def dense(rowptr, col, w, n):
    A = np.zeros((n, n), np.float64)
    for v in range(n):
        for e in range(rowptr[v], rowptr[v + 1]):
            A[v, col[e]] += w[e]
    return A


def q_dense(A, comm):
    """Textbook modularity with the reference's self-loop convention (a
    self-loop entry counts once in k_v and once in 2M)."""
    two_m = A.sum()
    k = A.sum(axis=1)
    same = comm[:, None] == comm[None, :]
    return (A[same].sum() - (np.outer(k, k)[same]).sum() / two_m) / two_m


def og_from(oracle, rowptr, col, w):
    return oracle.Graph(rowptr, col, w)


def test_two_triangles_closed_form(oracle):
    rowptr, col, w = graphs.two_triangles()
    g = og_from(oracle, rowptr, col, w)
    comm = np.array([0, 0, 0, 1, 1, 1], np.int32)
    assert oracle.modularity(g, comm) == pytest.approx(5.0 / 14.0, abs=1e-15)
    st, o2c, lv = oracle.run(g, 1e-6)
    assert st["modularity"] == pytest.approx(5.0 / 14.0, abs=1e-15)
    assert lv == 2
    assert list(o2c) == [0, 0, 0, 1, 1, 1]


def test_singletons_and_one_community_closed_forms(oracle):
    rowptr, col, w = graphs.random_graph(30, 80, 3, weighted=True)
    g = og_from(oracle, rowptr, col, w)
    k = oracle.strength(g)
    M = oracle.total_weight(g)
    q_single = oracle.modularity(g, np.arange(30, dtype=np.int32))
    assert q_single == pytest.approx(-(k * k).sum() / (4 * M * M), rel=1e-12)
    q_one = oracle.modularity(g, np.zeros(30, np.int32))
    assert q_one == pytest.approx(0.0, abs=1e-12)


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_modularity_matches_dense_definition(oracle, seed):
    rowptr, col, w = graphs.random_graph(25, 60, seed, weighted=True, self_loops=4)
    g = og_from(oracle, rowptr, col, w)
    A = dense(rowptr, col, w, 25)
    rng = np.random.default_rng(seed)
    for _ in range(5):
        comm = rng.integers(0, 6, 25).astype(np.int32)
        assert oracle.modularity(g, comm) == pytest.approx(q_dense(A, comm), abs=1e-12)


def test_gain_identity(oracle):
    """g(c) - K_cur/M equals the exact modularity change of moving v to c."""
    rng = np.random.default_rng(5)
    for _ in range(200):
        M = float(rng.uniform(10, 1000))
        k = float(rng.uniform(0.5, 20))
        scur = k + float(rng.uniform(0, 50))
        sc = float(rng.uniform(0, 50))
        K = float(rng.uniform(0, k))
        Kcur = float(rng.uniform(0, k - K))
        g = oracle.gain(np.float32(K), np.float32(M), np.float32(k), np.float32(scur), np.float32(sc))
        exact = (K - Kcur) / M + k * ((scur - k) - sc) / (2 * M * M)
        assert g - np.float32(Kcur) / np.float32(M) == pytest.approx(exact, abs=1e-5 * (abs(exact) + 1 / M))


def brute_force_decision(A, comm, v, size):
    """Best move of v by FP64 recomputation of Q after each single move.
    Returns (best community or cur, sorted list of (dQ, c))."""
    n = A.shape[0]
    cur = comm[v]
    base = q_dense(A, comm)
    nbr = {int(comm[u]) for u in range(n) if A[v, u] != 0 and u != v}
    cands = []
    for c in sorted(nbr):
        if c == cur:
            continue
        if not (c < cur or size[c] > 1 or size[cur] > 1):
            continue
        t = comm.copy()
        t[v] = c
        cands.append((q_dense(A, t) - base, c))
    return cands


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("state", ["singletons", "random"])
def test_decide_matches_brute_force(oracle, seed, state):
    n = 12
    rowptr, col, w = graphs.random_graph(n, 26, seed, weighted=(seed % 2 == 0), self_loops=2)
    g = og_from(oracle, rowptr, col, w)
    A = dense(rowptr, col, w, n)
    rng = np.random.default_rng(100 + seed)
    comm = np.arange(n, dtype=np.int32) if state == "singletons" else rng.integers(0, 4, n).astype(np.int32)
    size = np.bincount(comm, minlength=n)
    got = oracle.decide(g, comm, -1)
    for v in range(n):
        if rowptr[v + 1] == rowptr[v]:
            assert got[v] == comm[v]
            continue
        cands = brute_force_decision(A, comm, v, size)
        pos = [(dq, c) for dq, c in cands if dq > 1e-7]
        if not pos:
            assert got[v] == comm[v], f"v={v} moved without a positive gain"
            continue
        best = max(dq for dq, _ in pos)
        near = [c for dq, c in pos if best - dq <= 1e-6 * abs(best)]
        if len(near) > 1 and len(set(near)) > 1:
            # near-tie: the fp32 gain decides; it must pick one of them
            assert got[v] in near
        else:
            assert got[v] == near[0], f"v={v}: oracle {got[v]} brute force {near[0]} ({cands})"


def test_tie_breaks_to_lowest_id(oracle):
    # v=2 joined with equal weight to singletons 0 and 1 (both degree 1):
    # equal gains -> lowest id.  Singletons 0 and 1 may not join the
    # higher-id singleton 2 (singleton rule).
    rowptr, col, w = graphs.csr_from_edges(3, [(2, 1), (2, 0)])
    g = og_from(oracle, rowptr, col, w)
    comm = np.arange(3, dtype=np.int32)
    got = oracle.decide(g, comm, -1)
    assert got.tolist() == [0, 1, 0]
    # a singleton may join a lower-id singleton, never a higher one
    rowptr, col, w = graphs.csr_from_edges(3, [(2, 1)])
    g = og_from(oracle, rowptr, col, w)
    got = oracle.decide(g, comm, -1)
    assert got[2] == 1 and got[1] == 1


def test_planted_mistake_is_caught():
    """A dropped term in the gain (no -S_c) changes decisions that the brute
    force rejects, so the brute-force check has teeth."""
    n = 12
    caught = 0
    for seed in range(6):
        rowptr, col, w = graphs.random_graph(n, 26, seed, weighted=True)
        A = dense(rowptr, col, w, n)
        comm = np.arange(n, dtype=np.int32)
        size = np.ones(n, int)
        k = A.sum(axis=1)
        M = A.sum() / 2
        for v in range(n):
            cands = brute_force_decision(A, comm, v, size)
            pos = [(dq, c) for dq, c in cands if dq > 1e-9]
            if not pos:
                continue
            truth = max(pos)[1]
            wrong = max(((A[v, c] / M + k[v] * (k[v] - k[v]) / (2 * M * M)), c)
                        for _, c in cands)[1]  # S_c dropped
            caught += int(wrong != truth)
    assert caught > 0


@pytest.mark.parametrize("seed", range(5))
def test_aggregate_matches_dense_contraction(oracle, seed):
    n = 20
    rowptr, col, w = graphs.random_graph(n, 45, seed, weighted=True, self_loops=3)
    g = og_from(oracle, rowptr, col, w)
    A = dense(rowptr, col, w, n)
    rng = np.random.default_rng(seed)
    comm = rng.integers(0, 7, n).astype(np.int32)
    o2c = comm.copy()  # a phase leaves orig2comm composed with comm
    nxt, nid = oracle.aggregate(g, comm, o2c)
    present = sorted(set(comm.tolist()))
    assert nxt.V == len(present)
    assert [nid[c] for c in present] == list(range(len(present)))  # order-preserving compaction
    C = np.zeros((n, nxt.V))
    for v in range(n):
        C[v, nid[comm[v]]] = 1.0
    Ap = C.T @ A @ C
    rp, cl, ww = nxt.csr_arrays()
    for r in range(nxt.V):
        cols = cl[rp[r]:rp[r + 1]]
        assert list(cols) == sorted(cols) and len(set(cols)) == len(cols)
        expect = np.nonzero(Ap[r])[0]
        assert list(cols) == list(expect)
        np.testing.assert_allclose(ww[rp[r]:rp[r + 1]], Ap[r, expect].astype(np.float32), rtol=1e-7)
    assert np.array_equal(o2c, nid[comm])
    M = oracle.total_weight(g)
    q_before = oracle.modularity(g, comm, M)
    q_after = oracle.modularity(nxt, np.arange(nxt.V, dtype=np.int32), M)
    assert q_after == pytest.approx(q_before, abs=1e-6)


def test_phase_invariants(oracle, lvlib):
    csr = lvlib.gen_planted(3000, 10, 0.03, 0.001, seed=9)
    g = oracle.Graph.from_csr(csr)
    comm = np.arange(g.V, dtype=np.int32)
    o2c = np.arange(g.V, dtype=np.int32)
    q0 = oracle.modularity(g, comm)
    changed, sweeps = oracle.phase(g, comm, 1e-6, o2c)
    assert changed and sweeps >= 1
    assert oracle.modularity(g, comm) > q0
    assert np.array_equal(o2c, comm)
    size = np.bincount(comm, minlength=g.V)
    assert size.sum() == g.V and (size >= 0).all()
    k = oracle.strength(g)
    sigma = np.bincount(comm, weights=k, minlength=g.V)
    assert sigma.sum() == pytest.approx(2 * oracle.total_weight(g), rel=1e-12)


def test_run_orig2comm_in_range(oracle, lvlib):
    csr = lvlib.gen_rmat(12, 8, seed=3)
    st, o2c, lv = oracle.run(oracle.Graph.from_csr(csr), 1e-6)
    assert o2c.min() >= 0 and o2c.max() < lv
    assert -0.5 <= st["modularity"] <= 1.0


def test_thread_count_does_not_change_results(oracle, lvlib):
    csr = lvlib.gen_rmat(13, 16, seed=8)
    g = oracle.Graph.from_csr(csr)
    a = oracle.run(g, 1e-6, threads=1)
    b = oracle.run(g, 1e-6, threads=4)
    assert a[0]["modularity"] == b[0]["modularity"]
    assert np.array_equal(a[1], b[1])


def test_zero_min_gain_terminates(oracle):
    rowptr, col, w = graphs.two_triangles()
    st, o2c, lv = oracle.run(og_from(oracle, rowptr, col, w), 0.0)
    assert st["modularity"] == pytest.approx(5 / 14)


def test_group_bounds(oracle):
    # modularity_optimisation.cuh:8 buckets {0,4,8,16,32,84,319,INT_MAX}
    expect = {0: -1, 1: 0, 4: 0, 5: 1, 8: 1, 9: 2, 16: 2, 17: 3, 32: 3, 33: 4, 84: 4, 85: 5, 319: 5,
              320: 6, 100000: 6}
    for d, b in expect.items():
        assert oracle.group_of_degree(d) == b


@pytest.mark.parametrize("seed", range(4))
def test_decide_gap_matches_candidate_enumeration(oracle, seed):
    """orc_decide_group_gap: the relative gap between the two best options of
    each vertex ({K[cur]/M} and every eligible candidate's fp32 gain),
    recomputed here by enumerating the candidates with oracle.gain.  A
    planted mistake (dropping the stay option) is caught."""
    n = 14
    rowptr, col, w = graphs.random_graph(n, 30, seed, weighted=True, self_loops=2)
    g = og_from(oracle, rowptr, col, w)
    M = oracle.total_weight(g)
    M32 = np.float32(M)
    k = oracle.strength(g)
    rng = np.random.default_rng(200 + seed)
    comm = rng.integers(0, 5, n).astype(np.int32)
    size = np.bincount(comm, minlength=n)
    sigma = np.zeros(n)
    np.add.at(sigma, comm, k)
    got, gap = oracle.decide_gap(g, comm, -1)
    assert np.array_equal(got, oracle.decide(g, comm, -1))

    def options(v, with_stay=True):
        cur = comm[v]
        K = {}
        for e in range(rowptr[v], rowptr[v + 1]):
            u = col[e]
            if u != v:
                K[comm[u]] = K.get(comm[u], 0.0) + float(w[e])
        vals = [float(np.float32(K.get(cur, 0.0)) / M32)] if with_stay else []
        for c, Kc in K.items():
            if c == cur or not (c < cur or size[c] > 1 or size[cur] > 1):
                continue
            vals.append(oracle.gain(np.float32(Kc), M32, np.float32(k[v]), np.float32(sigma[cur]),
                                    np.float32(sigma[c])))
        return sorted(vals, reverse=True)

    def rel(vals):
        if len(vals) < 2:
            return np.inf
        mag = max(abs(vals[0]), abs(vals[1]))
        return 0.0 if mag == 0 else (vals[0] - vals[1]) / mag

    checked = planted_differs = 0
    for v in range(n):
        if rowptr[v + 1] == rowptr[v]:
            assert np.isnan(gap[v])
            continue
        want = rel(options(v))
        assert gap[v] == pytest.approx(want, rel=1e-12, abs=0) or (np.isinf(gap[v]) and np.isinf(want))
        checked += 1
        planted_differs += rel(options(v, with_stay=False)) != want
    assert checked > 5 and planted_differs > 0
