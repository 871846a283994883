"""The weight grid (no GPU needed).

Before a run the library rounds every weight to a grid of 2^-s on which
every FP64 sum of weights the engine forms is exact, so no result depends on
the order of atomics.  These tests pin the oracle's restatement of the rule
(orc_weight_shift / orc_quantize_weight) by hand-computed cases and exact
rational arithmetic, check that the library's host implementation
(lv_weight_grid / lv_quantize_weights) agrees with it, and check the input
checks lv_create makes before it touches a device.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import graphs


def test_shift_known_answers(oracle, lvlib):
    # T = 4: 2^c >= 4 * (1 + 2^-20) needs c = 3, s = 47
    cases = [([1.0, 1.0, 1.0, 1.0], 47),
             # T = 3: c = 2
             ([1.0, 2.0], 48),
             # T = 1: c = 1 (2^0 < 1 + 2^-20)
             ([1.0], 49),
             # T = 0.75: c = 0
             ([0.5, 0.25], 50),
             # signs do not matter: T = sum |w| = 6
             ([-3.0, 3.0], 47),
             ([], 0),
             ([0.0, 0.0], 0)]
    for w, s in cases:
        w = np.asarray(w, np.float32)
        assert oracle.weight_shift(w) == s, (w, s)
        assert lvlib.weight_grid(w)[0] == s, (w, s)


def test_shift_margin_near_powers_of_two(oracle, lvlib):
    # T = 2^k * (1 - 2^-19): T * (1 + 2^-20) < 2^k, so c = k; T = 2^k: c = k + 1
    for k in (-5, 0, 10, 26):
        assert oracle.weight_shift(np.asarray([2.0**k], np.float32)) == 50 - (k + 1)
        below = np.asarray([2.0**k * (1 - 2.0**-19)], np.float32)
        assert oracle.weight_shift(below) == 50 - k
        assert lvlib.weight_grid(below)[0] == 50 - k


def test_library_and_oracle_agree(oracle, lvlib):
    rng = np.random.default_rng(7)
    for n, scale in [(1, 1.0), (17, 1e-3), (1000, 1.0), (70000, 0.5), (200000, 1e3), (5000, 1e-30)]:
        w = (rng.random(n) * scale).astype(np.float32)
        w[rng.integers(0, n, max(1, n // 10))] *= np.float32(1e-6)
        s = oracle.weight_shift(w)
        assert lvlib.weight_grid(w)[0] == s
        q = lvlib.quantize_weights(w, s)
        sample = rng.integers(0, n, min(n, 2000))
        for i in sample:
            assert q[i] == np.float32(oracle.quantize_weight(float(w[i]), s))


def test_rounding_properties(oracle, lvlib):
    rng = np.random.default_rng(11)
    w = np.concatenate([rng.random(5000), rng.random(500) * 1e-7, [1e-30, 0.0, 1.0]]).astype(np.float32)
    s, changed = lvlib.weight_grid(w)
    q = lvlib.quantize_weights(w, s)
    assert changed == int(np.count_nonzero(q.view(np.uint32) != w.view(np.uint32)))
    assert changed > 0
    step = 2.0**-s
    # on the grid, within half a step, idempotent
    assert np.all(np.abs(q.astype(np.float64) - w.astype(np.float64)) <= step / 2)
    assert np.all(np.ldexp(q.astype(np.float64), s) == np.rint(np.ldexp(q.astype(np.float64), s)))
    assert np.array_equal(lvlib.quantize_weights(q, s), q)
    # the rounded weights' own grid is as fine or finer: rounding again is a no-op
    assert lvlib.weight_grid(q)[1] == 0


def test_sums_on_the_grid_are_exact(lvlib):
    """Every FP64 partial sum of rounded weights (signed, below 8T) equals the
    exact rational sum, in any order: the property that makes results
    independent of the order of atomics."""
    rng = np.random.default_rng(3)
    w = np.concatenate([rng.random(20000), rng.random(2000) * 1e-9]).astype(np.float32)
    s, _ = lvlib.weight_grid(w)
    q = lvlib.quantize_weights(w, s).astype(np.float64)
    exact = sum(Fraction(float(x)) for x in q)
    for seed in range(4):
        perm = np.random.default_rng(seed).permutation(len(q))
        acc = 0.0
        for x in q[perm]:
            acc += x
        assert Fraction(acc) == exact
    # signed partial sums (joins and leaves of a community total, corrections)
    signs = np.where(rng.random(len(q)) < 0.5, -1.0, 1.0)
    acc, ex = 0.0, Fraction(0)
    for x, sg in zip(q, signs):
        acc += sg * x
        ex += Fraction(sg * x)
        assert Fraction(acc) == ex
    # 8 T: the bound of every partial sum the engine forms
    big = 0.0
    for _ in range(8):
        for x in q:
            big += x
    assert Fraction(big) == 8 * exact


def test_generator_weights_are_on_the_grid(lvlib):
    for g in (lvlib.gen_rmat(14, 16, seed=14), lvlib.gen_rmat(12, 16, seed=1, unit_weights=True),
              lvlib.gen_planted(3000, 10, 0.02, 0.001, seed=2, unit_weights=False)):
        assert lvlib.weight_grid(g.w)[1] == 0


def test_real_weights_change_and_oracle_graph_matches(oracle, lvlib):
    g = lvlib.gen_rmat(10, 16, seed=5)
    w = graphs.real_weights(g.rowptr, g.col, seed=9, power=4.0)
    s, changed = lvlib.weight_grid(w)
    assert changed > 0
    og = oracle.Graph(g.rowptr, g.col, w, quantize=True)
    assert og.changed == changed
    assert np.array_equal(og.csr_arrays()[2].view(np.uint32), lvlib.quantize_weights(w, s).view(np.uint32))
    # symmetric: both directions of a pair carry the same (rounded) weight
    q = og.csr_arrays()[2]
    row = np.repeat(np.arange(g.V), np.diff(g.rowptr))
    fwd = {(int(a), int(b)): float(x) for a, b, x in zip(row, g.col, q)}
    assert all(fwd[(b, a)] == x for (a, b), x in fwd.items())


def _create_status(lvlib, rowptr, col, w, **opt):
    try:
        with lvlib.Graph(lvlib.CSR(np.asarray(rowptr, np.int32), np.asarray(col, np.int32),
                                   np.asarray(w, np.float32)), **opt):
            return 0, ""
    except lvlib.LouvainError as e:
        return e.status, str(e)


def test_create_rejects_asymmetric_csr(lvlib):
    # 0-1 stored in row 0 only
    st, msg = _create_status(lvlib, [0, 1, 1], [1], [1.0])
    assert st == lvlib.LV_ERR_ARG and "symmetric" in msg
    # both directions present but with different weights
    st, msg = _create_status(lvlib, [0, 1, 2], [1, 0], [1.0, 2.0])
    assert st == lvlib.LV_ERR_ARG and "symmetric" in msg


def test_create_rejects_non_finite_weights(lvlib):
    st, msg = _create_status(lvlib, [0, 1, 2], [1, 0], [np.inf, np.inf])
    assert st == lvlib.LV_ERR_ARG and "finite" in msg


def test_batch_create_checks_shape_before_ranges(lvlib):
    """A decreasing rowptr is reported, not walked (the range check reads
    col only after the shape checks)."""
    import ctypes
    lib = lvlib.load()
    lvlib._batch_sigs(lib)
    vptr = np.asarray([0, 2], np.int32)
    rowptr = np.asarray([0, 5, 2], np.int32)
    col = np.asarray([1, 0], np.int32)
    w = np.asarray([1.0, 1.0], np.float32)
    h = ctypes.c_void_p()
    opt = lvlib.Options()
    st = lib.lv_batch_create(1, lvlib._p(vptr, ctypes.c_int32), lvlib._p(rowptr, ctypes.c_int32),
                             lvlib._p(col, ctypes.c_int32), lvlib._p(w, ctypes.c_float), 2,
                             ctypes.byref(opt), ctypes.byref(h))
    assert st == lvlib.LV_ERR_ARG
    # asymmetric union
    rowptr = np.asarray([0, 1, 1], np.int32)
    st = lib.lv_batch_create(1, lvlib._p(vptr, ctypes.c_int32), lvlib._p(rowptr, ctypes.c_int32),
                             lvlib._p(col, ctypes.c_int32), lvlib._p(w, ctypes.c_float), 1,
                             ctypes.byref(opt), ctypes.byref(h))
    assert st == lvlib.LV_ERR_ARG
