"""lv_modularity_noalloc: modularity over caller-owned device arrays with no
device allocation (needs an MI355X).

Checked against the existing entry point lv_modularity of a handle holding
the same graph and partition (and the FP64 oracle), on two calls in a row,
the second with a smaller input, then on a partition with an out-of-range
community (the kernels' own range checks turn it into NaN; nothing outside
the arrays is touched).  The test's own device arrays are allocated through
the HIP runtime the library links (libamdhip64, plumbing only).
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

Q_TOL = 1e-12


def _sig(lib):
    lib.lv_modularity_noalloc.restype = ctypes.c_int32
    lib.lv_modularity_noalloc.argtypes = [ctypes.c_int32, ctypes.c_void_p] + [ctypes.c_void_p] * 4 + [
        ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p]


def _hip():
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    hip.hipSetDevice.argtypes = [ctypes.c_int]
    for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipDeviceSynchronize", "hipSetDevice"):
        getattr(hip, f).restype = ctypes.c_int
    return hip


H2D, D2H = 1, 2


def _noalloc_q(lib, hip, rowptr, col, w, comm):
    assert hip.hipSetDevice(0) == 0
    host = [np.ascontiguousarray(a) for a in (rowptr, col, w, comm)] + [np.full(1, -7.0)]
    ptrs = []
    try:
        for a in host:
            p = ctypes.c_void_p()
            assert hip.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
            ptrs.append(p)
            assert hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        st = lib.lv_modularity_noalloc(0, None, *ptrs[:4], len(rowptr) - 1, len(col), ptrs[4])
        assert hip.hipDeviceSynchronize() == 0
        q = np.zeros(1)
        assert hip.hipMemcpy(q.ctypes.data, ptrs[4], 8, D2H) == 0
        return st, float(q[0])
    finally:
        for p in ptrs:
            hip.hipFree(p)


def test_noalloc_modularity_matches_lv_modularity(lvlib, oracle):
    lib = lvlib.load()
    _sig(lib)
    hip = _hip()
    cases = [("rmat14", lvlib.gen_rmat(14, 16, seed=81), 40),
             ("planted3000", lvlib.gen_planted(3000, 12, 0.03, 0.002, seed=82, unit_weights=False), 12)]
    sizes = []
    for name, g, ncomm in cases:
        rng = np.random.default_rng(len(name))
        # vertices without edges stay alone; the others share ncomm labels
        # taken from their own ids
        nz = np.nonzero(np.diff(g.rowptr) > 0)[0]
        labels = rng.choice(nz, ncomm, replace=False)
        comm = np.arange(g.V, dtype=np.int32)
        comm[nz] = labels[rng.integers(0, ncomm, len(nz))]
        print(f"case {name} V={g.V} E={g.E}", flush=True)
        with lvlib.Graph(g) as G:
            G.set_communities(comm)
            q_ref = G.modularity()
            h = G.level_graph()  # the handle's (grid-rounded) weights
        st, q = _noalloc_q(lib, hip, h.rowptr, h.col, h.w, comm)
        assert st == lvlib.LV_OK, lvlib.load().lv_last_error()
        assert abs(q - q_ref) <= Q_TOL, (q, q_ref)
        og = oracle.Graph(h.rowptr, h.col, h.w)
        assert abs(q - oracle.modularity(og, comm)) <= Q_TOL
        sizes.append(g.V)
    assert sizes[1] < sizes[0]


def test_noalloc_flags_bad_community_and_rejects_oversize(lvlib):
    lib = lvlib.load()
    _sig(lib)
    hip = _hip()
    g = lvlib.gen_planted(2000, 8, 0.03, 0.002, seed=83)
    comm = np.zeros(g.V, np.int32)
    comm[5] = g.V + 3  # outside [0, V)
    st, q = _noalloc_q(lib, hip, g.rowptr, g.col, g.w, comm)
    assert st == lvlib.LV_OK and np.isnan(q)
    # a good call afterwards reuses the same scratch correctly
    comm[5] = 0
    st, q = _noalloc_q(lib, hip, g.rowptr, g.col, g.w, comm)
    assert st == lvlib.LV_OK and abs(q) <= 1e-12  # one community: Q = 0
    big = lvlib.LV_NOALLOC_MAX_V + 1
    st = lib.lv_modularity_noalloc(0, None, 1, 1, 1, 1, big, 0, 1)
    assert st == lvlib.LV_ERR_ARG
