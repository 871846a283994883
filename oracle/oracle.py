"""ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
benchmark's CPU-baseline leg, as the checker.  See oracle.h for the
reference file:line each function restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
SEQ_PATH = os.path.join(_HERE, "_build", "seqlouvain")

I32P = ctypes.POINTER(ctypes.c_int32)
F32P = ctypes.POINTER(ctypes.c_float)
F64P = ctypes.POINTER(ctypes.c_double)


class OrcCSR(ctypes.Structure):
    _fields_ = [("V", ctypes.c_int32), ("E", ctypes.c_int64), ("rowptr", I32P), ("col", I32P), ("w", F32P)]


class OrcStats(ctypes.Structure):
    _fields_ = [("levels", ctypes.c_int32), ("sweeps", ctypes.c_int32), ("final_V", ctypes.c_int32),
                ("modularity", ctypes.c_double), ("seconds", ctypes.c_double)]


_lib = None


def build() -> None:
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER(OrcCSR)
    sig = {
        "orc_read_mtx": (ctypes.c_int, [ctypes.c_char_p, P]),
        "orc_free_csr": (None, [P]),
        "orc_last_error": (ctypes.c_char_p, []),
        "orc_total_weight": (ctypes.c_double, [P]),
        "orc_strength": (None, [P, F64P]),
        "orc_group_of_degree": (ctypes.c_int, [ctypes.c_int32]),
        "orc_decide_group": (None, [P, ctypes.c_double, I32P, ctypes.c_int, I32P, ctypes.c_int]),
        "orc_gain": (ctypes.c_float, [ctypes.c_float] * 5),
        "orc_modularity": (ctypes.c_double, [P, ctypes.c_double, I32P]),
        "orc_sweep": (ctypes.c_int64, [P, ctypes.c_double, I32P, ctypes.c_int]),
        "orc_decide_group_obs": (None, [P, ctypes.c_double, I32P, ctypes.c_int, I32P, F32P, F32P, ctypes.c_int]),
        "orc_sweep_masked": (ctypes.c_int64, [P, ctypes.c_double, I32P, ctypes.POINTER(ctypes.c_uint8),
                                              ctypes.c_int]),
        "orc_phase": (ctypes.c_int, [P, ctypes.c_double, I32P, ctypes.c_float, I32P, ctypes.c_int32, I32P,
                                     ctypes.c_int]),
        "orc_aggregate": (ctypes.c_int, [P, I32P, P, I32P, ctypes.c_int32, I32P]),
        "orc_run": (ctypes.c_int, [P, ctypes.c_float, I32P, I32P, ctypes.POINTER(OrcStats), ctypes.c_int]),
        "orc_weight_shift": (ctypes.c_int32, [F32P, ctypes.c_int64]),
        "orc_quantize_weight": (ctypes.c_float, [ctypes.c_float, ctypes.c_int32]),
        "orc_quantize": (ctypes.c_int64, [P]),
        "orc_decide_group_gap": (None, [P, ctypes.c_double, I32P, ctypes.c_int, I32P, F32P, F32P, F64P,
                                        ctypes.c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


class Graph:
    """Holds numpy CSR arrays and an OrcCSR view of them.  quantize=True
    rounds the weights to the grid the HIP library rounds them to before a
    run (orc_quantize); the arrays are copied first."""

    def __init__(self, rowptr, col, w, quantize: bool = False):
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        self.col = np.ascontiguousarray(col if len(col) else np.zeros(1, np.int32), dtype=np.int32)
        self.w = np.array(w if len(w) else np.zeros(1, np.float32), dtype=np.float32, copy=True)
        self.E = int(len(col))
        self.V = int(len(rowptr) - 1)
        self.c = OrcCSR(self.V, self.E, _p(self.rowptr, ctypes.c_int32), _p(self.col, ctypes.c_int32),
                        _p(self.w, ctypes.c_float))
        self.changed = int(load().orc_quantize(ctypes.byref(self.c))) if quantize else 0

    @classmethod
    def from_csr(cls, g, quantize: bool = True):
        """The oracle's copy of a library CSR, on the library's weight grid
        (quantize=False keeps the weights as given)."""
        return cls(g.rowptr, g.col, g.w, quantize=quantize)

    def csr_arrays(self):
        return self.rowptr, self.col[: self.E], self.w[: self.E]

    @property
    def ref(self):
        return ctypes.byref(self.c)


def weight_shift(w: np.ndarray) -> int:
    w = np.ascontiguousarray(w, dtype=np.float32)
    return int(load().orc_weight_shift(_p(w, ctypes.c_float), len(w)))


def quantize_weight(w: float, shift: int) -> float:
    return float(load().orc_quantize_weight(w, shift))


def decide_gap(g: Graph, comm: np.ndarray, group: int, M: float | None = None, threads: int = 1):
    """decide() plus each decided vertex's relative gap between its two best
    options (orc_decide_group_gap; NaN for vertices outside the group)."""
    lib = load()
    M = total_weight(g) if M is None else M
    comm = np.ascontiguousarray(comm, dtype=np.int32)
    out = np.zeros(max(g.V, 1), np.int32)
    gap = np.full(max(g.V, 1), np.nan, np.float64)
    lib.orc_decide_group_gap(g.ref, M, _p(comm, ctypes.c_int32), group, _p(out, ctypes.c_int32), None, None,
                             _p(gap, ctypes.c_double), threads)
    return out[: g.V], gap[: g.V]


def read_mtx(path: str, quantize: bool = True) -> Graph:
    lib = load()
    c = OrcCSR()
    rc = lib.orc_read_mtx(path.encode(), ctypes.byref(c))
    if rc != 0:
        raise RuntimeError(lib.orc_last_error().decode())
    V, E = c.V, c.E
    rowptr = np.ctypeslib.as_array(c.rowptr, shape=(V + 1,)).copy()
    col = np.ctypeslib.as_array(c.col, shape=(max(E, 1),))[:E].copy()
    w = np.ctypeslib.as_array(c.w, shape=(max(E, 1),))[:E].copy()
    lib.orc_free_csr(ctypes.byref(c))
    return Graph(rowptr, col, w, quantize=quantize)


def total_weight(g: Graph) -> float:
    return load().orc_total_weight(g.ref)


def strength(g: Graph) -> np.ndarray:
    k = np.zeros(max(g.V, 1), np.float64)
    load().orc_strength(g.ref, _p(k, ctypes.c_double))
    return k[: g.V]


def group_of_degree(d: int) -> int:
    return load().orc_group_of_degree(d)


def gain(K: float, M32: float, k: float, scur: float, sc: float) -> float:
    return load().orc_gain(K, M32, k, scur, sc)


def decide(g: Graph, comm: np.ndarray, group: int, M: float | None = None, threads: int = 1) -> np.ndarray:
    lib = load()
    M = total_weight(g) if M is None else M
    comm = np.ascontiguousarray(comm, dtype=np.int32)
    out = np.zeros(max(g.V, 1), np.int32)
    lib.orc_decide_group(g.ref, M, _p(comm, ctypes.c_int32), group, _p(out, ctypes.c_int32), threads)
    return out[: g.V]


def decide_obs(g: Graph, comm: np.ndarray, group: int, M: float | None = None, threads: int = 1):
    """decide() plus the margin and fp32 K[cur] of each decided vertex (NaN elsewhere)."""
    lib = load()
    M = total_weight(g) if M is None else M
    comm = np.ascontiguousarray(comm, dtype=np.int32)
    out = np.zeros(max(g.V, 1), np.int32)
    margin = np.full(max(g.V, 1), np.nan, np.float32)
    kown = np.full(max(g.V, 1), np.nan, np.float32)
    lib.orc_decide_group_obs(g.ref, M, _p(comm, ctypes.c_int32), group, _p(out, ctypes.c_int32),
                             _p(margin, ctypes.c_float), _p(kown, ctypes.c_float), threads)
    return out[: g.V], margin[: g.V], kown[: g.V]


def sweep_masked(g: Graph, comm: np.ndarray, active: np.ndarray, M: float | None = None,
                 threads: int = 1) -> int:
    """One sweep in place deciding only vertices with active != 0."""
    M = total_weight(g) if M is None else M
    assert comm.dtype == np.int32 and comm.flags.c_contiguous
    active = np.ascontiguousarray(active, dtype=np.uint8)
    return int(load().orc_sweep_masked(g.ref, M, _p(comm, ctypes.c_int32), _p(active, ctypes.c_uint8),
                                       threads))


def modularity(g: Graph, comm: np.ndarray, M: float | None = None) -> float:
    M = total_weight(g) if M is None else M
    comm = np.ascontiguousarray(comm, dtype=np.int32)
    return load().orc_modularity(g.ref, M, _p(comm, ctypes.c_int32))


def sweep(g: Graph, comm: np.ndarray, M: float | None = None, threads: int = 1) -> int:
    """One sweep in place on `comm`; returns moved count."""
    M = total_weight(g) if M is None else M
    assert comm.dtype == np.int32 and comm.flags.c_contiguous
    return int(load().orc_sweep(g.ref, M, _p(comm, ctypes.c_int32), threads))


def phase(g: Graph, comm: np.ndarray, min_gain: float, orig2comm: np.ndarray, M: float | None = None,
          threads: int = 1):
    M = total_weight(g) if M is None else M
    sweeps = ctypes.c_int32()
    changed = load().orc_phase(g.ref, M, _p(comm, ctypes.c_int32), min_gain, _p(orig2comm, ctypes.c_int32),
                               len(orig2comm), ctypes.byref(sweeps), threads)
    return bool(changed), sweeps.value


def aggregate(g: Graph, comm: np.ndarray, orig2comm: np.ndarray | None = None):
    """Returns (next Graph, new_id array); renumbers orig2comm in place."""
    lib = load()
    comm = np.ascontiguousarray(comm, dtype=np.int32)
    if orig2comm is None:
        orig2comm = np.zeros(1, np.int32)
        n0 = 0
    else:
        n0 = len(orig2comm)
    out = OrcCSR()
    nid = np.zeros(max(g.V, 1), np.int32)
    lib.orc_aggregate(g.ref, _p(comm, ctypes.c_int32), ctypes.byref(out), _p(orig2comm, ctypes.c_int32), n0,
                      _p(nid, ctypes.c_int32))
    V, E = out.V, out.E
    rowptr = np.ctypeslib.as_array(out.rowptr, shape=(V + 1,)).copy()
    col = np.ctypeslib.as_array(out.col, shape=(max(E, 1),))[:E].copy()
    w = np.ctypeslib.as_array(out.w, shape=(max(E, 1),))[:E].copy()
    lib.orc_free_csr(ctypes.byref(out))
    return Graph(rowptr, col, w), nid[: g.V]


def run(g: Graph, min_gain: float, threads: int = 1):
    """Full run; returns (stats dict, orig2comm, level_V)."""
    lib = load()
    o2c = np.zeros(max(g.V, 1), np.int32)
    lv = ctypes.c_int32()
    st = OrcStats()
    lib.orc_run(g.ref, min_gain, _p(o2c, ctypes.c_int32), ctypes.byref(lv), ctypes.byref(st), threads)
    stats = {"levels": st.levels, "sweeps": st.sweeps, "final_V": st.final_V, "modularity": st.modularity,
             "seconds": st.seconds}
    return stats, o2c[: g.V], lv.value
