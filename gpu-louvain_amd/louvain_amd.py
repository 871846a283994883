"""Python binding of the MI355X Louvain library (ctypes over include/louvain.h).

Host-side mirror of the reference's operations, for tests and benchmarks:

    Graph(csr)            copyStructures + initM      (utils.cu:64-106, modularity_optimisation.cu:420-422)
    Graph.optimise(g)     optimiseModularity           (modularity_optimisation.cu:284-363)
    Graph.aggregate()     aggregateCommunities         (community_aggregation.cu:248-347)
    Graph.modularity()    calculateModularity          (modularity_optimisation.cu:395-401)
    Graph.run(g)          the level loop of main       (gpulouvain.cu:30-44)
    Graph.assignment()    data of printOriginalToCommunity (modularity_optimisation.cu:403-418)
    read_mtx(path)        readInputData                (utils.cu:9-62)

The learned frontier (no reference counterpart) adds Graph.run_pruned,
Graph.trace_phase, Graph.features, Graph.set_model, Graph.predict and the
standalone model kernels mlp_forward, mlp_grad and mlp_train.

Every call goes to the native library; there is no Python fallback.  A
missing library raises LouvainError at load time.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libgpulouvain.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "louvain.h")

LV_OK = 0
LV_ERR_ARG = -1
LV_ERR_OOM = -2
LV_ERR_HIP = -3
LV_ERR_IO = -4
LV_ERR_STATE = -5
LV_ERR_NODEVICE = -6
LV_NOALLOC_MAX_V = 1 << 20  # louvain.h LV_NOALLOC_MAX_V

K_MOVE_SUB, K_MOVE_WAVE, K_MOVE_WG, K_MOVE_GLOBAL, K_APPLY, K_MODULARITY, K_AGGREGATE = range(7)
KERNEL_NAMES = ["move_sub", "move_wave", "move_wg", "move_chunked", "apply", "modularity", "aggregate"]
# single HIP kernels (lv_kernel_timing_detail) and their device function names
# at the default options (sub_lanes picks the k_move_sub<lanes, entries>
# instance of buckets 1..3; wave_split and the workgroup tiers move rows
# between the others)
DETAIL_NAMES = ["move_b0", "move_b1", "move_b2", "move_b3", "move_b4", "move_b5", "move_b6", "move_b7",
                "chunk_aggregate", "chunk_merge", "chunk_final", "modularity_rows", "frontier",
                "modularity_xcd", "move_b8", "move_b9", "move_b10", "heavy_lds", "hinted"]
DETAIL_KERNELS = ["k_move_sub<4, 1>", "k_move_sub<4, 2>", "k_move_sub<4, 4>", "k_move_sub<4, 8>",
                  "k_move_wave<128, 3, 32>", "k_move_wave<192, 4, 32>", "k_move_wave<448, 0, 64>",
                  "k_move_wg<128, 0, false>",
                  "k_chunk_aggregate", "k_chunk_merge", "k_heavy_final", "k_modularity", "k_mlp<false>",
                  "k_modularity_xcd", "k_move_wg<256, 1, false>", "k_move_wg<256, 2, false>",
                  "k_move_wg<1024, 3, false>", "k_move_wg<1024, 4, false>", "k_move_hinted"]

I32P = ctypes.POINTER(ctypes.c_int32)
I64P = ctypes.POINTER(ctypes.c_int64)
F32P = ctypes.POINTER(ctypes.c_float)
F64P = ctypes.POINTER(ctypes.c_double)
U8P = ctypes.POINTER(ctypes.c_uint8)


class LouvainError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"louvain status {status}: {message}")
        self.status = status


class Options(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int32),
        ("timing", ctypes.c_int32),
        ("heavy_wg_max", ctypes.c_int32),
        ("chunk_edges", ctypes.c_int32),
        ("agg_lds_max", ctypes.c_int32),
        ("part_target", ctypes.c_int32),
        ("merge_limit", ctypes.c_int32),
        ("q_incremental", ctypes.c_int32),
        ("xcd_rows", ctypes.c_int32),
        ("agg_mode", ctypes.c_int32),
        ("compact", ctypes.c_int32),
        ("candidate_bounds", ctypes.c_int32),
        ("heavy_lds", ctypes.c_int32),
        ("exact_sums", ctypes.c_int32),
        ("wave_split", ctypes.c_int32),
        ("wg_bound_a", ctypes.c_int32),
        ("wg_bound_b", ctypes.c_int32),
        ("wg_bound_c", ctypes.c_int32),
        ("wg_lanes", ctypes.c_int32 * 4),
        ("table_pct", ctypes.c_int32),
        ("sub_lanes", ctypes.c_int32 * 3),
        ("col_order", ctypes.c_int32),
        ("hinted", ctypes.c_int32),
        ("nmv_pull", ctypes.c_int32),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("levels", ctypes.c_int32),
        ("sweeps", ctypes.c_int32),
        ("final_V", ctypes.c_int32),
        ("level0_sweeps", ctypes.c_int32),
        ("final_E", ctypes.c_int64),
        ("edges_swept", ctypes.c_int64),
        ("modularity", ctypes.c_double),
        ("algo_ms", ctypes.c_double),
        ("sweep_ms", ctypes.c_double),
        ("level0_sweep_ms", ctypes.c_double),
        ("aggregate_ms", ctypes.c_double),
        ("memory_ms", ctypes.c_double),
        ("evaluated", ctypes.c_int64),
        ("full_sweeps", ctypes.c_int32),
        ("pad_", ctypes.c_int32),
        ("reserved", ctypes.c_double * 4),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_
                if name not in ("reserved", "pad_")}


MLP_PARAMS = 1377
FEATURES = 8
HIDDEN = 32
FRONTIER_MODEL, FRONTIER_NEIGHBOR, FRONTIER_RANDOM, FRONTIER_ALL = range(4)


class Frontier(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("full_every", ctypes.c_int32),
        ("threshold", ctypes.c_float),
        ("random_p", ctypes.c_float),
        ("seed", ctypes.c_uint64),
        ("heavy_threshold", ctypes.c_float),
        ("heavy_degree", ctypes.c_int32),
        ("and_neighbor", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 5),
    ]


class TrainOptions(ctypes.Structure):
    _fields_ = [
        ("epochs", ctypes.c_int32),
        ("batch", ctypes.c_int32),
        ("lr", ctypes.c_float),
        ("pos_weight", ctypes.c_float),
        ("beta1", ctypes.c_float),
        ("beta2", ctypes.c_float),
        ("eps", ctypes.c_float),
        ("reserved", ctypes.c_int32 * 9),
    ]


def frontier(kind: int = FRONTIER_MODEL, full_every: int = 4, threshold: float = 0.5,
             random_p: float = 0.25, seed: int = 1, heavy_threshold: float = 0.0,
             heavy_degree: int = 0, and_neighbor: bool = False) -> Frontier:
    return Frontier(kind=kind, full_every=full_every, threshold=threshold, random_p=random_p,
                    seed=seed, heavy_threshold=heavy_threshold, heavy_degree=heavy_degree,
                    and_neighbor=int(and_neighbor))


class KernelTime(ctypes.Structure):
    _fields_ = [("launches", ctypes.c_int64), ("total_ms", ctypes.c_double), ("bytes", ctypes.c_double)]


_lib = None


def load() -> ctypes.CDLL:
    """Load libgpulouvain.so (built by `make -C gpu-louvain_amd`)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LouvainError(LV_ERR_STATE, f"native library missing: {LIB_PATH} (build it first)")
    lib = ctypes.CDLL(LIB_PATH)
    sig = {
        "lv_create": (ctypes.c_int32, [I32P, I32P, F32P, ctypes.c_int32, ctypes.c_int64,
                                       ctypes.POINTER(Options), ctypes.POINTER(ctypes.c_void_p)]),
        "lv_destroy": (None, [ctypes.c_void_p]),
        "lv_reset": (ctypes.c_int32, [ctypes.c_void_p]),
        "lv_optimise": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_float, I32P]),
        "lv_aggregate": (ctypes.c_int32, [ctypes.c_void_p]),
        "lv_modularity": (ctypes.c_int32, [ctypes.c_void_p, F64P]),
        "lv_run": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_float, ctypes.POINTER(Stats)]),
        "lv_assignment": (ctypes.c_int32, [ctypes.c_void_p, I32P, I32P]),
        "lv_level_size": (ctypes.c_int32, [ctypes.c_void_p, I32P, I64P]),
        "lv_level_graph": (ctypes.c_int32, [ctypes.c_void_p, I32P, I32P, F32P]),
        "lv_dump_level": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int32]),
        "lv_create_from_dump": (ctypes.c_int32, [ctypes.c_char_p, ctypes.POINTER(Options),
                                                 ctypes.POINTER(ctypes.c_void_p)]),
        "lv_get_communities": (ctypes.c_int32, [ctypes.c_void_p, I32P]),
        "lv_set_communities": (ctypes.c_int32, [ctypes.c_void_p, I32P]),
        "lv_decide": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, I32P]),
        "lv_sweep": (ctypes.c_int32, [ctypes.c_void_p, F64P, I64P]),
        "lv_kernel_timing": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(KernelTime)]),
        "lv_kernel_timing_reset": (ctypes.c_int32, [ctypes.c_void_p]),
        "lv_kernel_timing_detail": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(KernelTime)]),
        "lv_set_options": (ctypes.c_int32, [ctypes.c_void_p, ctypes.POINTER(Options)]),
        "lv_last_error": (ctypes.c_char_p, []),
        "lv_read_mtx": (ctypes.c_int32, [ctypes.c_char_p, ctypes.POINTER(I32P), ctypes.POINTER(I32P),
                                         ctypes.POINTER(F32P), I32P, I64P]),
        "lv_write_mtx": (ctypes.c_int32, [ctypes.c_char_p, I32P, I32P, F32P, ctypes.c_int32]),
        "lv_write_csr": (ctypes.c_int32, [ctypes.c_char_p, I32P, I32P, F32P, ctypes.c_int32, ctypes.c_int64]),
        "lv_read_csr": (ctypes.c_int32, [ctypes.c_char_p, ctypes.POINTER(I32P), ctypes.POINTER(I32P),
                                         ctypes.POINTER(F32P), I32P, I64P]),
        "lv_read_graph": (ctypes.c_int32, [ctypes.c_char_p, ctypes.POINTER(I32P), ctypes.POINTER(I32P),
                                           ctypes.POINTER(F32P), I32P, I64P]),
        "lv_gen_rmat": (ctypes.c_int32, [ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_uint64, ctypes.c_int32,
                                         ctypes.POINTER(I32P), ctypes.POINTER(I32P), ctypes.POINTER(F32P),
                                         I32P, I64P]),
        "lv_gen_planted": (ctypes.c_int32, [ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                            ctypes.c_uint64, ctypes.c_int32, ctypes.POINTER(I32P),
                                            ctypes.POINTER(I32P), ctypes.POINTER(F32P), I32P, I64P]),
        "lv_free_host": (None, [ctypes.c_void_p]),
        "lv_weight_grid": (ctypes.c_int32, [F32P, ctypes.c_int64, I32P, I64P]),
        "lv_quantize_weights": (ctypes.c_int32, [F32P, ctypes.c_int64, ctypes.c_int32, F32P]),
        "lv_run_pruned": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_float, ctypes.POINTER(Frontier),
                                           ctypes.POINTER(Stats)]),
        "lv_optimise_pruned": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_float, ctypes.POINTER(Frontier),
                                                I32P, I32P, I64P]),
        "lv_sweep_pruned": (ctypes.c_int32, [ctypes.c_void_p, ctypes.POINTER(Frontier), ctypes.c_int32,
                                             F64P, I64P, I64P]),
        "lv_trace_phase": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_float, F32P, U8P, ctypes.c_int32,
                                            I32P, I32P]),
        "lv_features": (ctypes.c_int32, [ctypes.c_void_p, F32P]),
        "lv_move_state": (ctypes.c_int32, [ctypes.c_void_p, F32P, F32P, U8P, I32P]),
        "lv_set_model": (ctypes.c_int32, [ctypes.c_void_p, F32P]),
        "lv_predict": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_float, F32P, U8P]),
        "lv_mlp_forward": (ctypes.c_int32, [ctypes.c_int32, F32P, F32P, ctypes.c_int64, F32P]),
        "lv_mlp_grad": (ctypes.c_int32, [ctypes.c_int32, F32P, F32P, F32P, ctypes.c_int64, ctypes.c_float,
                                         F32P, F32P]),
        "lv_mlp_train": (ctypes.c_int32, [ctypes.c_int32, F32P, F32P, ctypes.c_int64,
                                          ctypes.POINTER(TrainOptions), F32P, F32P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(status: int) -> None:
    if status != LV_OK:
        raise LouvainError(status, load().lv_last_error().decode(errors="replace"))


def _p(a: np.ndarray, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


@dataclass
class CSR:
    """Symmetric CSR graph: rowptr[V+1] int32, col[E] int32, w[E] float32."""

    rowptr: np.ndarray
    col: np.ndarray
    w: np.ndarray

    @property
    def V(self) -> int:
        return int(self.rowptr.shape[0] - 1)

    @property
    def E(self) -> int:
        return int(self.col.shape[0])

    def degrees(self) -> np.ndarray:
        return np.diff(self.rowptr)


def _take_host_csr(rp, cl, wt, V: ctypes.c_int32, E: ctypes.c_int64) -> CSR:
    lib = load()
    n, m = V.value, E.value
    rowptr = np.ctypeslib.as_array(rp, shape=(n + 1,)).copy()
    col = np.ctypeslib.as_array(cl, shape=(max(m, 1),))[:m].copy()
    w = np.ctypeslib.as_array(wt, shape=(max(m, 1),))[:m].copy()
    for ptr in (rp, cl, wt):
        lib.lv_free_host(ctypes.cast(ptr, ctypes.c_void_p))
    return CSR(rowptr, col, w)


def _read(fn: str, path: str) -> CSR:
    lib = load()
    rp, cl, wt = I32P(), I32P(), F32P()
    V, E = ctypes.c_int32(), ctypes.c_int64()
    _check(getattr(lib, fn)(path.encode(), ctypes.byref(rp), ctypes.byref(cl), ctypes.byref(wt),
                            ctypes.byref(V), ctypes.byref(E)))
    return _take_host_csr(rp, cl, wt, V, E)


def read_mtx(path: str) -> CSR:
    return _read("lv_read_mtx", path)


def read_csr(path: str) -> CSR:
    """Binary CSR file (lv_write_csr)."""
    return _read("lv_read_csr", path)


def read_graph(path: str) -> CSR:
    """Binary CSR or MatrixMarket file, by the file's magic."""
    return _read("lv_read_graph", path)


def write_csr(path: str, g: CSR) -> None:
    lib = load()
    rowptr = np.ascontiguousarray(g.rowptr, dtype=np.int32)
    col = np.ascontiguousarray(g.col, dtype=np.int32)
    w = np.ascontiguousarray(g.w, dtype=np.float32)
    _check(lib.lv_write_csr(path.encode(), _p(rowptr, ctypes.c_int32), _p(col, ctypes.c_int32),
                            _p(w, ctypes.c_float), g.V, g.E))


def write_mtx(path: str, g: CSR) -> None:
    lib = load()
    rowptr = np.ascontiguousarray(g.rowptr, dtype=np.int32)
    col = np.ascontiguousarray(g.col, dtype=np.int32)
    w = np.ascontiguousarray(g.w, dtype=np.float32)
    _check(lib.lv_write_mtx(path.encode(), _p(rowptr, ctypes.c_int32), _p(col, ctypes.c_int32),
                            _p(w, ctypes.c_float), g.V))


def weight_grid(w: np.ndarray) -> tuple[int, int]:
    """(shift, changed): the 2^-shift grid every weight is rounded to before a
    run (lv_weight_grid; host only) and how many weights it changes."""
    lib = load()
    w = np.ascontiguousarray(w, dtype=np.float32)
    shift, changed = ctypes.c_int32(), ctypes.c_int64()
    _check(lib.lv_weight_grid(_p(w, ctypes.c_float), len(w), ctypes.byref(shift), ctypes.byref(changed)))
    return shift.value, changed.value


def quantize_weights(w: np.ndarray, shift: int | None = None) -> np.ndarray:
    """Weights rounded to the grid (lv_quantize_weights; host only)."""
    lib = load()
    w = np.ascontiguousarray(w, dtype=np.float32)
    if shift is None:
        shift = weight_grid(w)[0]
    out = np.empty_like(w)
    _check(lib.lv_quantize_weights(_p(w, ctypes.c_float), len(w), shift, _p(out, ctypes.c_float)))
    return out


def gen_rmat(scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
             seed: int = 1, unit_weights: bool = False) -> CSR:
    lib = load()
    rp, cl, wt = I32P(), I32P(), F32P()
    V, E = ctypes.c_int32(), ctypes.c_int64()
    _check(lib.lv_gen_rmat(scale, edge_factor, a, b, c, seed, int(unit_weights), ctypes.byref(rp),
                           ctypes.byref(cl), ctypes.byref(wt), ctypes.byref(V), ctypes.byref(E)))
    return _take_host_csr(rp, cl, wt, V, E)


def gen_planted(n: int, blocks: int, p_in: float, p_out: float, seed: int = 1,
                unit_weights: bool = True) -> CSR:
    lib = load()
    rp, cl, wt = I32P(), I32P(), F32P()
    V, E = ctypes.c_int32(), ctypes.c_int64()
    _check(lib.lv_gen_planted(n, blocks, p_in, p_out, seed, int(unit_weights), ctypes.byref(rp),
                              ctypes.byref(cl), ctypes.byref(wt), ctypes.byref(V), ctypes.byref(E)))
    return _take_host_csr(rp, cl, wt, V, E)


class Graph:
    """A graph resident on one GPU with its community state (lv_graph)."""

    def __init__(self, g: CSR, device: int = 0, timing: bool = False, **options):
        """options: lv_options fields (heavy_wg_max, chunk_edges, q_incremental,
        exact_sums, wg_lanes=(a, b, c, d), ...); 0 = the library default."""
        lib = load()
        self._lib = lib
        self._rowptr = np.ascontiguousarray(g.rowptr, dtype=np.int32)
        self._col = np.ascontiguousarray(g.col, dtype=np.int32)
        self._w = np.ascontiguousarray(g.w, dtype=np.float32)
        self.V0 = int(self._rowptr.shape[0] - 1)
        self.device = device
        opt = options_struct(device, timing, **options)
        h = ctypes.c_void_p()
        col = self._col if self._col.size else np.zeros(1, np.int32)
        w = self._w if self._w.size else np.zeros(1, np.float32)
        _check(lib.lv_create(_p(self._rowptr, ctypes.c_int32), _p(col, ctypes.c_int32),
                             _p(w, ctypes.c_float), self.V0, int(self._col.size), ctypes.byref(opt),
                             ctypes.byref(h)))
        self._h = h
        self.Vorig = self.V0

    @classmethod
    def from_dump(cls, prefix: str, device: int = 0, timing: bool = False) -> "Graph":
        """A handle on a level written by dump_level (lv_create_from_dump)."""
        lib = load()
        self = cls.__new__(cls)
        self._lib = lib
        self.device = device
        opt = Options()
        opt.device = device
        opt.timing = int(timing)
        h = ctypes.c_void_p()
        _check(lib.lv_create_from_dump(prefix.encode(), ctypes.byref(opt), ctypes.byref(h)))
        self._h = h
        V, _ = self.level_size()
        self.V0 = V
        with open(prefix + ".o2c", "rb") as f:
            self.Vorig = int(np.frombuffer(f.read(24)[8:16], np.int64)[0])
        return self

    def dump_level(self, prefix: str, mtx: bool = False) -> None:
        """Writes prefix.lvcsr, prefix.o2c (and prefix.mtx) for the current level."""
        _check(self._lib.lv_dump_level(self._h, prefix.encode(), int(mtx)))

    # --- lifecycle
    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.lv_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self) -> None:
        _check(self._lib.lv_reset(self._h))

    # --- phases
    def optimise(self, min_gain: float) -> bool:
        changed = ctypes.c_int32()
        _check(self._lib.lv_optimise(self._h, min_gain, ctypes.byref(changed)))
        return bool(changed.value)

    def aggregate(self) -> None:
        _check(self._lib.lv_aggregate(self._h))

    def modularity(self) -> float:
        q = ctypes.c_double()
        _check(self._lib.lv_modularity(self._h, ctypes.byref(q)))
        return q.value

    def run(self, min_gain: float) -> dict:
        st = Stats()
        _check(self._lib.lv_run(self._h, min_gain, ctypes.byref(st)))
        return st.as_dict()

    def sweep(self) -> tuple[float, int]:
        """One sweep; returns (modularity after the sweep, vertices moved)."""
        q, moved = ctypes.c_double(), ctypes.c_int64()
        _check(self._lib.lv_sweep(self._h, ctypes.byref(q), ctypes.byref(moved)))
        return q.value, moved.value

    # --- state
    def level_size(self) -> tuple[int, int]:
        V, E = ctypes.c_int32(), ctypes.c_int64()
        _check(self._lib.lv_level_size(self._h, ctypes.byref(V), ctypes.byref(E)))
        return V.value, E.value

    def level_graph(self) -> CSR:
        V, E = self.level_size()
        rowptr = np.zeros(V + 1, np.int32)
        col = np.zeros(max(E, 1), np.int32)
        w = np.zeros(max(E, 1), np.float32)
        _check(self._lib.lv_level_graph(self._h, _p(rowptr, ctypes.c_int32), _p(col, ctypes.c_int32),
                                        _p(w, ctypes.c_float)))
        return CSR(rowptr, col[:E], w[:E])

    def assignment(self) -> tuple[np.ndarray, int]:
        o2c = np.zeros(max(self.Vorig, 1), np.int32)
        lv = ctypes.c_int32()
        _check(self._lib.lv_assignment(self._h, _p(o2c, ctypes.c_int32), ctypes.byref(lv)))
        return o2c[: self.Vorig], lv.value

    def communities(self) -> np.ndarray:
        V, _ = self.level_size()
        comm = np.zeros(max(V, 1), np.int32)
        _check(self._lib.lv_get_communities(self._h, _p(comm, ctypes.c_int32)))
        return comm[:V]

    def set_communities(self, comm: np.ndarray) -> None:
        comm = np.ascontiguousarray(comm, dtype=np.int32)
        _check(self._lib.lv_set_communities(self._h, _p(comm, ctypes.c_int32)))

    def decide(self, group: int) -> np.ndarray:
        V, _ = self.level_size()
        out = np.zeros(max(V, 1), np.int32)
        _check(self._lib.lv_decide(self._h, group, _p(out, ctypes.c_int32)))
        return out[:V]

    def kernel_timing(self, kernel: int) -> dict:
        kt = KernelTime()
        _check(self._lib.lv_kernel_timing(self._h, kernel, ctypes.byref(kt)))
        return {"launches": kt.launches, "total_ms": kt.total_ms, "bytes": kt.bytes}

    def set_options(self, **options) -> None:
        """New tier options for this handle (results do not depend on them);
        fields left out take the library default."""
        opt = options_struct(self.device, False, **options)
        _check(self._lib.lv_set_options(self._h, ctypes.byref(opt)))

    def kernel_timing_detail(self, kernel: int) -> dict:
        kt = KernelTime()
        _check(self._lib.lv_kernel_timing_detail(self._h, kernel, ctypes.byref(kt)))
        return {"launches": kt.launches, "total_ms": kt.total_ms, "bytes": kt.bytes}

    def kernel_timing_reset(self) -> None:
        _check(self._lib.lv_kernel_timing_reset(self._h))

    # --- learned frontier
    def run_pruned(self, min_gain: float, fr: Frontier) -> dict:
        st = Stats()
        _check(self._lib.lv_run_pruned(self._h, min_gain, ctypes.byref(fr), ctypes.byref(st)))
        return st.as_dict()

    def optimise_pruned(self, min_gain: float, fr: Frontier) -> tuple[bool, int, int]:
        """One pruned phase; returns (changed, sweeps, vertex evaluations)."""
        ch, sw, ev = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
        _check(self._lib.lv_optimise_pruned(self._h, min_gain, ctypes.byref(fr), ctypes.byref(ch),
                                            ctypes.byref(sw), ctypes.byref(ev)))
        return bool(ch.value), sw.value, ev.value

    def sweep_pruned(self, fr: Frontier, track: bool = True) -> tuple[float, int, int]:
        """One sweep over fr's frontier; returns (Q after, moved, evaluated)."""
        q, mv, ev = ctypes.c_double(), ctypes.c_int64(), ctypes.c_int64()
        _check(self._lib.lv_sweep_pruned(self._h, ctypes.byref(fr), int(track), ctypes.byref(q),
                                         ctypes.byref(mv), ctypes.byref(ev)))
        return q.value, mv.value, ev.value

    def trace_phase(self, min_gain: float, max_sweeps: int = 64) -> tuple[np.ndarray, np.ndarray, bool]:
        """One plain phase with traces: features [s, V, 8] before each sweep and
        moved flags [s, V] of each sweep."""
        V, _ = self.level_size()
        feat = np.zeros((max_sweeps, max(V, 1), FEATURES), np.float32)
        moved = np.zeros((max_sweeps, max(V, 1)), np.uint8)
        n, ch = ctypes.c_int32(), ctypes.c_int32()
        _check(self._lib.lv_trace_phase(self._h, min_gain, _p(feat, ctypes.c_float), _p(moved, ctypes.c_uint8),
                                        max_sweeps, ctypes.byref(n), ctypes.byref(ch)))
        return feat[: n.value, :V], moved[: n.value, :V], bool(ch.value)

    def features(self) -> np.ndarray:
        V, _ = self.level_size()
        feat = np.zeros((max(V, 1), FEATURES), np.float32)
        _check(self._lib.lv_features(self._h, _p(feat, ctypes.c_float)))
        return feat[:V]

    def move_state(self) -> dict:
        V, _ = self.level_size()
        n = max(V, 1)
        margin = np.zeros(n, np.float32)
        kown = np.zeros(n, np.float32)
        hist = np.zeros(n, np.uint8)
        nmv = np.zeros(n, np.int32)
        _check(self._lib.lv_move_state(self._h, _p(margin, ctypes.c_float), _p(kown, ctypes.c_float),
                                       _p(hist, ctypes.c_uint8), _p(nmv, ctypes.c_int32)))
        return {"margin": margin[:V], "kown": kown[:V], "hist": hist[:V], "nmv": nmv[:V]}

    def set_model(self, params: np.ndarray) -> None:
        params = np.ascontiguousarray(params, dtype=np.float32)
        if params.size != MLP_PARAMS:
            raise ValueError(f"model has {MLP_PARAMS} parameters, got {params.size}")
        _check(self._lib.lv_set_model(self._h, _p(params, ctypes.c_float)))

    def predict(self, threshold: float = 0.5) -> tuple[np.ndarray, np.ndarray]:
        V, _ = self.level_size()
        prob = np.zeros(max(V, 1), np.float32)
        act = np.zeros(max(V, 1), np.uint8)
        _check(self._lib.lv_predict(self._h, threshold, _p(prob, ctypes.c_float), _p(act, ctypes.c_uint8)))
        return prob[:V], act[:V]


def mlp_forward(params: np.ndarray, feat: np.ndarray, device: int = 0) -> np.ndarray:
    """MFMA inference: sigmoid(MLP(feat)) for feat [n, 8]."""
    lib = load()
    params = np.ascontiguousarray(params, dtype=np.float32)
    feat = np.ascontiguousarray(feat, dtype=np.float32).reshape(-1, FEATURES)
    prob = np.zeros(max(feat.shape[0], 1), np.float32)
    _check(lib.lv_mlp_forward(device, _p(params, ctypes.c_float), _p(feat, ctypes.c_float), feat.shape[0],
                              _p(prob, ctypes.c_float)))
    return prob[: feat.shape[0]]


def mlp_grad(params: np.ndarray, feat: np.ndarray, labels: np.ndarray, pos_weight: float = 1.0,
             device: int = 0) -> tuple[np.ndarray, float]:
    """Mean gradient and mean loss of the weighted BCE over (feat, labels)."""
    lib = load()
    params = np.ascontiguousarray(params, dtype=np.float32)
    feat = np.ascontiguousarray(feat, dtype=np.float32).reshape(-1, FEATURES)
    labels = np.ascontiguousarray(labels, dtype=np.float32)
    grad = np.zeros(MLP_PARAMS, np.float32)
    loss = ctypes.c_float()
    _check(lib.lv_mlp_grad(device, _p(params, ctypes.c_float), _p(feat, ctypes.c_float),
                           _p(labels, ctypes.c_float), feat.shape[0], pos_weight, _p(grad, ctypes.c_float),
                           ctypes.byref(loss)))
    return grad, loss.value


def mlp_train(feat: np.ndarray, labels: np.ndarray, params: np.ndarray, epochs: int, batch: int = 4096,
              lr: float = 1e-3, pos_weight: float = 1.0, device: int = 0) -> tuple[np.ndarray, np.ndarray]:
    """Adam training on the device; returns (trained params, mean loss per epoch)."""
    lib = load()
    params = np.array(params, dtype=np.float32, copy=True)
    feat = np.ascontiguousarray(feat, dtype=np.float32).reshape(-1, FEATURES)
    labels = np.ascontiguousarray(labels, dtype=np.float32)
    losses = np.zeros(max(epochs, 1), np.float32)
    opt = TrainOptions(epochs=epochs, batch=batch, lr=lr, pos_weight=pos_weight)
    _check(lib.lv_mlp_train(device, _p(feat, ctypes.c_float), _p(labels, ctypes.c_float), feat.shape[0],
                            ctypes.byref(opt), _p(params, ctypes.c_float), _p(losses, ctypes.c_float)))
    return params, losses[:epochs]


def mlp_init(seed: int = 0) -> np.ndarray:
    """He-initialised parameters in the library's layout."""
    rng = np.random.default_rng(seed)
    w1 = rng.normal(0.0, np.sqrt(2.0 / FEATURES), (HIDDEN, FEATURES))
    w2 = rng.normal(0.0, np.sqrt(2.0 / HIDDEN), (HIDDEN, HIDDEN))
    w3 = rng.normal(0.0, np.sqrt(1.0 / HIDDEN), HIDDEN)
    return np.concatenate([w1.ravel(), np.zeros(HIDDEN), w2.ravel(), np.zeros(HIDDEN), w3, [0.0]]).astype(np.float32)


class BatchResult(ctypes.Structure):
    _fields_ = [("modularity", ctypes.c_double), ("levels", ctypes.c_int32), ("sweeps", ctypes.c_int32),
                ("final_V", ctypes.c_int32), ("reserved", ctypes.c_int32)]


def options_struct(device: int = 0, timing: bool = False, **options) -> Options:
    """An lv_options with the given fields (wg_lanes takes a sequence of 4,
    sub_lanes one of 3)."""
    opt = Options()
    opt.device = device
    opt.timing = int(timing)
    fields = dict(Options._fields_)
    for k, v in options.items():
        if k not in fields or k in ("device", "timing", "reserved"):
            raise TypeError(f"unknown option {k}")
        if k in ("wg_lanes", "sub_lanes"):
            n = 4 if k == "wg_lanes" else 3
            v = list(v) + [0] * (n - len(v))
            setattr(opt, k, (ctypes.c_int32 * n)(*[int(x) for x in v[:n]]))
        else:
            setattr(opt, k, int(v))
    return opt


def union(graphs) -> tuple[np.ndarray, CSR]:
    """Block-diagonal union of CSRs: (vptr [G+1], union CSR with global ids)."""
    vptr = np.zeros(len(graphs) + 1, np.int64)
    eptr = np.zeros(len(graphs) + 1, np.int64)
    for i, g in enumerate(graphs):
        vptr[i + 1] = vptr[i] + g.V
        eptr[i + 1] = eptr[i] + g.E
    if vptr[-1] >= 2**31 or eptr[-1] >= 2**31:
        raise ValueError("union exceeds int32 indices")
    rowptr = np.zeros(int(vptr[-1]) + 1, np.int32)
    col = np.zeros(int(eptr[-1]), np.int32)
    w = np.zeros(int(eptr[-1]), np.float32)
    for i, g in enumerate(graphs):
        v0, e0 = int(vptr[i]), int(eptr[i])
        rowptr[v0 + 1:v0 + g.V + 1] = g.rowptr[1:] + e0
        col[e0:e0 + g.E] = g.col + v0
        w[e0:e0 + g.E] = g.w
    return vptr.astype(np.int32), CSR(rowptr, col, w)


class Batch:
    """Independent graphs solved together (lv_batch_*)."""

    def __init__(self, graphs, device: int = 0, timing: bool = False, **options):
        """options: lv_options fields (heavy_wg_max, q_incremental, ...)."""
        lib = load()
        _batch_sigs(lib)
        self._lib = lib
        self.vptr, u = union(graphs)
        self.G = len(graphs)
        self._keep = (u.rowptr, u.col if u.col.size else np.zeros(1, np.int32),
                      u.w if u.w.size else np.zeros(1, np.float32))
        self.device = device
        opt = options_struct(device, timing, **options)
        h = ctypes.c_void_p()
        _check(lib.lv_batch_create(self.G, _p(self.vptr, ctypes.c_int32), _p(self._keep[0], ctypes.c_int32),
                                   _p(self._keep[1], ctypes.c_int32), _p(self._keep[2], ctypes.c_float),
                                   int(u.col.size), ctypes.byref(opt), ctypes.byref(h)))
        self._h = h
        self.V = int(self.vptr[-1])

    def set_options(self, **options) -> None:
        """New tier options between runs (lv_batch_set_options)."""
        opt = options_struct(self.device, False, **options)
        _check(self._lib.lv_batch_set_options(self._h, ctypes.byref(opt)))

    def run(self, min_gain: float) -> tuple[list[dict], dict]:
        res = (BatchResult * self.G)()
        st = Stats()
        _check(self._lib.lv_batch_run(self._h, min_gain, res, ctypes.byref(st)))
        per = [{"modularity": r.modularity, "levels": r.levels, "sweeps": r.sweeps, "final_V": r.final_V}
               for r in res]
        return per, st.as_dict()

    def set_model(self, params: np.ndarray) -> None:
        params = np.ascontiguousarray(params, dtype=np.float32)
        if params.size != MLP_PARAMS:
            raise ValueError(f"model has {MLP_PARAMS} parameters")
        _check(self._lib.lv_batch_set_model(self._h, _p(params, ctypes.c_float)))

    def run_pruned(self, min_gain: float, fr: Frontier) -> tuple[list[dict], dict]:
        """Learned-pruned batched run (lv_batch_run_pruned)."""
        res = (BatchResult * self.G)()
        st = Stats()
        _check(self._lib.lv_batch_run_pruned(self._h, min_gain, ctypes.byref(fr), res, ctypes.byref(st)))
        per = [{"modularity": r.modularity, "levels": r.levels, "sweeps": r.sweeps, "final_V": r.final_V}
               for r in res]
        return per, st.as_dict()

    def assignment(self) -> list[np.ndarray]:
        comm = np.zeros(max(self.V, 1), np.int32)
        _check(self._lib.lv_batch_assignment(self._h, _p(comm, ctypes.c_int32)))
        return [comm[self.vptr[i]:self.vptr[i + 1]] for i in range(self.G)]

    def kernel_timing(self, kernel: int) -> dict:
        kt = KernelTime()
        _check(self._lib.lv_batch_kernel_timing(self._h, kernel, ctypes.byref(kt)))
        return {"launches": kt.launches, "total_ms": kt.total_ms, "bytes": kt.bytes}

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.lv_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_batch_done = False


def _batch_sigs(lib) -> None:
    global _batch_done
    if _batch_done:
        return
    sig = {
        "lv_batch_create": (ctypes.c_int32, [ctypes.c_int32, I32P, I32P, I32P, F32P, ctypes.c_int64,
                                             ctypes.POINTER(Options), ctypes.POINTER(ctypes.c_void_p)]),
        "lv_batch_destroy": (None, [ctypes.c_void_p]),
        "lv_batch_run": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_float, ctypes.POINTER(BatchResult),
                                          ctypes.POINTER(Stats)]),
        "lv_batch_assignment": (ctypes.c_int32, [ctypes.c_void_p, I32P]),
        "lv_batch_set_model": (ctypes.c_int32, [ctypes.c_void_p, F32P]),
        "lv_batch_run_pruned": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_float, ctypes.POINTER(Frontier),
                                                 ctypes.POINTER(BatchResult), ctypes.POINTER(Stats)]),
        "lv_batch_kernel_timing": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(KernelTime)]),
        "lv_batch_set_options": (ctypes.c_int32, [ctypes.c_void_p, ctypes.POINTER(Options)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _batch_done = True
