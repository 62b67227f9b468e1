"""Python binding of libmf, the MI355X (gfx950) push-relabel maximum-flow library.

This is a thin ctypes layer over the C ABI declared in ``include/mf.h``; every call goes to the
in-tree ``wbpr_amd/libmf.so`` (HIP kernels + host driver).  There is no Python or CPU fallback:
importing this module raises if the library has not been built.

The objects mirror the reference's host interface (maxflow-cuda/include/graph.h:28-118,
maxflow-cuda/src/main.cu:7-240): a host CSR graph loaded from SNAP txt/bin or DIMACS files or
made by a seeded generator, a residual layout (RCSR, BCSR, BCSR2) and an algorithm (TC, VC).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmf.so")
CLI_PATH = os.path.join(_HERE, "bin", "maxflow")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')"
    )
_lib = C.CDLL(LIB_PATH)

# ---- enums --------------------------------------------------------------------------------------
ALGO_TC, ALGO_VC = 0, 1
LAYOUT_RCSR, LAYOUT_BCSR, LAYOUT_BCSR2 = 0, 1, 2
ALGOS = {"tc": ALGO_TC, "vc": ALGO_VC}
LAYOUTS = {"rcsr": LAYOUT_RCSR, "bcsr": LAYOUT_BCSR, "bcsr2": LAYOUT_BCSR2}

STATUS = {
    0: "MF_OK",
    1: "MF_ERR_INVALID_ARGUMENT",
    2: "MF_ERR_IO",
    3: "MF_ERR_PARSE",
    4: "MF_ERR_OVERFLOW",
    5: "MF_ERR_DEVICE",
    6: "MF_ERR_NO_DEVICE",
    7: "MF_ERR_OUT_OF_MEMORY",
    8: "MF_ERR_ITERATION_LIMIT",
    9: "MF_ERR_INTERNAL",
    10: "MF_ERR_BUDGET",
    11: "MF_ERR_NOT_READY",
    12: "MF_ERR_NO_PATH",
}


class MFError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        self.name = STATUS.get(status, str(status))
        super().__init__(f"{what}: {self.name} ({_lib.mf_status_string(status).decode()})")


# ---- C structs ----------------------------------------------------------------------------------
_i32p = C.POINTER(C.c_int32)


class _Csr(C.Structure):
    _fields_ = [
        ("n", C.c_int32),
        ("m", C.c_int64),
        ("offsets", _i32p),
        ("dst", _i32p),
        ("cap", _i32p),
        ("s", C.c_int32),
        ("t", C.c_int32),
    ]


class _HostGraph(C.Structure):
    _fields_ = [
        ("n", C.c_int32),
        ("m", C.c_int64),
        ("offsets", _i32p),
        ("dst", _i32p),
        ("cap", _i32p),
        ("s", C.c_int32),
        ("t", C.c_int32),
        ("m_header", C.c_int64),
        ("m_processed", C.c_int64),
    ]


class _Options(C.Structure):
    _fields_ = [
        ("cycles_per_relabel", C.c_int32),
        ("tile_lanes", C.c_int32),
        ("max_rounds", C.c_int64),
        ("verbose", C.c_int32),
        ("profile_kernels", C.c_int32),
        ("bfs_block_frontier", C.c_int32),
        ("relabel_time_ratio", C.c_float),
        ("batch_algo", C.c_int32),
        ("batch_mode", C.c_int32),
        ("bfs_direction", C.c_int32),
        ("class_lanes", C.c_uint16 * 5),
        ("pad0", C.c_uint16),
        ("forward_hops", C.c_int32),
        ("record_workload", C.c_int32),
        ("cycle_graphs", C.c_int32),
        ("bfs_cooperative", C.c_int32),
        ("reserved", C.c_int32 * 1),
    ]

    def to_dict(self) -> dict:
        """The tunable fields, in the keyword form options() accepts."""
        return {"cycles_per_relabel": self.cycles_per_relabel, "tile_lanes": self.tile_lanes,
                "bfs_block_frontier": self.bfs_block_frontier,
                "relabel_time_ratio": round(float(self.relabel_time_ratio), 4),
                "batch_algo": self.batch_algo, "batch_mode": self.batch_mode,
                "bfs_direction": self.bfs_direction, "class_lanes": list(self.class_lanes),
                "forward_hops": self.forward_hops, "cycle_graphs": self.cycle_graphs,
                "bfs_cooperative": self.bfs_cooperative}


class _Result(C.Structure):
    _fields_ = [
        ("value", C.c_int64),
        ("source_outflow", C.c_int64),
        ("solve_ms", C.c_double),
        ("kernel_ms", C.c_double),
        ("relabel_ms", C.c_double),
        ("upload_ms", C.c_double),
        ("rounds", C.c_int64),
        ("cycles", C.c_int64),
        ("active_total", C.c_int64),
        ("slots_scanned", C.c_int64),
        ("pushes", C.c_int64),
        ("relabels", C.c_int64),
        ("bfs_slots", C.c_int64),
        ("bytes_algorithmic", C.c_int64),
        ("bytes_bfs", C.c_int64),
        ("profiled_kernel_ms", C.c_double),
        ("profiled_launches", C.c_int64),
        ("bfs_levels", C.c_int64),
        ("bfs_grid_levels", C.c_int64),
        ("status", C.c_int32),
        ("bfs_bottom_up_levels", C.c_int32),
        ("reserved", C.c_int32 * 6),
        ("graph_launches", C.c_int64),
    ]


class _LayoutArrays(C.Structure):
    _fields_ = [
        ("n", C.c_int32),
        ("slots", C.c_int64),
        ("a", _i32p),
        ("b", _i32p),
        ("c", _i32p),
        ("d", _i32p),
        ("e", _i32p),
    ]


class _ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_int64),
        ("admitted", C.c_int64),
        ("completed", C.c_int64),
        ("rejected", C.c_int64),
        ("launches", C.c_int64),
        ("arena_bytes", C.c_int64),
        ("arena_peak_bytes", C.c_int64),
        ("batched", C.c_int64),
        ("solo", C.c_int64),
        ("serve_kernel_launches", C.c_int64),
    ]


class _ServerTimes(C.Structure):
    _fields_ = [
        ("submit_ms", C.c_double),
        ("admit_ms", C.c_double),
        ("start_ms", C.c_double),
        ("finish_ms", C.c_double),
        ("device_ms", C.c_double),
        ("batched", C.c_int32),
        ("reserved", C.c_int32),
    ]


_vp = C.c_void_p
_hgpp = C.POINTER(C.POINTER(_HostGraph))


def _sig(name, restype, *argtypes):
    f = getattr(_lib, name)
    f.restype = restype
    f.argtypes = list(argtypes)
    return f


_sig("mf_status_string", C.c_char_p, C.c_int)
_sig("mf_version", C.c_char_p)
_sig("mf_options_default", None, C.POINTER(_Options))
_sig("mf_context_create", C.c_int, C.c_int, C.POINTER(_vp))
_sig("mf_context_destroy", C.c_int, _vp)
_sig("mf_device_synchronize", C.c_int, _vp)
_sig("mf_device_info", C.c_int, _vp, _i32p, C.c_char_p, C.c_size_t)
_sig("mf_load_snap_txt", C.c_int, C.c_char_p, _hgpp)
_sig("mf_load_snap_bin", C.c_int, C.c_char_p, _hgpp)
_sig("mf_load_dimacs", C.c_int, C.c_char_p, _hgpp)
_sig("mf_load_konect_bipartite", C.c_int, C.c_char_p, _hgpp)
_sig("mf_super_terminals", C.c_int, C.POINTER(_Csr), _i32p, C.c_int32, _i32p, C.c_int32, _hgpp)
_sig("mf_bfs_terminal_pairs", C.c_int, C.POINTER(_Csr), C.c_int32, C.c_uint64, _i32p, _i32p,
     _hgpp)
_sig("mf_save_snap_bin", C.c_int, C.c_char_p, C.POINTER(_Csr))
_sig("mf_save_dimacs", C.c_int, C.c_char_p, C.POINTER(_Csr))
_sig("mf_host_graph_free", None, C.POINTER(_HostGraph))
_sig("mf_gen_washington_rlg", C.c_int, C.c_int32, C.c_int32, C.c_int32, C.c_uint64, _hgpp)
_sig("mf_gen_genrmf", C.c_int, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_uint64, _hgpp)
_sig("mf_gen_rmat", C.c_int, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_int32,
     C.c_uint64, _hgpp)
_sig("mf_gen_bipartite", C.c_int, C.c_int32, C.c_int32, C.c_double, C.c_uint64, _hgpp)
_sig("mf_build_layout", C.c_int, C.POINTER(_Csr), C.c_int, C.POINTER(_LayoutArrays))
_sig("mf_layout_free", None, C.POINTER(_LayoutArrays))
_sig("mf_graph_upload", C.c_int, _vp, C.POINTER(_Csr), C.c_int, C.POINTER(_vp))
_sig("mf_graph_solve", C.c_int, _vp, _vp, C.c_int, C.POINTER(_Options), C.POINTER(_Result))
_sig("mf_graph_read_flow", C.c_int, _vp, _vp, _i32p, C.c_int32)
_sig("mf_graph_read_cut", C.c_int, _vp, _vp, C.POINTER(C.c_uint8))
_sig("mf_graph_read_state", C.c_int, _vp, _vp, _i32p, _i32p)
_sig("mf_graph_read_layout", C.c_int, _vp, _vp, C.POINTER(_LayoutArrays))
_sig("mf_graph_read_workload", C.c_int, _vp, _vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint8),
     C.c_int64, C.POINTER(C.c_int64))
_sig("mf_graph_destroy", C.c_int, _vp, _vp)
_sig("mf_solve", C.c_int, _vp, C.POINTER(_Csr), C.c_int, C.c_int, C.POINTER(_Options),
     C.POINTER(_Result), _i32p)
_sig("mf_solve_noalloc", C.c_int, _vp, C.POINTER(_Csr), C.c_int, C.c_int, C.POINTER(_Options),
     C.POINTER(_Result), _i32p)
_sig("mf_rcsr_push_relabel", C.c_int, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
     _i32p, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p,
     C.POINTER(_Options), C.POINTER(_Result))
_sig("mf_batch_solve", C.c_int, _vp, C.POINTER(_Csr), C.c_int32, C.POINTER(_Options),
     C.POINTER(C.c_int64), C.POINTER(_Result))
_sig("mf_server_create", C.c_int, _vp, C.c_int64, C.c_int32, C.POINTER(_vp))
_sig("mf_server_submit", C.c_int, _vp, C.POINTER(_Csr), C.POINTER(C.c_int64))
_sig("mf_server_progress", C.c_int, _vp, C.c_int32)
_sig("mf_server_result", C.c_int, _vp, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_double))
_sig("mf_server_completed", C.c_int, _vp, C.POINTER(C.c_int64), C.c_int32, C.POINTER(C.c_int32))
_sig("mf_server_stats_get", C.c_int, _vp, C.POINTER(_ServerStats))
_sig("mf_server_request_times", C.c_int, _vp, C.c_int64, C.POINTER(_ServerTimes))
_sig("mf_server_latency", C.c_int, _vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
     C.POINTER(C.c_double))
_sig("mf_server_destroy", C.c_int, _vp)
_sig("mf_footprint_bytes", C.c_int64, C.c_int32, C.c_int64)
_sig("mf_flow_return_excess", C.c_int, C.POINTER(_Csr), _i32p)
_sig("mf_autotune", C.c_int, _vp, _vp, C.c_int, C.POINTER(_Options), C.POINTER(_Options),
     C.POINTER(C.c_double), C.POINTER(C.c_int32))
_sig("mf_batch_autotune", C.c_int, _vp, C.POINTER(_Csr), C.c_int32, C.c_int32,
     C.POINTER(_Options), C.POINTER(_Options), C.POINTER(C.c_double), C.POINTER(C.c_int32))

#: every function the C header declares (checked by the export test)
EXPORTED = [
    "mf_status_string", "mf_version", "mf_options_default", "mf_context_create",
    "mf_context_destroy", "mf_device_synchronize", "mf_device_info", "mf_load_snap_txt",
    "mf_load_snap_bin", "mf_save_snap_bin", "mf_load_dimacs", "mf_save_dimacs",
    "mf_host_graph_free", "mf_host_graph_view", "mf_gen_washington_rlg", "mf_gen_genrmf",
    "mf_gen_rmat", "mf_gen_bipartite", "mf_build_layout", "mf_layout_free", "mf_graph_upload",
    "mf_graph_solve", "mf_graph_read_flow", "mf_graph_read_cut", "mf_graph_read_state",
    "mf_graph_destroy", "mf_solve", "mf_rcsr_push_relabel", "mf_batch_solve",
    "mf_server_create", "mf_server_submit", "mf_server_progress", "mf_server_result",
    "mf_server_stats_get", "mf_server_destroy", "mf_footprint_bytes", "mf_flow_return_excess",
    "mf_autotune", "mf_batch_autotune", "mf_graph_read_layout", "mf_graph_read_workload",
    "mf_load_konect_bipartite", "mf_super_terminals", "mf_bfs_terminal_pairs",
    "mf_server_request_times", "mf_server_latency", "mf_solve_noalloc", "mf_server_completed",
]


def _check(status: int, what: str):
    if status != 0:
        raise MFError(status, what)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_i32p)


def version() -> str:
    return _lib.mf_version().decode()


# ---- host graphs --------------------------------------------------------------------------------
@dataclass
class Graph:
    """Directed graph in CSR form with int32 capacities (CSRGraph of the reference)."""

    n: int
    offsets: np.ndarray
    dst: np.ndarray
    cap: np.ndarray
    s: int = -1
    t: int = -1
    m_header: int = -1
    m_processed: int = -1

    def __post_init__(self):
        self.offsets = np.ascontiguousarray(self.offsets, dtype=np.int32)
        self.dst = np.ascontiguousarray(self.dst, dtype=np.int32)
        self.cap = np.ascontiguousarray(self.cap, dtype=np.int32)

    @property
    def m(self) -> int:
        return int(self.dst.shape[0])

    def tails(self) -> np.ndarray:
        return np.repeat(np.arange(self.n, dtype=np.int32), np.diff(self.offsets))

    def with_terminals(self, s: int, t: int) -> "Graph":
        return Graph(self.n, self.offsets, self.dst, self.cap, s, t, self.m_header, self.m_processed)

    def csr(self) -> _Csr:
        c = _Csr()
        c.n = self.n
        c.m = self.m
        c.offsets = _ptr(self.offsets)
        c.dst = _ptr(self.dst)
        c.cap = _ptr(self.cap)
        c.s = self.s
        c.t = self.t
        c._keep = (self.offsets, self.dst, self.cap)
        return c

    @staticmethod
    def from_arcs(n: int, u, v, cap, s: int, t: int) -> "Graph":
        """CSR from an arc list, stable in arc order within each tail."""
        u = np.asarray(u, dtype=np.int64)
        order = np.argsort(u, kind="stable")
        offsets = np.zeros(n + 1, dtype=np.int64)
        np.add.at(offsets, u + 1, 1)
        offsets = np.cumsum(offsets)
        return Graph(n, offsets.astype(np.int32), np.asarray(v)[order], np.asarray(cap)[order], s, t)

    def save_dimacs(self, path: str):
        c = self.csr()
        _check(_lib.mf_save_dimacs(path.encode(), C.byref(c)), "save_dimacs")

    def save_snap_bin(self, path: str):
        c = self.csr()
        _check(_lib.mf_save_snap_bin(path.encode(), C.byref(c)), "save_snap_bin")


def _take(hg_ptr) -> Graph:
    h = hg_ptr.contents
    n, m = h.n, h.m
    g = Graph(
        n,
        np.ctypeslib.as_array(h.offsets, shape=(n + 1,)).copy(),
        np.ctypeslib.as_array(h.dst, shape=(max(m, 1),))[:m].copy(),
        np.ctypeslib.as_array(h.cap, shape=(max(m, 1),))[:m].copy(),
        h.s,
        h.t,
        h.m_header,
        h.m_processed,
    )
    _lib.mf_host_graph_free(hg_ptr)
    return g


def _load(fn, path: str, what: str) -> Graph:
    out = C.POINTER(_HostGraph)()
    _check(fn(os.fsencode(path), C.byref(out)), f"{what}({path})")
    return _take(out)


def load_snap_txt(path: str) -> Graph:
    """CSRGraph::buildFromTxtFile (maxflow-cuda/src/graph.cpp:28-67)."""
    return _load(_lib.mf_load_snap_txt, path, "load_snap_txt")


def load_snap_bin(path: str) -> Graph:
    """CSRGraph::loadFromBinary (maxflow-cuda/src/graph.cpp:292-314)."""
    return _load(_lib.mf_load_snap_bin, path, "load_snap_bin")


def load_dimacs(path: str) -> Graph:
    """CSRGraph::buildFromDIMACSFile (maxflow-cuda/src/graph.cpp:151-214)."""
    return _load(_lib.mf_load_dimacs, path, "load_dimacs")


def load_konect_bipartite(path: str) -> Graph:
    """KONECT bipartite edge list -> unit-capacity matching network (the paper's p.4 setup)."""
    return _load(_lib.mf_load_konect_bipartite, path, "load_konect_bipartite")


def load(path: str, file_type: int) -> Graph:
    """-v flag of the CLI: 0 SNAP txt, 1 SNAP binary, 2 DIMACS (the reference's), 3 KONECT
    bipartite."""
    return {0: load_snap_txt, 1: load_snap_bin, 2: load_dimacs,
            3: load_konect_bipartite}[file_type](path)


def super_terminals(g: Graph, sources, sinks) -> Graph:
    """Super-source n -> sources, sinks -> super-sink n+1 (capacities that never bind)."""
    c = g.csr()
    src = np.ascontiguousarray(sources, dtype=np.int32)
    snk = np.ascontiguousarray(sinks, dtype=np.int32)
    out = C.POINTER(_HostGraph)()
    _check(_lib.mf_super_terminals(C.byref(c), _ptr(src), len(src), _ptr(snk), len(snk),
                                   C.byref(out)), "super_terminals")
    return _take(out)


def bfs_terminal_pairs(g: Graph, pairs: int = 20, seed: int = 1):
    """The paper's SNAP transform: `pairs` BFS-chosen source/sink pairs joined to a super-source
    and super-sink.  Returns (graph, sources, sinks)."""
    c = g.csr()
    src = np.zeros(pairs, dtype=np.int32)
    snk = np.zeros(pairs, dtype=np.int32)
    out = C.POINTER(_HostGraph)()
    _check(_lib.mf_bfs_terminal_pairs(C.byref(c), pairs, seed, _ptr(src), _ptr(snk),
                                      C.byref(out)), "bfs_terminal_pairs")
    return _take(out), src, snk


def _gen(fn, *args) -> Graph:
    out = C.POINTER(_HostGraph)()
    _check(fn(*args, C.byref(out)), fn.__name__)
    return _take(out)


def gen_washington_rlg(rows: int, levels: int, r: int = 10000, seed: int = 1) -> Graph:
    return _gen(_lib.mf_gen_washington_rlg, rows, levels, r, seed)


def gen_genrmf(a: int, b: int, c1: int = 1, c2: int = 10000, seed: int = 1) -> Graph:
    return _gen(_lib.mf_gen_genrmf, a, b, c1, c2, seed)


def gen_rmat(scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
             cap_max: int = 1000, seed: int = 1) -> Graph:
    return _gen(_lib.mf_gen_rmat, scale, edge_factor, a, b, c, cap_max, seed)


def gen_bipartite(L: int, R: int, mean_deg: float = 8.0, seed: int = 1) -> Graph:
    return _gen(_lib.mf_gen_bipartite, L, R, mean_deg, seed)


def _layout_dict(arr, layout: int, m: int) -> dict:
    n, S = arr.n, arr.slots

    def cp(p, k):
        return np.ctypeslib.as_array(p, shape=(max(k, 1),))[:k].copy() if p else None

    if layout == LAYOUT_RCSR:
        out = {"roffsets": cp(arr.a, n + 1), "rdst": cp(arr.b, S), "flow_index": cp(arr.c, S)}
    else:
        out = {"offsets": cp(arr.a, n + 1), "dst": cp(arr.b, S), "rev": cp(arr.c, S),
               "residual": cp(arr.d, S), "fslot": cp(arr.e, m)}
    _lib.mf_layout_free(C.byref(arr))
    return out


def build_layout(g: Graph, layout: int) -> dict:
    """Host residual layout arrays (ResidualGraph::buildFromCSRGraph) for layout tests."""
    c = g.csr()
    arr = _LayoutArrays()
    _check(_lib.mf_build_layout(C.byref(c), layout, C.byref(arr)), "build_layout")
    return _layout_dict(arr, layout, g.m)


def return_excess(g: Graph, flow: np.ndarray) -> np.ndarray:
    """Preflow -> flow of the same value (host repair used by DeviceGraph.read_flow)."""
    fl = np.ascontiguousarray(flow, dtype=np.int32).copy()
    c = g.csr()
    _check(_lib.mf_flow_return_excess(C.byref(c), _ptr(fl)), "flow_return_excess")
    return fl


def footprint_bytes(n: int, m: int) -> int:
    return int(_lib.mf_footprint_bytes(n, m))


# ---- device solve -------------------------------------------------------------------------------
@dataclass
class Result:
    value: int
    source_outflow: int
    solve_ms: float
    kernel_ms: float
    relabel_ms: float
    upload_ms: float
    rounds: int
    cycles: int
    active_total: int
    slots_scanned: int
    pushes: int
    relabels: int
    bfs_slots: int
    bytes_algorithmic: int
    bytes_bfs: int
    profiled_kernel_ms: float
    profiled_launches: int
    bfs_levels: int
    bfs_grid_levels: int
    bfs_bottom_up_levels: int
    graph_launches: int

    @staticmethod
    def _from(r: _Result) -> "Result":
        return Result(*(getattr(r, f) for f in Result.__dataclass_fields__))


def options(**kw) -> _Options:
    o = _Options()
    _lib.mf_options_default(C.byref(o))
    for k, v in kw.items():
        if k == "relabel_time_ratio":
            o.relabel_time_ratio = float(v)
        elif k == "class_lanes":
            for i, x in enumerate(list(v)[:5]):
                o.class_lanes[i] = int(x)
        else:
            setattr(o, k, int(v))
    return o


class Context:
    """One HIP stream on one GPU (mf_context)."""

    def __init__(self, device: int = 0):
        self._p = _vp()
        _check(_lib.mf_context_create(device, C.byref(self._p)), "context_create")

    def close(self):
        if self._p:
            _lib.mf_context_destroy(self._p)
            self._p = _vp()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_info(self):
        ncu = C.c_int32()
        name = C.create_string_buffer(300)
        _check(_lib.mf_device_info(self._p, C.byref(ncu), name, 300), "device_info")
        return name.value.decode(), ncu.value

    def synchronize(self):
        _check(_lib.mf_device_synchronize(self._p), "synchronize")

    def upload(self, g: Graph, layout: int = LAYOUT_BCSR) -> "DeviceGraph":
        c = g.csr()
        p = _vp()
        _check(_lib.mf_graph_upload(self._p, C.byref(c), layout, C.byref(p)), "graph_upload")
        return DeviceGraph(self, p, g, layout)

    def solve(self, g: Graph, algo: int = ALGO_VC, layout: int = LAYOUT_BCSR, flow: bool = False,
              **opts):
        c = g.csr()
        o = options(**opts)
        r = _Result()
        fl = np.zeros(max(g.m, 1), dtype=np.int32) if flow else None
        _check(_lib.mf_solve(self._p, C.byref(c), algo, layout, C.byref(o), C.byref(r),
                             _ptr(fl) if flow else None), "solve")
        res = Result._from(r)
        return (res, fl[: g.m]) if flow else res

    def solve_noalloc(self, g: Graph, algo: int = ALGO_VC, flow: bool = False, **opts):
        """mf_solve_noalloc: BCSR solve whose device arrays live in the library's static scratch
        array (nothing is allocated on the device)."""
        c = g.csr()
        o = options(**opts)
        r = _Result()
        fl = np.zeros(max(g.m, 1), dtype=np.int32) if flow else None
        _check(_lib.mf_solve_noalloc(self._p, C.byref(c), algo, LAYOUT_BCSR, C.byref(o),
                                     C.byref(r), _ptr(fl) if flow else None), "solve_noalloc")
        res = Result._from(r)
        return (res, fl[: g.m]) if flow else res

    def rcsr_push_relabel(self, algo_type, V, E, source, sink, height, excess, offsets, dst, cap,
                          fflows, bflows, roffsets, rdst, flow_idx, **opts):
        """Reference-shaped entry (preflow + push_relabel of maxflow-cuda) on RCSR host arrays."""
        et = C.c_int()
        o = options(**opts)
        r = _Result()
        _check(_lib.mf_rcsr_push_relabel(self._p, algo_type, V, E, source, sink, _ptr(height),
                                         _ptr(excess), _ptr(offsets), _ptr(dst), _ptr(cap),
                                         _ptr(fflows), _ptr(bflows), _ptr(roffsets), _ptr(rdst),
                                         _ptr(flow_idx), C.byref(et), C.byref(o), C.byref(r)),
               "rcsr_push_relabel")
        return et.value, Result._from(r)

    def batch_solve(self, graphs, **opts):
        k = len(graphs)
        arr = (_Csr * k)()
        keep = []
        for i, g in enumerate(graphs):
            c = g.csr()
            keep.append(c)
            arr[i] = c
        vals = (C.c_int64 * k)()
        o = options(**opts)
        r = _Result()
        _check(_lib.mf_batch_solve(self._p, arr, k, C.byref(o), vals, C.byref(r)), "batch_solve")
        return np.array(vals[:], dtype=np.int64), Result._from(r)

    def batch_autotune(self, graphs, sample: int = 256, **base):
        """Tuned batch options (dict for batch_solve), best sample time in ms, trials run."""
        k = len(graphs)
        arr = (_Csr * k)()
        keep = []
        for i, g in enumerate(graphs):
            c = g.csr()
            keep.append(c)
            arr[i] = c
        o, best = options(**base), _Options()
        ms, n = C.c_double(), C.c_int32()
        _check(_lib.mf_batch_autotune(self._p, arr, k, sample, C.byref(o), C.byref(best),
                                      C.byref(ms), C.byref(n)), "batch_autotune")
        return best.to_dict(), ms.value, n.value


class DeviceGraph:
    """Residual graph resident in HBM (mf_graph)."""

    def __init__(self, ctx: Context, p, g: Graph, layout: int = LAYOUT_BCSR):
        self.ctx, self._p, self.graph, self.layout = ctx, p, g, layout

    def solve(self, algo: int = ALGO_VC, **opts) -> Result:
        o = options(**opts)
        r = _Result()
        _check(_lib.mf_graph_solve(self.ctx._p, self._p, algo, C.byref(o), C.byref(r)), "graph_solve")
        return Result._from(r)

    def autotune(self, algo: int = ALGO_VC, **base):
        """Tuned options (dict for solve), best solve time in ms, trial solves run."""
        o, best = options(**base), _Options()
        ms, n = C.c_double(), C.c_int32()
        _check(_lib.mf_autotune(self.ctx._p, self._p, algo, C.byref(o), C.byref(best),
                                C.byref(ms), C.byref(n)), "autotune")
        return best.to_dict(), ms.value, n.value

    def read_flow(self, return_excess: bool = True) -> np.ndarray:
        fl = np.zeros(max(self.graph.m, 1), dtype=np.int32)
        _check(_lib.mf_graph_read_flow(self.ctx._p, self._p, _ptr(fl), int(return_excess)),
               "read_flow")
        return fl[: self.graph.m]

    def read_cut(self) -> np.ndarray:
        side = np.zeros(self.graph.n, dtype=np.uint8)
        _check(_lib.mf_graph_read_cut(self.ctx._p, self._p,
                                      side.ctypes.data_as(C.POINTER(C.c_uint8))), "read_cut")
        return side

    def read_layout(self) -> dict:
        """The resident residual layout (built on the device), in build_layout's format."""
        arr = _LayoutArrays()
        _check(_lib.mf_graph_read_layout(self.ctx._p, self._p, C.byref(arr)), "read_layout")
        return _layout_dict(arr, self.layout, self.graph.m)

    def read_workload(self):
        """Per-wave busy cycles of the last solve run with record_workload=1, and the degree
        class of each wave's block (empty arrays when nothing was recorded)."""
        n = C.c_int64()
        _check(_lib.mf_graph_read_workload(self.ctx._p, self._p, None, None, 0, C.byref(n)),
               "read_workload")
        cyc = np.zeros(max(n.value, 1), dtype=np.uint64)
        cls = np.zeros(max(n.value, 1), dtype=np.uint8)
        if n.value:
            _check(_lib.mf_graph_read_workload(
                self.ctx._p, self._p, cyc.ctypes.data_as(C.POINTER(C.c_uint64)),
                cls.ctypes.data_as(C.POINTER(C.c_uint8)), n.value, C.byref(n)), "read_workload")
        return cyc[: n.value], cls[: n.value]

    def read_state(self):
        h = np.zeros(self.graph.n, dtype=np.int32)
        e = np.zeros(self.graph.n, dtype=np.int32)
        _check(_lib.mf_graph_read_state(self.ctx._p, self._p, _ptr(h), _ptr(e)), "read_state")
        return h, e

    def close(self):
        if self._p:
            _lib.mf_graph_destroy(self.ctx._p, self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request loop over a device-resident arena (mf_server)."""

    def __init__(self, ctx: Context, hbm_budget_bytes: int, max_inflight_launches: int = 4):
        self.ctx = ctx
        self._p = _vp()
        _check(_lib.mf_server_create(ctx._p, hbm_budget_bytes, max_inflight_launches,
                                     C.byref(self._p)), "server_create")

    def submit(self, g: Graph) -> int:
        c = g.csr()
        rid = C.c_int64()
        _check(_lib.mf_server_submit(self._p, C.byref(c), C.byref(rid)), "server_submit")
        return rid.value

    def progress(self, wait: bool = False):
        _check(_lib.mf_server_progress(self._p, int(wait)), "server_progress")

    def result(self, rid: int):
        """(value, device_ms) or None while the request is still running."""
        v = C.c_int64()
        ms = C.c_double()
        st = _lib.mf_server_result(self._p, rid, C.byref(v), C.byref(ms))
        if st == 11:
            return None
        _check(st, "server_result")
        return v.value, ms.value

    def completed(self, cap: int = 4096):
        """Ids of the requests that completed since the previous call, in completion order."""
        ids = (C.c_int64 * cap)()
        k = C.c_int32()
        out = []
        while True:
            _check(_lib.mf_server_completed(self._p, ids, cap, C.byref(k)), "server_completed")
            out.extend(ids[i] for i in range(k.value))
            if k.value < cap:
                return out

    def times(self, rid: int):
        """Host times (ms since the server was created) of a completed request, or None."""
        t = _ServerTimes()
        st = _lib.mf_server_request_times(self._p, rid, C.byref(t))
        if st == 11:
            return None
        _check(st, "server_request_times")
        return {f: getattr(t, f) for f, _ in _ServerTimes._fields_ if f != "reserved"}

    def latency(self):
        """(p50 ms, p99 ms, solves/s) over the completed requests, or None before the first."""
        p50, p99, sps = C.c_double(), C.c_double(), C.c_double()
        st = _lib.mf_server_latency(self._p, C.byref(p50), C.byref(p99), C.byref(sps))
        if st == 11:
            return None
        _check(st, "server_latency")
        return p50.value, p99.value, sps.value

    def stats(self) -> dict:
        s = _ServerStats()
        _check(_lib.mf_server_stats_get(self._p, C.byref(s)), "server_stats")
        return {f: getattr(s, f) for f, _ in _ServerStats._fields_}

    def close(self):
        if self._p:
            _lib.mf_server_destroy(self._p)
            self._p = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
