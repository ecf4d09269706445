"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker or the timed CPU
baseline -- never as the product path.  The functions restate the reference's
CPU solver (src/Sequential.cpp, src/SequentialParameterized.cpp,
src/helperFunctions.cpp, src/auxFunctions.cpp); see vc_oracle.h for which
check pins each one.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
REF_DRIVER = os.path.join(_HERE, "_ref", "ref_driver")

_u32p = C.POINTER(C.c_uint32)
_i32p = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)
_u64p = C.POINTER(C.c_uint64)


class _OrGraph(C.Structure):
    _fields_ = [("n", C.c_uint32), ("m", C.c_uint32), ("srcPtr", _u32p), ("dst", _u32p), ("degree", _i32p)]


class _OrStats(C.Structure):
    _fields_ = [
        ("max_nodes", C.c_uint64),
        ("nodes", C.c_uint64),
        ("leaf_calls", C.c_uint64),
        ("tri_calls", C.c_uint64),
        ("high_calls", C.c_uint64),
        ("square_calls", C.c_uint64),
        ("delete_calls", C.c_uint64),
        ("branches", C.c_uint64),
        ("max_top", C.c_int32),
        ("completed", C.c_int32),
    ]


def build() -> None:
    """Compile liboracle.so (and oracle/_ref when the reference is present)."""
    subprocess.run(["make", "-C", _HERE, "all"], check=True, capture_output=True)
    subprocess.run(["make", "-C", _HERE, "ref"], check=True, capture_output=True)


_lib = None


def load_library(path: str):
    """ctypes handle of an oracle build at `path` with every signature set."""
    L = C.CDLL(path)
    _configure(L)
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = load_library(_LIB_PATH)
    return _lib


def _configure(L):
    if True:
        L.or_graph_load.argtypes = [C.c_char_p, C.POINTER(_OrGraph)]
        L.or_graph_from_edges.argtypes = [C.c_uint32, C.c_uint32, _u32p, _u32p, C.POINTER(_OrGraph)]
        L.or_graph_free.argtypes = [C.POINTER(_OrGraph)]
        L.or_check_graph.argtypes = [C.POINTER(_OrGraph)]
        L.or_remove_max_approx.argtypes = [C.POINTER(_OrGraph)]
        L.or_remove_max_approx.restype = C.c_uint32
        L.or_remove_edge_approx.argtypes = [C.POINTER(_OrGraph), C.c_uint32]
        L.or_remove_edge_approx.restype = C.c_uint32
        L.or_sequential.argtypes = [C.POINTER(_OrGraph), C.c_uint32, C.POINTER(_OrStats), _u8p, _i32p]
        L.or_sequential.restype = C.c_uint32
        L.or_sequential_parameterized.argtypes = [
            C.POINTER(_OrGraph), C.c_uint32, C.c_uint32, _u32p, C.POINTER(_OrStats), _u8p]
        L.or_sequential_parameterized.restype = C.c_uint32
        L.or_cover_is_valid.argtypes = [C.POINTER(_OrGraph), _u8p]
        L.or_solve_batch_mt.argtypes = [C.c_int, C.POINTER(_OrGraph), C.c_int, C.c_uint64, _u32p, _u64p, _i32p]
        L.or_solve_batch_mt.restype = C.c_double


@dataclass
class Graph:
    """CSR graph in the reference's layout (src/CSRGraphRep.h:4-18)."""
    n: int
    m: int
    srcPtr: np.ndarray  # uint32[n+1]
    dst: np.ndarray     # uint32[2m], sorted per vertex
    degree: np.ndarray  # int32[n]

    def _c(self) -> _OrGraph:
        return _OrGraph(self.n, self.m, self.srcPtr.ctypes.data_as(_u32p), self.dst.ctypes.data_as(_u32p),
                        self.degree.ctypes.data_as(_i32p))


def _from_c(cg: _OrGraph) -> Graph:
    n, m = cg.n, cg.m
    g = Graph(n, m,
              np.ctypeslib.as_array(cg.srcPtr, (n + 1,)).copy(),
              np.ctypeslib.as_array(cg.dst, (max(2 * m, 1),))[: 2 * m].copy() if m else np.zeros(0, np.uint32),
              np.ctypeslib.as_array(cg.degree, (max(n, 1),))[:n].copy() if n else np.zeros(0, np.int32))
    lib().or_graph_free(C.byref(cg))
    return g


def load(path: str) -> Graph:
    cg = _OrGraph()
    rc = lib().or_graph_load(path.encode(), C.byref(cg))
    if rc != 0:
        raise ValueError(f"or_graph_load({path}) failed: {rc}")
    return _from_c(cg)


def from_edges(n: int, edges) -> Graph:
    e = np.asarray(edges, dtype=np.uint32).reshape(-1, 2)
    eu = np.ascontiguousarray(e[:, 0]) if len(e) else np.zeros(1, np.uint32)
    ev = np.ascontiguousarray(e[:, 1]) if len(e) else np.zeros(1, np.uint32)
    cg = _OrGraph()
    rc = lib().or_graph_from_edges(n, len(e), eu.ctypes.data_as(_u32p), ev.ctypes.data_as(_u32p), C.byref(cg))
    if rc != 0:
        raise ValueError(f"or_graph_from_edges failed: {rc}")
    return _from_c(cg)


def check_graph(g: Graph) -> bool:
    cg = g._c()
    return bool(lib().or_check_graph(C.byref(cg)))


def remove_max(g: Graph) -> int:
    cg = g._c()
    return int(lib().or_remove_max_approx(C.byref(cg)))


def remove_edge(g: Graph, seed: int) -> int:
    cg = g._c()
    return int(lib().or_remove_edge_approx(C.byref(cg), seed))


def _stats(st: _OrStats) -> dict:
    return {f: int(getattr(st, f)) for f, _ in _OrStats._fields_}


def sequential(g: Graph, minimum: int, max_nodes: int = 0):
    """Sequential(graph, minimum) -> (result, stats, cover or None)."""
    cg = g._c()
    st = _OrStats()
    st.max_nodes = max_nodes
    cover = np.zeros(max(g.n, 1), np.uint8)
    found = C.c_int32(0)
    r = lib().or_sequential(C.byref(cg), minimum, C.byref(st), cover.ctypes.data_as(_u8p), C.byref(found))
    return int(r), _stats(st), (cover[: g.n].astype(bool) if found.value else None)


def sequential_parameterized(g: Graph, minimum: int, k: int, max_nodes: int = 0):
    """SequentialParameterized -> (kFound, stats, cover or None)."""
    cg = g._c()
    st = _OrStats()
    st.max_nodes = max_nodes
    cover = np.zeros(max(g.n, 1), np.uint8)
    kf = C.c_uint32(0)
    lib().or_sequential_parameterized(C.byref(cg), minimum, k, C.byref(kf), C.byref(st),
                                      cover.ctypes.data_as(_u8p))
    return int(kf.value), _stats(st), (cover[: g.n].astype(bool) if kf.value else None)


def cover_is_valid(g: Graph, cover) -> bool:
    c = np.ascontiguousarray(np.asarray(cover, dtype=np.uint8))
    if g.n == 0:
        return True
    cg = g._c()
    return bool(lib().or_cover_is_valid(C.byref(cg), c.ctypes.data_as(_u8p)))


def mvc(g: Graph, max_nodes: int = 0):
    """Exact MVC the way the reference driver computes it with VC 1
    (src/main.cu:54-73 restricted to the deterministic RemoveMax bound,
    then src/Sequential.cpp).  Returns (mvc, stats, cover)."""
    U = remove_max(g)
    r, st, cover = sequential(g, U, max_nodes)
    if cover is None and g.n > 0:
        # No cover smaller than U exists: U is optimal.  The greedy cover of
        # size U is not materialised by the reference; callers that need a
        # witness use the GPU result or brute force.
        pass
    return r, st, cover


def solve_batch_mt(graphs, threads: int, max_nodes: int = 0):
    """CPU baseline: pthread pool over the graphs (RemoveMax U + Sequential).
    Returns (wall_seconds, results, nodes, done)."""
    keep = list(graphs)
    arr = (_OrGraph * len(keep))(*[g._c() for g in keep])
    res = np.zeros(len(keep), np.uint32)
    nodes = np.zeros(len(keep), np.uint64)
    done = np.zeros(len(keep), np.int32)
    secs = lib().or_solve_batch_mt(len(keep), arr, threads, max_nodes, res.ctypes.data_as(_u32p),
                                   nodes.ctypes.data_as(_u64p), done.ctypes.data_as(_i32p))
    return float(secs), res, nodes, done


def ref_available() -> bool:
    return os.path.exists(REF_DRIVER)


def ref_run(path: str, pvc_k: int | None = None, minimum: int | None = None, seed: int | None = None,
            approx_only: bool = False, timeout: float = 120.0) -> dict:
    """Run the reference's own CPU code (oracle/_ref/ref_driver) on a graph file."""
    import json
    cmd = [REF_DRIVER, path]
    if pvc_k is not None:
        cmd += ["--pvc", str(pvc_k)]
    if minimum is not None:
        cmd += ["--min", str(minimum)]
    if seed is not None:
        cmd += ["--seed", str(seed)]
    if approx_only:
        cmd += ["--approx-only"]
    out = subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=timeout).stdout
    return json.loads(out)
