"""Python mirror of the solver's C ABI (include/vcgpu.h) over ctypes.

The product path is libvcgpu.so (hand-written gfx950 HIP kernels).  This
module only marshals arguments; there is no CPU fallback: if the library is
missing, or no GPU is present, the solve calls raise.

The high-level functions follow the reference driver's call sequence
(src/main.cu:49-98 with upstream VC 1 semantics): load / build the CSR graph,
check it, take U = min(RemoveMax, RemoveEdge) as the upper bound, then run the
device solve for MVC or PVC.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# VCGPU_LIBRARY selects another build of the same library (A/B measurements).
LIB_PATH = os.environ.get("VCGPU_LIBRARY") or os.path.join(_HERE, "libvcgpu.so")

VC_OK, VC_EINVAL, VC_ENOMEM, VC_EDEVICE, VC_ETIMEOUT, VC_ECANCELLED, VC_ENOTSUP, VC_EBADGRAPH = range(8)
VC_MVC, VC_PVC = 0, 1
VC_STACK_ONLY, VC_HYBRID = 0, 1

_u32p = C.POINTER(C.c_uint32)
_i32p = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)


class _Csr(C.Structure):
    _fields_ = [("vertexNum", C.c_uint32), ("edgeNum", C.c_uint32), ("srcPtr", _u32p), ("dst", _u32p),
                ("degree", _i32p)]


class _CsrOwned(C.Structure):
    _fields_ = [("csr", _Csr), ("storage", C.c_void_p)]


class _Job(C.Structure):
    _fields_ = [("instance", C.c_uint32), ("k", C.c_uint32), ("upper_bound", C.c_uint32), ("reserved", C.c_uint32),
                ("max_nodes", C.c_uint64)]


class _Params(C.Structure):
    _fields_ = [("version", C.c_uint32), ("block_dim", C.c_uint32), ("num_blocks", C.c_uint32),
                ("use_global_memory", C.c_int32), ("list_size", C.c_uint32), ("list_threshold", C.c_float),
                ("start_depth", C.c_uint32), ("time_limit_ms", C.c_uint32), ("flags", C.c_uint32)]


class _Result(C.Structure):
    _fields_ = [("minimum", C.c_uint32), ("k_found", C.c_uint32), ("solved", C.c_uint32),
                ("cover_from_gpu", C.c_uint32), ("tree_nodes", C.c_uint64), ("status", C.c_int32),
                ("done_us", C.c_uint32)]


class _Info(C.Structure):
    _fields_ = [("num_blocks", C.c_uint32), ("block_dim", C.c_uint32), ("workers", C.c_uint32),
                ("blocks_per_cu", C.c_uint32), ("use_global_memory", C.c_uint32), ("list_size", C.c_uint32),
                ("list_threshold", C.c_uint32), ("stack_depth", C.c_uint32), ("kernel_ms", C.c_float),
                ("cu_count", C.c_uint32), ("tree_nodes", C.c_uint64), ("donations", C.c_uint64),
                ("steals", C.c_uint64), ("idle_fraction", C.c_float), ("active_ms", C.c_float),
                ("state_bytes", C.c_uint64)]


class _ServerConfig(C.Structure):
    _fields_ = [("max_vertices", C.c_uint32), ("slots", C.c_uint32), ("arena_bytes", C.c_uint64),
                ("num_blocks", C.c_uint32), ("block_dim", C.c_uint32), ("stack_depth", C.c_uint32),
                ("time_limit_ms", C.c_uint32)]


# Every symbol include/vcgpu.h declares.
EXPORTS = [
    "vc_ctx_create", "vc_ctx_destroy", "vc_graphs_upload", "vc_graphs_free", "vc_solve", "vc_occupancy",
    "vc_load_graph_file", "vc_load_graph_file_indexed", "vc_csr_from_edges", "vc_csr_free", "vc_check_graph",
    "vc_approx_remove_max", "vc_approx_remove_edge", "vc_status_string", "vc_server_start", "vc_server_submit", "vc_server_poll",
    "vc_server_wait", "vc_server_stop", "vc_now_ns", "vc_autotune", "vc_solve_noalloc",
]
VC_EAGAIN = 8


class VCError(RuntimeError):
    def __init__(self, status: int, what: str):
        self.status = status
        super().__init__(f"{what}: {status_string(status)} ({status})")


_lib = None


def lib():
    """Load libvcgpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() (make -C "
                               f"{_HERE}); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        L.vc_ctx_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
        L.vc_ctx_destroy.argtypes = [C.c_void_p]
        L.vc_graphs_upload.argtypes = [C.c_void_p, C.c_int, C.POINTER(_Csr), C.POINTER(C.c_void_p)]
        L.vc_graphs_free.argtypes = [C.c_void_p, C.c_void_p]
        L.vc_solve.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(_Job), C.POINTER(_Params), C.POINTER(_Result),
                               C.POINTER(_u8p), C.POINTER(_Info)]
        L.vc_solve_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(_Job), C.POINTER(_Params),
                                       C.POINTER(_Result), C.POINTER(_u8p), C.POINTER(_Info)]
        L.vc_occupancy.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(_Params), _i32p, _i32p, _i32p]
        L.vc_autotune.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(_Job), C.POINTER(_Params), C.POINTER(_Params),
                                  C.POINTER(C.c_float)]
        L.vc_load_graph_file.argtypes = [C.c_char_p, C.POINTER(_CsrOwned)]
        L.vc_load_graph_file_indexed.argtypes = [C.c_char_p, C.c_uint32, C.POINTER(_CsrOwned)]
        L.vc_csr_from_edges.argtypes = [C.c_uint32, C.c_uint32, _u32p, _u32p, C.POINTER(_CsrOwned)]
        L.vc_csr_free.argtypes = [C.POINTER(_CsrOwned)]
        L.vc_check_graph.argtypes = [C.POINTER(_Csr)]
        L.vc_approx_remove_max.argtypes = [C.POINTER(_Csr), _u8p]
        L.vc_approx_remove_max.restype = C.c_uint32
        L.vc_approx_remove_edge.argtypes = [C.POINTER(_Csr), C.c_uint32, _u8p]
        L.vc_approx_remove_edge.restype = C.c_uint32
        L.vc_status_string.argtypes = [C.c_int]
        L.vc_status_string.restype = C.c_char_p
        L.vc_server_start.argtypes = [C.c_void_p, C.POINTER(_ServerConfig), C.POINTER(C.c_void_p)]
        L.vc_server_submit.argtypes = [C.c_void_p, C.POINTER(_Csr), C.POINTER(_Job), C.POINTER(C.c_uint64)]
        L.vc_server_poll.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(_Result), _u8p, C.POINTER(C.c_uint64)]
        L.vc_server_wait.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.POINTER(_Result), _u8p,
                                     C.POINTER(C.c_uint64)]
        L.vc_server_stop.argtypes = [C.c_void_p, C.POINTER(_Info)]
        L.vc_now_ns.argtypes = []
        L.vc_now_ns.restype = C.c_uint64
        _lib = L
    return _lib


def status_string(s: int) -> str:
    return lib().vc_status_string(s).decode()


def _check(s: int, what: str) -> None:
    if s != VC_OK:
        raise VCError(s, what)


@dataclass
class CSR:
    """CSR graph in the reference layout (src/CSRGraphRep.h:4-18)."""
    n: int
    m: int
    srcPtr: np.ndarray
    dst: np.ndarray
    degree: np.ndarray

    def _c(self) -> _Csr:
        return _Csr(self.n, self.m, self.srcPtr.ctypes.data_as(_u32p), self.dst.ctypes.data_as(_u32p),
                    self.degree.ctypes.data_as(_i32p))

    def edges(self) -> np.ndarray:
        """Undirected edge list (u < v), shape (m, 2)."""
        src = np.repeat(np.arange(self.n, dtype=np.uint32), np.diff(self.srcPtr.astype(np.int64)))
        keep = src < self.dst
        return np.stack([src[keep], self.dst[keep]], axis=1)


def _own(o: _CsrOwned) -> CSR:
    n, m = o.csr.vertexNum, o.csr.edgeNum
    src = np.ctypeslib.as_array(o.csr.srcPtr, (n + 1,)).copy()
    dst = np.ctypeslib.as_array(o.csr.dst, (2 * m + 1,))[: 2 * m].copy()
    deg = np.ctypeslib.as_array(o.csr.degree, (n + 1,))[:n].copy()
    lib().vc_csr_free(C.byref(o))
    return CSR(n, m, src.astype(np.uint32), dst.astype(np.uint32), deg.astype(np.int32))


def load_graph_file(path: str, index_base: int = 1) -> CSR:
    """createCSRGraphFromFile (src/auxFunctions.cpp:129-209).  index_base 1 is
    the reference's format; 0 reads its 0-indexed files (formated_*,
    lastfm-asia, *-cell)."""
    o = _CsrOwned()
    if index_base == 1:
        _check(lib().vc_load_graph_file(path.encode(), C.byref(o)), f"load {path}")
    else:
        _check(lib().vc_load_graph_file_indexed(path.encode(), index_base, C.byref(o)), f"load {path}")
    return _own(o)


def csr_from_edges(n: int, edges) -> CSR:
    e = np.ascontiguousarray(np.asarray(edges, dtype=np.uint32).reshape(-1, 2))
    eu = np.ascontiguousarray(e[:, 0]) if len(e) else np.zeros(1, np.uint32)
    ev = np.ascontiguousarray(e[:, 1]) if len(e) else np.zeros(1, np.uint32)
    o = _CsrOwned()
    _check(lib().vc_csr_from_edges(n, len(e), eu.ctypes.data_as(_u32p), ev.ctypes.data_as(_u32p), C.byref(o)),
           "csr_from_edges")
    return _own(o)


def check_graph(g: CSR) -> bool:
    c = g._c()
    return bool(lib().vc_check_graph(C.byref(c)))


def approx_remove_max(g: CSR):
    """RemoveMaxApproximateMVC (src/auxFunctions.cpp:211-256) -> (U, cover)."""
    cover = np.zeros(max(g.n, 1), np.uint8)
    c = g._c()
    u = lib().vc_approx_remove_max(C.byref(c), cover.ctypes.data_as(_u8p))
    return int(u), cover[: g.n].astype(bool)


def approx_remove_edge(g: CSR, seed: int):
    """RemoveEdgeApproximateMVC (src/auxFunctions.cpp:265-306), explicit seed -> (U, cover)."""
    cover = np.zeros(max(g.n, 1), np.uint8)
    c = g._c()
    u = lib().vc_approx_remove_edge(C.byref(c), seed & 0xFFFFFFFF, cover.ctypes.data_as(_u8p))
    return int(u), cover[: g.n].astype(bool)


def upper_bound(g: CSR, seed: int | None = 0):
    """U = min(RemoveMax, RemoveEdge) as in src/main.cu:54-73, with the cover
    that achieves it.  seed=None uses RemoveMax only."""
    u1, c1 = approx_remove_max(g)
    if seed is None:
        return u1, c1
    u2, c2 = approx_remove_edge(g, seed)
    return (u1, c1) if u1 <= u2 else (u2, c2)


@dataclass
class Result:
    minimum: int
    k_found: int
    solved: bool
    tree_nodes: int
    status: int
    cover: np.ndarray | None
    done_us: int = 0  # vc_solve: search end, microseconds after the launch's first worker started


@dataclass
class LaunchInfo:
    num_blocks: int
    block_dim: int
    workers: int
    blocks_per_cu: int
    use_global_memory: int
    list_size: int
    list_threshold: int
    stack_depth: int
    kernel_ms: float
    cu_count: int
    tree_nodes: int
    donations: int
    steals: int
    idle_fraction: float
    active_ms: float  # first worker start to last worker exit (device wall clock)
    state_bytes: int  # HBM bytes of node state moved by stack and worklist copies


@dataclass
class Job:
    instance: int = VC_MVC
    k: int = 0
    upper_bound: int = 0
    max_nodes: int = 0


@dataclass
class Params:
    version: int = VC_HYBRID
    block_dim: int = 0
    num_blocks: int = 0
    use_global_memory: int = -1
    list_size: int = 0
    list_threshold: float = 0.0
    start_depth: int = 4
    time_limit_ms: int = 0
    flags: int = 0

    def _c(self) -> _Params:
        return _Params(self.version, self.block_dim, self.num_blocks, self.use_global_memory, self.list_size,
                       self.list_threshold, self.start_depth, self.time_limit_ms, self.flags)

    @staticmethod
    def _from_c(p: "_Params") -> "Params":
        return Params(p.version, p.block_dim, p.num_blocks, p.use_global_memory, p.list_size, p.list_threshold,
                      p.start_depth, p.time_limit_ms, p.flags)


VC_FLAG_GRAPH = 1
VC_FLAG_WIDE = 2


class GraphSet:
    """Graphs resident in HBM (vc_graphs_upload)."""

    def __init__(self, ctx: "Context", graphs: list[CSR]):
        self.ctx = ctx
        self.graphs = list(graphs)
        arr = (_Csr * len(self.graphs))(*[g._c() for g in self.graphs])
        h = C.c_void_p()
        _check(lib().vc_graphs_upload(ctx.handle, len(self.graphs), arr, C.byref(h)), "vc_graphs_upload")
        self.handle = h

    def __len__(self):
        return len(self.graphs)

    def free(self):
        if self.handle:
            lib().vc_graphs_free(self.ctx.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """vc_ctx on one GPU."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(lib().vc_ctx_create(device, C.byref(h)), "vc_ctx_create")
        self.handle = h

    def close(self):
        if self.handle:
            lib().vc_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload(self, graphs) -> GraphSet:
        if isinstance(graphs, CSR):
            graphs = [graphs]
        return GraphSet(self, graphs)

    def occupancy(self, gset: GraphSet, params: Params | None = None):
        p = (params or Params())._c()
        b, d, g = C.c_int32(), C.c_int32(), C.c_int32()
        _check(lib().vc_occupancy(self.handle, gset.handle, C.byref(p), C.byref(b), C.byref(d), C.byref(g)),
               "vc_occupancy")
        return b.value, d.value, g.value

    def autotune(self, gset: GraphSet, jobs, base: Params | None = None):
        """vc_autotune: (best Params, best kernel ms per solved graph or 0 if cached)."""
        G = len(gset)
        if isinstance(jobs, Job):
            jobs = [jobs] * G
        cj = (_Job * G)(*[_Job(j.instance, j.k, j.upper_bound, 0, j.max_nodes) for j in jobs])
        b = (base or Params())._c()
        out = _Params()
        ms = C.c_float()
        _check(lib().vc_autotune(self.handle, gset.handle, cj, C.byref(b), C.byref(out), C.byref(ms)), "vc_autotune")
        return Params._from_c(out), float(ms.value)

    def solve(self, gset: GraphSet, jobs, params: Params | None = None, want_covers: bool = True,
              raise_on_error: bool = True, noalloc: bool = False):
        """One persistent launch over every graph of the set (vc_solve, or
        vc_solve_noalloc with noalloc=True: launch scratch in the library's
        fixed device array, no allocation).
        Returns (list[Result], LaunchInfo, overall_status)."""
        G = len(gset)
        if isinstance(jobs, Job):
            jobs = [jobs] * G
        cj = (_Job * G)(*[_Job(j.instance, j.k, j.upper_bound, 0, j.max_nodes) for j in jobs])
        res = (_Result * G)()
        info = _Info()
        p = (params or Params())._c()
        bufs = [np.zeros(max(g.n, 1), np.uint8) for g in gset.graphs] if want_covers else None
        cov = (_u8p * G)(*[b.ctypes.data_as(_u8p) for b in bufs]) if want_covers else None
        fn = lib().vc_solve_noalloc if noalloc else lib().vc_solve
        st = fn(self.handle, gset.handle, cj, C.byref(p), res, cov, C.byref(info))
        if st != VC_OK and raise_on_error:
            raise VCError(st, "vc_solve_noalloc" if noalloc else "vc_solve")
        out = []
        for i in range(G):
            r = res[i]
            cover = bufs[i][: gset.graphs[i].n].astype(bool) if (want_covers and r.cover_from_gpu) else None
            out.append(Result(int(r.minimum), int(r.k_found), bool(r.solved), int(r.tree_nodes), int(r.status), cover,
                              int(r.done_us)))
        li = LaunchInfo(*[getattr(info, f) for f, _ in _Info._fields_])
        return out, li, st


def solve_mvc(ctx: Context, g: CSR, params: Params | None = None, seed: int | None = 0, max_nodes: int = 0):
    """Exact MVC of one graph: U from the host approximations, then one device
    solve.  Returns (mvc, cover, Result, LaunchInfo)."""
    U, ucover = upper_bound(g, seed)
    gs = ctx.upload(g)
    try:
        (r,), info, _ = ctx.solve(gs, Job(VC_MVC, 0, U, max_nodes), params)
    finally:
        gs.free()
    cover = r.cover if r.cover is not None else ucover
    return r.minimum, cover, r, info


def solve_pvc(ctx: Context, g: CSR, k: int, params: Params | None = None, seed: int | None = 0, max_nodes: int = 0):
    """Is there a cover of size <= k?  Returns (k_found, cover|None, Result, LaunchInfo)."""
    U, _ = upper_bound(g, seed)
    gs = ctx.upload(g)
    try:
        (r,), info, _ = ctx.solve(gs, Job(VC_PVC, k, U, max_nodes), params)
    finally:
        gs.free()
    return r.k_found, r.cover, r, info


def now_ns() -> int:
    """The host monotonic clock the server stamps completions with."""
    return int(lib().vc_now_ns())


class Server:
    """Continuous-batching server (vc_server_*): one persistent launch,
    requests admitted as slots and arena memory free up."""

    def __init__(self, ctx: Context, max_vertices: int, slots: int = 0, arena_bytes: int = 0, num_blocks: int = 0,
                 block_dim: int = 0, stack_depth: int = 0, time_limit_ms: int = 0):
        cfg = _ServerConfig(max_vertices, slots, arena_bytes, num_blocks, block_dim, stack_depth, time_limit_ms)
        h = C.c_void_p()
        _check(lib().vc_server_start(ctx.handle, C.byref(cfg), C.byref(h)), "vc_server_start")
        self.handle = h
        self._n = {}

    def submit(self, g: CSR, job: Job) -> int:
        c = g._c()
        j = _Job(job.instance, job.k, job.upper_bound, 0, job.max_nodes)
        t = C.c_uint64()
        _check(lib().vc_server_submit(self.handle, C.byref(c), C.byref(j), C.byref(t)), "vc_server_submit")
        self._n[t.value] = g.n
        return t.value

    def _take(self, fn, ticket, *extra, want_cover=True):
        n = self._n.get(ticket, 0)
        buf = np.zeros(max(n, 1), np.uint8)
        r = _Result()
        done = C.c_uint64()
        st = fn(self.handle, ticket, *extra, C.byref(r), buf.ctypes.data_as(_u8p) if want_cover else None,
                C.byref(done))
        if st == VC_EAGAIN:
            return None
        _check(st, "vc_server_poll")
        self._n.pop(ticket, None)
        cover = buf[:n].astype(bool) if (want_cover and r.cover_from_gpu) else None
        res = Result(int(r.minimum), int(r.k_found), bool(r.solved), int(r.tree_nodes), int(r.status), cover)
        return res, int(done.value)

    def poll(self, ticket: int, want_cover: bool = True):
        """(Result, done_ns) or None while pending."""
        return self._take(lib().vc_server_poll, ticket, want_cover=want_cover)

    def wait(self, ticket: int, timeout_ms: int = 60000, want_cover: bool = True):
        """(Result, done_ns) or None on timeout."""
        return self._take(lib().vc_server_wait, ticket, timeout_ms, want_cover=want_cover)

    def stop(self) -> LaunchInfo:
        info = _Info()
        st = lib().vc_server_stop(self.handle, C.byref(info))
        self.handle = None
        _check(st, "vc_server_stop")
        return LaunchInfo(*[getattr(info, f) for f, _ in _Info._fields_])
